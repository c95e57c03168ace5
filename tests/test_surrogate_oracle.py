"""Pins of the surrogate MLP oracle (oracle/mlp_oracle.py): BF16 rounding on known values, the FP64
backward pass against central finite differences of the loss, and one AdamW step written out."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_known_values():
    one = 1.0
    assert M.bf16(one) == 1.0
    assert M.bf16(1 + 2.0 ** -9) == 1.0                 # below half a BF16 ulp (2^-7 at 1)
    assert M.bf16(1 + 2.0 ** -8) == 1.0                 # tie, mantissa 0 is even
    assert M.bf16(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6   # tie between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
    assert M.bf16(1 + 2.0 ** -7) == 1 + 2.0 ** -7       # representable
    assert M.bf16(-3.0) == -3.0
    x = np.random.default_rng(0).standard_normal(1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-12))


def _net(rng, dims, E):
    Ws = [rng.standard_normal((E, a, b)) / np.sqrt(a) for a, b in zip(dims[:-1], dims[1:])]
    bs = [0.1 * rng.standard_normal((E, b)) for b in dims[1:]]
    return Ws, bs


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    dims, E, n = (2, 7, 6, 3), 2, 5
    Ws, bs = _net(rng, dims, E)
    X = rng.standard_normal((n, dims[0]))
    T = rng.standard_normal((n, dims[-1]))
    w = np.array([3.0, 1.0, 0.5])

    def total(Ws_, bs_):
        out, _ = M.forward(Ws_, bs_, X)
        return M.loss_and_dout(out, T, w)[0].sum()

    out, cache = M.forward(Ws, bs, X)
    _, dout = M.loss_and_dout(out, T, w)
    dWs, dbs = M.backward(Ws, cache, dout)
    h = 1e-6
    for l in range(len(Ws)):
        for idx in [(0, 0, 0), (1, dims[l] - 1, dims[l + 1] - 1), (1, 1, 2 % dims[l + 1])]:
            Wp = [W.copy() for W in Ws]
            Wm = [W.copy() for W in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (total(Wp, bs) - total(Wm, bs)) / (2 * h)
            assert abs(fd - dWs[l][idx]) < 1e-7 * max(1.0, abs(fd)), (l, idx, fd, dWs[l][idx])
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][1, 0] += h
        bm[l][1, 0] -= h
        fd = (total(Ws, bp) - total(Ws, bm)) / (2 * h)
        assert abs(fd - dbs[l][1, 0]) < 1e-7 * max(1.0, abs(fd))


def test_adamw_step_written_out():
    w, g, m, v = np.array([1.0]), np.array([0.5]), np.array([0.1]), np.array([0.2])
    lr, b1, b2, eps, wd, t = 0.01, 0.9, 0.999, 1e-8, 0.1, 3
    w2, m2, v2 = M.adamw(w, g, m, v, t, lr, b1, b2, eps, wd)
    mm = 0.9 * 0.1 + 0.1 * 0.5
    vv = 0.999 * 0.2 + 0.001 * 0.25
    ref = 1.0 * (1 - 0.001) - 0.01 * (mm / (1 - 0.9 ** 3)) / (np.sqrt(vv / (1 - 0.999 ** 3)) + 1e-8)
    assert abs(w2[0] - ref) < 1e-15 and abs(m2[0] - mm) < 1e-15 and abs(v2[0] - vv) < 1e-15
