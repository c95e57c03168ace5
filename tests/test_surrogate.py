"""The surrogate module (grassmanntn_amd.surrogate) on the CPU through the numpy kernel emulator:
its forward and backward passes against the MLP oracle's device arithmetic (BF16 weights, split
BF16 activations), a training run that lowers the loss, and the fallback rule of Surrogate.query
(points whose ensemble spread exceeds the tolerance take the TRG value)."""
import numpy as np

from kernel_emulator import emulate
from oracle import mlp_oracle as M
from oracle import gtn_oracle as O


def test_forward_backward_match_oracle():
    rng = np.random.default_rng(3)
    dims, E, n = (2, 16, 16, 5), 3, 24
    X = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    T = rng.standard_normal((n, dims[-1])).astype(np.float32)
    w = np.array([4, 1, 1, 1, 1], dtype=np.float32) / 8
    with emulate():
        from grassmanntn_amd._device import DeviceArray
        from grassmanntn_amd.surrogate import MLPEnsemble
        ens = MLPEnsemble(dims, E, seed=5)
        Ws, bs = ens.host_params()
        bs = [rng.standard_normal(b.shape) * 0.1 for b in bs]
        h = ens.params.numpy()
        for l in range(ens.L):
            _, ob = ens.offsets[l]
            h[ob:ob + bs[l].size] = bs[l].reshape(-1)
        ens.params = DeviceArray.from_numpy(h)
        Xd = DeviceArray.from_numpy(X)
        out = ens.forward(Xd).numpy()
        dout = ens.loss_grad(ens.forward(Xd), DeviceArray.from_numpy(T), DeviceArray.from_numpy(w))
        ens.backward(Xd, dout)
        g = ens.grads.numpy()
        gW = [ens.W(l, DeviceArray.from_numpy(g)).numpy() for l in range(ens.L)]
        gb = [ens.b(l, DeviceArray.from_numpy(g)).numpy() for l in range(ens.L)]
    bs = [b.astype(np.float32).astype(np.float64) for b in bs]
    ref, cache = M.forward(Ws, bs, X.astype(np.float64), mode="device")
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-6)
    _, dref = M.loss_and_dout(ref, T.astype(np.float64), w.astype(np.float64))
    dWs, dbs = M.backward(Ws, cache, dref, mode="device")
    for l in range(len(dims) - 1):
        sc = np.abs(dWs[l]).max()
        np.testing.assert_allclose(gW[l], dWs[l], rtol=0, atol=1e-4 * sc)
        np.testing.assert_allclose(gb[l], dbs[l], rtol=0, atol=1e-4 * np.abs(dbs[l]).max())


def test_training_lowers_the_loss():
    rng = np.random.default_rng(4)
    n = 64
    X = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 1], X[:, 0] ** 2], 1).astype(np.float32)
    with emulate():
        from grassmanntn_amd._device import DeviceArray
        from grassmanntn_amd.surrogate import MLPEnsemble
        ens = MLPEnsemble((2, 16, 16, 2), 2, seed=1)
        Xd, Td, wd = DeviceArray.from_numpy(X), DeviceArray.from_numpy(T), DeviceArray.from_numpy(np.full(2, 0.5, np.float32))
        first = ens.fit(Xd, Td, wd, 1, lr=1e-2)
        last = ens.fit(Xd, Td, wd, 150, lr=1e-2, lr_min=1e-3)
    assert np.all(last < 0.25 * first), (first, last)


def test_query_falls_back_to_trg_above_tolerance():
    from sweep_helpers import small_sweep_labels
    beta, m, ref = small_sweep_labels()
    with emulate():
        from grassmanntn_amd.sweep import CouplingSweep
        from grassmanntn_amd.surrogate import Surrogate
        sw = CouplingSweep(D=8, steps=2, gen_seeds=(1, 2, 3), max_batch=lambda c: 8)
        sur = Surrogate(sw, chi=8, members=2, dims=(2, 8, 8, 17), tol=np.inf)
        Y = sur.labels(beta, m)
        np.testing.assert_allclose(Y[:, 0], ref.real, rtol=0, atol=1e-12)
        assert np.all(Y[:, 1] == 1.0) and np.all((Y[:, 1:] >= 0) & (Y[:, 1:] <= 1))
        sur.train(beta, m, Y, steps=3)
        lnz, fb, _, _ = sur.query(beta, m)
        assert not fb.any()
        sur.tol = 0.0
        lnz, fb, _, _ = sur.query(beta, m)
        assert fb.all()
        np.testing.assert_allclose(lnz, ref.real, rtol=0, atol=1e-12)
        sw.close()
