"""CPU checks of the coupling-sweep workload (grassmanntn_amd.sweep): its site-tensor generator is the
oracle's, its point draws are those of the committed golden, and a small sweep through the numpy
kernel emulator matches the oracle point by point (ln Z per site, example.py:186-193)."""
import numpy as np

from kernel_emulator import emulate
from oracle import gtn_oracle as O


def test_site_array_equals_oracle_generator():
    from grassmanntn_amd.sweep import site_array
    for D, seed in ((8, 0), (16, 3), (32, 10)):
        assert np.array_equal(site_array(D, seed), O.site_tensor(D, seed).data)


def test_coupling_points_are_the_golden_draws():
    import json
    import os
    from grassmanntn_amd.sweep import coupling_points
    b, m, chi = coupling_points(512, 100)
    assert b.min() >= 0.1 and b.max() <= 2.0 and m.min() >= 0.0 and m.max() <= 1.0
    assert set(int(c) for c in chi) == {16, 24, 32, 48, 64}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sweep_golden.json")
    if os.path.exists(path):
        for p in json.load(open(path))["points"]:
            i = p["index"]
            assert (b[i], m[i], int(chi[i])) == (p["beta"], p["m"], p["chi"])
            assert i == int(np.nonzero(chi == p["chi"])[0][0])


def test_small_sweep_matches_oracle():
    """Leg dim 8 generators, seven points over bonds 4 (three points) and 8 (four points), 3 steps,
    FP64, with a batch limit of 2: two batches of two for bond 8, and for bond 4 a batch of two and
    a padded one."""
    from grassmanntn_amd.sweep import coupling_points
    beta, m, chi = coupling_points(7, 5, chis=(4, 8))
    A = [O.site_tensor(8, s).data for s in (1, 2, 3)]
    ref = []
    for i in range(7):
        T = O.dense(A[0] + beta[i] * A[1] + m[i] * A[2], encoder="parity-preserving", statistics=(1, 1, -1, -1))
        ref.append(O.trg_lnz(T, int(chi[i]), 3)[0])
    with emulate():
        from grassmanntn_amd.sweep import CouplingSweep
        sw = CouplingSweep(D=8, steps=3, gen_seeds=(1, 2, 3), max_batch=lambda c: 2)
        assert [(c, P, len(i)) for c, P, i in sw.batches(beta, m, chi)] == [(4, 2, 2), (4, 2, 1), (8, 2, 2), (8, 2, 2)]
        sw.prepare(beta, m, chi)
        got = sw.run(beta, m, chi)
        sw.close()
    np.testing.assert_allclose(got, np.array(ref), rtol=0, atol=1e-12)


def test_batch_sizes_bound_the_padding():
    """A batch of n points gets n rounded up to a multiple of max(8, q / 16), q the next power of two
    (powers of two up to 8 kept): never fewer slots than points, less than one granule of padding,
    at most 16 sizes between consecutive powers of two, and the sweep's point counts per chi
    (89-116) padded by at most 7 slots each."""
    from grassmanntn_amd.sweep import CouplingSweep, coupling_points
    sizes = {}
    for n in range(1, 1025):
        P = CouplingSweep.batch_size(n)
        q = 1 << int(np.ceil(np.log2(n)))
        assert n <= P <= q
        assert P - n < max(8, q // 16) or q <= 8
        sizes.setdefault(q, set()).add(P)
    assert all(len(v) <= 16 for v in sizes.values())
    _, _, chi = coupling_points(512, 100)
    for c in set(int(x) for x in chi):
        n = int(np.sum(chi == c))
        assert CouplingSweep.batch_size(n) - n <= 7
