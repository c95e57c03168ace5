"""The batched TRG driver (grassmanntn_amd.batch) on the CPU, through the numpy kernel emulator
(tests/kernel_emulator.py), whose graph capture records the calls issued while a capture is open
and replays them on launch.  Checked against the oracle's TRG run of every point separately
(gauge2d.py:1647-1755 and the ln Z rule of example.py:186-193), eager and graph-replayed."""
import numpy as np
import pytest

from conftest import diagonal_site
from kernel_emulator import emulate
from oracle import gtn_oracle as O


def _points(D, seeds):
    return np.stack([O.site_tensor(D, s).data for s in seeds])


def _oracle_lnz(T0s, dcut, steps):
    out = []
    for T in T0s:
        F, _ = O.trg_lnz(O.dense(T, encoder="parity-preserving", statistics=(1, 1, -1, -1)), dcut, steps)
        out.append(F)
    return np.array(out)


@pytest.mark.parametrize("D,dcut,steps", [(8, 8, 3), (8, 4, 4), (4, 12, 3)])
@pytest.mark.parametrize("graphs", [False, True])
def test_batched_trg_matches_oracle_per_point(D, dcut, steps, graphs):
    """(8, 8): the bond stays 8; (8, 4): it drops to 4 after the first step (a first-step graph and a
    steady-state graph); (4, 12): int(12/2) = 6 values padded to a bond of 16 (zero singular values,
    the reference's power-of-two padding)."""
    T0s = _points(D, [3, 4, 5])
    ref = _oracle_lnz(T0s, dcut, steps)
    with emulate():
        from grassmanntn_amd.batch import BatchedTRG
        run = BatchedTRG(3, D, dcut, steps, graphs=graphs)
        run.prepare(T0s)
        got = run.run(T0s)
        run.close()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def test_graph_replay_is_repeatable_and_matches_eager():
    """Two runs of the same recorded graphs from different inputs, and the eager run: the second
    run's results do not depend on the first (the graphs' buffers are fully rewritten), and replay
    equals eager execution bit for bit (the emulator replays the same numpy operations)."""
    A, B = _points(8, [6, 7]), _points(8, [8, 9])
    with emulate():
        from grassmanntn_amd.batch import BatchedTRG
        g = BatchedTRG(2, 8, 8, 3, graphs=True)
        g.prepare(A)
        ra = g.run(A)
        rb = g.run(B)
        ra2 = g.run(A)
        g.close()
        e = BatchedTRG(2, 8, 8, 3, graphs=False)
        ea, eb = e.run(A), e.run(B)
    assert np.array_equal(ra, ra2)
    assert np.array_equal(ra, ea) and np.array_equal(rb, eb)
    np.testing.assert_allclose(rb, _oracle_lnz(B, 8, 3), rtol=0, atol=1e-12)


def test_batched_rank_deficient_points_match_oracle():
    """A point whose parity blocks keep fewer than int(dcut/2) values (the diagonal tensor of three
    terms: one even value, no odd ones) next to a full-rank point: the batch keeps the full bond with
    zero singular values where the reference shrinks it; norms and ln Z agree."""
    T0s = np.stack([diagonal_site(8, 7).data, O.site_tensor(8, 2).data])
    ref = _oracle_lnz(T0s, 8, 2)
    with emulate():
        from grassmanntn_amd.batch import BatchedTRG
        run = BatchedTRG(2, 8, 8, 2, graphs=True)
        run.prepare(T0s)
        got = run.run(T0s)
        norms = run.log_norms
        run.close()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
    T = O.dense(T0s[0], encoder="parity-preserving", statistics=(1, 1, -1, -1))
    _, n1 = O.trg(T, 8)
    assert abs(norms[0, 0] - np.log(n1)) < 1e-12


def test_capture_rejects_host_readback():
    """A synchronous device->host copy inside a capture is an error (the emulated runtime refuses it
    as HIP does), so a readback slipping into the recorded step cannot go unnoticed."""
    with emulate():
        from grassmanntn_amd import _lib as L
        from grassmanntn_amd._device import DeviceArray
        from grassmanntn_amd.batch import Graph
        x = DeviceArray.zeros(4)
        with pytest.raises(L.GrassmannDeviceError):
            Graph(lambda: x.numpy(), pool=99)
