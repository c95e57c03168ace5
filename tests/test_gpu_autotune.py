"""GPU checks of the grouped GEMM's autotuner: a first eager call of a signature times the candidate
launch configurations and keeps the fastest (never slower than the heuristic configuration, which
is among the candidates); results stay correct; a call issued while the stream is capturing a graph
times nothing (no new signature) and its replays are correct."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")


def _gemm(lib, L, A, B, Cd, M, N, K, batch):
    from grassmanntn_amd.core import _gemm_desc
    d = _gemm_desc(A.data_ptr(), K, 1, B.data_ptr(), N, 1, Cd.data_ptr(), N, 1, M, N, K, batch=batch,
                   sab=M * K, sbb=K * N, scb=M * N)
    arr = (L.GemmDesc * 1)(d)
    L.check(lib.gtn_gemm_grouped(arr, 1, L.GTN_F64, L.PREC["f64"], L.stream_ptr()), "gemm")


def _report(lib, L):
    n = lib.gtn_autotune_report(None, 0)
    arr = (L.TuneEntry * max(n, 1))()
    lib.gtn_autotune_report(arr, n)
    return [arr[i] for i in range(n)]


def test_tuned_configuration_is_kept_and_correct():
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray
    lib = L.lib()
    lib.gtn_autotune_set(1)
    rng = np.random.default_rng(2)
    M, N, K, batch = 640, 192, 3072, 3
    A, B = rng.standard_normal((batch, M, K)), rng.standard_normal((batch, K, N))
    dA, dB, dC = DeviceArray.from_numpy(A), DeviceArray.from_numpy(B), DeviceArray.empty((batch, M, N))
    before = len(_report(lib, L))
    _gemm(lib, L, dA, dB, dC, M, N, K, batch)
    ents = [e for e in _report(lib, L) if (e.M, e.N, e.K, e.batch) == (M, N, K, batch)]
    assert len(_report(lib, L)) == before + 1 and len(ents) == 1
    e = ents[0]
    print("tuned: tile %d, split target %d, %.3f ms (heuristic %.3f ms)" % (e.tile, e.split_target, e.ms_best,
                                                                          e.ms_default))
    assert e.tile in (64, 128) and e.split_target in (0, 512, 1024, 2048)
    assert e.ms_best <= e.ms_default * 1.02
    _gemm(lib, L, dA, dB, dC, M, N, K, batch)
    assert len(_report(lib, L)) == before + 1
    ref = np.matmul(A, B)
    assert np.abs(dC.numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_no_timing_inside_a_capture():
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray
    from grassmanntn_amd.batch import Graph, _new_pool
    lib = L.lib()
    lib.gtn_autotune_set(1)
    rng = np.random.default_rng(3)
    M, N, K, batch = 320, 96, 1536, 2
    A, B = rng.standard_normal((batch, M, K)), rng.standard_normal((batch, K, N))
    dA, dB, dC = DeviceArray.from_numpy(A), DeviceArray.from_numpy(B), DeviceArray.zeros((batch, M, N))
    before = len(_report(lib, L))
    g = Graph(lambda: _gemm(lib, L, dA, dB, dC, M, N, K, batch), _new_pool())
    assert len(_report(lib, L)) == before
    for _ in range(3):
        dC.fill_(0.0)
        g.launch()
        ref = np.matmul(A, B)
        assert np.abs(dC.numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    g.destroy()
