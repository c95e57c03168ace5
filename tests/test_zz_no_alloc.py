"""The allocation-free grouped GEMM against the allocating one, through the C ABI.

gtn_gemm_grouped_noalloc takes the same descriptors as gtn_gemm_grouped and must give the same
result bit for bit: the shapes below have few output tiles and long reductions, so both entry points
split the reduction and sum the partials -- the allocating one in a workspace it allocates, the
allocation-free one in its fixed scratch array.  Two calls in a row, the second smaller, so that the
second reuses scratch holding the first call's partials; alpha and beta are both exercised.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")


def _call(entry, A, B, C0, M, N, K, batch, prec, alpha, beta):
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray
    from grassmanntn_amd.core import _gemm_desc
    dA, dB, dC = DeviceArray.from_numpy(A), DeviceArray.from_numpy(B), DeviceArray.from_numpy(C0)
    d = _gemm_desc(dA.data_ptr(), K, 1, dB.data_ptr(), N, 1, dC.data_ptr(), N, 1, M, N, K, batch=batch,
                   sab=M * K, sbb=K * N, scb=M * N, alpha=alpha, beta=beta)
    arr = (L.GemmDesc * 1)(d)
    lib = L.lib()
    rc = getattr(lib, entry)(arr, 1, L.dtype_code(dC), L.PREC[prec], L.stream_ptr())
    L.check(rc, entry)
    return dC.numpy()


@pytest.mark.parametrize("dtype,prec", [(np.float64, "f64"), (np.float64, "f32"), (np.float32, "f32")])
def test_noalloc_gemm_matches_allocating_gemm(dtype, prec):
    from grassmanntn_amd import _lib as L
    L.lib()
    assert L.lib().gtn_gemm_noalloc_max_scratch() == 64 << 20
    rng = np.random.default_rng(7)
    # (M, N, K, batch): 32 output tiles split 16 ways (16.8 MB of FP64 partials), then 4 tiles
    for M, N, K, batch in ((256, 256, 8192, 2), (96, 128, 4096, 1)):
        print("case M=%d N=%d K=%d batch=%d %s/%s" % (M, N, K, batch, np.dtype(dtype).name, prec), flush=True)
        A = rng.standard_normal((batch, M, K)).astype(dtype)
        B = rng.standard_normal((batch, K, N)).astype(dtype)
        C0 = rng.standard_normal((batch, M, N)).astype(dtype)
        ref = _call("gtn_gemm_grouped", A, B, C0, M, N, K, batch, prec, 1.25, -0.5)
        got = _call("gtn_gemm_grouped_noalloc", A, B, C0, M, N, K, batch, prec, 1.25, -0.5)
        assert np.array_equal(got, ref)
        exact = 1.25 * np.matmul(A.astype(np.float64), B.astype(np.float64)) - 0.5 * C0
        tol = 1e-12 if (dtype == np.float64 and prec == "f64") else 1e-3
        assert np.abs(got - exact).max() <= tol * np.abs(exact).max()
