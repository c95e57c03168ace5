"""Matrix-core lane layouts the GEMM kernels assume, checked on the device (gtn_mfma_selftest)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")


def expected(form):
    """FP64 16x16x4 returns one row per register (row = lane/16 + 4v); the FP32 16x16x4 and BF16
    16x16x32 forms return four consecutive rows per lane (row = 4 (lane/16) + v)."""
    lane = np.arange(64)[:, None]
    v = np.arange(4)[None, :]
    row = lane // 16 + 4 * v if form == 0 else 4 * (lane // 16) + v
    col = lane % 16
    return (16 * row + col + 1).astype(np.float64).reshape(-1)


@pytest.mark.parametrize("form,name", [(0, "f64 16x16x4"), (1, "f32 16x16x4"), (2, "bf16 16x16x32")])
def test_mfma_lane_layout(form, name):
    import ctypes as C
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray
    out = DeviceArray.zeros((768,), np.float64)
    L.check(L.lib().gtn_mfma_selftest(C.c_void_p(out.data_ptr()), L.stream_ptr()), "gtn_mfma_selftest")
    got = out.numpy()[256 * form:256 * (form + 1)]
    exp = expected(form)
    if not np.array_equal(got, exp):
        # decode what each (lane, register) holds: value - 1 = 16 row + col
        dec = [(int(x) - 1) // 16 if x == int(x) else -1 for x in got]
        print(name, "observed rows per (lane, v):", dec[:32])
        print(name, "observed values:", got[:32])
    np.testing.assert_array_equal(got, exp, err_msg=name)
