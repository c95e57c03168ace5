"""Matrix-core lane layouts the GEMM kernels assume, checked on the device (gtn_mfma_selftest), and the
split-reduction path of gtn_gemm_grouped against numpy."""
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


def _gemm_case(rng, M, N, K, batch, a_rowmajor, scm_pad, alpha, beta):
    """Host operands of one gtn_gemm_grouped group: A (batch, M, K) stored row-major or transposed,
    B (batch, K, N), C (batch, M, N + scm_pad) holding values that beta scales."""
    A = rng.standard_normal((batch, M, K)) if a_rowmajor else rng.standard_normal((batch, K, M))
    B = rng.standard_normal((batch, K, N))
    C = rng.standard_normal((batch, M, N + scm_pad))
    Aop = A if a_rowmajor else A.transpose(0, 2, 1)
    ref = C.copy()
    ref[:, :, :N] = alpha * (Aop @ B) + beta * C[:, :, :N]
    return A, B, C, ref


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 1e-5), ("bf16x6", 1e-5)])
def test_gemm_split_reduction(prec, tol):
    """Launches with few tiles and long reductions are cut into K chunks (ragged last chunk) whose
    partial products are summed afterwards; checked with alpha, beta, batch strides, a padded C and
    a transposed A against numpy."""
    import ctypes as C
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd.core import _gemm_desc, _gemm
    from grassmanntn_amd._device import DeviceArray
    rng = np.random.default_rng(11)
    cases = [(192, 80, 3000, 3, True, 5, 0.5, -2.0), (100, 64, 1000, 1, False, 0, 1.0, 0.0)]
    groups, keep = [], []
    for (M, N, K, batch, rowm, pad, alpha, beta) in cases:
        A, B, Cm, ref = _gemm_case(rng, M, N, K, batch, rowm, pad, alpha, beta)
        dA, dB, dC = DeviceArray.from_numpy(A), DeviceArray.from_numpy(B), DeviceArray.from_numpy(Cm)
        sam, sak = (K, 1) if rowm else (1, M)
        groups.append(_gemm_desc(dA.data_ptr(), sam, sak, dB.data_ptr(), N, 1, dC.data_ptr(), N + pad, 1, M, N, K,
                                 batch=batch, sab=M * K, sbb=K * N, scb=M * (N + pad), alpha=alpha, beta=beta))
        keep.append((dA, dB, dC, ref, N))
    _gemm(groups, L.dtype_code(keep[0][0]), prec)
    for dA, dB, dC, ref, N in keep:
        got = dC.numpy()
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got[:, :, N:], ref[:, :, N:], rtol=0, atol=0)  # padding untouched
        assert np.abs(got - ref).max() <= tol * scale, (prec, np.abs(got - ref).max() / scale)


def test_int8_mfma_lane_layout():
    """v_mfma_i32_16x16x64_i8 (the FP64-by-int8 range finder's product): lane l feeds
    A[l % 16][16 (l / 16) + t] and B[16 (l / 16) + t][l % 16] (bytes t = 0..15) and receives
    D[4 (l / 16) + v][l % 16] in result register v."""
    import ctypes as C
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray
    out = DeviceArray.empty((256,), np.int32)
    L.check(L.lib().gtn_mfma_selftest_i8(C.c_void_p(out.data_ptr()), L.stream_ptr()), "gtn_mfma_selftest_i8")
    got = out.numpy().reshape(64, 4)
    i, k, j = np.arange(16), np.arange(64), np.arange(16)
    A = (i[:, None] + 3 * k[None, :]) % 7 - 3
    B = (2 * j[None, :] + k[:, None]) % 5 - 2
    D = A @ B
    exp = np.array([[D[4 * (l // 16) + v][l % 16] for v in range(4)] for l in range(64)])
    if not np.array_equal(got, exp):
        print("device result registers (lane x register):\n", got)
    np.testing.assert_array_equal(got, exp)


def _layout_case(rng, M, N, K, batch, a_rowmajor, b_rowmajor, dt=np.float64):
    """Operands (element type dt) of one group with A stored (M, K) or (K, M) and B stored (K, N) or
    (N, K), and their FP64 product."""
    A = rng.standard_normal((batch, M, K) if a_rowmajor else (batch, K, M)).astype(dt)
    B = rng.standard_normal((batch, K, N) if b_rowmajor else (batch, N, K)).astype(dt)
    Aop = A if a_rowmajor else A.transpose(0, 2, 1)
    Bop = B if b_rowmajor else B.transpose(0, 2, 1)
    return A, B, Aop.astype(np.float64) @ Bop.astype(np.float64)


@pytest.mark.parametrize("prec,tol,dt", [("f64", 1e-13, np.float64), ("f32", 1e-5, np.float64),
                                         ("f32", 1e-5, np.float32), ("f64", 1e-5, np.float32)])
@pytest.mark.parametrize("tile_launch", ["128", "64"])
def test_gemm_interior_and_edge_panels(prec, tol, dt, tile_launch):
    """Tiles whose operand slabs lie wholly in range (rows inside, K a multiple of the panel depth)
    read their panels through buffer loads at per-slot offsets; edge tiles and ragged reductions
    take the range-checked loads.  Each launch mixes both kinds over all four storage orders of A and
    B (beta = 1 on a zero C, so the launch follows the tile heuristics: 128 x 128 tiles when every
    group is at least 128 x 128 and the launch has 256 such tiles), for FP64 and FP32 data, against
    numpy."""
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd.core import _gemm_desc, _gemm
    from grassmanntn_amd._device import DeviceArray
    rng = np.random.default_rng(23)
    if tile_launch == "128":  # 4 x 4 x 8 + 2 x 2 x 2 x 6 + ... >= 256 tiles of 128 x 128
        cases = [(512, 512, 32, 8, True, True), (256, 256, 48, 6, False, True), (256, 384, 32, 8, True, False),
                 (200, 136, 40, 2, False, False), (384, 256, 16, 8, True, True)]
    else:
        cases = [(192, 128, 48, 2, True, True), (128, 192, 32, 2, False, False), (100, 72, 24, 1, True, False),
                 (64, 64, 16, 3, False, True)]
    groups, keep = [], []
    for (M, N, K, batch, arow, brow) in cases:
        A, B, ref = _layout_case(rng, M, N, K, batch, arow, brow, dt)
        dA, dB = DeviceArray.from_numpy(A), DeviceArray.from_numpy(B)
        dC = DeviceArray.zeros((batch, M, N), dt)
        sam, sak = (K, 1) if arow else (1, M)
        sbk, sbn = (N, 1) if brow else (1, K)
        groups.append(_gemm_desc(dA.data_ptr(), sam, sak, dB.data_ptr(), sbk, sbn, dC.data_ptr(), N, 1, M, N, K,
                                 batch=batch, sab=M * K, sbb=K * N, scb=M * N, alpha=1.0, beta=1.0))
        keep.append((dC, ref, dA, dB))  # the operands stay allocated until the launch has run
    _gemm(groups, L.dtype_code(keep[0][0]), prec)
    for i, (dC, ref, _a, _b) in enumerate(keep):
        err = np.abs(dC.numpy() - ref).max() / np.abs(ref).max()
        assert err <= tol, (cases[i], prec, err)
