"""CPU validation of the library's launch geometry in host dry-run mode (no GPU needed).

In dry-run mode the library launches nothing and checks that every GEMM operand range and every
buffer a kernel would touch lies inside the registered regions; placeholder addresses are used.
"""
import ctypes as C

import pytest

from grassmanntn_amd import _lib as L

BASE = 1 << 40
GAP = 1 << 34


class Regions:
    def __init__(self):
        self.items = []

    def add(self, nbytes):
        addr = BASE + GAP * len(self.items)
        self.items.append((addr, max(int(nbytes), 8)))
        return addr

    def install(self, lib):
        n = len(self.items)
        bases = (C.c_uint64 * n)(*[a for a, _ in self.items])
        sizes = (C.c_int64 * n)(*[b for _, b in self.items])
        assert lib.gtn_set_dry_run(n, bases, sizes) == 0


@pytest.fixture
def lib():
    lib = L.load_library()
    yield lib
    lib.gtn_set_dry_run(0, None, None)


SHAPES = [(8, 8, 8), (8, 8, 4), (128, 128, 64), (128, 128, 8), (512, 512, 16), (2048, 2048, 32),
          (2048, 2048, 48), (256, 64, 32), (64, 256, 32), (300, 200, 20), (130, 130, 65), (1, 8, 1),
          (4096, 512, 64)]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("precision", [0, 2])
def test_svd_blocks_geometry(lib, mode, precision):
    shapes = [s for s in SHAPES if mode == 0 or s[0] == s[1]]
    R = Regions()
    arr = (L.SvdBlock * len(shapes))()
    for i, (m, n, k) in enumerate(shapes):
        lda, ldu, ldv = n + 3, k + 1, n + 5
        arr[i].M = R.add(8 * ((m - 1) * lda + n))
        arr[i].U = R.add(8 * ((m - 1) * ldu + k))
        arr[i].S = R.add(8 * k)
        arr[i].Vt = R.add(8 * ((k - 1) * ldv + n))
        arr[i].m, arr[i].n, arr[i].lda, arr[i].ldu, arr[i].ldv, arr[i].k = m, n, lda, ldu, ldv, k
    opts = L.SvdOpts(96, 3, precision, mode, 7)
    ws = lib.gtn_svd_workspace_bytes(arr, len(shapes), C.byref(opts))
    wsa = R.add(ws)
    R.install(lib)
    rc = lib.gtn_svd_blocks(arr, len(shapes), C.byref(opts), C.c_void_p(wsa), ws, None)
    assert rc == 0, lib.gtn_last_error().decode()


def test_dry_run_catches_a_short_buffer(lib):
    R = Regions()
    arr = (L.SvdBlock * 1)()
    m = n = 256
    k = 32
    arr[0].M = R.add(8 * m * n)
    arr[0].U = R.add(8 * m * k - 8)  # one element short
    arr[0].S = R.add(8 * k)
    arr[0].Vt = R.add(8 * k * n)
    arr[0].m, arr[0].n, arr[0].lda, arr[0].ldu, arr[0].ldv, arr[0].k = m, n, n, k, n, k
    opts = L.SvdOpts(96, 2, 0, 0, 7)
    ws = lib.gtn_svd_workspace_bytes(arr, 1, C.byref(opts))
    wsa = R.add(ws)
    R.install(lib)
    assert lib.gtn_svd_blocks(arr, 1, C.byref(opts), C.c_void_p(wsa), ws, None) != 0


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("m,n,k,count", [(2048, 2048, 32, 4), (128, 128, 8, 70), (512, 512, 16, 65),
                                         (256, 64, 32, 3), (64, 256, 32, 2), (300, 200, 20, 5)])
def test_svd_blocks_uniform_runs_geometry(lib, mode, m, n, k, count):
    """Equal blocks at constant strides form one run; every batched launch stays inside the buffers."""
    if mode == 1 and m != n:
        pytest.skip("eig needs square blocks")
    R = Regions()
    arr = (L.SvdBlock * count)()
    lda, ldu, ldv = n, k, n
    sM, sU, sV = m * n, m * k, k * n
    M = R.add(8 * sM * count)
    U = R.add(8 * sU * count)
    S = R.add(8 * k * count)
    V = R.add(8 * sV * count)
    for i in range(count):
        arr[i].M, arr[i].U, arr[i].S, arr[i].Vt = M + 8 * sM * i, U + 8 * sU * i, S + 8 * k * i, V + 8 * sV * i
        arr[i].m, arr[i].n, arr[i].lda, arr[i].ldu, arr[i].ldv, arr[i].k = m, n, lda, ldu, ldv, k
    opts = L.SvdOpts(96, 2, 0, mode, 7)
    ws = lib.gtn_svd_workspace_bytes(arr, count, C.byref(opts))
    wsa = R.add(ws)
    R.install(lib)
    rc = lib.gtn_svd_blocks(arr, count, C.byref(opts), C.c_void_p(wsa), ws, None)
    assert rc == 0, lib.gtn_last_error().decode()


def test_dry_run_catches_a_short_buffer_at_the_end_of_a_run(lib):
    R = Regions()
    count, m, n, k = 6, 256, 256, 32
    arr = (L.SvdBlock * count)()
    M = R.add(8 * m * n * count)
    U = R.add(8 * m * k * count - 8)  # the last block's U is one element short
    S = R.add(8 * k * count)
    V = R.add(8 * k * n * count)
    for i in range(count):
        arr[i].M, arr[i].U, arr[i].S, arr[i].Vt = M + 8 * m * n * i, U + 8 * m * k * i, S + 8 * k * i, V + 8 * k * n * i
        arr[i].m, arr[i].n, arr[i].lda, arr[i].ldu, arr[i].ldv, arr[i].k = m, n, n, k, n, k
    opts = L.SvdOpts(96, 2, 0, 0, 7)
    ws = lib.gtn_svd_workspace_bytes(arr, count, C.byref(opts))
    wsa = R.add(ws)
    R.install(lib)
    assert lib.gtn_svd_blocks(arr, count, C.byref(opts), C.c_void_p(wsa), ws, None) != 0
