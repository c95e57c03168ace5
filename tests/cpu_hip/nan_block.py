"""Run by tests/test_cpu_hip_nan.py under AddressSanitizer (LD_PRELOAD of the ASan runtime): parity
blocks holding NaN through gtn_svd_blocks on the CPU build of the library.  The Jacobi sort must
stay a permutation (no out-of-bounds access for ASan to catch) and the host must raise its
non-finite error.  Prints one line per case; exits 0 when every case raised that error."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(os.path.dirname(HERE)), os.path.dirname(HERE)]

import numpy as np  # noqa: E402
from kernel_emulator import emulate  # noqa: E402


def main():
    lib = os.path.join(HERE, "_build", "libgtn_cpu.so")
    failures = 0
    with emulate(real_lib_path=lib):
        from grassmanntn_amd import core, _lib as L
        from grassmanntn_amd._device import DeviceArray
        rng = np.random.default_rng(0)
        # blocks of at most 128 rows or columns: Gram matrix of the short side, then k_jacobi
        for (m, n, k, where) in [(24, 20, 8, "all"), (32, 32, 16, "one"), (40, 64, 12, "one")]:
            M = rng.standard_normal((m, n))
            if where == "all":
                M[:] = np.nan
            else:
                M[m // 2, n // 3] = np.nan
            try:
                core._svd_blocks_raw([(DeviceArray.from_numpy(M), k)], 0)
                print("case %dx%d (%s NaN): no error raised" % (m, n, where), flush=True)
                failures += 1
            except L.GrassmannDeviceError as e:
                print("case %dx%d (%s NaN): raised: %s" % (m, n, where, str(e)[:80]), flush=True)
    sys.exit(1 if failures else 0)


if __name__ == "__main__":
    main()
