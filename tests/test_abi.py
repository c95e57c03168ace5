"""CPU checks of the C ABI: the library loads without a GPU and exports every declared symbol."""
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "gtn_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(gtn_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    from grassmanntn_amd import _lib
    lib = _lib.load_library()
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert lib.gtn_version() >= 100


def test_product_fails_loudly_without_gpu():
    import ctypes as C
    from grassmanntn_amd import _lib
    if _lib.load_library().gtn_device_info(None, 0, C.byref(C.c_int())) == 0:
        pytest.skip("a gfx950 GPU is present")
    import grassmanntn_amd as gtn
    with pytest.raises(gtn.GrassmannDeviceError):
        gtn.dense([[1.0]], statistics=(1, -1))
