"""ctypes binding of libgtn_hip.so (the C ABI declared in include/gtn_hip.h).

The library is loaded on first use and brings its own device memory management, so a process that
uses grassmanntn_amd needs no other GPU framework.  There is no CPU fallback: if the library is
missing, no GPU is visible, or the GPU is not a gfx950, every operation raises GrassmannDeviceError.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgtn_hip.so")

GTN_F64, GTN_F32, GTN_C128, GTN_I8 = 0, 1, 2, 3
GTN_ACT_NONE, GTN_ACT_GELU = 0, 1
PREC = {"f64": 0, "f32": 1, "bf16x3": 2, "bf16x6": 3, "bf16": 4, "bf16_split_a": 5, "f64_int8": 6, "f64_gram": 7}
MAX_AX, MAX_COMP, MAX_BITS = 10, 16, 10


class GrassmannDeviceError(RuntimeError):
    """Raised when the HIP library is unavailable or a kernel call fails."""


class PermDesc(C.Structure):
    _fields_ = [
        ("dtype", C.c_int32), ("n_out", C.c_int32), ("n_virt", C.c_int32), ("n_comp", C.c_int32),
        ("n_in", C.c_int32), ("n_bits", C.c_int32),
        ("out_dim", C.c_int64 * MAX_AX),
        ("out_virt", (C.c_int32 * 2) * MAX_AX),
        ("out_vmul", (C.c_int64 * 2) * MAX_AX),
        ("virt_dim", C.c_int64 * MAX_AX),
        ("virt_enc", C.c_int32 * MAX_AX),
        ("virt_sigma", C.c_int32 * MAX_AX),
        ("virt_ncomp", C.c_int32 * MAX_AX),
        ("comp_dim", C.c_int64 * MAX_COMP),
        ("comp_bit", C.c_int32 * MAX_COMP),
        ("comp_sigma", C.c_int32 * MAX_COMP),
        ("comp_in", (C.c_int32 * 2) * MAX_COMP),
        ("comp_mult", (C.c_int64 * 2) * MAX_COMP),
        ("in_dim", C.c_int64 * MAX_AX),
        ("in_enc", C.c_int32 * MAX_AX),
        ("in_sigma", C.c_int32 * MAX_AX),
        ("in_split", C.c_int32 * MAX_AX),
        ("in_stride", C.c_int64 * MAX_AX),
        ("in_stride_lo", C.c_int64 * MAX_AX),
        ("table", C.c_uint32 * (2 * (1 << MAX_BITS) // 32)),
        ("alpha_re", C.c_double), ("alpha_im", C.c_double),
        ("src_elems", C.c_int64),
    ]


class GemmDesc(C.Structure):
    _fields_ = [
        ("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p),
        ("M", C.c_int64), ("N", C.c_int64), ("K", C.c_int64), ("batch", C.c_int64),
        ("sam", C.c_int64), ("sak", C.c_int64), ("sbk", C.c_int64), ("sbn", C.c_int64),
        ("scm", C.c_int64), ("scn", C.c_int64),
        ("sab", C.c_int64), ("sbb", C.c_int64), ("scb", C.c_int64),
        ("alpha", C.c_double), ("beta", C.c_double),
    ]


class SvdBlock(C.Structure):
    _fields_ = [
        ("M", C.c_void_p), ("U", C.c_void_p), ("S", C.c_void_p), ("Vt", C.c_void_p),
        ("m", C.c_int64), ("n", C.c_int64), ("lda", C.c_int64), ("ldu", C.c_int64), ("ldv", C.c_int64),
        ("k", C.c_int32),
    ]


class SvdOpts(C.Structure):
    _fields_ = [
        ("oversample", C.c_int32), ("iters", C.c_int32), ("precision", C.c_int32), ("mode", C.c_int32),
        ("seed", C.c_uint64), ("orth_every", C.c_int32), ("f64_tail", C.c_int32),
    ]


class AdamwOpts(C.Structure):
    _fields_ = [
        ("lr_max", C.c_double), ("lr_min", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double),
        ("eps", C.c_double), ("weight_decay", C.c_double), ("total_steps", C.c_int64),
    ]


class TuneEntry(C.Structure):
    _fields_ = [
        ("dtype", C.c_int32), ("precision", C.c_int32), ("n_groups", C.c_int32), ("tile", C.c_int32),
        ("split_target", C.c_int32), ("pad_", C.c_int32), ("M", C.c_int64), ("N", C.c_int64), ("K", C.c_int64),
        ("batch", C.c_int64), ("ms_best", C.c_float), ("ms_default", C.c_float),
    ]


_lib = None
_checked = False


def _bind(lib):
    vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int, C.c_double
    sig = {
        "gtn_version": (C.c_int, []),
        "gtn_last_error": (C.c_char_p, []),
        "gtn_set_dry_run": (C.c_int, [i32, C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]),
        "gtn_device_info": (C.c_int, [C.c_char_p, i32, C.POINTER(C.c_int)]),
        "gtn_malloc": (C.c_int, [i64, C.POINTER(C.c_void_p)]),
        "gtn_free": (C.c_int, [vp]),
        "gtn_mem_info": (C.c_int, [C.POINTER(C.c_int64)] * 4),
        "gtn_memcpy": (C.c_int, [vp, vp, i64, i32, vp]),
        "gtn_fill": (C.c_int, [vp, i64, i32, d, vp]),
        "gtn_copy2d": (C.c_int, [vp, i64, vp, i64, i64, i64, i32, vp]),
        "gtn_synchronize": (C.c_int, []),
        "gtn_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
        "gtn_event_record": (C.c_int, [vp, vp]),
        "gtn_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
        "gtn_event_destroy": (C.c_int, [vp]),
        "gtn_index_tables": (C.c_int, [i64, vp, vp, vp, vp]),
        "gtn_permute": (C.c_int, [C.POINTER(PermDesc), vp, vp, vp]),
        "gtn_permute2": (C.c_int, [C.POINTER(PermDesc), C.POINTER(PermDesc), vp, vp, vp]),
        "gtn_sign_tensor": (C.c_int, [i32, C.POINTER(C.c_int64), C.POINTER(C.c_uint32), vp, vp]),
        "gtn_gemm_grouped": (C.c_int, [C.POINTER(GemmDesc), i32, i32, i32, vp]),
        "gtn_gemm_grouped_noalloc": (C.c_int, [C.POINTER(GemmDesc), i32, i32, i32, vp]),
        "gtn_gemm_noalloc_max_scratch": (C.c_int64, []),
        "gtn_mfma_selftest": (C.c_int, [vp, vp]),
        "gtn_mfma_selftest_i8": (C.c_int, [vp, vp]),
        "gtn_autotune_set": (C.c_int, [i32]),
        "gtn_autotune_clear": (C.c_int, []),
        "gtn_autotune_report": (C.c_int, [C.POINTER(TuneEntry), i32]),
        "gtn_sumsq": (C.c_int, [vp, i64, i32, vp, vp]),
        "gtn_scale_by_norm": (C.c_int, [vp, i64, i32, vp, vp]),
        "gtn_reduce_last": (C.c_int, [vp, vp, i64, i64, i32, vp]),
        "gtn_lognorm_update": (C.c_int, [vp, vp, vp]),
        "gtn_power_diag": (C.c_int, [vp, i64, i64, i32, d, d, vp]),
        "gtn_elementwise": (C.c_int, [vp, i64, i32, i32, d, d, vp]),
        "gtn_axpby": (C.c_int, [vp, vp, vp, i64, i32, d, d, d, d, vp]),
        "gtn_odd_norm": (C.c_int, [vp, i32, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32), i32, vp, vp]),
        "gtn_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
        "gtn_stream_destroy": (C.c_int, [vp]),
        "gtn_stream_synchronize": (C.c_int, [vp]),
        "gtn_graph_begin": (C.c_int, [vp, i32]),
        "gtn_graph_end": (C.c_int, [vp, C.POINTER(C.c_void_p)]),
        "gtn_graph_launch": (C.c_int, [vp, vp]),
        "gtn_graph_destroy": (C.c_int, [vp]),
        "gtn_pool_release": (C.c_int, [i32]),
        "gtn_normalize_batch": (C.c_int, [vp, vp, i64, i64, i32, vp, vp, vp, i32, vp, vp]),
        "gtn_diag_interleave": (C.c_int, [vp, i64, i64, vp, i32, vp]),
        "gtn_mlp_bias_act": (C.c_int, [vp, i64, i64, i64, vp, i32, vp, vp]),
        "gtn_mlp_act_backward": (C.c_int, [vp, vp, i64, i32, vp]),
        "gtn_mlp_colsum": (C.c_int, [vp, i64, i64, i64, vp, vp, vp]),
        "gtn_mlp_colsum_workspace": (C.c_int64, [i64, i64]),
        "gtn_mlp_mse": (C.c_int, [vp, vp, vp, i64, i64, i64, vp, vp, vp]),
        "gtn_adamw": (C.c_int, [vp, vp, vp, vp, i64, C.POINTER(AdamwOpts), vp, vp]),
        "gtn_ensemble_stats": (C.c_int, [vp, i64, i64, vp, vp, vp]),
        "gtn_svd_workspace_bytes": (C.c_int64, [C.POINTER(SvdBlock), i32, C.POINTER(SvdOpts)]),
        "gtn_svd_blocks": (C.c_int, [C.POINTER(SvdBlock), i32, C.POINTER(SvdOpts), vp, i64, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load_library():
    """Load the shared library without touching the GPU (used by CPU tests of the ABI)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GrassmannDeviceError(
                "libgtn_hip.so is not built (%s); run `python -m grassmanntn_amd.build`" % LIB_PATH)
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def lib():
    """The library, after checking (once) that a gfx950 GPU is current."""
    global _checked
    L = load_library()
    if not _checked:
        name = C.create_string_buffer(128)
        ncu = C.c_int(0)
        rc = L.gtn_device_info(name, 128, C.byref(ncu))
        if rc != 0:
            raise GrassmannDeviceError("grassmanntn_amd needs an MI355X (gfx950) GPU: " + L.gtn_last_error().decode())
        _checked = True
    return L


def check(rc, what=""):
    if rc != 0:
        msg = _lib.gtn_last_error().decode() if _lib is not None else "unknown"
        raise GrassmannDeviceError("%s failed (code %d): %s" % (what, rc, msg))


_stream = None


def stream_ptr():
    """The stream all host-layer work is ordered on: a non-blocking stream the library creates on
    first use (graph capture needs a stream other than the null stream)."""
    global _stream
    if _stream is None:
        s = C.c_void_p()
        check(lib().gtn_stream_create(C.byref(s)), "gtn_stream_create")
        _stream = s
    return _stream


def ptr(t):
    return C.c_void_p(t.data_ptr())


def dtype_code(t):
    import numpy as np
    dt = np.dtype(t.dtype)
    if dt == np.float64:
        return GTN_F64
    if dt == np.float32:
        return GTN_F32
    if dt == np.complex128:
        return GTN_C128
    if dt == np.int8:
        return GTN_I8
    raise GrassmannDeviceError("unsupported dtype %s" % dt)


def mem_info():
    """{live, cached, device_free, device_total}: bytes held by live allocations, bytes the allocator
    keeps for reuse, and the device's free and total memory (gtn_mem_info)."""
    v = [C.c_int64(0) for _ in range(4)]
    check(lib().gtn_mem_info(*[C.byref(x) for x in v]), "gtn_mem_info")
    return dict(zip(("live", "cached", "device_free", "device_total"), (x.value for x in v)))
