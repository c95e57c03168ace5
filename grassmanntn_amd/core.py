"""Grassmann tensors on the GPU and the operations of grassmanntn's hot path.

Mirrors the reference's Python API (dense tensors with statistics / encoder / format, einsum,
join_legs / split_legs, svd / eig with '(ij)(kl)' grouping, hconjugate, sqrt / power).  Tensor data
lives in HBM as DeviceArrays owned by libgtn_hip.so; every numeric step runs in its HIP kernels.
"""
import ctypes as C
import math
import sys

import numpy as np

from . import _lib as L
from ._device import DeviceArray, copy2d
from ._perm import Perm, strides, run_composed
from ._plan import (BOSE, FERMI, HYBRID, SEPARATORS, Contraction, GrassmannError, contraction, denumerate, fail,
                    fresh_char, group_info, parse_partition, sign_table)

ENCODERS = ("canonical", "parity-preserving")
FORMATS = ("standard", "matrix")
ALLOWED_STAT = (0, 1, -1, HYBRID)
NUMER_CUTOFF = 1.0e-14


def _pkg_flag(name, default):
    pkg = sys.modules.get("grassmanntn_amd")
    return getattr(pkg, name, default) if pkg is not None else default


# --------------------------------------------------------------------------------------------------
# precision of the contraction GEMMs
# --------------------------------------------------------------------------------------------------
class _Settings:
    precision = "f64"       # GEMM arithmetic of einsum: f64 | f32 | bf16x3 | bf16x6
    # range-finder arithmetic of the block SVD (svd_spec): FP64 subspace iteration on the Gram matrix
    # M M^T, one FP64 product per iteration (the alternating M^T Y, M Z products: "f64")
    svd_precision = "f64_gram"
    svd_oversample = 96
    svd_iters = 20          # subspace iterations: error ~ (sigma_{p+1} / sigma_k)^(2 iters)
    # CholeskyQR of the basis every this many iterations.  Every iteration: between two
    # orthonormalisations a direction grows by (sigma_i / sigma_0)^(2 every), and one whose squared
    # relative size falls below the Cholesky's 1e-14 pivot floor is dropped as dependent; every
    # second iteration already loses directions with sigma_i / sigma_0 < 0.018, which the kept
    # values of late TRG steps reach (0.0085 at leg dim 64, step 20)
    svd_orth_every = 1
    seed = 12345


settings = _Settings()


def svd_iters_for(k):
    """Subspace iterations of the range finder for blocks keeping k values.  The block width is
    p = min(k + oversample, 128) (128: the p x p problems live in one workgroup's LDS); the subspace
    error falls like (sigma_{p+1} / sigma_k)^(2 iterations), so when the cap shrinks p / k below 4 the
    count grows as (4 k / p)^2: 20 at k <= 32 (leg dim 64, chi 64: ln Z per site within ~5e-11 of
    LAPACK after 20 steps), 45 at k = 48 (chi 96, 8192 x 8192 blocks: 20 iterations leave 1.7e-6
    after three steps, 40 leave 8e-11, 80 leave 2e-14)."""
    p = min(k + settings.svd_oversample, 128)
    r = 4.0 * k / p
    return int(math.ceil(settings.svd_iters * max(1.0, r * r)))


SVD_ONLY = ("f64_int8", "f64_gram")  # range-finder arithmetic that the contraction GEMM does not have


def svd_spec(spec):
    """(range-finder GEMM precision code, FP64 tail) of a decomposition arithmetic: a precision name
    ("f64", "f32", "bf16x3", ...) runs every range-finder iteration in that arithmetic;
    "<name>:<t>" runs the last t iterations in FP64 instead.  Subspace iteration is a fixed-point
    map, so the FP64 iterations remove the reduced iterations' rounding error outside the top-p
    subspace (by (sigma_{p+1} / sigma_k)^2 per iteration); the Rayleigh-Ritz step is FP64 either way."""
    name, _, tail = str(spec).partition(":")
    if name not in L.PREC or (tail and not tail.isdigit()):
        raise ValueError("svd arithmetic must be one of %s, optionally followed by ':<FP64 iterations>'"
                         % list(L.PREC))
    return L.PREC[name], int(tail or 0)


class precision:
    """Context manager: ``with precision('bf16x3'): ...`` sets the contraction GEMM arithmetic."""

    def __init__(self, gemm="f64", svd=None):
        if svd is not None:
            svd_spec(svd)
        if gemm not in L.PREC or gemm in SVD_ONLY:
            raise ValueError("precision must be one of %s" % [k for k in L.PREC if k not in SVD_ONLY])
        self.gemm, self.svd = gemm, svd

    def __enter__(self):
        self._old = (settings.precision, settings.svd_precision)
        settings.precision = self.gemm
        if self.svd is not None:
            settings.svd_precision = self.svd
        return self

    def __exit__(self, *a):
        settings.precision, settings.svd_precision = self._old


MIXED_FP64_STEPS = 5


def mixed_schedule(step, contraction="f32", svd=None, fp64_steps=MIXED_FP64_STEPS):
    """(contraction, svd) arithmetic of coarse-graining step `step` (0-based) in the mixed-precision
    schedule.  Contractions: FP64 in the first `fp64_steps` steps, then `contraction`.  The FP64
    leading steps are there because ln Z per site weighs the normalisation of step n by 2^-n
    (example.py:191-193): FP32 rounding (relative ~1e-7 in a step's norm) costs ~1e-8 when it starts
    at step 4 and ~4e-9 from step 6 on (measured at leg dims 16 and 64).  Decompositions: `svd` (see
    svd_spec; default settings.svd_precision, FP64 throughout); the normalisation and the log-norm
    sum are FP64.  Other range-finder arithmetic measured at leg dim 64, chi 64 with FP64
    contractions, largest |d ln Z/site| over four seeds: split BF16 with its last three iterations
    FP64 ('bf16x3:3') 2.6e-10 (two: 1.8e-9, six: 2.4e-11); FP64 on int8 slices ('f64_int8') and the
    alternating FP64 products ('f64') as the Gram iteration."""
    return ("f64" if step < fp64_steps else contraction, svd if svd is not None else settings.svd_precision)


def _to_device(x):
    if isinstance(x, DeviceArray):
        return x.clone()
    a = np.asarray(x)
    if a.dtype.kind == "c":
        a = a.astype(np.complex128)
    elif a.dtype not in (np.float64, np.float32):
        a = a.astype(np.float64)
    if a.ndim == 0:
        a = a.reshape(1)
    return DeviceArray.from_numpy(a)


def _is_pp(T, axis):
    return T.encoder == "parity-preserving" and T.statistics[axis] in FERMI


# --------------------------------------------------------------------------------------------------
# dense (__init__.py:814-1195)
# --------------------------------------------------------------------------------------------------
class dense:
    """Dense Grassmann tensor: coefficient array in HBM plus per-leg statistics (+1 fermion,
    -1 conjugated fermion, 0 boson, '*' hybrid), encoder and format (__init__.py:818-870)."""

    __array_priority__ = 1000000

    def __init__(self, data=None, encoder="canonical", format="standard", statistics=None):
        if encoder not in ENCODERS:
            fail("Error[dense]: Unknown encoder.")
        if format not in FORMATS:
            fail("Error[dense]: Unknown format.")
        self.data = None
        self.statistics = None
        self.encoder = encoder
        self.format = format
        self._even = None
        given = True
        if isinstance(data, dense):
            self.data = data.data.clone()
            self.statistics = data.statistics
            self.encoder = data.encoder
            self.format = data.format
            self._even = data._even
        elif isinstance(data, sparse):
            self.data = _to_device(data.toarray())
            self.statistics, self.encoder, self.format = data.statistics, data.encoder, data.format
        elif data is None:
            given = False
        else:
            from .block import block
            if isinstance(data, block):
                T = data.todense()
                self.data, self.statistics, self.encoder, self.format = T.data, T.statistics, T.encoder, T.format
            else:
                self.data = _to_device(data)
        if statistics is not None:
            self.statistics = tuple(statistics) if not np.isscalar(statistics) else (statistics,)
        if given and not _pkg_flag("skip_power_of_two_check", False):
            for i, d in enumerate(self.data.shape):
                if self.statistics[i] in FERMI and d != 2 ** int(math.floor(math.log2(d))):
                    fail("Error[dense]: Some of the fermionic tensor shapes are not a power of two."
                         "\n              Have you added the <statistics> argument when calling this function?")

    # ---- array-like properties ----------------------------------------------------------------
    def __getitem__(self, index):
        return self.data.numpy()[index]

    def __setitem__(self, index, value):
        h = self.data.numpy()
        h[index] = value
        self.data = DeviceArray.from_numpy(h)
        self._even = None

    shape = property(lambda self: tuple(self.data.shape))
    size = property(lambda self: self.data.numel())
    ndim = property(lambda self: self.data.ndim)

    @property
    def norm(self):
        return math.sqrt(_sumsq(self.data))

    @property
    def nnz(self):
        return int((np.abs(self.data.numpy()) > _pkg_flag("numer_display_cutoff", 1000 * NUMER_CUTOFF)).sum())

    def numpy(self):
        return self.data.numpy()

    def copy(self):
        r = dense.__new__(dense)
        r.data = self.data.clone()
        r.statistics, r.encoder, r.format, r._even = self.statistics, self.encoder, self.format, self._even
        return r

    def _like(self, data, statistics=None, encoder=None, format=None, even=None):
        r = dense.__new__(dense)
        r.data = data
        r.statistics = self.statistics if statistics is None else tuple(statistics)
        r.encoder = self.encoder if encoder is None else encoder
        r.format = self.format if format is None else format
        r._even = even
        return r

    # ---- arithmetic (__init__.py:933-1000) -----------------------------------------------------
    def __add__(self, other):
        if (self.shape != other.shape or self.statistics != other.statistics or self.format != other.format
                or self.encoder != other.encoder):
            fail("Error[dense.+]: Inconsistent object properties")
        out = DeviceArray.empty(self.data.shape, self.data.dtype)
        _axpby(self.data, other.data, out, 1.0, 1.0)
        ev = True if (self._even and other._even) else None
        return self._like(out, even=ev)

    def __mul__(self, other):
        if not np.isscalar(other):
            fail("Error[dense.*]: Only scalar multiplication is allowed.")
        data = self.data
        if isinstance(other, complex) and not data.is_complex():
            data = data.astype_complex()
        out = DeviceArray.empty(data.shape, data.dtype)
        _axpby(data, None, out, other, 0.0)
        return self._like(out, even=self._even)

    def __truediv__(self, other):
        if not np.isscalar(other):
            fail("Error[dense./]: Only scalar division is allowed.")
        return self * (1.0 / other)

    def __pos__(self):
        return self

    def __neg__(self):
        return self * (-1)

    def __radd__(self, other):
        return self + other

    def __sub__(self, other):
        return self + (-1) * other

    def __rsub__(self, other):
        return other + (-1) * self

    def __rmul__(self, other):
        return self * other

    def __len__(self):
        return len(self.data)

    def __str__(self):
        return str(self.numpy())

    def __repr__(self):
        return repr(self.numpy())

    def info(self, name=None):
        print()
        if name is not None:
            print("        name:", name)
        print("  array type: dense")
        print("       shape:", self.shape)
        print("  statistics:", self.statistics)
        print("      format:", self.format)
        print("     encoder:", self.encoder)
        print("        norm:", self.norm)
        print()

    display = info

    # ---- encoder / format / parity switches (__init__.py:1011-1170) ----------------------------
    def switch_encoder(self, save_memory=False):
        enc = "parity-preserving" if self.encoder == "canonical" else "canonical"
        out = _relayout(self, list(range(self.ndim)), out_encoder=enc)
        return self._like(out, encoder=enc, even=self._even)

    def switch_format(self, save_memory=False):
        for s in self.statistics:
            if s == HYBRID:
                fail("Error[switch_format]: Cannot switch format with a hybrid index.\n"
                     "                      Split them into bosonic and fermionic ones first!")
        sig = [s == -1 for s in self.statistics]
        out = _relayout(self, list(range(self.ndim)), out_encoder=self.encoder, sigma=sig)
        fmt = "matrix" if self.format == "standard" else "standard"
        return self._like(out, format=fmt, even=self._even)

    def switch_parity(self, save_memory=False):
        for s in self.statistics:
            if s == HYBRID:
                fail("Error[switch_parity]: Cannot switch format with a hybrid index.\n"
                     "                      Split them into bosonic and fermionic ones first!")
        lin = [s == 1 for s in self.statistics]
        out = _relayout(self, list(range(self.ndim)), out_encoder=self.encoder, parity_lin=lin)
        return self._like(out, even=self._even)

    def force_encoder(self, target="canonical"):
        if target not in ENCODERS:
            fail("Error[dense.force_encoder]: Unrecognized target encoder.")
        return self.switch_encoder() if target != self.encoder else self.copy()

    def force_format(self, target="standard"):
        if target not in FORMATS:
            fail("Error[dense.force_format]: Unrecognized target format.")
        return self.switch_format() if target != self.format else self.copy()

    def is_even(self):
        """Grassmann evenness (cached): no entry of odd total fermionic parity is nonzero."""
        if self._even is None:
            self._even = is_grassmann_even(self)
        return self._even

    # ---- methods that forward to functions -----------------------------------------------------
    def join_legs(self, string_inp, make_format="standard", intermediate_stat=None, save_memory=False):
        return join_legs(self, string_inp, make_format, intermediate_stat)

    def split_legs(self, string_inp, final_stat, final_shape, intermediate_stat=None, save_memory=False):
        return split_legs(self, string_inp, final_stat, final_shape, intermediate_stat)

    def hconjugate(self, input_string, save_memory=False):
        return hconjugate(self, input_string)

    def svd(self, string_inp, cutoff=None, save_memory=False):
        return svd(self, string_inp, cutoff)

    def eig(self, string_inp, cutoff=None, debug_mode=False, save_memory=False):
        return eig(self, string_inp, cutoff)

    def toblock(self):
        from .block import block
        return block(self)


class sparse:
    """Sparse (COO) Grassmann tensor (__init__.py:1196-1478): coordinates and values on the host,
    with the same statistics / encoder / format properties as ``dense``.  Numeric operations upload
    the tensor to the device as a dense tensor and return a sparse result when the reference does
    (the type of the first operand)."""

    __array_priority__ = 1000000

    def __init__(self, data=None, encoder="canonical", format="standard", statistics=None):
        if encoder not in ENCODERS:
            fail("Error[sparse]: Unknown encoder.")
        if format not in FORMATS:
            fail("Error[sparse]: Unknown format.")
        self.statistics = None
        self.encoder, self.format = encoder, format
        self._c = np.zeros((0, 0), dtype=np.int64)
        self._v = np.zeros(0)
        self._shape = ()
        given = True
        from .block import block
        if isinstance(data, block):
            data = data.todense()
        if isinstance(data, sparse):
            self._c, self._v, self._shape = data._c.copy(), data._v.copy(), data._shape
            self.statistics, self.encoder, self.format = data.statistics, data.encoder, data.format
        elif isinstance(data, dense):
            self._from_array(data.data.numpy())
            self.statistics, self.encoder, self.format = data.statistics, data.encoder, data.format
        elif isinstance(data, (np.ndarray, list)):
            self._from_array(np.array(data))
        elif data is None:
            given = False
        else:
            fail("Error[sparse]: Invalid initialized data")
        if statistics is not None:
            self.statistics = tuple(statistics) if not np.isscalar(statistics) else (statistics,)
        if given and not _pkg_flag("skip_power_of_two_check", False):
            for i, d in enumerate(self._shape):
                if self.statistics[i] in FERMI and d != 2 ** int(math.floor(math.log2(d))):
                    fail("Error[sparse]: Some of the fermionic tensor shapes are not a power of two."
                         "\n               Have you added the <statistics> argument when calling this function?")

    def _from_array(self, a):
        idx = np.nonzero(a)
        self._c = np.array(idx, dtype=np.int64).reshape(a.ndim, -1)
        self._v = a[idx].copy()
        self._shape = tuple(a.shape)

    def toarray(self):
        """The coefficients as a host numpy array."""
        a = np.zeros(self._shape, dtype=self._v.dtype if self._v.size else np.float64)
        if self._v.size:
            a[tuple(self._c)] = self._v
        return a

    shape = property(lambda self: self._shape)
    ndim = property(lambda self: len(self._shape))
    size = property(lambda self: int(np.prod(self._shape)) if self._shape else 1)
    nnz = property(lambda self: int(self._v.size))
    coords = property(lambda self: [tuple(int(x) for x in c) for c in self._c.T])
    value = property(lambda self: list(self._v))

    @property
    def norm(self):
        return float(np.linalg.norm(self._v))

    def _like(self, coords, values):
        r = sparse.__new__(sparse)
        r._c, r._v, r._shape = coords, values, self._shape
        r.statistics, r.encoder, r.format = self.statistics, self.encoder, self.format
        return r

    def copy(self):
        return self._like(self._c.copy(), self._v.copy())

    def remove_zeros(self):
        keep = self._v != 0
        return self._like(self._c[:, keep], self._v[keep])

    def set_value(self, entry, value):
        self._v[entry] = value

    def remove_entry(self, entry):
        keep = np.ones(self.nnz, dtype=bool)
        keep[entry] = False
        self._c, self._v = self._c[:, keep], self._v[keep]

    def append_entry(self, coords, value):
        self._c = np.concatenate([self._c, np.array(coords, dtype=np.int64).reshape(-1, 1)], axis=1)
        self._v = np.concatenate([self._v, np.array([value], dtype=np.result_type(self._v, value))])

    # arithmetic (__init__.py:1377-1405): through the dense form
    def __add__(self, other):
        return sparse(dense(self) + dense(other))

    def __mul__(self, other):
        if not np.isscalar(other):
            fail("Error[sparse.*]: Only scalar multiplication is allowed.")
        return self._like(self._c.copy(), self._v * other)

    def __truediv__(self, other):
        if not np.isscalar(other):
            fail("Error[sparse./]: Only scalar division is allowed.")
        return self._like(self._c.copy(), self._v / other)

    def __pos__(self):
        return self

    def __neg__(self):
        return self * (-1)

    def __radd__(self, other):
        return self + other

    def __sub__(self, other):
        return self + (-1) * other

    def __rsub__(self, other):
        return other + (-1) * self

    def __rmul__(self, other):
        return self * other

    # encoder / format / structure through the device (__init__.py:1426-1478)
    def switch_format(self, save_memory=False):
        return sparse(dense(self).switch_format())

    def switch_encoder(self, save_memory=False):
        return sparse(dense(self).switch_encoder())

    def switch_parity(self, save_memory=False):
        return sparse(dense(self).switch_parity())

    def force_encoder(self, target="canonical"):
        return self.switch_encoder() if target != self.encoder else self.copy()

    def force_format(self, target="standard"):
        return self.switch_format() if target != self.format else self.copy()

    def join_legs(self, string_inp, make_format="standard", intermediate_stat=None, save_memory=False):
        return join_legs(self, string_inp, make_format, intermediate_stat)

    def split_legs(self, string_inp, final_stat, final_shape, intermediate_stat=None, save_memory=False):
        return split_legs(self, string_inp, final_stat, final_shape, intermediate_stat)

    def hconjugate(self, input_string, save_memory=False):
        return hconjugate(self, input_string)

    def svd(self, string_inp, cutoff=None, save_memory=False):
        return svd(self, string_inp, cutoff)

    def eig(self, string_inp, cutoff=None, debug_mode=False, save_memory=False):
        return eig(self, string_inp, cutoff)

    def toblock(self):
        from .block import block
        return block(dense(self))

    def todense(self):
        return dense(self)

    def pinv(self, string, rcond=1e-10):
        return pinv(self, string, rcond=rcond)


def _sparse_in(fn):
    """Run `fn` on the dense form of a sparse first argument and return sparse results (the
    reference's this_type == sparse branches)."""
    import functools

    @functools.wraps(fn)
    def wrapped(T, *a, **k):
        if not isinstance(T, sparse):
            return fn(T, *a, **k)
        out = fn(dense(T), *a, **k)
        if isinstance(out, tuple):
            return tuple(sparse(x) if isinstance(x, dense) else x for x in out)
        return sparse(out) if isinstance(out, dense) else out
    return wrapped


# --------------------------------------------------------------------------------------------------
# device helpers
# --------------------------------------------------------------------------------------------------
def _sumsq(t):
    lib = L.lib()
    out = DeviceArray.empty(1, np.float64)
    L.check(lib.gtn_sumsq(L.ptr(t), t.numel(), L.dtype_code(t), L.ptr(out), L.stream_ptr()), "gtn_sumsq")
    return float(out.item())


def _axpby(x, y, out, a, b):
    lib = L.lib()
    a, b = complex(a), complex(b)
    L.check(lib.gtn_axpby(L.ptr(x), L.ptr(y) if y is not None else None, L.ptr(out), x.numel(), L.dtype_code(x),
                          a.real, a.imag, b.real, b.imag, L.stream_ptr()), "gtn_axpby")


def is_grassmann_even(T):
    """(__init__.py:3900-3925) True iff every entry of odd total fermionic parity is exactly zero."""
    from .block import block
    if isinstance(T, block):
        return T.is_even()
    lib = L.lib()
    nd = T.ndim
    dims = (C.c_int64 * max(nd, 1))(*T.shape)
    fer = (C.c_int32 * max(nd, 1))(*[1 if s in FERMI else 0 for s in T.statistics])
    pp = (C.c_int32 * max(nd, 1))(*[1 if T.encoder == "parity-preserving" else 0] * nd)
    out = DeviceArray.empty(1, np.float64)
    L.check(lib.gtn_odd_norm(L.ptr(T.data), nd, dims, fer, pp, L.dtype_code(T.data), L.ptr(out), L.stream_ptr()),
            "gtn_odd_norm")
    return float(out.item()) == 0.0


def _relayout(T, order, out_encoder, sigma=None, parity_lin=None, table_chars=None, codes=None, alpha=1.0):
    """out axis a = T axis order[a], in `out_encoder`; sigma[a]: multiply sigma(canonical index) of
    source axis a; parity_lin[a]: multiply (-1)^p of source axis a."""
    P = Perm()
    ins = []
    st = strides(T.shape)
    for a, d in enumerate(T.shape):
        ins.append(P.add_in(d, st[a], enc=_is_pp(T, a)))
    bits = {}
    lin_axes = [a for a in range(T.ndim) if parity_lin and parity_lin[a]]
    for k, a in enumerate(lin_axes):
        bits[a] = k
    for a in order:
        d = T.shape[a]
        pp = out_encoder == "parity-preserving" and T.statistics[a] in FERMI
        v = P.add_virt(d, enc=pp)
        P.add_comp(v, d, [(ins[a], 1)], bit=bits.get(a, -1), sigma=bool(sigma and sigma[a]))
        P.add_out(d, [(v, 1)])
    if lin_axes:
        n = len(lin_axes)
        P.set_table([bin(m).count("1") & 1 for m in range(1 << n)])
    return P.run(T.data, alpha=alpha)


# --------------------------------------------------------------------------------------------------
# einsum (__init__.py:2308-2344 dispatcher; einsum_ds :1633-2306)
# --------------------------------------------------------------------------------------------------
def einsum(*args, ignore_anticommutation=False):
    subscripts = denumerate(args[0].replace(" ", ""))
    if subscripts.count("->") > 1:
        fail("Error[einsum]: There should be no more than one '->' indicator in the subscript input.")
    inp = subscripts.split("->")[0]
    nobj = len(inp.split(","))
    ops = list(args[1:1 + nobj])
    from .block import block, einsum_block
    if isinstance(ops[0], block):
        return einsum_block(subscripts, *ops, ignore_anticommutation=ignore_anticommutation)
    if isinstance(ops[0], sparse):
        # computed dense on the device; the result takes the first operand's type (__init__.py:2265)
        res = einsum(args[0], *[dense(o) if isinstance(o, sparse) else o for o in ops],
                     ignore_anticommutation=ignore_anticommutation)
        return sparse(res) if isinstance(res, dense) else res
    ops = [o if isinstance(o, dense) else dense(o) for o in ops]
    first = ops[0]
    if nobj >= 3:
        return _einsum_pairwise(subscripts, ops, ignore_anticommutation)
    plan = contraction(subscripts, [(o.statistics, o.shape) for o in ops])
    if nobj == 1:
        res = _einsum1(plan, ops[0], first, ignore_anticommutation)
    else:
        res = _einsum2(plan, ops[0], ops[1], first, ignore_anticommutation)
    return res


def _einsum_pairwise(subscripts, ops, ignore):
    """Contract the first two operands into an intermediate whose legs keep their order of
    appearance, then recurse (Grassmann contraction is associative under this ordering)."""
    inp, out = (subscripts.split("->") + [""])[:2] if "->" in subscripts else (subscripts, None)
    strs = inp.split(",")
    s1, s2, rest = strs[0], strs[1], strs[2:]
    later = "".join(rest) + (out or "")
    keep = []
    for c in s1 + s2:
        if c in keep:
            continue
        n12 = (s1 + s2).count(c)
        if c in later or n12 == 1:
            keep.append(c)
    inter = "".join(keep)
    mid = einsum(s1 + "," + s2 + "->" + inter, ops[0], ops[1], ignore_anticommutation=ignore)
    sub2 = ",".join([inter] + rest) + ("->" + out if out is not None else "")
    return einsum(sub2, mid, *ops[2:], ignore_anticommutation=ignore)


def _fermionic(plan, c):
    return any(t in FERMI for t in plan.stat_of[c])


def _prod(xs):
    r = 1
    for x in xs:
        r *= x
    return r


def _out_encoder_pp(first, stat):
    return first.encoder == "parity-preserving" and stat in FERMI


class _PermutedView(dense):
    """A pure leg permutation of a dense tensor (an einsum such as 'ijkl->jkli') whose gather is
    deferred: `join_legs` reads the source directly with the permutation's sign table folded into its
    own gather; any other use materialises the permuted tensor on first access of `.data`."""

    def __init__(self, src, order, codes, fbits, stats):
        self._src, self._order, self._codes, self._fbits = src, tuple(order), list(codes), list(fbits)
        self._data = None
        self.statistics = tuple(stats)
        self.encoder, self.format = src.encoder, src.format
        self._even = src._even

    @property
    def data(self):
        if self._data is None:
            self._data = _gather_permuted(self._src, self._order, self._codes, self._fbits)
        return self._data

    @data.setter
    def data(self, value):
        self._data = value

    shape = property(lambda self: tuple(self._src.shape[a] for a in self._order) if self._data is None
                     else tuple(self._data.shape))
    ndim = property(lambda self: len(self._order))


class _PendingOutput(dense):
    """A contraction result whose output gather is deferred (einsum inside `deferred_outputs()`):
    `_src` holds the GEMM's parity-block result and `_perm` the gather that makes the dense tensor.
    The operand gather of a following contraction composes with it (one pass over the data,
    gtn_permute2); any other use materialises the tensor on first access of `.data`."""

    def __init__(self, perm, src, stats, encoder, fmt, even):
        self._perm, self._src = perm, src
        self._shape = tuple(perm.out_shape())
        self._data = None
        self.statistics = tuple(stats)
        self.encoder, self.format = encoder, fmt
        self._even = even

    @property
    def data(self):
        if self._data is None:
            self._data = self._perm.run(self._src)
        return self._data

    @data.setter
    def data(self, value):
        self._data = value

    shape = property(lambda self: self._shape if self._data is None else tuple(self._data.shape))
    ndim = property(lambda self: len(self.shape))


_defer = [False]


class deferred_outputs:
    """Context manager: two-operand contractions inside return _PendingOutput results, whose output
    gather merges into the next contraction's operand gather."""

    def __enter__(self):
        self._old = _defer[0]
        _defer[0] = True
        return self

    def __exit__(self, *a):
        _defer[0] = self._old


def _source_dtype(T):
    return T._src.dtype if isinstance(T, _PendingOutput) and T._data is None else T.data.dtype


def _gather_permuted(src, order, codes, fbits):
    """out[x] = code(parities of x) * src[x permuted]: out axis a is source axis order[a]; bit k of the
    table index is the parity of out axis fbits[k]."""
    P = Perm()
    st = strides(src.shape)
    ins = [P.add_in(src.shape[b], st[b], enc=_is_pp(src, b), sigma=(src.format == "matrix" and src.statistics[b] == -1))
           for b in range(src.ndim)]
    for a, b in enumerate(order):
        pp = _is_pp(src, b)
        sig = src.format == "matrix" and src.statistics[b] == -1
        v = P.add_virt(src.shape[b], enc=pp, sigma=sig)
        P.add_comp(v, src.shape[b], [(ins[b], 1)], bit=fbits.index(a) if a in fbits else -1)
        P.add_out(src.shape[b], [(v, 1)])
    P.set_table(codes)
    return P.run(src.data)


def permuted_view(T, subscripts):
    """einsum(subscripts, T) for a pure permutation of T's legs (each index once on both sides), with
    the gather deferred (see _PermutedView); other subscripts run einsum."""
    sub = denumerate(subscripts.replace(" ", ""))
    if "->" not in sub or not isinstance(T, dense) or isinstance(T, _PermutedView) or T.data.is_complex():
        return einsum(subscripts, T)
    s, out = sub.split("->")
    if len(set(s)) != len(s) or sorted(s) != sorted(out) or len(s) != T.ndim:
        return einsum(subscripts, T)
    plan = contraction(s + "->" + out, [(T.statistics, T.shape)])
    if plan.pairs:
        return einsum(subscripts, T)
    fa = [c for c in dict.fromkeys(s) if _fermionic(plan, c)]
    lin, quad = plan.quadratic_form()
    codes = sign_table(fa, lin, quad)
    order = [s.index(c) for c in out]
    fbits = [out.index(c) for c in fa]
    return _PermutedView(T, order, codes, fbits, plan.out_stats)


def _einsum1(plan, A, first, ignore):
    s = plan.inputs[0]
    out = plan.output
    dims = plan.dims
    summed = [c for c in dict.fromkeys(s) if c not in out]
    fa = [c for c in dict.fromkeys(s) if _fermionic(plan, c)]
    lin, quad = ({}, {}) if ignore else plan.quadratic_form()
    codes = sign_table(fa, lin, quad)
    P = Perm()
    st = strides(A.shape)
    ins = [P.add_in(A.shape[a], st[a], enc=_is_pp(A, a), sigma=(A.format == "matrix" and A.statistics[a] == -1))
           for a in range(A.ndim)]
    where = {}
    for a, c in enumerate(s):
        where.setdefault(c, []).append(ins[a])
    pairs = set(plan.pairs) if not ignore else set()
    for c in list(out) + summed:
        tg = [(i, 1) for i in where[c]]
        if len(tg) > 2:
            raise NotImplementedError("index '%s' repeated more than twice on one operand" % c)
        pp = c in out and _out_encoder_pp(first, plan.stat_of[c][0])
        sig_out = c in out and first.format == "matrix" and plan.stat_of[c][0] == -1
        v = P.add_virt(dims[c], enc=pp, sigma=sig_out)
        P.add_comp(v, dims[c], tg, bit=fa.index(c) if c in fa else -1, sigma=c in pairs)
        P.add_out(dims[c], [(v, 1)])
    P.set_table(codes)
    g = P.run(A.data)
    red = _prod(dims[c] for c in summed)
    if summed:
        rows = g.numel() // red
        r = DeviceArray.empty(rows, g.dtype)
        L.check(L.lib().gtn_reduce_last(L.ptr(g), L.ptr(r), rows, red, L.dtype_code(g), L.stream_ptr()), "reduce")
        g = r.reshape(tuple(dims[c] for c in out))
    if not plan.has_output:
        return g.numpy().reshape(-1)[0].item()
    res = dense.__new__(dense)
    res.data = g.reshape(tuple(dims[c] for c in out))
    res.statistics = plan.out_stats
    res.encoder, res.format = first.encoder, first.format
    res._even = True if A._even else None
    return res


class GemmProfile:
    """Collects HIP-event durations of every grouped-GEMM launch while active (bench only):
    key (M, N, K, groups, batch, precision) -> list of milliseconds."""

    active = None

    def __init__(self):
        self.pending = []
        self.ms = {}

    def __enter__(self):
        GemmProfile.active = self
        return self

    def __exit__(self, *a):
        GemmProfile.active = None
        self.resolve()

    def resolve(self):
        for key, e0, e1 in self.pending:
            self.ms.setdefault(key, []).append(e0.elapsed_ms(e1))
        self.pending = []


def _gemm(groups, dtype_code, prec):
    lib = L.lib()
    arr = (L.GemmDesc * len(groups))(*groups)
    prof = GemmProfile.active
    if prof is not None:
        from ._device import Event
        e0, e1 = Event(), Event()
        e0.record()
    L.check(lib.gtn_gemm_grouped(arr, len(groups), dtype_code, L.PREC[prec], L.stream_ptr()), "gtn_gemm_grouped")
    if prof is not None:
        e1.record()
        g = groups[0]
        prof.pending.append(((int(g.M), int(g.N), int(g.K), len(groups), int(g.batch), prec), e0, e1))


def _gemm_desc(A, sam, sak, B, sbk, sbn, Cp, scm, scn, M, N, K, batch=1, sab=0, sbb=0, scb=0, alpha=1.0, beta=0.0):
    d = L.GemmDesc()
    d.A, d.B, d.C = A, B, Cp
    d.M, d.N, d.K, d.batch = M, N, K, batch
    d.sam, d.sak, d.sbk, d.sbn, d.scm, d.scn = sam, sak, sbk, sbn, scm, scn
    d.sab, d.sbb, d.scb = sab, sbb, scb
    d.alpha, d.beta = alpha, beta
    return d


def _operand_gather(T, s, plan, layout, fchars, codes, sig_chars):
    """Gather operand T (index string s) into the GEMM layout.

    layout: list of raw axes, each ('b', char) bosonic/batch leg, ('P',) parity, ('F', group) the
    half (blocked) or full fused fermionic index of a group; groups: dict name -> list of chars."""
    P = Perm()
    st = strides(T.shape)
    ins = [P.add_in(T.shape[a], st[a], enc=_is_pp(T, a), sigma=(T.format == "matrix" and T.statistics[a] == -1))
           for a in range(T.ndim)]
    where = {}
    for a, c in enumerate(s):
        where.setdefault(c, []).append(ins[a])
    dims = plan.dims
    virt_of = {}
    for item in layout["axes"]:
        if item[0] == "b":
            c = item[1]
            v = P.add_virt(dims[c])
            P.add_comp(v, dims[c], [(i, 1) for i in where[c]], bit=fchars.index(c) if c in fchars else -1,
                       sigma=c in sig_chars)
            virt_of[("b", c)] = v
        elif item[0] == "F":
            grp = layout["groups"][item[1]]
            v = P.add_virt(_prod(dims[c] for c in grp), enc=True)
            for c in grp:
                tg = [(i, 1) for i in where[c]]
                if len(tg) > 2:
                    raise NotImplementedError("index '%s' repeated more than twice on one operand" % c)
                P.add_comp(v, dims[c], tg, bit=fchars.index(c) if c in fchars else -1, sigma=c in sig_chars)
            virt_of[("F", item[1])] = v
    blocked = layout["blocked"]
    for item in layout["axes"]:
        if item[0] == "b":
            P.add_out(dims[item[1]], [(virt_of[("b", item[1])], 1)])
        elif item[0] == "P":
            P.add_out(2, [(virt_of[("F", g)], 1) for g in layout["pgroups"]])
        else:
            full = _prod(dims[c] for c in layout["groups"][item[1]])
            if blocked:
                P.add_out(full // 2, [(virt_of[("F", item[1])], 2)])
            else:
                P.add_out(full, [(virt_of[("F", item[1])], 1)])
    P.set_table(codes)
    if isinstance(T, _PendingOutput) and T._data is None:
        return run_composed(P, T._perm, T._src)
    return P.run(T.data)


def _einsum2(plan, A, B, first, ignore):
    sA, sB = plan.inputs
    out = plan.output
    dims = plan.dims
    setA, setB, setO = set(sA), set(sB), set(out)
    uA, uB = list(dict.fromkeys(sA)), list(dict.fromkeys(sB))
    dA, dB = np.dtype(_source_dtype(A)), np.dtype(_source_dtype(B))
    if not all(d in (np.dtype(np.float64), np.dtype(np.complex128)) for d in (dA, dB)) and dA != dB:
        raise NotImplementedError("operands of different dtypes")
    isF = lambda c: _fermionic(plan, c)
    batch = [c for c in uA if c in setB and c in setO]
    contr = [c for c in uA if c in setB and c not in setO]
    freeA = [c for c in out if c in setA and c not in setB]
    freeB = [c for c in out if c in setB and c not in setA]
    redA = [c for c in uA if c not in setB and c not in setO]
    redB = [c for c in uB if c not in setA and c not in setO]
    RB, RF = [c for c in freeA if not isF(c)], [c for c in freeA if isF(c)]
    KB, KF = [c for c in contr if not isF(c)], [c for c in contr if isF(c)]
    NB, NF = [c for c in freeB if not isF(c)], [c for c in freeB if isF(c)]
    for c in batch:
        if isF(c):
            fail("Error[einsum]: Inconsistent index statistics.")
    # a fermionic group of total dimension 1 carries no parity
    hasR = _prod(dims[c] for c in RF) >= 2
    hasK = _prod(dims[c] for c in KF) >= 2
    hasN = _prod(dims[c] for c in NF) >= 2
    blocked = A.is_even() and B.is_even()
    split = blocked and hasR and hasK and hasN
    fa = [c for c in uA if isF(c)]
    fb = [c for c in uB if isF(c)]
    fo = [c for c in out if isF(c)]
    if ignore:
        partA = partB = partO = ({}, {})
    else:
        partA, partB, partO = plan.split_form([set(fa), set(fb), set(fo)])
    sigA = set() if ignore else {c for c in plan.pairs if c in setA}
    sigB = set() if ignore else {c for c in plan.pairs if c not in setA}
    codesA = sign_table(fa, *partA)
    codesB = sign_table(fb, *partB)
    codesO = sign_table(fo, *partO, zero_if_odd=blocked)
    if blocked and not split and hasR and hasN:
        # an outer product of even operands: only the even-parity row and column blocks exist (the
        # GEMM computes that one block), so entries with an odd fused row or column are zero
        rb = [fo.index(c) for c in RF]
        for m in range(len(codesO)):
            if sum((m >> b) & 1 for b in rb) & 1:
                codesO[m] = 2

    def layout_for(rowsB, rowsF, hasRows, colsB, colsF, hasCols, red, gname):
        axes = [("b", c) for c in batch]
        if split:
            axes.append(("P",))
        axes += [("b", c) for c in rowsB]
        if hasRows:
            axes.append(("F", gname[0]))
        axes += [("b", c) for c in colsB]
        if hasCols:
            axes.append(("F", gname[1]))
        axes += [("b", c) for c in red]
        groups = {gname[0]: rowsF, gname[1]: colsF}
        return {"axes": axes, "groups": groups, "blocked": blocked, "pgroups": list(gname)}

    # fermionic groups of dimension 1 are gathered as plain legs
    def plainify(Fs, has):
        return ([], Fs) if has else (Fs, [])

    RFb, RFf = plainify(RF, hasR)
    KFb, KFf = plainify(KF, hasK)
    NFb, NFf = plainify(NF, hasN)
    layA = layout_for(RB + RFb, RFf, hasR, KB + KFb, KFf, hasK, redA, ("R", "K"))
    layB = layout_for(KB + KFb, KFf, hasK, NB + NFb, NFf, hasN, redB, ("K", "N"))
    Ag = _operand_gather(A, sA, plan, layA, fa, codesA, sigA)
    Bg = _operand_gather(B, sB, plan, layB, fb, codesB, sigB)
    half = 2 if blocked else 1
    nb = _prod(dims[c] for c in batch)
    nP = 2 if split else 1
    M = _prod(dims[c] for c in RB + RFb) * (_prod(dims[c] for c in RFf) // half if hasR else 1)
    K = _prod(dims[c] for c in KB + KFb) * (_prod(dims[c] for c in KFf) // half if hasK else 1)
    N = _prod(dims[c] for c in NB + NFb) * (_prod(dims[c] for c in NFf) // half if hasN else 1)
    ra, rb = _prod(dims[c] for c in redA), _prod(dims[c] for c in redB)
    lib = L.lib()
    if ra > 1:
        t = DeviceArray.empty(Ag.numel() // ra, Ag.dtype)
        L.check(lib.gtn_reduce_last(L.ptr(Ag), L.ptr(t), t.numel(), ra, L.dtype_code(Ag), L.stream_ptr()), "reduce")
        Ag = t
    if rb > 1:
        t = DeviceArray.empty(Bg.numel() // rb, Bg.dtype)
        L.check(lib.gtn_reduce_last(L.ptr(Bg), L.ptr(t), t.numel(), rb, L.dtype_code(Bg), L.stream_ptr()), "reduce")
        Bg = t
    Cg = _contract_blocks(Ag, Bg, nb, nP, M, N, K)
    if not plan.has_output:
        return Cg.numpy().reshape(-1)[0].item()
    # output gather from the GEMM layout [batch][P][RB][RFh][NB][NFw]
    shapeC = [dims[c] for c in batch] + ([2] if split else []) + [dims[c] for c in RB + RFb]
    if hasR:
        shapeC.append(_prod(dims[c] for c in RFf) // half)
    shapeC += [dims[c] for c in NB + NFb]
    if hasN:
        shapeC.append(_prod(dims[c] for c in NFf) // half)
    stC = strides(shapeC)
    P = Perm()
    pos = 0
    src = {}
    for c in batch:
        src[c] = (P.add_in(dims[c], stC[pos]), 1)
        pos += 1
    p_stride = 0
    if split:
        p_stride = stC[pos]
        pos += 1
    for c in RB + RFb:
        src[c] = (P.add_in(dims[c], stC[pos]), 1)
        pos += 1
    if hasR:
        full = _prod(dims[c] for c in RFf)
        i = P.add_in(full, stC[pos], enc=True, split=blocked, stride_lo=p_stride)
        pos += 1
        m = full
        for c in RFf:
            m //= dims[c]
            src[c] = (i, m)
    for c in NB + NFb:
        src[c] = (P.add_in(dims[c], stC[pos]), 1)
        pos += 1
    if hasN:
        full = _prod(dims[c] for c in NFf)
        i = P.add_in(full, stC[pos], enc=True, split=blocked, stride_lo=0)
        pos += 1
        m = full
        for c in NFf:
            m //= dims[c]
            src[c] = (i, m)
    for c in out:
        st = plan.stat_of[c][0]
        v = P.add_virt(dims[c], enc=_out_encoder_pp(first, st), sigma=(first.format == "matrix" and st == -1))
        i, m = src[c]
        P.add_comp(v, dims[c], [(i, m)], bit=fo.index(c) if c in fo else -1)
        P.add_out(dims[c], [(v, 1)])
    P.set_table(codesO)
    if _defer[0]:
        return _PendingOutput(P, Cg, plan.out_stats, first.encoder, first.format, True if blocked else None)
    res = dense.__new__(dense)
    res.data = P.run(Cg)
    res.statistics = plan.out_stats
    res.encoder, res.format = first.encoder, first.format
    res._even = True if blocked else None
    return res


def _planes_of_complex(X):
    """(2, n) FP64 device array: the real then the imaginary parts of complex128 X (n values), one
    gather over X's storage read as 2n doubles."""
    n = X.numel()
    raw = DeviceArray(X.storage, X.offset, (2 * n,), np.float64)
    P = Perm()
    i_el, i_part = P.add_in(n, 2), P.add_in(2, 1)
    v_part, v_el = P.add_virt(2), P.add_virt(n)
    P.add_comp(v_part, 2, [(i_part, 1)])
    P.add_comp(v_el, n, [(i_el, 1)])
    P.add_out(2, [(v_part, 1)])
    P.add_out(n, [(v_el, 1)])
    return P.run(raw)


def _complex_of_planes(R, n):
    """complex128 array of n values from the (2, n) FP64 planes R (real, imaginary)."""
    out = DeviceArray.empty((n,), np.complex128)
    raw = DeviceArray(out.storage, out.offset, (2 * n,), np.float64)
    P = Perm()
    i_part, i_el = P.add_in(2, n), P.add_in(n, 1)
    v_el, v_part = P.add_virt(n), P.add_virt(2)
    P.add_comp(v_el, n, [(i_el, 1)])
    P.add_comp(v_part, 2, [(i_part, 1)])
    P.add_out(n, [(v_el, 1)])
    P.add_out(2, [(v_part, 1)])
    P.run(R, dst=raw)
    return out


def _contract_blocks_complex(Ag, Bg, nb, nP, M, N, K):
    """C[b][p] = A[b][p] B[b][p] with complex128 blocks (one operand may be real FP64): the complex
    operands are gathered into real and imaginary planes, combined by real GEMMs -- Re C = Re A Re B
    - Im A Im B, Im C = Re A Im B + Im A Re B, the terms of a real operand's zero imaginary part left
    out -- in two launches (the second accumulates, beta = 1), and interleaved back."""
    ca, cb = Ag.dtype.kind == "c", Bg.dtype.kind == "c"
    na, nbn = Ag.numel(), Bg.numel()
    Ap = _planes_of_complex(Ag) if ca else Ag
    Bp = _planes_of_complex(Bg) if cb else Bg
    nc = nb * nP * M * N
    Cp = DeviceArray.empty((2, nc))
    es = 8

    def desc(imag_a, imag_b, imag_c, alpha, beta, p):
        a = Ap.data_ptr() + es * ((na if imag_a else 0) + p * M * K)
        b = Bp.data_ptr() + es * ((nbn if imag_b else 0) + p * K * N)
        c = Cp.data_ptr() + es * ((nc if imag_c else 0) + p * M * N)
        return _gemm_desc(a, K, 1, b, N, 1, c, N, 1, M, N, K, batch=nb, sab=nP * M * K, sbb=nP * K * N,
                          scb=nP * M * N, alpha=alpha, beta=beta)

    first, second = [], []
    for p in range(nP):
        first.append(desc(False, False, False, 1.0, 0.0, p))
        first.append(desc(False, True, True, 1.0, 0.0, p) if cb else desc(True, False, True, 1.0, 0.0, p))
        if ca and cb:
            second.append(desc(True, True, False, -1.0, 1.0, p))
            second.append(desc(True, False, True, 1.0, 1.0, p))
    _gemm(first, L.GTN_F64, settings.precision)
    if second:
        _gemm(second, L.GTN_F64, settings.precision)
    return _complex_of_planes(Cp, nc)


def _contract_blocks(Ag, Bg, nb, nP, M, N, K):
    """C[b][p] = A[b][p] (M x K) * B[b][p] (K x N) for every batch b and parity block p."""
    if Ag.dtype.kind == "c" or Bg.dtype.kind == "c":
        return _contract_blocks_complex(Ag, Bg, nb, nP, M, N, K)
    dt = L.dtype_code(Ag)
    prec = settings.precision if Ag.dtype == np.float64 else "f32"
    Cg = DeviceArray.empty(nb * nP * M * N, Ag.dtype)
    small = M * N * nb * nP <= 256 and K >= 8192
    if small:
        # split K so that a dot-product-shaped contraction spreads over many workgroups
        S = 1
        while S < 256 and K % (2 * S) == 0 and K // (2 * S) >= 512:
            S *= 2
        part = DeviceArray.empty(nb * nP * M * N * S, Ag.dtype)
        groups = []
        for b in range(nb):
            for p in range(nP):
                base = (b * nP + p)
                es = Ag.element_size()
                groups.append(_gemm_desc(Ag.data_ptr() + es * base * M * K, K, 1, Bg.data_ptr() + es * base * K * N, N, 1,
                                         part.data_ptr() + es * base * M * N * S, N * S, S, M, N, K // S, batch=S,
                                         sab=K // S, sbb=(K // S) * N, scb=1))
        _gemm(groups, dt, prec)
        L.check(L.lib().gtn_reduce_last(L.ptr(part), L.ptr(Cg), nb * nP * M * N, S, dt, L.stream_ptr()), "reduce")
        return Cg
    esz = Ag.element_size()
    groups = []
    for p in range(nP):
        groups.append(_gemm_desc(Ag.data_ptr() + esz * p * M * K, K, 1, Bg.data_ptr() + esz * p * K * N, N, 1,
                                 Cg.data_ptr() + esz * p * M * N, N, 1, M, N, K, batch=nb,
                                 sab=nP * M * K, sbb=nP * K * N, scb=nP * M * N))
    _gemm(groups, dt, prec)
    return Cg


# --------------------------------------------------------------------------------------------------
# join_legs / split_legs (__init__.py:2947-3316)
# --------------------------------------------------------------------------------------------------
def _grouping_sign_axes(srt, inter):
    """positions (in the original leg order) that get (-1)^p: +1 legs in a group of intermediate -1"""
    if len(srt) != len(inter):
        fail("Error[get_grouping_sign_factors]: Inconsistent number of intermediate_stat and groupings!")
    out = set()
    for g, (chars, st, dm, pos) in enumerate(srt):
        for t, p in zip(st, pos):
            if t == 1 and inter[g] == -1:
                out.add(p)
    return out


def _intermediate(srt, inter):
    """Per group: (boson dim, fermion dim, stats of the fused legs, final stat)."""
    res = []
    for g, (chars, st, dm, pos) in enumerate(srt):
        bd = _prod(d for t, d in zip(st, dm) if t in BOSE)
        fd = _prod(d for t, d in zip(st, dm) if t in FERMI)
        bn = sum(1 for t in st if t in BOSE)
        fn = sum(1 for t in st if t in FERMI)
        final = HYBRID if (bn and fn) else inter[g]
        res.append((bd, fd, bn, fn, final))
    return res


@_sparse_in
def join_legs(T, string_inp, make_format="standard", intermediate_stat=None, save_memory=False):
    """Fuse groups of legs into one leg each; the result is parity-preserving (__init__.py:2947-3068).
    Bosons move left inside a group, +1 legs in a -1 group pick up (-1)^p, the fermionic legs fuse in
    C order (first leg = high bits), matrix format applies sigma to each fused -1 leg, then the
    encoder switches on the fused leg."""
    P, src_data, fin_shape, fin_stats = _join_gather(T, string_inp, make_format, intermediate_stat)
    out = T._like(P.run(src_data), statistics=fin_stats, encoder="parity-preserving", format=make_format,
                  even=T._even)
    out.data = out.data.reshape(fin_shape)
    return out


def _join_gather(T, string_inp, make_format="standard", intermediate_stat=None, blocks=False):
    """The gather of join_legs: (Perm, source data, fused shape, fused statistics).  With blocks=True
    (a bosonic first group, then two purely fermionic groups: the batched matrices M[b] of a
    decomposition) the gather writes the two parity blocks M[b][p::2, p::2] instead, as a
    (B, 2, m/2, n/2) array (BlockSVD's ME / MO, __init__.py:3990-3996): the parity p is an output axis
    of its own feeding both fused legs."""
    string_inp = denumerate(string_inp.replace(" ", ""))
    inter = tuple(intermediate_stat) if not np.isscalar(intermediate_stat) else (intermediate_stat,)
    groups, srt = group_info(string_inp, T.statistics, T.shape)
    sgn_pos = _grouping_sign_axes(srt, inter)
    info = _intermediate(srt, inter)
    P = Perm()
    view = (isinstance(T, _PermutedView) and T._data is None and T.format == "standard"
            and len(set(sgn_pos) | set(T._fbits)) <= 10)
    std = T.format == "matrix"  # input must first go to standard: sigma on its -1 legs
    if view:
        # read the permuted view's source directly: T axis a is source axis order[a]; the view's
        # sign table joins the grouping sign in this gather's table
        st_src = strides(T._src.shape)
        ins = [P.add_in(T.shape[a], st_src[T._order[a]], enc=_is_pp(T, a)) for a in range(T.ndim)]
        src_data = T._src.data
    else:
        st = strides(T.shape)
        ins = [P.add_in(T.shape[a], st[a], enc=_is_pp(T, a), sigma=(std and T.statistics[a] == -1))
               for a in range(T.ndim)]
        src_data = T.data
    bit_axes = sorted(set(sgn_pos) | (set(T._fbits) if view else set()))
    nb = len(bit_axes)
    bitpos = {p: k for k, p in enumerate(bit_axes)}
    fin_shape, fin_stats = [], []
    if blocks:
        ok = (len(srt) == 3 and info[0][3] == 0 and all(info[g][2] == 0 and info[g][3] and info[g][1] % 2 == 0
                                                          for g in (1, 2)))
        if not ok:
            raise ValueError("_join_gather: blocks=True needs a bosonic group and two fermionic groups")
        fvirt = []
    for g, (chars, stg, dm, pos) in enumerate(srt):
        bd, fd, bn, fn, final = info[g]
        # the fused leg: bosons (high) then fermions; the fermionic part is switched to the
        # parity-preserving encoder over its own dimension, matrix format applies sigma to it
        if bn and fn:
            vb = P.add_virt(bd)
            for t, d, p in zip(stg, dm, pos):
                if t in BOSE:
                    P.add_comp(vb, d, [(ins[p], 1)])
            vf = P.add_virt(fd, enc=True, sigma=(make_format == "matrix" and inter[g] == -1))
            for t, d, p in zip(stg, dm, pos):
                if t in FERMI:
                    P.add_comp(vf, d, [(ins[p], 1)], bit=bitpos.get(p, -1))
            # two raw axes (boson high, fermion low) whose C-order product is the fused leg
            P.add_out(bd, [(vb, 1)])
            P.add_out(fd, [(vf, 1)])
        elif fn:
            vf = P.add_virt(fd, enc=inter[g] in FERMI, sigma=(make_format == "matrix" and inter[g] == -1))
            for t, d, p in zip(stg, dm, pos):
                P.add_comp(vf, d, [(ins[p], 1)], bit=bitpos.get(p, -1))
            if blocks:
                fvirt.append((vf, fd))
                if len(fvirt) == 2:  # (parity, row / 2, column / 2): index = 2 half + parity
                    (vr, mr), (vc, nc) = fvirt
                    P.add_out(2, [(vr, 1), (vc, 1)])
                    P.add_out(mr // 2, [(vr, 2)])
                    P.add_out(nc // 2, [(vc, 2)])
            else:
                P.add_out(fd, [(vf, 1)])
        else:
            # a purely bosonic group is neither encoded nor sigma-scaled (its intermediate leg has
            # statistic 0) but ends up labelled with the requested intermediate statistic
            vb = P.add_virt(bd)
            for t, d, p in zip(stg, dm, pos):
                P.add_comp(vb, d, [(ins[p], 1)])
            P.add_out(bd, [(vb, 1)])
        fin_shape.append(bd * fd)
        fin_stats.append(final)
    if nb:
        table = []
        for m in range(1 << nb):
            code = sum((m >> bitpos[p]) & 1 for p in sgn_pos) & 1
            if view:
                t = T._codes[sum(((m >> bitpos[a]) & 1) << k for k, a in enumerate(T._fbits))]
                code = 2 if t == 2 else code ^ t
            table.append(code)
        P.set_table(table)
    return P, src_data, fin_shape, fin_stats


@_sparse_in
def split_legs(T, string_inp, final_stat, final_shape, intermediate_stat=None, save_memory=False):
    """Inverse of join_legs (__init__.py:3070-3170), including its early return for matrix-format
    input: such input comes back canonical and in matrix format."""
    string_inp = denumerate(string_inp.replace(" ", ""))
    inter = tuple(intermediate_stat) if not np.isscalar(intermediate_stat) else (intermediate_stat,)
    final_stat = tuple(final_stat)
    final_shape = tuple(final_shape)
    groups, srt = group_info(string_inp, final_stat, final_shape)
    sgn_pos = _grouping_sign_axes(srt, inter)
    info = _intermediate(srt, inter)
    fmt, enc = T.format, T.encoder
    matrix = fmt == "matrix"
    out_enc_pp = (not matrix) and enc == "parity-preserving"
    P = Perm()
    st = strides(T.shape)
    # source: fused legs of T; for a hybrid leg the bosonic part is the high factor
    src_of = {}
    for g, (chars, stg, dm, pos) in enumerate(srt):
        bd, fd, bn, fn, final = info[g]
        fused_pp = enc == "parity-preserving"
        if bn and fn:
            ib = P.add_in(bd, st[g] * fd)
            i_f = P.add_in(fd, st[g], enc=fused_pp, sigma=(matrix and inter[g] == -1))
            mb, mf = bd, fd
            for c, t, d, p in zip(chars, stg, dm, pos):
                if t in BOSE:
                    mb //= d
                    src_of[p] = (ib, mb)
                else:
                    mf //= d
                    src_of[p] = (i_f, mf)
        else:
            full = bd * fd
            i_ = P.add_in(full, st[g], enc=fused_pp and fn > 0 and inter[g] in FERMI,
                          sigma=(matrix and fn > 0 and inter[g] == -1))
            m = full
            for c, t, d, p in zip(chars, stg, dm, pos):
                m //= d
                src_of[p] = (i_, m)
    bitpos = {p: k for k, p in enumerate(sorted(sgn_pos))}
    for p, (t, d) in enumerate(zip(final_stat, final_shape)):
        # output leg p: encoder of the result, matrix format re-applied on final -1 legs
        v = P.add_virt(d, enc=out_enc_pp and t in FERMI, sigma=(matrix and t == -1))
        i_, m = src_of[p]
        P.add_comp(v, d, [(i_, m)], bit=bitpos.get(p, -1))
        P.add_out(d, [(v, 1)])
    if sgn_pos:
        P.set_table([bin(m).count("1") & 1 for m in range(1 << len(sgn_pos))])
    data = P.run(T.data)
    res = T._like(data, statistics=final_stat, encoder=("canonical" if matrix else enc), format=fmt, even=T._even)
    return res


# --------------------------------------------------------------------------------------------------
# svd / eig (__init__.py:3931-4698)
# --------------------------------------------------------------------------------------------------
def _block_prepare(Mdata, cutoff, rule="dense"):
    """Parity blocks ME, MO of a parity-preserving (-1,+1) matrix and the number of values to compute
    in each: int(cutoff/2) per block (BlockSVD / BlockEig, __init__.py:3953-4029, 4347-4422), or
    ceil(cutoff/2) even and floor(cutoff/2) odd with rule="block" (decompose_block,
    __init__.py:5076-5085)."""
    m, n = Mdata.shape
    if cutoff is None:
        caps = (None, None)
    elif rule == "block":
        caps = (int(math.ceil(cutoff / 2)), int(math.floor(cutoff / 2)))
    else:
        caps = (int(cutoff / 2), int(cutoff / 2))
    specs = []
    for par, cap in enumerate(caps):
        k = min(m // 2, n // 2)
        if cap is not None:
            k = min(k, cap)
        specs.append((Mdata, par, k))
    return specs


def _block_assemble(outs, kk, m, n, dt, kept=None):
    """Full interleaved (U, S, Vt) from the two blocks' results, both padded to the next power of two
    of the larger kept count kk (the rank rule of SortedSVD: the nonzero values read back);
    `kept` (a list) receives that count."""
    d = max(kk)
    if kept is not None:
        kept.append(d)
    d = int(2 ** math.ceil(math.log2(d))) if d > 0 else 1
    mh, nh = m // 2, n // 2
    # stack the kept columns / rows of both blocks, zero-padded to d, then interleave by parity
    SU = DeviceArray.zeros((2, mh, d), dt)
    SV = DeviceArray.zeros((2, d, nh), dt)
    sv = DeviceArray.zeros((2 * d,), dt)
    for par, (U_b, S_b, V_b, k) in enumerate(outs):
        kp = kk[par]
        if kp == 0:
            continue
        copy2d(SU, d, U_b, k, mh, kp, dst_off=par * mh * d)
        copy2d(SV, nh, V_b, n // 2, kp, nh, dst_off=par * d * nh)
        copy2d(sv, 2, S_b, 1, kp, 1, dst_off=par)  # sv[2j + par] = s_par[j]
    U = _interleave(SU, m, 2 * d, row_stride=d, par_stride=mh * d)
    Vt = _interleave(SV, 2 * d, n, row_stride=nh, par_stride=d * nh)
    S = DeviceArray.zeros((2 * d, 2 * d), dt)
    copy2d(S, 2 * d + 1, sv, 1, 2 * d, 1)  # the diagonal
    return U, S, Vt


def _interleave(stack, rows, cols, row_stride, par_stride):
    """full[r][c] = stack[r % 2][r // 2][c // 2] when r and c have equal parity, else 0 (both dims
    even): the interleaving of get_full_matrix (__init__.py:4017-4028)."""
    P = Perm()
    ir = P.add_in(rows, row_stride, enc=True, split=True, stride_lo=par_stride)
    ic = P.add_in(cols, 1, enc=True, split=True, stride_lo=0)
    vr = P.add_virt(rows, enc=True)
    P.add_comp(vr, rows, [(ir, 1)], bit=0)
    vc = P.add_virt(cols, enc=True)
    P.add_comp(vc, cols, [(ic, 1)], bit=1)
    P.add_out(rows, [(vr, 1)])
    P.add_out(cols, [(vc, 1)])
    P.set_table([0, 2, 2, 0])  # zero unless both indices have the same parity
    return P.run(stack)


def _parity_leak_dev(M):
    """Squared Frobenius norm of the parity-odd entries M[i, j], (i + j) odd, of a parity-preserving
    matrix, left on the device (read back with the decomposition's results, so that the check costs no
    extra synchronisation).  The reference sums |M_ij| (__init__.py:3977-3984); both are zero exactly
    for an even tensor."""
    lib = L.lib()
    dims = (C.c_int64 * 2)(*M.shape)
    one = (C.c_int32 * 2)(1, 1)
    out = DeviceArray.empty(1, np.float64)
    L.check(lib.gtn_odd_norm(L.ptr(M), 2, dims, one, one, L.dtype_code(M), L.ptr(out), L.stream_ptr()), "gtn_odd_norm")
    return out


def _check_parity_leak(out, m, n):
    if math.sqrt(float(out.item())) / (m * n / 2) > 1.0e-14:
        fail("Error[BlockSVD]: This matrix is not constructed from a Grassmann-even tensor.")


def _strided_block(M, par, out=None):
    """Contiguous copy of M[par::2, par::2] (the ME / MO extraction of __init__.py:3990-3996), into
    `out` when given."""
    m, n = M.shape
    P = Perm()
    i0 = P.add_in(max(m - 1, 1), n)
    i1 = P.add_in(max(n - 1, 1), 1)
    v0 = P.add_virt(m // 2)
    P.add_comp(v0, m // 2, [(i0, 2)])
    v1 = P.add_virt(n // 2)
    P.add_comp(v1, n // 2, [(i1, 2)])
    P.add_out(m // 2, [(v0, 1)])
    P.add_out(n // 2, [(v1, 1)])
    off = par * n + par
    return P.run(M.flat(off, M.numel() - off), dst=out)


def _runs(keys):
    """Split a list into maximal runs of equal consecutive keys: [(start, count), ...]."""
    out = []
    for i, k in enumerate(keys):
        if out and keys[out[-1][0]] == k:
            out[-1][1] += 1
        else:
            out.append([i, 1])
    return out


def _materialize(specs):
    """specs: [(matrix, parity or None, k)] -> [(contiguous matrix, k)].  Parity blocks of equal shape
    that follow each other are gathered into one buffer, so that the device decomposition sees them at
    a constant stride and batches them."""
    keys = [(None if par is None else (M.shape[0] // 2, M.shape[1] // 2), i if par is None else -1)
            for i, (M, par, k) in enumerate(specs)]
    res = []
    for st, cnt in _runs(keys):
        M0, par0, _ = specs[st]
        if par0 is None:
            res.append((M0, specs[st][2]))
            continue
        mh, nh = M0.shape[0] // 2, M0.shape[1] // 2
        buf = DeviceArray.empty((cnt, mh, nh), M0.dtype)
        for j in range(cnt):
            M, par, k = specs[st + j]
            view = buf.flat(j * mh * nh, mh * nh, (mh, nh))
            _strided_block(M, par, out=view)
            res.append((view, k))
    return res


def _svd_blocks_raw(blocks, mode, readback=True):
    """blocks: list of (contiguous FP64 matrix, k).  Returns (per block (U (m x k), S (k), Vt (k x n), k),
    per block the number of nonzero values in S).  Consecutive blocks of equal shape get their outputs
    in shared buffers at constant strides.  With readback=False nothing is read back: the counts are
    taken to be k and the third value returned, (groups, S buffers, keys), lets _Pending check them
    later."""
    lib = L.lib()
    nblk = len(blocks)
    arr = (L.SvdBlock * nblk)()
    outs = [None] * nblk
    counts = [0] * nblk
    keys = []
    for Mb, k in blocks:
        mb, nb_ = Mb.shape
        keys.append((mb, nb_, k if (mb and nb_) else 0))
    groups = _runs(keys)
    sbufs = []
    for st, cnt in groups:
        mb, nb_, kk = keys[st]
        k1 = max(kk, 1)
        U = DeviceArray.zeros((cnt, mb, k1), np.float64)
        S = DeviceArray.zeros((cnt, k1), np.float64)
        V = DeviceArray.zeros((cnt, k1, nb_), np.float64)
        sbufs.append(S)
        for j in range(cnt):
            i = st + j
            Mb = blocks[i][0]
            u = U.flat(j * mb * k1, mb * k1, (mb, k1))
            sv = S.flat(j * k1, k1, (k1,))
            v = V.flat(j * k1 * nb_, k1 * nb_, (k1, nb_))
            outs[i] = (u, sv, v, kk)
            arr[i].M = Mb.data_ptr() if Mb.numel() else None
            arr[i].U, arr[i].S, arr[i].Vt = u.data_ptr(), sv.data_ptr(), v.data_ptr()
            arr[i].m, arr[i].n, arr[i].lda, arr[i].ldu, arr[i].ldv = mb, nb_, max(nb_, 1), k1, max(nb_, 1)
            arr[i].k = kk
    opts = L.SvdOpts()
    opts.oversample = settings.svd_oversample
    opts.iters = svd_iters_for(max([kk for (_, _, kk) in keys] + [1]))
    opts.precision, opts.f64_tail = svd_spec(settings.svd_precision)
    opts.mode = mode
    opts.seed = settings.seed
    opts.orth_every = settings.svd_orth_every
    ws_bytes = lib.gtn_svd_workspace_bytes(arr, nblk, C.byref(opts))
    ws = DeviceArray.empty(max(ws_bytes // 8, 1), np.float64)
    L.check(lib.gtn_svd_blocks(arr, nblk, C.byref(opts), L.ptr(ws), ws_bytes, L.stream_ptr()), "gtn_svd_blocks")
    if not readback:
        return outs, [keys[i][2] for i in range(nblk)], (groups, sbufs, keys)
    return outs, _rank_counts(groups, sbufs, keys)


def _rank_counts(groups, sbufs, keys):
    """The rank rule's count of nonzero values per block, one read-back per group of equal blocks."""
    counts = [0] * len(keys)
    for (st, cnt), S in zip(groups, sbufs):
        kk = keys[st][2]
        h = S.numpy()
        if not np.all(np.isfinite(h)):
            raise L.GrassmannDeviceError("gtn_svd_blocks returned non-finite singular values (input holds "
                                         "inf/nan or the decomposition broke down)")
        for j in range(cnt):
            counts[st + j] = int(np.count_nonzero(h[j][:kk])) if kk else 0
    return counts


class _Pending:
    """Checks of a decomposition run without read-back (svd_many(..., speculate=True)): the rank
    counts assumed equal to the requested k, the finiteness of the values and the parity leaks."""

    def __init__(self, raw, assumed, leaks):
        self.raw, self.assumed, self.leaks = raw, assumed, leaks

    def verify(self):
        """True when every count was the assumed one; raises what the synchronous path raises."""
        counts = _rank_counts(*self.raw)
        self.counts = counts
        for leak in self.leaks:
            _check_parity_leak(*leak)
        return counts == self.assumed


class _Job:
    """One svd / eig call in flight: the joined matrix, its blocks and what the split needs."""
    __slots__ = ("T", "left", "right", "cutoff", "M", "blocks", "blocked", "kept", "done", "leak")


def _decompose_prepare(T, string, cutoff, mode, rule):
    name = "svd" if mode == 0 else "eig"
    if not isinstance(T, dense):
        fail("Error[%s]: Object type must only be dense or sparse!" % name)
    Td = T._src.data if isinstance(T, _PermutedView) and T._data is None else T.data
    if Td.is_complex() or Td.dtype != np.float64:
        raise NotImplementedError("%s of non-FP64-real tensors is not provided by this build" % name)
    for st in T.statistics:
        if st not in ALLOWED_STAT:
            fail("Error[%s]: The input object contains illegal statistics. (0, 1, -1, or * only)" % name)
    job = _Job()
    job.T, job.cutoff, job.kept, job.done, job.leak = T, cutoff, [], None, None
    left, right = parse_partition(string, name)
    job.left, job.right = left, right
    fmt, enc = T.format, T.encoder
    xstats, xshape = T.statistics, T.shape
    nl = len(left)
    sl, sr = xstats[:nl], xstats[nl:]
    shl, shr = xshape[:nl], xshape[nl:]
    lb1 = list(sl) in ([-1], [1]) and list(shl) == [1]
    rb1 = list(sr) in ([-1], [1]) and list(shr) == [1]
    if lb1 or rb1:
        # a single fermionic leg of dimension one (__init__.py:4132-4170)
        nrm = T.norm
        M = T.force_format("matrix")
        if lb1:
            U = dense([[1.0]], statistics=(*sl, 1), format="matrix")
            S = dense([[nrm]], statistics=(-1, 1), format="matrix")
            V = M / nrm
            V.statistics = (-1,) + tuple(V.statistics[1:])
        else:
            U = M / nrm
            S = dense([[nrm]], statistics=(-1, 1), format="matrix")
            V = dense([[1.0]], statistics=(-1, *sr), format="matrix")
            U.statistics = tuple(U.statistics[:-1]) + (1,)
        out = [U, S, V]
        if fmt == "standard":
            out = [x.switch_format() for x in out]
        if enc == "parity-preserving":
            out = [x.switch_encoder() for x in out]
        job.done = tuple(out)
        job.blocks = []
        return job
    inter = (0 if all(s == 0 for s in sl) else -1, 0 if all(s == 0 for s in sr) else 1)
    M = join_legs(T, "(" + left + ")(" + right + ")", "matrix", intermediate_stat=inter)
    M = M.force_encoder("parity-preserving")
    job.M = M
    Md = M.data
    if M.statistics[0] == 0 or M.statistics[1] == 0:
        mm, nn = Md.shape
        k = min(mm, nn) if cutoff is None else min(mm, nn, cutoff)
        job.blocked = False
        job.blocks = [(Md, None, k)]
    else:
        job.blocked = True
        job.blocks = _block_prepare(Md, cutoff, rule)
        if not _pkg_flag("skip_parity_blocking_check", False):
            job.leak = (_parity_leak_dev(Md), Md.shape[0], Md.shape[1])
    return job


def _decompose_finish(job, outs, counts):
    T, M = job.T, job.M
    Md = M.data
    fmt, enc = T.format, T.encoder
    xstats, xshape = T.statistics, T.shape
    left, right = job.left, job.right
    nl = len(left)
    if not job.blocked:
        mm, nn = Md.shape
        U_b, S_b, V_b, k = outs[0]
        nz = counts[0]
        U = DeviceArray.zeros((mm, max(nz, 1)), np.float64)
        Vt = DeviceArray.zeros((max(nz, 1), nn), np.float64)
        S = DeviceArray.zeros((max(nz, 1), max(nz, 1)), np.float64)
        if nz:
            copy2d(U, nz, U_b, k, mm, nz)
            copy2d(Vt, nn, V_b, nn, nz, nn)
            copy2d(S, nz + 1, S_b, 1, nz, 1)
    else:
        U, S, Vt = _block_assemble(outs, counts, Md.shape[0], Md.shape[1], Md.dtype, job.kept)
    ls, rs = -1, 1
    s0, s1 = M.statistics

    def mk(data, stats):
        r = dense.__new__(dense)
        r.data, r.statistics, r.encoder, r.format, r._even = data, tuple(stats), "parity-preserving", "matrix", True
        return r

    if s0 == 0:
        U, S, V = mk(U, (0, 0)), mk(S, (0, 0)), mk(Vt, (0, s1))
        ls = rs = 0
    elif s1 == 0:
        U, S, V = mk(U, (s0, 0)), mk(S, (0, 0)), mk(Vt, (0, 0))
        ls = rs = 0
    else:
        U, S, V = mk(U, (s0, 1)), mk(S, (-1, 1)), mk(Vt, (-1, s1))
    dl = S.shape[0]
    ustats = tuple(xstats[:nl]) + (-ls,)
    vstats = (-rs,) + tuple(xstats[nl:])
    ushape = tuple(xshape[:nl]) + (dl,)
    vshape = (dl,) + tuple(xshape[nl:])
    nc = fresh_char(left + right)
    U = split_legs(U, "(" + left + ")" + nc, ustats, ushape, intermediate_stat=(-1, 1))
    S = S.switch_encoder()
    V = split_legs(V, nc + "(" + right + ")", vstats, vshape, intermediate_stat=(-1, 1))
    out = [U, S, V]
    if fmt == "standard":
        out = [x.switch_format() for x in out]
    if enc == "parity-preserving":
        out = [x.switch_encoder() for x in out]
    for x in out:
        x._even = True
    return tuple(out)


def _decompose_many(calls, mode, rule="dense", speculate=False):
    """Several svd / eig calls [(T, string, cutoff), ...] with all their parity blocks decomposed in
    one gtn_svd_blocks call.  Returns [(U, S, V, kept), ...] (kept: the larger per-parity count).
    speculate=True assumes every block keeps its requested count instead of reading the counts back
    (no synchronisation) and returns (results, _Pending) for the caller to verify."""
    jobs = [_decompose_prepare(T, string, cutoff, mode, rule) for T, string, cutoff in calls]
    specs = [b for j in jobs for b in j.blocks]
    pending = None
    if speculate and specs:
        outs, counts, raw = _svd_blocks_raw(_materialize(specs), mode, readback=False)
        pending = _Pending(raw, list(counts), [j.leak for j in jobs if j.leak is not None])
    else:
        outs, counts = _svd_blocks_raw(_materialize(specs), mode) if specs else ([], [])
        for j in jobs:
            if j.leak is not None:
                _check_parity_leak(*j.leak)
    res, at = [], 0
    for j in jobs:
        nb = len(j.blocks)
        r = j.done if j.done is not None else _decompose_finish(j, outs[at:at + nb], counts[at:at + nb])
        at += nb
        res.append(tuple(r) + (j.kept[0] if j.kept else None,))
    return (res, pending) if speculate else res


def _decompose(T, string, cutoff, mode, rule="dense", kept=None):
    U, S, V, k = _decompose_many([(T, string, cutoff)], mode, rule)[0]
    if kept is not None and k is not None:
        kept.append(k)
    return U, S, V


def svd_many(calls, mode=0, speculate=False):
    """[(T, string, cutoff), ...] -> [(U, S, V), ...]: several dense svd calls (eig with mode=1) whose
    parity blocks share one batched device decomposition.  speculate=True returns (results, pending):
    the results assume every parity block keeps its requested number of values, and pending.verify()
    (which synchronises) says whether that held; when it did not the results must be recomputed."""
    if speculate:
        res, pending = _decompose_many(calls, mode, speculate=True)
        return [r[:3] for r in res], pending
    return [r[:3] for r in _decompose_many(calls, mode)]


@_sparse_in
def svd(InpObj, string, cutoff=None, save_memory=False):
    """(__init__.py:4033-4308) U, S, V with T = einsum(U S V); parity blocks kept int(cutoff/2) each."""
    return _decompose(InpObj, string, cutoff, 0)


@_sparse_in
def eig(InpObj, string, cutoff=None, debug_mode=False, save_memory=False):
    """(__init__.py:4425-4698) U, Lambda, U^+ of a Hermitian tensor; eigenvalues ordered by |lambda|."""
    M = InpObj
    return _decompose(M, string, cutoff, 1)


# --------------------------------------------------------------------------------------------------
# hconjugate (__init__.py:5300-5493)
# --------------------------------------------------------------------------------------------------
@_sparse_in
def hconjugate(InpObj, string, save_memory=False):
    from .block import block, hconjugate_block
    if isinstance(InpObj, block):
        return hconjugate_block(InpObj, string)
    T = InpObj
    if T.data.is_complex():
        raise NotImplementedError("complex hconjugate is not provided by this build")
    left, right = parse_partition(string, "hconjugate")
    fmt, enc = T.format, T.encoder
    xstats, xshape = T.statistics, T.shape
    nl = len(left)
    M = join_legs(T, "(" + left + ")(" + right + ")", "matrix", intermediate_stat=(-1, 1))
    M = M.force_encoder("parity-preserving")
    # conjugate transpose (real data: transpose) as a gather
    Mt = _relayout(M, [1, 0], out_encoder=M.encoder)
    s0, s1 = M.statistics
    M = M._like(Mt, statistics=((-s1 if s1 in FERMI else s1), (-s0 if s0 in FERMI else s0)))
    new_stats = list(xstats[nl:]) + list(xstats[:nl])
    new_stats = tuple(-s if s in FERMI else s for s in new_stats)
    new_shape = tuple(xshape[nl:]) + tuple(xshape[:nl])
    O = split_legs(M, "(" + right + ")(" + left + ")", new_stats, new_shape, intermediate_stat=(-1, 1))
    if fmt != O.format:
        O = O.switch_format()
    if enc != O.encoder:
        O = O.switch_encoder()
    O._even = T._even
    return O


# --------------------------------------------------------------------------------------------------
# power / sqrt (__init__.py:6050-6069)
# --------------------------------------------------------------------------------------------------
@_sparse_in
def power(T, p, rcond=1e-10):
    from .block import block, power_block
    if isinstance(T, block):
        return power_block(T, p)
    fmt, enc = T.format, T.encoder
    M = T.force_format("matrix").force_encoder("canonical")
    lib = L.lib()
    L.check(lib.gtn_power_diag(L.ptr(M.data), M.data.numel(), 1, L.dtype_code(M.data), float(p), float(rcond),
                               L.stream_ptr()), "gtn_power_diag")
    return M.force_format(fmt).force_encoder(enc)


def sqrt(T, rcond=1e-10):
    return power(T, 0.5)


def _elementwise(T, op, rcond):
    """x -> (|x| > rcond) ? f(x) : 0 on the matrix-format canonical coefficients (the common shape of
    power_ds / inv_nz_ds / exp_ds, __init__.py:5992-6106); blocks act on their storage."""
    from .block import block
    if isinstance(T, block):
        fmt = T.format
        B = T.force_format("matrix")
        t = B._t.copy()
        L.check(L.lib().gtn_elementwise(L.ptr(t.data), t.data.numel(), L.dtype_code(t.data), op, 0.0, float(rcond),
                                        L.stream_ptr()), "gtn_elementwise")
        return block._wrap(t, B.even_shape, B.odd_shape, B.shape, B.marked_as_joined).force_format(fmt)
    fmt, enc = T.format, T.encoder
    M = T.force_format("matrix").force_encoder("canonical")
    if M is T:
        M = M.copy()
    L.check(L.lib().gtn_elementwise(L.ptr(M.data), M.data.numel(), L.dtype_code(M.data), op, 0.0, float(rcond),
                                    L.stream_ptr()), "gtn_elementwise")
    M._even = None
    return M.force_format(fmt).force_encoder(enc)


@_sparse_in
def inv_nz(T, rcond=1e-10):
    """Reciprocal of every coefficient above rcond, zero elsewhere (inv_nz, __init__.py:5987-6019)."""
    return _elementwise(T, 1, rcond)


@_sparse_in
def exp(T, rcond=1e-10):
    """exp of every coefficient above rcond, zero elsewhere (exp_ds / exp_block, __init__.py:6086-6120);
    Grassmann numbers (arith) are not provided by this build."""
    return _elementwise(T, 2, rcond)


@_sparse_in
def pinv(InpObj, string, rcond=1e-10):
    """Pseudo-inverse through the SVD (pinv, __init__.py:5961-5983): V^+ inv_nz(S^+) U^+ with the
    legs of the right group first."""
    string = denumerate(string.replace(" ", ""))
    stra = strb = None
    for sep in SEPARATORS:
        if sep in string:
            stra, strb = string.split(sep)
            break
    if stra is None:
        fail("Error[pinv]: the leg grouping needs a separator such as '|'.")
    strba = strb + stra
    strx = fresh_char(strba)
    stry = fresh_char(strba + strx)
    U, S, V = InpObj.svd(string)
    Uinv = U.hconjugate(stra + "|" + strx)
    Sinv = inv_nz(S.hconjugate(strx + "|" + stry), rcond=rcond)
    Vinv = V.hconjugate(strx + "|" + strb)
    ret = einsum(strb + strx + "," + strx + stry + "->" + strb + stry, Vinv, Sinv)
    return einsum(strb + stry + "," + stry + stra + "->" + strba, ret, Uinv)
