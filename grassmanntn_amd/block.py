"""Block Grassmann tensors (the reference's ``block`` format, __init__.py:242-809) on the device.

The reference keeps one array ("cell") per assignment of parities to the fermionic legs: a
parity-preserving index 2 i + p of a leg sits in that leg's parity-p block at position i, and every
leg carries sign vectors sgn[p][i] = sigma(encoder(2 i + p)) (__init__.py:288-337).  A leg's even and
odd extents may differ from half its dimension after truncation.

Here the cells are never separated: a block tensor is one dense parity-preserving tensor in HBM whose
fermionic legs have a power-of-two storage extent 2 * 2^ceil(log2 max(even, odd)), with the entries
beyond a leg's even / odd extent held at zero.  Every block operation is then the dense device
operation on that storage (the per-parity loop of einsum_block is exactly the parity-sector
contraction the dense path runs), plus bookkeeping of the even / odd extents and of the "physical"
shape the reference reports.  Sign vectors and cells are produced on demand for inspection.
"""
import math

import numpy as np

from . import core
from ._device import DeviceArray, copy2d
from ._plan import denumerate, fail, parse_partition
from .core import dense, sparse, FERMI, ALLOWED_STAT


def _pow2(n):
    return 1 if n <= 1 else int(2 ** math.ceil(math.log2(n)))


def _store_dim(stat, d, e, o):
    return d if stat not in FERMI else 2 * _pow2(max(e, o, 1))


def _resize(T, new_shape):
    """Dense tensor of new_shape holding T's entries in the common corner, zero elsewhere (one 2-D copy
    per changed axis, on the device)."""
    data = T.data
    shape = list(data.shape)
    for a, m in enumerate(new_shape):
        n = shape[a]
        if n == m:
            continue
        outer = int(np.prod(shape[:a])) if a else 1
        inner = int(np.prod(shape[a + 1:])) if a + 1 < len(shape) else 1
        nshape = list(shape)
        nshape[a] = m
        out = DeviceArray.zeros(tuple(nshape), data.dtype)
        copy2d(out, m * inner, data, n * inner, outer, min(n, m) * inner)
        data, shape = out, nshape
    return T._like(data)


def _sign_vector(p, n):
    from .param import sgn, encoder
    return np.array([sgn(encoder(2 * i + p)) for i in range(n)], dtype=np.int64)


class block:
    """Z2-graded Grassmann tensor with per-leg even / odd extents (__init__.py:242-809)."""

    __array_priority__ = 1000000

    def __init__(self, data=None, is_zero=False):
        if isinstance(data, block):
            o = data.copy()
            self._t, self._even_shape, self._odd_shape = o._t, o._even_shape, o._odd_shape
            self.shape, self.marked_as_joined = o.shape, o.marked_as_joined
            return
        if isinstance(data, sparse):
            data = dense(data)
        if not isinstance(data, dense):
            fail("Error[block]: a block tensor is built from a dense tensor (dense(...).toblock()).")
        T = data.force_encoder("parity-preserving")
        stats = T.statistics
        es = tuple(max(1, d // 2) if s in FERMI else d for d, s in zip(T.shape, stats))
        os_ = tuple((d // 2 if d > 1 else 0) if s in FERMI else d for d, s in zip(T.shape, stats))
        store = tuple(_store_dim(s, d, e, o) for s, d, e, o in zip(stats, T.shape, es, os_))
        if store != T.shape:
            T = _resize(T, store)
        if is_zero:
            T.data.fill_(0.0)
            T._even = True
        self._t = T
        self._even_shape, self._odd_shape = es, os_
        self.shape = tuple(data.shape)
        self.marked_as_joined = False

    @staticmethod
    def _wrap(t, even_shape, odd_shape, shape, joined=False):
        b = block.__new__(block)
        b._t = t
        b._even_shape, b._odd_shape = tuple(even_shape), tuple(odd_shape)
        b.shape = tuple(shape)
        b.marked_as_joined = joined
        return b

    # ---- properties (__init__.py:431-475) -------------------------------------------------------
    statistics = property(lambda self: self._t.statistics)
    format = property(lambda self: self._t.format)
    even_shape = property(lambda self: self._even_shape)
    odd_shape = property(lambda self: self._odd_shape)
    ndim = property(lambda self: len(self.shape))
    encoder = "block"

    @property
    def effective_shape(self):
        return tuple(d if s not in FERMI else e + o
                     for d, s, e, o in zip(self.shape, self.statistics, self._even_shape, self._odd_shape))

    @property
    def dtype(self):
        return self._t.data.dtype

    @property
    def norm(self):
        return self._t.norm

    @property
    def sgn(self):
        """[even sign vectors per leg, odd sign vectors per leg] (host arrays)."""
        out = [[], []]
        for a, s in enumerate(self.statistics):
            for p, n in ((0, self._even_shape[a]), (1, self._odd_shape[a])):
                out[p].append(np.ones(n, dtype=np.int64) if s not in FERMI else _sign_vector(p, n))
        return out

    @property
    def data(self):
        """The cells as a host object array indexed by the fermionic legs' parities."""
        st = self._t.data.numpy()
        fax = [a for a, s in enumerate(self.statistics) if s in FERMI]
        cells = np.empty((2,) * len(fax), dtype=object)
        for par in np.ndindex(*cells.shape):
            pa = dict(zip(fax, par))
            sl = tuple(slice(pa[a], 2 * (self._even_shape[a] if pa[a] == 0 else self._odd_shape[a]), 2)
                       if a in pa else slice(None) for a in range(self.ndim))
            cells[par] = st[sl].copy()
        return cells

    def copy(self):
        return block._wrap(self._t.copy(), self._even_shape, self._odd_shape, self.shape, self.marked_as_joined)

    def _with(self, t):
        return block._wrap(t, self._even_shape, self._odd_shape, self.shape, self.marked_as_joined)

    def info(self, name=None, indent_size=0):
        ind = " " * indent_size
        print()
        if name is not None:
            print(ind + "            name:", name)
        print(ind + "      array type: block")
        print(ind + "     total shape:", self.shape)
        print(ind + " effective shape:", self.effective_shape)
        print(ind + "      even shape:", self.even_shape)
        print(ind + "       odd shape:", self.odd_shape)
        print(ind + "      statistics:", self.statistics)
        print(ind + "          format:", self.format)
        print(ind + "         encoder: block")
        print(ind + "            norm:", self.norm)
        print()

    display = info

    # ---- arithmetic (__init__.py:489-549) -------------------------------------------------------
    def __add__(self, other):
        if (type(self) != type(other) or self.even_shape != other.even_shape or self.odd_shape != other.odd_shape
                or self.shape != other.shape or self.statistics != other.statistics or self.format != other.format):
            fail("Error[block.+]: Inconsistent object properties")
        if self.marked_as_joined or other.marked_as_joined:
            fail("Error[block.+]: You cannot add a joined object to ther object.")
        return self._with(self._t + other._t)

    def __mul__(self, other):
        if not np.isscalar(other):
            fail("Error[block.*]: Only scalar multiplication is allowed.")
        return self._with(self._t * other)

    def __truediv__(self, other):
        if not np.isscalar(other):
            fail("Error[block./]: Only scalar division is allowed.")
        return self._with(self._t * (1.0 / other))

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

    # ---- switches (__init__.py:551-603): sigma / parity of the stored index ---------------------
    def switch_format(self):
        return self._with(self._t.switch_format())

    def force_format(self, format):
        return self.copy() if self.format == format else self.switch_format()

    def switch_parity(self):
        return self._with(self._t.switch_parity())

    # ---- conversions (__init__.py:603-729) ------------------------------------------------------
    def todense(self, encoder="canonical", skip_joined_check=False):
        if not skip_joined_check and self.marked_as_joined:
            fail("Error[todense]: Split the legs first!")
        for a, s in enumerate(self.statistics):
            e, o = self._even_shape[a], self._odd_shape[a]
            top = max(2 * e - 1 if e else 0, 2 * o if o else 0)  # one past the last live index
            if s in FERMI and top > self.shape[a]:
                fail("Error[todense]: a block extent does not fit the tensor's shape.")
        t = self._t if self._t.shape == self.shape else _resize(self._t, self.shape)
        return t.force_encoder(encoder)

    def tosparse(self, encoder="canonical"):
        return sparse(self.todense(encoder))

    def is_even(self):
        return self._t.is_even()

    def join_legs(self, string_inp, final_stat):
        raise NotImplementedError("block join_legs / split_legs are internal steps of svd / eig / hconjugate "
                                  "(__init__.py:3376-3864) and are not exposed by this build")

    def split_legs(self, *a, **k):
        raise NotImplementedError("block join_legs / split_legs are internal steps of svd / eig / hconjugate "
                                  "(__init__.py:3376-3864) and are not exposed by this build")

    def hconjugate(self, string, save_memory=False):
        return hconjugate_block(self, string)

    def svd(self, string, cutoff=None, save_memory=False):
        return svd_block(self, string, cutoff)

    def eig(self, string, cutoff=None, save_memory=False):
        return eig_block(self, string, cutoff)


def zero_block(effective_shape, statistics, format="standard", dtype=float):
    """(__init__.py:632-644): extents round(d/2) per parity, physical shape the next power of two."""
    shape = [_pow2(d) if s in FERMI else d for d, s in zip(effective_shape, statistics)]
    oe = [max(1, int(round(d / 2))) if s in FERMI else d for d, s in zip(effective_shape, statistics)]
    return zero_block_eo(oe, oe, statistics, format, dtype, shape)


def zero_block_eo(even_shape, odd_shape, statistics, format="standard", dtype=float, shape=None):
    """(__init__.py:646-672)"""
    eff = [e + o if s in FERMI else e for e, o, s in zip(even_shape, odd_shape, statistics)]
    if shape is None:
        shape = [_pow2(d) if s in FERMI else d for d, s in zip(eff, statistics)]
    store = tuple(_store_dim(s, d, e, o) for s, d, e, o in zip(statistics, shape, even_shape, odd_shape))
    dt = np.complex128 if dtype in (complex, np.complex128) else np.float64
    t = dense.__new__(dense)
    t.data = DeviceArray.zeros(store, dt)
    t.statistics, t.encoder, t.format, t._even = tuple(statistics), "parity-preserving", format, True
    return block._wrap(t, even_shape, odd_shape, shape)


def _reorder(src, dst, xs):
    return [xs[src.index(c)] for c in dst]


def einsum_block(*args, ignore_anticommutation=False):
    """(__init__.py:2346-2941) The per-parity-assignment loop of the reference is the parity-sector
    contraction of the dense path on the block storage; the extents of the output legs are those of the
    input legs they come from, and the physical shape is the next power of two of the effective one."""
    subscripts = denumerate(args[0].replace(" ", ""))
    if subscripts.count("->") > 1:
        fail("Error[einsum_block]: There should be no more than one '->' indicator in the subscript input.")
    has_output = "->" in subscripts
    before, after = subscripts.split("->") if has_output else (subscripts, "")
    strings = before.split(",")
    ops = list(args[1:1 + len(strings)])
    for o in ops:
        if not isinstance(o, block):
            fail("Error[einsum_block]: This function only works with block data format.")
    summand = before.replace(",", "")
    stats = sum([list(o.statistics) for o in ops], [])
    eff = sum([list(o.effective_shape) for o in ops], [])
    esh = sum([list(o.even_shape) for o in ops], [])
    osh = sum([list(o.odd_shape) for o in ops], [])
    for c in set(summand):
        occ = [i for i, x in enumerate(summand) if x == c]
        if stats[occ[0]] in FERMI and any((esh[i], osh[i]) != (esh[occ[0]], osh[occ[0]]) for i in occ[1:]):
            fail("Error[einsum_block]: some sign factors are not consistent.")
    res = core.einsum(subscripts, *[o._t for o in ops], ignore_anticommutation=ignore_anticommutation)
    if not has_output:
        return res
    out_stats = _reorder(summand, after, stats)
    out_eff = _reorder(summand, after, eff)
    shape = tuple(d if s not in FERMI else _pow2(d) for d, s in zip(out_eff, out_stats))
    return block._wrap(res, _reorder(summand, after, esh), _reorder(summand, after, osh), shape)


def _decompose_block(B, string, cutoff, mode):
    """decompose_block (__init__.py:4704-5289): ceil(cutoff/2) even and floor(cutoff/2) odd values, both
    parities padded to the larger kept count n, new bond of physical dimension 2^(1+ceil(log2 n))."""
    name = "svd_block" if mode == 0 else "eig_block"
    for s in B.statistics:
        if s not in ALLOWED_STAT:
            fail("Error[%s]: The input object contains illegal statistics. (0, 1, -1, or * only)" % name)
    left, right = parse_partition(string, name)
    nl = len(left)
    if all(s not in FERMI for s in B.statistics[:nl]) or all(s not in FERMI for s in B.statistics[nl:]):
        fail("Error[SVD/EigV_block]: This function do not support the case where one of the decomposed "
             "branches is totally bosonic.\nPlease do it in the dense/sparse format instead.")
    if not B._t.is_even():
        fail("Error[SVD/EigV_block]: This function only works for even tensors!")
    kept = []
    U, S, V = core._decompose(B._t, string, cutoff, mode, rule="block", kept=kept)
    n = kept[0]
    nfull = 2 ** int(1 + math.ceil(math.log2(max(n, 1))))
    es, os_ = B.even_shape, B.odd_shape
    Ub = block._wrap(U, es[:nl] + (n,), os_[:nl] + (n,), B.shape[:nl] + (nfull,))
    Sb = block._wrap(S, (n, n), (n, n), (nfull, nfull))
    Vb = block._wrap(V, (n,) + es[nl:], (n,) + os_[nl:], (nfull,) + B.shape[nl:])
    return Ub, Sb, Vb


def svd_block(InpObj, string, cutoff=None, save_memory=False):
    return _decompose_block(InpObj, string, cutoff, 0)


def eig_block(InpObj, string, cutoff=None, save_memory=False):
    return _decompose_block(InpObj, string, cutoff, 1)


def hconjugate_block(InpObj, string, save_memory=False):
    """(__init__.py:5495-5955) the dense conjugate transpose of the storage; extents follow the legs."""
    left, right = parse_partition(string, "hconjugate_block")
    nl = len(left)
    H = core.hconjugate(InpObj._t, string)
    perm = list(range(nl, InpObj.ndim)) + list(range(nl))
    return block._wrap(H, [InpObj.even_shape[a] for a in perm], [InpObj.odd_shape[a] for a in perm],
                       [InpObj.shape[a] for a in perm])


def power_block(T, p, rcond=1e-10):
    """Elementwise power of the matrix-format coefficients where |x| > rcond (__init__.py:6071-6082;
    the result is kept, as the documented block TRG numbers require)."""
    return T._with(core.power(T._t, p, rcond))
