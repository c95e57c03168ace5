"""Builder for gtn_permute descriptors (the signed-gather kernel of include/gtn_hip.h).

A gather is described in terms of *components*: canonical indices of single legs.  Output raw
axes feed virtual axes (fused canonical or parity-preserving indices, possibly assembled from
several raw axes such as (parity, half-index)); virtual axes split into components; components
add into the fused indices of the source's axes.
"""
import ctypes as C

from . import _lib as L
from ._device import DeviceArray


class Perm:
    def __init__(self):
        self.outs = []    # (dim, [(virt, mul)])
        self.virts = []   # [dim, enc, sigma, [comp ids]]
        self.comps = []   # [dim, bit, sigma, [(in, mult)]]
        self.ins = []     # [dim, enc, sigma, split, stride, stride_lo]
        self.codes = [0]
        self.n_bits = 0

    def add_in(self, dim, stride, enc=False, sigma=False, split=False, stride_lo=0):
        self.ins.append([int(dim), bool(enc), bool(sigma), bool(split), int(stride), int(stride_lo)])
        return len(self.ins) - 1

    def add_virt(self, dim, enc=False, sigma=False):
        self.virts.append([int(dim), bool(enc), bool(sigma), []])
        return len(self.virts) - 1

    def add_comp(self, virt, dim, targets, bit=-1, sigma=False):
        """Append a component to `virt` (call in C order: most significant first)."""
        self.comps.append([int(dim), int(bit), bool(sigma), [(int(i), int(m)) for i, m in targets]])
        cid = len(self.comps) - 1
        self.virts[virt][3].append(cid)
        return cid

    def add_out(self, dim, feeds):
        self.outs.append((int(dim), [(int(v), int(m)) for v, m in feeds]))
        return len(self.outs) - 1

    def set_table(self, codes):
        n = max(0, (len(codes) - 1).bit_length())
        if len(codes) != (1 << n):
            raise ValueError("table length must be a power of two")
        self.codes = list(codes)
        self.n_bits = n

    def desc(self, dtype, src_elems, alpha=1.0):
        d = L.PermDesc()
        if len(self.outs) > L.MAX_AX or len(self.virts) > L.MAX_AX or len(self.ins) > L.MAX_AX:
            raise NotImplementedError("gather with more than %d axes" % L.MAX_AX)
        if len(self.comps) > L.MAX_COMP:
            raise NotImplementedError("gather with more than %d legs" % L.MAX_COMP)
        d.dtype = dtype
        d.n_out, d.n_virt, d.n_in, d.n_bits = len(self.outs), len(self.virts), len(self.ins), self.n_bits
        for a, (dim, feeds) in enumerate(self.outs):
            d.out_dim[a] = dim
            d.out_virt[a][0] = d.out_virt[a][1] = -1
            for t, (v, m) in enumerate(feeds):
                d.out_virt[a][t] = v
                d.out_vmul[a][t] = m
        order = []
        for v, (dim, enc, sig, comps) in enumerate(self.virts):
            d.virt_dim[v] = dim
            d.virt_enc[v] = enc
            d.virt_sigma[v] = sig
            d.virt_ncomp[v] = len(comps)
            order += comps
        d.n_comp = len(order)
        for j, cid in enumerate(order):
            dim, bit, sig, tg = self.comps[cid]
            d.comp_dim[j] = dim
            d.comp_bit[j] = bit
            d.comp_sigma[j] = sig
            d.comp_in[j][0] = d.comp_in[j][1] = -1
            for t, (i, m) in enumerate(tg):
                d.comp_in[j][t] = i
                d.comp_mult[j][t] = m
        for b, (dim, enc, sig, split, st, stlo) in enumerate(self.ins):
            d.in_dim[b] = dim
            d.in_enc[b] = enc
            d.in_sigma[b] = sig
            d.in_split[b] = split
            d.in_stride[b] = st
            d.in_stride_lo[b] = stlo
        words = [0] * len(d.table)
        for m, code in enumerate(self.codes):
            words[m >> 4] |= (code & 3) << ((m & 15) * 2)
        for i, w in enumerate(words):
            d.table[i] = w
        if isinstance(alpha, complex):
            d.alpha_re, d.alpha_im = alpha.real, alpha.imag
        else:
            d.alpha_re, d.alpha_im = float(alpha), 0.0
        d.src_elems = int(src_elems)
        return d

    def out_shape(self):
        return tuple(dim for dim, _ in self.outs)

    def run(self, src, dst=None, alpha=1.0):
        """Launch on the current stream; returns dst (allocated when None)."""
        lib = L.lib()
        shape = self.out_shape()
        if dst is None:
            dst = DeviceArray.empty(shape, src.dtype)
        if dst.numel() == 0:
            return dst
        if not (src.is_contiguous() and dst.is_contiguous()):
            raise ValueError("gather operands must be contiguous")
        d = self.desc(L.dtype_code(src), src.numel(), alpha)
        L.check(lib.gtn_permute(C.byref(d), L.ptr(src), L.ptr(dst), L.stream_ptr()), "gtn_permute")
        return dst


def run_composed(outer, inner, src, dst=None, alpha=1.0):
    """outer(inner(src)) in one pass (gtn_permute2): `outer` is built against the contiguous tensor
    that `inner` would write."""
    lib = L.lib()
    if dst is None:
        dst = DeviceArray.empty(outer.out_shape(), src.dtype)
    if dst.numel() == 0:
        return dst
    if not (src.is_contiguous() and dst.is_contiguous()):
        raise ValueError("gather operands must be contiguous")
    mid = 1
    for d in inner.out_shape():
        mid *= d
    do = outer.desc(L.dtype_code(src), mid, alpha)
    dn = inner.desc(L.dtype_code(src), src.numel())
    L.check(lib.gtn_permute2(C.byref(do), C.byref(dn), L.ptr(src), L.ptr(dst), L.stream_ptr()), "gtn_permute2")
    return dst


def strides(shape):
    st, acc = [0] * len(shape), 1
    for i in range(len(shape) - 1, -1, -1):
        st[i] = acc
        acc *= shape[i]
    return st


def source_axes(perm, shape, encs, sigmas=None, base_stride=None):
    """Register a plain C-ordered source tensor (one leg per axis); returns the input ids."""
    st = strides(shape) if base_stride is None else base_stride
    ids = []
    for a, dim in enumerate(shape):
        ids.append(perm.add_in(dim, st[a], enc=encs[a], sigma=(sigmas[a] if sigmas else False)))
    return ids
