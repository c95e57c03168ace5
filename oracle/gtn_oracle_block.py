"""CPU FP64 oracle of the block Grassmann format -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of grassmanntn's ``block`` container and the operations on it: the container and
its sign vectors (__init__.py:242-337), todense (__init__.py:690-729), switch_format / switch_parity
(__init__.py:537-603), einsum_block (__init__.py:2346-2941), decompose_block = svd_block / eig_block
(__init__.py:4704-5289), hconjugate_block (__init__.py:5495-5955) and power_block (__init__.py:6071-6082,
with the documented square root: the snapshot's power_block discards its result).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py`` may import it; the product never routes through it.

The container is kept exactly as the reference keeps it: one numpy array ("cell") per assignment of
parities to the fermionic legs, cell[p][i] holding the parity-preserving entry 2 i + p, and per leg and
parity a sign vector sgn[p][axis][i] = sigma(encoder(2 i + p)).  einsum_block is restated literally
(loop over the parity assignments of the fermionic characters, S1 * S3 from the same relative-sign
windows as einsum_ds, sigma vectors taken from the operands' sign vectors).

decompose_block and hconjugate_block are restated through the dense parity-preserving form: the
reference's block join orders each joined leg's sub-blocks by component parities and gives it the
product of the component sign vectors, while the dense join orders it by the fused parity-preserving
index with sigma of the fused index.  The two matrices differ by a row/column permutation and a
diagonal +-1 that the split undoes, so U, S, V differ only by the gauge of U / V; the truncation
(ceil / floor halves, both parities padded to the larger kept count, bond 2^(1+ceil(log2 kept))) is
the block one.

Pinning (tests/test_oracle_block.py): every cell and sign vector against the parity-preserving slices
and the param tables; einsum_block against the dense oracle einsum (itself pinned by brute-force
Grassmann algebra) on todense'd operands for every contraction pattern of the TRG / ATRG drivers and
scalar outputs; format / parity switches against the dense ones; decompositions by reconstruction,
spectra against numpy.linalg.svd of the parity blocks and the truncation / shape rules; block TRG
against dense TRG at power-of-two bonds (docs/_sources/schwinger.rst.txt:920-922).
Parity unpinned: as for the dense oracle, the LAPACK gauge of U / V.
"""
import itertools
import math

import numpy as np

from . import gtn_oracle as O

FERMI, BOSE = O.FERMI, O.BOSE


def _pow2_ceil(n):
    return 1 if n <= 1 else int(2 ** math.ceil(math.log2(n)))


def _sign_vector(p, n):
    """sigma(encoder(2 i + p)) for i < n: the sign vector of block.__init__ (__init__.py:288-301)."""
    i = 2 * np.arange(n) + p
    return O.sgn(O.encoder(i)).astype(np.int64)


class block:
    """Cells + sign vectors + statistics / format / physical shape (__init__.py:242-337)."""

    def __init__(self, cells, sgn, statistics, format="standard", shape=None, joined=False):
        self.cells = cells            # dict: fermionic parity tuple -> ndarray
        self.sgn = sgn                # [even vectors per axis, odd vectors per axis]
        self.statistics = tuple(statistics)
        self.format = format
        self.shape = tuple(shape)
        self.marked_as_joined = joined

    # ---- construction from dense (block.__init__, __init__.py:247-337) -------------------------
    @staticmethod
    def from_dense(T, even_shape=None, odd_shape=None, shape=None):
        """The reference's block(dense) conversion; optional even/odd shapes trim the cells (for
        tensors whose live extent is smaller than half the stored dimension)."""
        P = T.force_encoder("parity-preserving")
        stats = P.statistics
        fax = [a for a, s in enumerate(stats) if s in FERMI]
        base = [max(1, d // 2) if s in FERMI else d for d, s in zip(P.shape, stats)]
        sgn = [[], []]
        for a, (d, s) in enumerate(zip(P.shape, stats)):
            for p in (0, 1):
                if s in BOSE:
                    sgn[p].append(np.ones(d, dtype=np.int64))
                else:
                    v = np.zeros(base[a], dtype=np.int64)
                    live = len(range(p, d, 2))
                    v[:live] = _sign_vector(p, live)
                    sgn[p].append(v)
        cells = {}
        for par in itertools.product((0, 1), repeat=len(fax)):
            cell = np.zeros(base, dtype=P.data.dtype)
            src, dst = [], []
            pa = dict(zip(fax, par))
            for a, d in enumerate(P.shape):
                if a in pa:
                    idx = np.arange(pa[a], d, 2)
                    src.append(idx)
                    dst.append(np.arange(len(idx)))
                else:
                    src.append(np.arange(d))
                    dst.append(np.arange(d))
            cell[np.ix_(*dst)] = P.data[np.ix_(*src)]
            cells[par] = cell
        B = block(cells, sgn, stats, P.format, P.shape if shape is None else shape)
        if even_shape is not None:
            B = B._trim(even_shape, odd_shape)
        return B

    def _trim(self, even_shape, odd_shape):
        fax = self.fermionic_axes()
        cells = {}
        for par, c in self.cells.items():
            pa = dict(zip(fax, par))
            sl = tuple(slice(0, (even_shape[a] if pa[a] == 0 else odd_shape[a])) if a in pa else slice(None)
                       for a in range(self.ndim))
            cells[par] = c[sl].copy()
        sgn = [[v[:(even_shape[a] if self.statistics[a] in FERMI else len(v))].copy() for a, v in enumerate(self.sgn[0])],
               [v[:(odd_shape[a] if self.statistics[a] in FERMI else len(v))].copy() for a, v in enumerate(self.sgn[1])]]
        return block(cells, sgn, self.statistics, self.format, self.shape, self.marked_as_joined)

    # ---- properties (__init__.py:431-475) -------------------------------------------------------
    def fermionic_axes(self):
        return [a for a, s in enumerate(self.statistics) if s in FERMI]

    @property
    def ndim(self):
        return len(self.statistics)

    @property
    def even_shape(self):
        return tuple(len(v) if len(v) and v[0] != 0 else 0 for v in self.sgn[0])

    @property
    def odd_shape(self):
        return tuple(len(v) if len(v) and v[0] != 0 else 0 for v in self.sgn[1])

    @property
    def effective_shape(self):
        return tuple(self.shape[a] if s in BOSE else e + o
                     for a, (s, e, o) in enumerate(zip(self.statistics, self.even_shape, self.odd_shape)))

    @property
    def norm(self):
        return float(math.sqrt(sum(float(np.linalg.norm(c)) ** 2 for c in self.cells.values())))

    def copy(self):
        return block({k: v.copy() for k, v in self.cells.items()}, [[v.copy() for v in s] for s in self.sgn],
                     self.statistics, self.format, self.shape, self.marked_as_joined)

    # ---- arithmetic (__init__.py:489-549) -------------------------------------------------------
    def __add__(self, other):
        if (self.even_shape != other.even_shape or self.odd_shape != other.odd_shape or self.shape != other.shape
                or self.statistics != other.statistics or self.format != other.format):
            O._fail("Error[block.+]: Inconsistent object properties")
        r = self.copy()
        for k in r.cells:
            r.cells[k] = self.cells[k] + other.cells[k]
        return r

    def __mul__(self, x):
        if not np.isscalar(x):
            O._fail("Error[block.*]: Only scalar multiplication is allowed.")
        r = self.copy()
        for k in r.cells:
            r.cells[k] = r.cells[k] * x
        return r

    __rmul__ = __mul__

    def __truediv__(self, x):
        return self * (1.0 / x)

    def __neg__(self):
        return self * (-1)

    def __sub__(self, other):
        return self + (-1) * other

    # ---- switches (__init__.py:551-603) ---------------------------------------------------------
    def switch_format(self):
        r = self.copy()
        fax = self.fermionic_axes()
        for par, c in r.cells.items():
            pa = dict(zip(fax, par))
            for a, s in enumerate(self.statistics):
                if s == -1:
                    v = self.sgn[pa[a]][a]
                    shp = [1] * c.ndim
                    shp[a] = c.shape[a]
                    c = c * v[:c.shape[a]].reshape(shp)
            r.cells[par] = c
        r.format = "matrix" if self.format == "standard" else "standard"
        return r

    def force_format(self, fmt):
        return self.copy() if fmt == self.format else self.switch_format()

    def switch_parity(self):
        r = self.copy()
        fax = self.fermionic_axes()
        for par, c in r.cells.items():
            s = 1
            for a, p in zip(fax, par):
                if self.statistics[a] == 1 and p == 1:
                    s = -s
            r.cells[par] = c * s
        return r

    # ---- dense views (todense, __init__.py:690-729) ---------------------------------------------
    def todense(self, encoder="canonical", shape=None):
        shape = self.shape if shape is None else shape
        dt = np.result_type(*[c.dtype for c in self.cells.values()])
        data = np.zeros(shape, dtype=dt)
        fax = self.fermionic_axes()
        for par, c in self.cells.items():
            pa = dict(zip(fax, par))
            idx = [2 * np.arange(c.shape[a]) + pa[a] if a in pa else np.arange(c.shape[a]) for a in range(c.ndim)]
            for a, ix in enumerate(idx):
                if len(ix) and ix[-1] >= shape[a]:
                    O._fail("Error[todense]: block entry outside the physical shape")
            data[np.ix_(*idx)] = c
        T = O.dense(data, encoder="parity-preserving", format=self.format, statistics=self.statistics,
                    skip_power_of_two_check=True)
        return T.force_encoder(encoder)

    def storage_shape(self):
        """Power-of-two parity-preserving extent that holds every live entry of each leg."""
        return tuple(d if s in BOSE else 2 * _pow2_ceil(max(e, o, 1))
                     for d, s, e, o in zip(self.shape, self.statistics, self.even_shape, self.odd_shape))

    def svd(self, string, cutoff=None):
        return svd_block(self, string, cutoff)

    def eig(self, string, cutoff=None):
        return eig_block(self, string, cutoff)

    def hconjugate(self, string):
        return hconjugate_block(self, string)


def zero_block_eo(even_shape, odd_shape, statistics, format="standard", dtype=float):
    """(__init__.py:646-672)"""
    eff = [e + o if s in FERMI else e for e, o, s in zip(even_shape, odd_shape, statistics)]
    shape = [_pow2_ceil(d) if s in FERMI else d for d, s in zip(eff, statistics)]
    fax = [a for a, s in enumerate(statistics) if s in FERMI]
    cells = {}
    for par in itertools.product((0, 1), repeat=len(fax)):
        pa = dict(zip(fax, par))
        cells[par] = np.zeros([(even_shape[a] if pa[a] == 0 else odd_shape[a]) if a in pa else odd_shape[a]
                               for a in range(len(statistics))], dtype=dtype)
    sgn = [[], []]
    for a, s in enumerate(statistics):
        for p, n in ((0, even_shape[a]), (1, odd_shape[a])):
            sgn[p].append(np.ones(n, dtype=np.int64) if s in BOSE else _sign_vector(p, n))
    return block(cells, sgn, statistics, format, shape)


def _reorder(src, dst, xs):
    return [xs[src.index(c)] for c in dst]


# --------------------------------------------------------------------------------------------------
# einsum_block (__init__.py:2346-2941)
# --------------------------------------------------------------------------------------------------
def einsum_block(subscripts, *operands):
    sub = O.denumerate(subscripts.replace(" ", ""))
    if sub.count("->") > 1:
        O._fail("Error[einsum_block]: There should be no more than one '->' indicator in the subscript input.")
    has_output = "->" in sub
    before, after = sub.split("->") if has_output else (sub, "")
    strings = before.split(",")
    ops = list(operands[:len(strings)])
    for o in ops:
        if not isinstance(o, block):
            O._fail("Error[einsum_block]: This function only works with block data format.")
    first_format = ops[0].format
    ops = [o.force_format("standard") for o in ops]
    summand = before.replace(",", "")
    stats = sum([list(o.statistics) for o in ops], [])
    eff = sum([list(o.effective_shape) for o in ops], [])
    esh = sum([list(o.even_shape) for o in ops], [])
    osh = sum([list(o.odd_shape) for o in ops], [])
    # sign windows and statistic checks are einsum_ds's (the block code copies them, :2490-2650)
    plan = O.EinsumPlan(sub, stats, eff)
    all_sgn = [sum([list(o.sgn[p]) for o in ops], []) for p in (0, 1)]
    uniq = sorted(set(c for c, s in zip(summand, stats) if s in FERMI))
    obj_f = [[c for c in s if c in uniq] for s in strings]
    pairs = [c for c, _ in plan.pairs]
    dt = np.result_type(*[c.dtype for o in ops for c in o.cells.values()])
    if has_output:
        out_stats = _reorder(summand, after, stats)
        out_e = _reorder(summand, after, esh)
        out_o = _reorder(summand, after, osh)
        out_eff = _reorder(summand, after, eff)
        ret = zero_block_eo(out_e, out_o, out_stats, "standard", dt)
        # output sign vectors come from the first operand holding each output index (:2420-2440)
        for k, c in enumerate(after):
            j = next(n for n, s in enumerate(strings) if c in s)
            ret.sgn[0][k] = ops[j].sgn[0][strings[j].index(c)].copy()
            ret.sgn[1][k] = ops[j].sgn[1][strings[j].index(c)].copy()
        f_after = [c for c in after if c in uniq]
        ret.shape = tuple(d if s in BOSE else _pow2_ceil(d) for d, s in zip(out_eff, out_stats))
        scalar = None
    else:
        scalar = 0.0
    spec = before + "".join("," + c for c in pairs) + "->" + after
    for bits in itertools.product((0, 1), repeat=len(uniq)):
        par = dict(zip(uniq, bits))
        cells = [o.cells[tuple(par[c] for c in f)] for o, f in zip(ops, obj_f)]
        sig = []
        for c in uniq:
            occ = [i for i, x in enumerate(summand) if x == c]
            v0 = all_sgn[par[c]][occ[0]]
            for i in occ[1:]:
                if np.linalg.norm(v0 - all_sgn[par[c]][i]) > O.NUMER_CUTOFF:
                    O._fail("Error[einsum_block]: some sign factors are not consistent.")
        for c in pairs:
            sig.append(all_sgn[par[c]][summand.index(c)])
        val = plan.sign(par) * np.einsum(spec, *cells, *sig)
        if has_output:
            ret.cells[tuple(par[c] for c in f_after)] += val
        else:
            scalar = scalar + val
    if not has_output:
        return scalar
    return ret.force_format(first_format)


# --------------------------------------------------------------------------------------------------
# decompose_block (__init__.py:4704-5289), hconjugate_block (__init__.py:5495-5955)
# --------------------------------------------------------------------------------------------------
def _as_dense(B):
    """The block as a dense parity-preserving tensor of its power-of-two storage shape."""
    return B.todense("parity-preserving", B.storage_shape())


def _decompose_block(B, string, cutoff, kind):
    name = "svd_block" if kind == "svd" else "eig_block"
    for s in B.statistics:
        if s not in O.ALLOWED_STAT:
            O._fail("Error[%s]: The input object contains illegal statistics. (0, 1, -1, or * only)" % name)
    string = O._parse_partition(string, name)
    left = string.split([c for c in string if c in O.SEPARATORS][0])[0]
    nl = len(left)
    sl, sr = B.statistics[:nl], B.statistics[nl:]
    if all(s in BOSE for s in sl) or all(s in BOSE for s in sr):
        O._fail("Error[SVD/EigV_block]: This function do not support the case where one of the decomposed "
                "branches is totally bosonic.")
    Bs = B.force_format("standard")
    for par, c in Bs.cells.items():
        if sum(par) % 2 == 1 and np.linalg.norm(c) > O.NUMER_CUTOFF:
            O._fail("Error[SVD/EigV_block]: This function only works for even tensors!")
    kept = []
    U, S, V = O._decompose(_as_dense(Bs), string, cutoff, kind, rule="block", kept=kept)
    n = kept[0]
    nfull = 2 ** int(1 + math.ceil(math.log2(max(n, 1))))
    es, os_ = B.even_shape, B.odd_shape
    Ub = block.from_dense(U, es[:nl] + (n,), os_[:nl] + (n,), B.shape[:nl] + (nfull,))
    Sb = block.from_dense(S, (n, n), (n, n), (nfull, nfull))
    Vb = block.from_dense(V, (n,) + es[nl:], (n,) + os_[nl:], (nfull,) + B.shape[nl:])
    return tuple(x.force_format(B.format) for x in (Ub, Sb, Vb))


def svd_block(B, string, cutoff=None):
    return _decompose_block(B, string, cutoff, "svd")


def eig_block(B, string, cutoff=None):
    return _decompose_block(B, string, cutoff, "eig")


def hconjugate_block(B, string):
    string = O._parse_partition(string, "hconjugate_block")
    left = string.split([c for c in string if c in O.SEPARATORS][0])[0]
    nl = len(left)
    H = O.hconjugate(_as_dense(B), string)
    perm = list(range(nl, B.ndim)) + list(range(nl))
    es = tuple(B.even_shape[a] for a in perm)
    os_ = tuple(B.odd_shape[a] for a in perm)
    return block.from_dense(H, es, os_, tuple(B.shape[a] for a in perm)).force_format(B.format)


def power_block(B, p, rcond=1e-10):
    """Elementwise power of the matrix-format cells where |x| > rcond (__init__.py:6071-6082, with the
    result kept as the docs' numbers require)."""
    fmt = B.format
    M = B.force_format("matrix")
    for k, c in M.cells.items():
        with np.errstate(invalid="ignore", divide="ignore"):
            M.cells[k] = np.where(np.abs(c) > rcond, np.power(c.astype(np.result_type(c, float)), p), 0.0)
    return M.force_format(fmt)


def sqrt_block(B):
    return power_block(B, 0.5)


# --------------------------------------------------------------------------------------------------
# block TRG driver and logZ (gauge2d_block.py:1601-1760)
# --------------------------------------------------------------------------------------------------
def trg_block(T, dcut=64, error_test=False):
    if [T.shape[0], T.shape[1]] != [T.shape[2], T.shape[3]]:
        O._fail("Error[trg]: The shape must be of the form (m,n,m,n)!")
    if list(T.statistics) != [1, 1, -1, -1]:
        O._fail("Error[trg]: The statistics must be (1,1,-1,-1)!")
    T1 = einsum_block("ijkl->jkli", T)
    T2 = einsum_block("ijkl->klij", T)
    U1, S1, V1 = svd_block(T1, "ab|cd", dcut)
    U2, S2, V2 = svd_block(T2, "ab|cd", dcut)
    r = sqrt_block(S1)
    U1 = einsum_block("abx,xc->abc", U1, r)
    V1 = einsum_block("ax,xbc->abc", r, V1)
    r = sqrt_block(S2)
    U2 = einsum_block("abx,xc->abc", U2, r)
    V2 = einsum_block("ax,xbc->abc", r, V2)
    VV = einsum_block("kwz,lxw->lxzk", V1, V2)
    UU = einsum_block("yxi,zyj->jzxi", U1, U2)
    Tn = einsum_block("lxzk,jzxi->ijkl", VV, UU)
    err = None
    if error_test:
        Z1 = einsum_block("ijkl,klij", T, T)
        Z2 = einsum_block("ijij", Tn)
        err = abs(1 - Z2 / Z1)
    nrm = Tn.norm
    Tn = Tn / nrm
    return (Tn, nrm, err) if error_test else (Tn, nrm)


def logZ_block(T):
    """ln of einsum('IJIJ'); the block driver applies no anti-periodic signs (gauge2d_block.py:1601-1615)."""
    Z = einsum_block("IJIJ", T)
    return np.log(Z + 0j) if np.isreal(Z) and Z < 0 else np.log(Z)
