"""CPU FP64 oracle for the Grassmann-tensor hot path of grassmanntn -- TEST INFRASTRUCTURE ONLY.

This module is a from-scratch numpy restatement of the reference algorithm.  It is the checker that
the HIP product path is compared against; it is never the thing measured or shipped.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it, and
the product package (``grassmanntn_amd``) must never route through it.

Each function names the reference file:line it follows (paths relative to the grassmanntn tree).

Pinning (what checks each part independently of this restatement; see tests/test_oracle_*.py):
  * param closed forms      -- the three 65,536-entry literal tables of param.py:4-6 (fixture
                               tests/golden/param_tables.npz, written by tests/golden/make_param_tables.py
                               which parses the tables as text) and the docs' closed forms.
  * einsum signs            -- brute-force exact Grassmann algebra (oracle/grassmann_algebra.py) on
                               1-2 bit legs, the docs' closed form for 'ilk,jl->ijk'
                               (docs/_sources/grassmanntn/einsum.rst.txt:51-107), and agreement of the
                               literal S1/S2/S3 contraction with the per-parity-sector contraction.
  * join/split, encoder     -- round trips (join->split = id, switch_encoder twice = id) and the closed
                               form of the C-order reshape.
  * svd / eig / hconjugate  -- reconstruction identities (U.S.V = T, U.L.U^+ = H, H^+ = H for H = A A^+),
                               block orthonormality, singular values against numpy.linalg.svd of the
                               parity blocks, and the truncation/padding rules.
  * trg / logZ              -- the driver's own error_test identity Z(T') ~ Z(T,T) at full bond, and the
                               Frobenius normalisation.
Parity unpinned: the exact LAPACK gauge of U/V (no GPU method reproduces it; only gauge-invariant
quantities are compared) and tie-breaking of degenerate singular values at the truncation edge.
"""

import math
import numpy as np

# --------------------------------------------------------------------------------------------------
# globals (__init__.py:19-37)
# --------------------------------------------------------------------------------------------------
HYBRID = "*"
SEPARATORS = ("|", ":", ";", ",", ".")
DIGITS = tuple("0123456789")
FERMI = (1, -1)
BOSE = (0, HYBRID)
ALLOWED_STAT = (0, 1, -1, HYBRID)
ENCODERS = ("canonical", "parity-preserving")
FORMATS = ("standard", "matrix")
NUMER_CUTOFF = 1.0e-14
# index alphabet used when fresh characters are needed (__init__.py:31-36)
CHARS = tuple("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ") + tuple(
    "αβΓγΔδεζηΘθικλμνΞξΠπρΣσςτυΦϕφχΨψΩω")


class GrassmannError(NameError):
    """The reference's error() prints and then raises NameError on purpose (__init__.py:57-63)."""


def _fail(msg):
    raise GrassmannError(msg)


# --------------------------------------------------------------------------------------------------
# A1: parity, sigma and encoder (param.py:50-83, closed forms of the tables at param.py:4-6)
# --------------------------------------------------------------------------------------------------
def popcount(x):
    x = np.asarray(x, dtype=np.int64)
    c = np.zeros_like(x)
    y = x.copy()
    while np.any(y):
        c += y & 1
        y >>= 1
    return c


def gparity(i):
    """Number of occupied Grassmann bits of canonical index i (param.py:50-54, 61-65)."""
    return popcount(i)


def sgn(i):
    """sigma(i) = (-1)^(p(p-1)/2) with p = popcount(i) (param.py:56-59, 67-71)."""
    p = popcount(i)
    return np.where((p * (p - 1) // 2) % 2 == 0, 1, -1).astype(np.int64)


def encoder(i):
    """Canonical <-> parity-preserving switch; an involution whose LSB is the parity (param.py:34-47, 73-83)."""
    i = np.asarray(i, dtype=np.int64)
    return i ^ (popcount(i >> 1) & 1)


def parity_of_canonical(i):
    return popcount(i) & 1


# --------------------------------------------------------------------------------------------------
# A2: sign of a reordering in which only Grassmann-odd elements anticommute (__init__.py:1483-1591)
# --------------------------------------------------------------------------------------------------
def perm_sign(src, dst, odd):
    """Sign picked up when the sequence ``src`` is reordered into ``dst``.

    ``odd`` maps each element to its parity; the sign is (-1)^(number of pairs of odd elements whose
    relative order differs between src and dst).  Elements that occur more than once in ``src`` are
    dropped first, as relative_sign does for summed pairs (__init__.py:1580-1590).
    """
    src = [c for c in src if src.count(c) == 1]
    dst = [c for c in dst if c in src]
    pos = {c: k for k, c in enumerate(dst)}
    o = [c for c in src if odd[c] % 2 == 1]
    inv = 0
    for a in range(len(o)):
        for b in range(a + 1, len(o)):
            if pos[o[a]] > pos[o[b]]:
                inv += 1
    return -1 if inv % 2 else 1


# --------------------------------------------------------------------------------------------------
# string helpers
# --------------------------------------------------------------------------------------------------
def fresh_char(used):
    for c in CHARS:
        if c not in used:
            return c
    _fail("Error[get_char]: Running out of index character!")


def denumerate(string):
    """Numbered indices ('i12') become single fresh characters (__init__.py:1610-1631)."""
    tokens = []
    for c in string:
        if c in DIGITS and tokens:
            tokens[-1] += c
        else:
            tokens.append(c)
    multi = sorted({t for t in tokens if len(t) > 1}, key=len, reverse=True)
    used = string
    repl = []
    for t in multi:
        c = fresh_char(used)
        used += c
        repl.append((t, c))
    for t, c in repl:
        string = string.replace(t, c)
    return string


# --------------------------------------------------------------------------------------------------
# dense container (__init__.py:814-1195)
# --------------------------------------------------------------------------------------------------
class dense:
    def __init__(self, data=None, encoder="canonical", format="standard", statistics=None,
                 skip_power_of_two_check=False):
        if encoder not in ENCODERS:
            _fail("Error[dense]: Unknown encoder.")
        if format not in FORMATS:
            _fail("Error[dense]: Unknown format.")
        self.encoder = encoder
        self.format = format
        self.statistics = None
        self.data = None
        if isinstance(data, dense):
            self.data = data.data.copy()
            self.statistics = data.statistics
            self.encoder = data.encoder
            self.format = data.format
        elif data is not None:
            self.data = np.array(data)
            if self.data.ndim == 0:
                self.data = self.data.reshape(1)
        if statistics is not None:
            self.statistics = tuple(statistics) if not np.isscalar(statistics) else (statistics,)
        if self.data is not None and not skip_power_of_two_check:
            for d, s in zip(self.data.shape, self.statistics):
                if s in FERMI and d != 2 ** int(math.floor(math.log2(d))):
                    _fail("Error[dense]: Some of the fermionic tensor shapes are not a power of two.")

    shape = property(lambda self: self.data.shape)
    ndim = property(lambda self: self.data.ndim)
    size = property(lambda self: self.data.size)
    norm = property(lambda self: float(np.linalg.norm(self.data)))

    def copy(self):
        r = dense.__new__(dense)
        r.data = self.data.copy()
        r.statistics = self.statistics
        r.encoder = self.encoder
        r.format = self.format
        return r

    def _axis_scale(self, which, vec_of):
        d = self.data
        for ax, s in enumerate(self.statistics):
            if s == HYBRID:
                _fail("Error[switch_format]: Cannot switch format with a hybrid index.")
            if s == which:
                v = vec_of(np.arange(d.shape[ax]))
                shp = [1] * d.ndim
                shp[ax] = d.shape[ax]
                d = d * v.reshape(shp)
        return d

    def switch_encoder(self):
        """Gather by the encoder permutation along every fermionic axis (__init__.py:1070-1102)."""
        r = self.copy()
        for ax, s in enumerate(r.statistics):
            if s in FERMI:
                r.data = np.take(r.data, encoder(np.arange(r.data.shape[ax])), axis=ax)
        r.encoder = "parity-preserving" if r.encoder == "canonical" else "canonical"
        return r

    def switch_format(self):
        """Multiply sigma(i) along every -1 axis, in the canonical encoder (__init__.py:1011-1068)."""
        r = self.copy()
        pp = r.encoder == "parity-preserving"
        if pp:
            r = r.switch_encoder()
        r.data = r._axis_scale(-1, sgn)
        r.format = "matrix" if r.format == "standard" else "standard"
        if pp:
            r = r.switch_encoder()
        return r

    def switch_parity(self):
        """Multiply (-1)^p(i) along every +1 axis, in the canonical encoder (__init__.py:1104-1154)."""
        r = self.copy()
        pp = r.encoder == "parity-preserving"
        if pp:
            r = r.switch_encoder()
        r.data = r._axis_scale(1, lambda i: 1 - 2 * parity_of_canonical(i))
        if pp:
            r = r.switch_encoder()
        return r

    def force_encoder(self, target="canonical"):
        if target not in ENCODERS:
            _fail("Error[dense.force_encoder]: Unrecognized target encoder.")
        return self.switch_encoder() if target != self.encoder else self.copy()

    def force_format(self, target="standard"):
        if target not in FORMATS:
            _fail("Error[dense.force_format]: Unrecognized target format.")
        return self.switch_format() if target != self.format else self.copy()

    def __add__(self, other):
        if (self.shape != other.shape or self.statistics != other.statistics
                or self.format != other.format or self.encoder != other.encoder):
            _fail("Error[dense.+]: Inconsistent object properties")
        r = self.copy()
        r.data = r.data + other.data
        return r

    def __mul__(self, x):
        if not np.isscalar(x):
            _fail("Error[dense.*]: Only scalar multiplication is allowed.")
        r = self.copy()
        r.data = r.data * x
        return r

    __rmul__ = __mul__

    def __truediv__(self, x):
        if not np.isscalar(x):
            _fail("Error[dense./]: Only scalar division is allowed.")
        r = self.copy()
        r.data = r.data / x
        return r

    def __neg__(self):
        return self * (-1)

    def __sub__(self, other):
        return self + (-1) * other

    def svd(self, string, cutoff=None):
        return svd(self, string, cutoff)

    def eig(self, string, cutoff=None):
        return eig(self, string, cutoff)

    def hconjugate(self, string):
        return hconjugate(self, string)

    def join_legs(self, string, make_format="standard", intermediate_stat=None):
        return join_legs(self, string, make_format, intermediate_stat)

    def split_legs(self, string, final_stat, final_shape, intermediate_stat=None):
        return split_legs(self, string, final_stat, final_shape, intermediate_stat)


def is_grassmann_even(T):
    """True iff every exactly-nonzero entry has even total fermionic parity (__init__.py:3900-3925)."""
    P = T.force_encoder("parity-preserving")
    par = np.zeros(P.shape, dtype=np.int64)
    for ax, s in enumerate(P.statistics):
        if s in FERMI:
            shp = [1] * P.ndim
            shp[ax] = P.shape[ax]
            par = par + (np.arange(P.shape[ax]) & 1).reshape(shp)
    return not np.any((par % 2 == 1) & (P.data != 0))


# --------------------------------------------------------------------------------------------------
# A3-A5: einsum (dispatcher __init__.py:2308-2344, dense path einsum_ds __init__.py:1633-2306)
# --------------------------------------------------------------------------------------------------
class EinsumPlan:
    """Sign bookkeeping of one einsum call, restated from einsum_ds steps 0-4 (__init__.py:1655-2126).

    Attributes
      chars      : fermionic characters (original names) whose parities the sign depends on
      pairs      : list of (c, c_conj_renamed) for contracted fermionic pairs; sigma(c) enters (S2)
      s1_chars   : axes of the reference's S1 tensor (trimmed window, __init__.py:1804-1940)
      s3_chars   : axes of the reference's S3 tensor (trimmed window, __init__.py:2045-2085)
      out_stats  : statistics of the result (__init__.py:2216-2221)
    """

    def __init__(self, subscripts, stats_list, shape_list):
        sub = denumerate(subscripts.replace(" ", ""))
        if sub.count("->") > 1:
            _fail("Error[einsum]: Only one arrow is allowed in the input string!")
        self.has_output = "->" in sub
        if self.has_output:
            inp, out = sub.split("->")
            if "," in out:
                _fail("Error[einsum]: Output string must not contain commas (',') !")
        else:
            inp, out = sub, ""
        self.input_string, self.output_string = inp, out
        self.operand_strings = inp.split(",")
        summand = inp.replace(",", "")
        self.summand = summand
        if len(summand) != len(stats_list):
            _fail("Error[einsum]: subscripts do not match the operands' dimensions.")
        for st in stats_list:
            if st == HYBRID:
                raise NotImplementedError("einsum over hybrid ('*') indices is not supported")
        self.stats_list = list(stats_list)
        self.shape_list = list(shape_list)
        dim = {}
        for c, d in zip(summand, shape_list):
            if c in dim and dim[c] != d:
                _fail("Error[einsum]: inconsistent dimensions for index '%s'" % c)
            dim[c] = d
        self.dim = dim

        # statistic checks on fermionic indices (__init__.py:1691-1741)
        seen = {}
        for i, c in enumerate(summand):
            if stats_list[i] in BOSE:
                continue
            lf = inp.count(c)
            if self.has_output:
                rf = out.count(c)
                if lf > 2 or rf > 1 or (lf + rf) % 2 == 1:
                    _fail("Error[einsum]: Inconsistent index statistics.")
            elif lf > 2:
                _fail("Error[einsum]: Inconsistent index statistics.")
            if lf == 2:
                seen.setdefault(c, []).append(stats_list[i])
        for c, st in seen.items():
            if len(st) == 2 and sorted(st) != [-1, 1]:
                _fail("Error[einsum]: The contracted indices have inconsistent statistics!")

        # step 1: fermionic string; rename the conjugated member of each contracted pair
        used = summand
        fs, fs_u, fstats = [], [], []
        partner = {}
        for i, c in enumerate(summand):
            if stats_list[i] == 0:
                continue
            fs.append(c)
            fstats.append(stats_list[i])
            if stats_list[i] == -1 and summand.count(c) > 1:
                n = fresh_char(used)
                used += n
                partner[c] = n
                fs_u.append(n)
            else:
                fs_u.append(c)
        order = []
        for c in summand:
            if c in partner and c not in order:
                order.append(c)
        self.pairs = [(c, partner[c]) for c in order]
        orig = {c: c for c in fs_u}
        for c, n in self.pairs:
            orig[n] = c
        self.orig = orig

        # step 2: move each renamed conjugate to the right of its partner (__init__.py:1791-1800)
        fs_ur = [c for c in fs_u if c not in partner.values()]
        for c, n in self.pairs:
            k = fs_ur.index(c)
            fs_ur.insert(k + 1, n)
        self.fs_u, self.fs_ur = fs_u, fs_ur

        # reference's S1 window: trim the common prefix and suffix (__init__.py:1809-1834)
        nall = len(fs_u)
        nl = 0
        while nl < nall and fs_u[nl] == fs_ur[nl]:
            nl += 1
        ns = 0
        while ns < nall and fs_u[nall - 1 - ns] == fs_ur[nall - 1 - ns]:
            ns += 1
        nr = nall - ns
        self.skip_s1 = nl > nr
        win = fs_u[nl:nr] if not self.skip_s1 else []
        self.s1_window = (win, fs_ur[nl:nr] if not self.skip_s1 else [])
        # drop a renamed conjugate whose partner is inside the window (__init__.py:1855-1871)
        utx = [c for c in win if not (c in partner.values() and orig[c] in win)]
        self.s1_chars = [orig[c] for c in utx]

        # step 4: remaining fermions after removing contracted pairs, to the output order
        rest = [c for c in fs_ur if c not in partner and c not in partner.values()]
        self.rest = rest
        if self.has_output:
            fout = [c for c in out if stats_list[summand.index(c)] != 0]
            n3 = len(rest)
            a = 0
            while a < n3 and rest[a] == fout[a]:
                a += 1
            b = 0
            while b < n3 and rest[n3 - 1 - b] == fout[n3 - 1 - b]:
                b += 1
            b = n3 - b
            self.skip_s3 = a > b
            self.s3_window = (rest[a:b], fout[a:b]) if not self.skip_s3 else ([], [])
            self.s3_chars = list(rest[a:b]) if not self.skip_s3 else []
            self.fout = fout
        else:
            self.skip_s3 = True
            self.s3_window = ([], [])
            self.s3_chars = []
            self.fout = []

        chars = []
        for c in fs:
            if c not in chars:
                chars.append(c)
        self.chars = chars
        self.out_stats = tuple(stats_list[summand.index(c)] for c in out) if self.has_output else ()

    # sign as a function of the parities of self.chars ----------------------------------------
    def s1(self, par):
        """par: dict char -> parity (0/1); renamed conjugates share their partner's parity."""
        if self.skip_s1:
            return 1
        odd = {c: par[self.orig[c]] for c in self.fs_u}
        return perm_sign(self.s1_window[0], self.s1_window[1], odd)

    def s3(self, par):
        if self.skip_s3:
            return 1
        return perm_sign(self.s3_window[0], self.s3_window[1], {c: par[c] for c in self.rest})

    def sign(self, par):
        return self.s1(par) * self.s3(par)

    def sign_tensor(self, which):
        """The reference's S1 or S3 int tensor (values +-1) over its window axes (canonical indices)."""
        chars = self.s1_chars if which == 1 else self.s3_chars
        if not chars:
            return None, chars
        shape = tuple(self.dim[c] for c in chars)
        n = len(chars)
        table = np.empty((2,) * n, dtype=np.int8)
        fn = self.s1 if which == 1 else self.s3
        for bits in np.ndindex(*((2,) * n)):
            par = {c: 0 for c in self.chars}
            for c, b in zip(chars, bits):
                par[c] = b
            table[bits] = fn(par)
        idx = [parity_of_canonical(np.arange(d)) for d in shape]
        grids = np.ix_(*idx)
        return table[grids].astype(np.int64), chars

    def s2_vectors(self):
        return [(c, sgn(np.arange(self.dim[c]))) for c, _ in self.pairs]


def _prepare_operands(operands, nobj):
    ops = list(operands[:nobj])
    for o in ops:
        if not isinstance(o, dense):
            _fail("Error[einsum]: operands must be dense tensors.")
    first = ops[0]
    canon = [o.force_encoder("canonical").force_format("standard") for o in ops]
    return first, canon


def einsum(subscripts, *operands, mode="sector"):
    """Grassmann einsum with the reference's semantics (__init__.py:2308-2344 and 1633-2306).

    mode='literal' contracts with the reference's S1/S2/S3 tensors in one numpy einsum, exactly the
    structure of the call at __init__.py:2295.  mode='sector' loops over parity assignments of the
    fermionic characters (the formulation of einsum_block, __init__.py:2680-2790) and skips sectors
    whose operand slices are exactly zero.  Both return identical values up to summation order.
    """
    sub = denumerate(subscripts.replace(" ", ""))
    inp = sub.split("->")[0]
    nobj = len(inp.split(","))
    first, canon = _prepare_operands(operands, nobj)
    stats = sum([list(o.statistics) for o in canon], [])
    shapes = sum([list(o.shape) for o in canon], [])
    plan = EinsumPlan(sub, stats, shapes)
    if mode == "literal":
        val = _contract_literal(plan, canon)
    elif mode == "sector":
        val = _contract_sector(plan, canon)
    else:
        raise ValueError(mode)
    if not plan.has_output:
        return np.array(val).flatten()[0]
    r = dense(val, statistics=plan.out_stats, skip_power_of_two_check=True)
    r = r.force_encoder(first.encoder)
    return r.force_format(first.format)


def _contract_literal(plan, canon):
    spec = plan.input_string
    arrs = [o.data for o in canon]
    S1, c1 = plan.sign_tensor(1)
    if S1 is not None:
        spec += "," + "".join(c1)
        arrs.append(S1)
    for c, v in plan.s2_vectors():
        spec += "," + c
        arrs.append(v)
    S3, c3 = plan.sign_tensor(3)
    if S3 is not None:
        spec += "," + "".join(c3)
        arrs.append(S3)
    spec += "->" + plan.output_string
    return np.einsum(spec, *arrs, optimize=True)


def _contract_sector(plan, canon):
    chars = plan.chars
    dim = plan.dim
    cls = {c: [np.nonzero(parity_of_canonical(np.arange(dim[c])) == b)[0] for b in (0, 1)]
           for c in chars}
    sig = {c: v for c, v in plan.s2_vectors()}
    out_shape = tuple(dim[c] for c in plan.output_string)
    acc = None
    dtype = np.result_type(*[o.data.dtype for o in canon], np.float64)
    if plan.has_output:
        acc = np.zeros(out_shape, dtype=dtype)
    else:
        acc = np.zeros((), dtype=dtype)
    for bits in np.ndindex(*((2,) * len(chars))):
        par = dict(zip(chars, bits))
        slices = []
        empty = False
        for s, o in zip(plan.operand_strings, canon):
            sl = tuple(cls[c][par[c]] if c in par else slice(None) for c in s)
            a = o.data[np.ix_(*[x if not isinstance(x, slice) else np.arange(o.data.shape[k])
                                for k, x in enumerate(sl)])] if o.data.ndim else o.data
            if a.size == 0 or not np.any(a):
                empty = True
                break
            slices.append(a)
        if empty:
            continue
        s = plan.sign(par)
        spec = plan.input_string
        arrs = list(slices)
        for c, _ in plan.pairs:
            spec += "," + c
            arrs.append(sig[c][cls[c][par[c]]])
        spec += "->" + plan.output_string
        val = np.einsum(spec, *arrs, optimize=True) * s
        if plan.has_output:
            tgt = np.ix_(*[cls[c][par[c]] if c in par else np.arange(dim[c])
                           for c in plan.output_string])
            acc[tgt] += val
        else:
            acc = acc + val
    return acc


# --------------------------------------------------------------------------------------------------
# A8: join_legs / split_legs (__init__.py:2947-3316)
# --------------------------------------------------------------------------------------------------
def _group_info(string, stats, shape):
    """Parse '(ab)c(de)' or 'ab|cd'; bosons move to the left inside each group (__init__.py:3172-3243)."""
    s = string
    if "(" not in s and ")" not in s:
        for sep in SEPARATORS:
            s = s.replace(sep, ")(")
        s = "(" + s + ")"
    elif any(sep in s for sep in SEPARATORS):
        _fail("Error[get_grouping_info]: Do not mix the string format.")
    groups = []
    inside = False
    loc = 0
    for ch in s:
        if ch == "(" and not inside:
            inside = True
            groups.append(["", [], []])
            continue
        if ch == ")" and inside:
            inside = False
            continue
        if ch in "()":
            _fail("Error[get_grouping_info]: No nested parenthesis allowed!")
        if not inside:
            groups.append([ch, [stats[loc]], [shape[loc]]])
        else:
            groups[-1][0] += ch
            groups[-1][1].append(stats[loc])
            groups[-1][2].append(shape[loc])
        loc += 1
    srt = []
    for idx, st, sh in groups:
        b = [(c, t, d) for c, t, d in zip(idx, st, sh) if t in BOSE]
        f = [(c, t, d) for c, t, d in zip(idx, st, sh) if t not in BOSE]
        b.reverse()
        items = b + f
        srt.append(["".join(x[0] for x in items), [x[1] for x in items], [x[2] for x in items]])
    return groups, srt


def _grouping_signs(srt, inter):
    """(-1)^p(a) for every +1 leg a joined into a group of intermediate statistic -1 (__init__.py:3245-3275)."""
    if len(srt) != len(inter):
        _fail("Error[get_grouping_sign_factors]: Inconsistent number of intermediate_stat and groupings!")
    out = []
    for g, (idx, st, sh) in enumerate(srt):
        for c, t, d in zip(idx, st, sh):
            if t == 1 and inter[g] == -1:
                out.append((c, 1 - 2 * parity_of_canonical(np.arange(d))))
    return out


def _intermediate_info(srt, inter):
    """Shapes/statistics after fusing bosons and fermions of each group (__init__.py:3277-3316)."""
    new_shape, new_stats, fin_shape, fin_stats = [], [], [], []
    for g, (idx, st, sh) in enumerate(srt):
        fd = bd = 1
        fn = bn = 0
        for t, d in zip(st, sh):
            if t in FERMI:
                fd *= d
                fn += 1
            else:
                bd *= d
                bn += 1
        if bn:
            new_shape.append(bd)
            new_stats.append(0)
        if fn:
            new_shape.append(fd)
            new_stats.append(inter[g])
        fin_shape.append(bd * fd)
        fin_stats.append(HYBRID if (bn and fn) else inter[g])
    return tuple(new_stats), tuple(new_shape), tuple(fin_stats), tuple(fin_shape)


def _apply_signs(data, string, signs, target):
    spec = string + "".join("," + c for c, _ in signs) + "->" + target
    return np.einsum(spec, data, *[v for _, v in signs])


def join_legs(T, string, make_format="standard", intermediate_stat=None):
    """Fuse groups of legs; output is always parity-preserving (__init__.py:2947-3068)."""
    string = denumerate(string.replace(" ", ""))
    inter = tuple(intermediate_stat) if not np.isscalar(intermediate_stat) else (intermediate_stat,)
    O = T.copy()
    if O.format == "matrix":
        O = O.switch_format()
    if O.encoder == "parity-preserving":
        O = O.switch_encoder()
    groups, srt = _group_info(string, O.statistics, O.shape)
    if sum(len(g[0]) for g in groups) != O.ndim:
        _fail("Error[join_legs]: The number of indices is not consistent with the object's shape.")
    signs = _grouping_signs(srt, inter)
    uns = "".join(g[0] for g in groups)
    srs = "".join(g[0] for g in srt)
    O.data = _apply_signs(O.data, uns, signs, srs)
    O.statistics = tuple(sum([list(g[1]) for g in srt], []))
    new_stats, new_shape, fin_stats, fin_shape = _intermediate_info(srt, inter)
    O.data = O.data.reshape(new_shape)
    O.statistics = new_stats
    if make_format == "matrix":
        O = O.switch_format()
    O = O.switch_encoder()
    O.data = O.data.reshape(fin_shape)
    O.statistics = fin_stats
    return O


def split_legs(T, string, final_stat, final_shape, intermediate_stat=None):
    """Inverse of join_legs, including its early return for matrix-format input (__init__.py:3070-3170)."""
    string = denumerate(string.replace(" ", ""))
    inter = tuple(intermediate_stat) if not np.isscalar(intermediate_stat) else (intermediate_stat,)
    final_stat = tuple(final_stat)
    final_shape = tuple(final_shape)
    O = T.copy()
    fmt, enc = O.format, O.encoder
    groups, srt = _group_info(string, final_stat, final_shape)
    new_stats, new_shape, _, _ = _intermediate_info(srt, inter)
    signs = _grouping_signs(srt, inter)
    O.data = O.data.reshape(new_shape)
    O.statistics = new_stats
    if enc == "parity-preserving":
        O = O.switch_encoder()
    if fmt == "matrix":
        O = O.switch_format()
    st = tuple(sum([list(g[1]) for g in srt], []))
    sh = tuple(sum([list(g[2]) for g in srt], []))
    O.data = O.data.reshape(sh)
    O.statistics = st
    uns = "".join(g[0] for g in groups)
    srs = "".join(g[0] for g in srt)
    O.data = _apply_signs(O.data, srs, signs, uns)
    O.statistics = final_stat
    if fmt == "matrix":
        return O.switch_format()
    if enc == "parity-preserving":
        O = O.switch_encoder()
    return O


# --------------------------------------------------------------------------------------------------
# A10-A11: svd / eig (__init__.py:3931-4698)
# --------------------------------------------------------------------------------------------------
def sorted_svd(M, cutoff=None):
    """LAPACK SVD with the relative 1e-14 rank rule and an optional hard cap (__init__.py:3931-3951)."""
    U, s, Vh = np.linalg.svd(M, full_matrices=False)
    nnz = int(np.sum(np.abs(s / (np.abs(s[0]) + NUMER_CUTOFF)) > NUMER_CUTOFF)) if s.size else 0
    if cutoff is not None and cutoff < nnz:
        nnz = cutoff
    return U[:, :nnz], s[:nnz], Vh[:nnz, :]


def sorted_eig(M, cutoff=None):
    """Hermitian eigendecomposition via SVD, eigenvalues ordered by |lambda| (__init__.py:4314-4345)."""
    nrm = np.linalg.norm(M)
    herm = 0.0 if nrm < NUMER_CUTOFF else np.linalg.norm(M - M.conj().T) / nrm
    if herm > NUMER_CUTOFF:
        _fail("Error[SortedEig]: The input matrix is not Hermitian!")
    U, s, Vh = sorted_svd(M, cutoff)
    lam = np.einsum("i,ij,jk->k", s, Vh, U)
    return U, lam, U.conj().T


def parity_blocks(M):
    """ME = M[0::2,0::2], MO = M[1::2,1::2] of a parity-preserving matrix, after the parity-leak
    check of __init__.py:3977-3984 (index map of __init__.py:3990-3996)."""
    m, n = M.shape
    leak = np.abs(M[0::2, 1::2]).sum() + np.abs(M[1::2, 0::2]).sum()
    if leak / (m * n / 2) > 1.0e-14:
        _fail("Error[BlockSVD]: This matrix is not constructed from a Grassmann-even tensor.")
    return M[0::2, 0::2], M[1::2, 1::2]


def _interleave(AE, AO):
    mh, nh = AE.shape
    A = np.zeros((2 * mh, 2 * nh), dtype=np.result_type(AE, AO))
    A[0::2, 0::2] = AE
    A[1::2, 1::2] = AO
    return A


def block_decompose(M, cutoff=None, kind="svd", block_fn=None, rule="dense", kept=None):
    """BlockSVD / BlockEig: per-parity-block decomposition, int(cutoff/2) kept per block, both blocks
    padded to the next power of two and interleaved (__init__.py:3953-4029, 4347-4422).

    rule="block" keeps ceil(cutoff/2) even and floor(cutoff/2) odd values, the truncation of
    decompose_block (__init__.py:5076-5085); the block format records max(kept) live values per
    parity (__init__.py:5092-5100) and the storage here is still padded to a power of two.
    ``kept`` (a list) receives that max(kept).
    ``block_fn`` may replace the per-block LAPACK call (used to apply a gauge convention)."""
    ME, MO = parity_blocks(M)
    if cutoff is None:
        cE = cO = None
    elif rule == "block":
        cE, cO = int(math.ceil(cutoff / 2)), int(math.floor(cutoff / 2))
    else:
        cE = cO = int(cutoff / 2)
    fn = block_fn or (sorted_svd if kind == "svd" else sorted_eig)
    UE, sE, VE = fn(ME, cE)
    UO, sO, VO = fn(MO, cO)
    d = max(len(sE), len(sO))
    if kept is not None:
        kept.append(d)
    d = int(2 ** math.ceil(math.log2(d))) if d > 0 else 1

    def pad(U, s, V):
        p = d - len(s)
        return (np.pad(U, ((0, 0), (0, p))), np.diag(np.pad(s, (0, p))), np.pad(V, ((0, p), (0, 0))))

    UE, SE, VE = pad(UE, sE, VE)
    UO, SO, VO = pad(UO, sO, VO)
    return _interleave(UE, UO), _interleave(SE, SO), _interleave(VE, VO)


def _parse_partition(string, name):
    string = denumerate(string.replace(" ", ""))
    if string.count("(") == string.count(")") and string.count("(") > 0:
        string = string.replace(")(", "|")
        if string.count("(") > 1 or string.count(")") < 1:
            _fail("Error[%s]: Parentheses don't match" % name)
        string = string.replace(")", "").replace("(", "")
    if sum(string.count(p) for p in SEPARATORS) != 1:
        _fail("Error[%s]: The input string must contain one and only one partition." % name)
    return string


def _decompose(T, string, cutoff, kind, block_fn=None, rule="dense", kept=None):
    """Shared body of svd (__init__.py:4033-4308) and eig (__init__.py:4425-4698)."""
    name = kind
    string = _parse_partition(string, name)
    O = T.copy()
    fmt, enc = O.format, O.encoder
    xstats, xshape = O.statistics, O.shape
    for st in O.statistics:
        if st not in ALLOWED_STAT:
            _fail("Error[%s]: The input object contains illegal statistics." % name)
    left = ""
    right = ""
    seen = False
    for ch in string:
        if ch in SEPARATORS:
            seen = True
            continue
        if seen:
            right += ch
        else:
            left += ch
    nl = len(left)
    js = "(" + left + ")(" + right + ")"
    sl, sr = O.statistics[:nl], O.statistics[nl:]
    shl, shr = O.shape[:nl], O.shape[nl:]
    lb1 = list(sl) in ([-1], [1]) and list(shl) == [1]
    rb1 = list(sr) in ([-1], [1]) and list(shr) == [1]
    if lb1 or rb1:
        # a single fermionic leg of dimension one (__init__.py:4132-4170)
        nrm = O.norm
        M = O.force_format("matrix")
        if lb1:
            U = dense([[1]], statistics=(*sl, 1), format="matrix")
            L = dense([[nrm]], statistics=(-1, 1), format="matrix")
            V = M / nrm
            V.statistics = (-1,) + tuple(V.statistics[1:])
        else:
            U = M / nrm
            L = dense([[nrm]], statistics=(-1, 1), format="matrix")
            V = dense([[1]], statistics=(-1, *sr), format="matrix")
            U.statistics = tuple(U.statistics[:-1]) + (1,)
        out = [U, L, V]
        if fmt == "standard":
            out = [x.switch_format() for x in out]
        if enc == "parity-preserving":
            out = [x.switch_encoder() for x in out]
        return tuple(out)

    inter = (0 if all(s == 0 for s in sl) else -1, 0 if all(s == 0 for s in sr) else 1)
    O = join_legs(O, js, "matrix", intermediate_stat=inter)
    O = O.force_encoder("parity-preserving")
    if O.statistics[0] == 0 or O.statistics[1] == 0:
        fn = block_fn or (sorted_svd if kind == "svd" else sorted_eig)
        U, s, V = fn(O.data, cutoff)
        S = np.diag(s)
    else:
        U, S, V = block_decompose(O.data, cutoff, kind, block_fn, rule, kept)
    ls, rs = -1, 1
    kw = dict(encoder="parity-preserving", format="matrix", skip_power_of_two_check=True)
    if O.statistics[0] == 0:
        U = dense(U, statistics=(0, 0), **kw)
        S = dense(S, statistics=(0, 0), **kw)
        V = dense(V, statistics=(0, O.statistics[1]), **kw)
        ls = rs = 0
    elif O.statistics[1] == 0:
        U = dense(U, statistics=(O.statistics[0], 0), **kw)
        S = dense(S, statistics=(0, 0), **kw)
        V = dense(V, statistics=(0, 0), **kw)
        ls = rs = 0
    else:
        U = dense(U, statistics=(O.statistics[0], 1), **kw)
        S = dense(S, statistics=(-1, 1), **kw)
        V = dense(V, statistics=(-1, O.statistics[1]), **kw)
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
    return tuple(out)


def svd(T, string, cutoff=None, block_fn=None):
    return _decompose(T, string, cutoff, "svd", block_fn)


def eig(T, string, cutoff=None, block_fn=None):
    return _decompose(T, string, cutoff, "eig", block_fn)


# --------------------------------------------------------------------------------------------------
# A13: hconjugate (dense) (__init__.py:5300-5493)
# --------------------------------------------------------------------------------------------------
def hconjugate(T, string):
    string = _parse_partition(string, "hconjugate")
    O = T.copy()
    fmt, enc = O.format, O.encoder
    xstats, xshape = O.statistics, O.shape
    left, right = [], []
    seen = False
    for ch in string:
        if ch in SEPARATORS:
            seen = True
            continue
        (right if seen else left).append(ch)
    left, right = "".join(left), "".join(right)
    nl = len(left)
    O = join_legs(O, "(" + left + ")(" + right + ")", "matrix", intermediate_stat=(-1, 1))
    O = O.force_encoder("parity-preserving")
    O.data = np.conjugate(O.data.T)
    s0, s1 = O.statistics
    O.statistics = ((-s1 if s1 in FERMI else s1), (-s0 if s0 in FERMI else s0))
    new_stats = list(xstats[nl:]) + list(xstats[:nl])
    new_stats = tuple(-s if s in FERMI else s for s in new_stats)
    new_shape = tuple(xshape[nl:]) + tuple(xshape[:nl])
    O = split_legs(O, "(" + right + ")(" + left + ")", new_stats, new_shape, intermediate_stat=(-1, 1))
    if fmt != O.format:
        O = O.switch_format()
    if enc != O.encoder:
        O = O.switch_encoder()
    return O


# --------------------------------------------------------------------------------------------------
# A14: power / sqrt on matrix-format coefficients (__init__.py:6050-6069)
# --------------------------------------------------------------------------------------------------
def power(T, p, rcond=1e-10):
    fmt, enc = T.format, T.encoder
    M = T.force_format("matrix").force_encoder("canonical")
    with np.errstate(invalid="ignore", divide="ignore"):
        M.data = np.where(np.abs(M.data) > rcond, np.power(M.data.astype(np.result_type(M.data, float)), p), 0.0)
    return M.force_format(fmt).force_encoder(enc)


def sqrt(T):
    return power(T, 0.5)


def _elementwise(T, f, rcond):
    """(|x| > rcond) ? f(x) : 0 on the matrix-format canonical coefficients (__init__.py:5992-6106)."""
    fmt, enc = T.format, T.encoder
    M = T.force_format("matrix").force_encoder("canonical")
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        M.data = np.where(np.abs(M.data) > rcond, f(M.data.astype(np.result_type(M.data, float))), 0.0)
    return M.force_format(fmt).force_encoder(enc)


def inv_nz(T, rcond=1e-10):
    """inv_nz_ds (__init__.py:5992-6004)."""
    return _elementwise(T, lambda x: 1.0 / x, rcond)


def exp(T, rcond=1e-10):
    """exp_ds (__init__.py:6093-6106)."""
    return _elementwise(T, np.exp, rcond)


def pinv(T, string, rcond=1e-10):
    """pinv (__init__.py:5961-5983): V^+ inv_nz(S^+) U^+, legs of the right group first."""
    stra, strb = string.split("|")
    strba = strb + stra
    used = set(strba)
    free = [c for c in "xyzwvutsrqponmlkjihgfedcbaXYZ" if c not in used]
    strx, stry = free[0], free[1]
    U, S, V = svd(T, string)
    Uinv = hconjugate(U, stra + "|" + strx)
    Sinv = inv_nz(hconjugate(S, strx + "|" + stry), rcond=rcond)
    Vinv = hconjugate(V, strx + "|" + strb)
    ret = einsum(strb + strx + "," + strx + stry + "->" + strb + stry, Vinv, Sinv)
    return einsum(strb + stry + "," + stry + stra + "->" + strba, ret, Uinv)


# --------------------------------------------------------------------------------------------------
# A15-A17: TRG, ATRG, logZ, zcap (gauge2d.py:1591-1885) and the driver's accumulation (example.py:162-196)
# --------------------------------------------------------------------------------------------------
def trg(T, dcut=64, error_test=False, mode="sector", block_fn=None):
    """One Levin-Nave step (gauge2d.py:1647-1755)."""
    if [T.shape[0], T.shape[1]] != [T.shape[2], T.shape[3]]:
        _fail("Error[trg]: The shape must be of the form (m,n,m,n)!")
    if list(T.statistics) != [1, 1, -1, -1]:
        _fail("Error[trg]: The statistics must be (1,1,-1,-1)!")
    T1 = einsum("ijkl->jkli", T, mode=mode)
    T2 = einsum("ijkl->klij", T, mode=mode)
    U1, S1, V1 = svd(T1, "ab|cd", dcut, block_fn=block_fn)
    U2, S2, V2 = svd(T2, "ab|cd", dcut, block_fn=block_fn)
    r = sqrt(S1)
    U1 = einsum("abx,xc->abc", U1, r, mode=mode)
    V1 = einsum("ax,xbc->abc", r, V1, mode=mode)
    r = sqrt(S2)
    U2 = einsum("abx,xc->abc", U2, r, mode=mode)
    V2 = einsum("ax,xbc->abc", r, V2, mode=mode)
    VV = einsum("kwz,lxw->lxzk", V1, V2, mode=mode)
    UU = einsum("yxi,zyj->jzxi", U1, U2, mode=mode)
    Tn = einsum("lxzk,jzxi->ijkl", VV, UU, mode=mode)
    err = None
    if error_test:
        Z1 = einsum("ijkl,klij", T, T, mode=mode)
        Z2 = einsum("ijij", Tn, mode=mode)
        err = abs(1 - Z2 / Z1)
    nrm = Tn.norm
    Tn = Tn / nrm
    return (Tn, nrm, err) if error_test else (Tn, nrm)


def atrg2dy(T1, T2, dcut=64, intermediate_dcut=None, error_test=False, mode="sector", block_fn=None):
    """Anisotropic TRG along y (gauge2d.py:1761-1869)."""
    T1o, T2o = T1, T2
    if intermediate_dcut is None:
        intermediate_dcut = dcut
    T1 = einsum("ijkl->lijk", T1, mode=mode)
    T2 = einsum("ijkl->lijk", T2, mode=mode)
    U1, S1, V1 = svd(T1, "li|jk", intermediate_dcut, block_fn=block_fn)
    U2, S2, V2 = svd(T2, "li|jk", intermediate_dcut, block_fn=block_fn)
    A = V1
    B = einsum("lia,ab->lib", U1, S1, mode=mode)
    C = einsum("ab,bjk->ajk", S2, V2, mode=mode)
    D = U2
    M = einsum("ajk,jib->aibk", C, B, mode=mode)
    U, S, V = svd(M, "ai|bk", intermediate_dcut, block_fn=block_fn)
    r = sqrt(S)
    Y = einsum("abx,xc->abc", U, r, mode=mode)
    X = einsum("ax,xbc->abc", r, V, mode=mode)
    Q1 = einsum("iax,xbj->ijab", D, Y, mode=mode)
    Q2 = einsum("kya,ylb->abkl", X, A, mode=mode)
    Q = einsum("ijab,abkl->ijkl", Q1, Q2, mode=mode)
    U, S, V = svd(Q, "ij|kl", dcut, block_fn=block_fn)
    r = sqrt(S)
    H = einsum("abx,xc->abc", U, r, mode=mode)
    G = einsum("ax,xbc->abc", r, V, mode=mode)
    H = einsum("lai->ila", H, mode=mode)
    G = einsum("kaj->ajk", G, mode=mode)
    T = einsum("ila,ajk->ijkl", H, G, mode=mode)
    err = None
    if error_test:
        Z1 = einsum("IJIK,iKiJ", T1o, T2o, mode=mode)
        Z2 = einsum("IJIJ", T, mode=mode)
        err = abs(1 - Z2 / Z1)
    nrm = T.norm
    T = T / nrm
    return (T, nrm, err) if error_test else (T, nrm)


def atrg2dx(T1, T2, dcut=64, intermediate_dcut=None, error_test=False, mode="sector", block_fn=None):
    """atrg2dy with the x and y legs exchanged (gauge2d.py:1871-1885)."""
    def swap(T):
        return einsum("jikl->jilk", einsum("ijkl->jikl", T, mode=mode), mode=mode)
    res = atrg2dy(swap(T1), swap(T2), dcut, intermediate_dcut, error_test, mode, block_fn)
    T = swap(res[0])
    return (T,) + tuple(res[1:])


def hotrg3dz(T1, T2, dcut=64, intermediate_dcut=None, error_test=False, mode="sector", block_fn=None):
    """One 3-D HOTRG step along z (gauge2d.py:1891-2145): each site tensor (i1 i2 i3 i4 m n) is split
    by two SVDs into X (x legs), P, Q and Y (y legs); the isometries Ux, Uy are the eig of the
    squared four-leg blocks (the one of the two candidates with fewer kept values, as the
    reference's S.shape[0] comparison picks), and T' = (X' Y') with legs (t1 t2 t3 t4 m n).  The
    sequence of contractions and transpositions is the reference's, so that every sign matches."""
    T1o, T2o = T1, T2
    if intermediate_dcut is None:
        intermediate_dcut = dcut
    e = lambda sub, *ops: einsum(sub, *ops, mode=mode)

    def split(T):
        T = e("i1 i2 i3 i4 mn-> i1 i3 mn i2 i4", T)
        X, S, Y = svd(T, "(i1 i3 m)(n i2 i4)", intermediate_dcut, block_fn=block_fn)
        r = sqrt(S)
        X = e("i1 i3 m a,ab->i1 i3 b m", X, r)
        Y = e("ab,b n i2 i4->n a i2 i4", r, Y)
        X, S, P = svd(X, "(i1 i3)(a m)", intermediate_dcut, block_fn=block_fn)
        X = e("i1 i3 x, xy -> i1 i3 y", X, S)
        Q, S, Y = svd(Y, "(na)(i2 i4)", intermediate_dcut, block_fn=block_fn)
        Y = e("xy, y i2 i4 -> x i2 i4", S, Y)
        return X, P, Q, Y

    X1, P1, Q1, Y1 = split(T1)
    X2, P2, Q2, Y2 = split(T2)
    XX = e("i1 i3 a, j1 j3 b -> i1 i3 ab j1 j3", X1, X2)
    XX = e("i1 i3 ab j1 j3 -> i1 i3 ab j3 j1 ", XX)
    XX = e("i1 i3 ab j3 j1 -> i1 i3 j3 ab j1 ", XX)
    XX = e("i1 i3 j3 ab j1 -> i3 j3 i1 ab j1 ", XX)
    XX = e("i3 j3 i1 ab j1 -> i3 j3 ab i1 j1 ", XX)
    c1 = hconjugate(XX, "(i3 j3 ab)(i1 j1)")
    M1 = e(" I1 J1 i3 j3 ab, i3 j3 ab i1 j1 -> I1 J1 i1 j1", c1, XX)
    U1, S1, _ = eig(M1, "(I J)(i j)", dcut, block_fn=block_fn)
    c3 = hconjugate(XX, "(i3 j3)(ab i1 j1)")
    M3 = e(" I3 J3 ab i1 j1, ab i1 j1 i3 j3  -> I3 J3 i3 j3", XX, c3)
    U3, S3, _ = eig(M3, "(I J)(i j)", dcut, block_fn=block_fn)
    Ux = U1 if S1.shape[0] < S3.shape[0] else U3
    cUx = hconjugate(Ux, "ij|a")
    XX = e("i3 j3 ab i1 j1 -> j3 i3 ab i1 j1", XX)
    XX = e("j3 i3 ab i1 j1 -> j3 i3 ab j1 i1", XX)
    Xp = e("s i3 j3,j3 i3 kl j1 i1 -> s kl j1 i1", cUx, XX)
    Xp = e("s kl j1 i1, i1 j1 t -> s kl t", Xp, Ux)
    Xp = e("s kl t -> t s kl", Xp)
    Xp = e("t s kl , kam -> t s al m", Xp, P1)
    Xp = e("t s al m , lbm -> t s ab m", Xp, P2)
    YY = e("b j2 j4, a i2 i4 -> j2 j4 ba i2 i4 ", Y2, Y1)
    YY = e("j2 j4 ba i2 i4 -> j2 j4 i4 ba i2 ", YY)
    YY = e("j2 j4 i4 ba i2 -> j4 i4 j2 ba i2 ", YY)
    YY = e("j4 i4 j2 ba i2 -> j4 i4 ba i2 j2 ", YY)
    YY = e("j4 i4 ba i2 j2 -> i4 j4 ba i2 j2 ", YY)
    c2 = hconjugate(YY, "(i4 j4 ba)(i2 j2)")
    M2 = e(" I2 J2 i4 j4 ba, i4 j4 ba i2 j2 -> I2 J2 i2 j2", c2, YY)
    U2, S2, _ = eig(M2, "(I J)(i j)", dcut, block_fn=block_fn)
    c4 = hconjugate(YY, "(i4 j4)(ba i2 j2)")
    M4 = e(" I4 J4 ba i2 j2, ba i2 j2 i4 j4  -> I4 J4 i4 j4", YY, c4)
    U4, S4, _ = eig(M4, "(I J)(i j)", dcut, block_fn=block_fn)
    Uy = U2 if S2.shape[0] < S4.shape[0] else U4
    cUy = hconjugate(Uy, "ij|a")
    YY = e("i4 j4 b a i2 j2 -> j4 i4 b a i2 j2", YY)
    YY = e("j4 i4 b a i2 j2 -> j4 i4 b a j2 i2", YY)
    Yp = e("s i4 j4,j4 i4 lk j2 i2 -> s lk j2 i2", cUy, YY)
    Yp = e("s lk j2 i2, i2 j2 t -> s lk t", Yp, Uy)
    Yp = e("s lk t -> lk t s", Yp)
    Yp = e("nak, lk t s -> n la t s", Q1, Yp)
    Yp = e("nbl, n la t s -> n ba t s", Q2, Yp)
    T = e(" t1 t3 kl m, n lk t2 t4 -> t1 t2 t3 t4 mn ", Xp, Yp)
    err = None
    if error_test:
        Z1 = e("IJIJmn,KLKLmn->mn", T1o, T2o)
        Z2 = e("IJIJmn->mn", T)
        err = (Z1 - Z2).norm / Z1.norm
    nrm = T.norm
    T = T / nrm
    return (T, nrm, err) if error_test else (T, nrm)


def logZhotrg3dz(T1, T2, boundary_conditions="periodic", mode="sector"):
    """ln of einsum('IJIJmn,KLKLmn') (gauge2d.py:1617-1641); anti-periodic signs (-1)^popcount of
    the stored index along leg J of both tensors."""
    T1, T2 = T1.copy(), T2.copy()
    if boundary_conditions == "anti-periodic":
        d = T1.shape[1]
        s = (-1) ** gparity(np.arange(d))
        T1.data = np.einsum("IJKLmn,J->IJKLmn", T1.data, s)
        T2.data = np.einsum("IJKLmn,J->IJKLmn", T2.data, s)
    Z = einsum("IJIJmn,KLKLmn", T1, T2, mode=mode)
    return np.log(Z + 0j) if np.isreal(Z) and Z < 0 else np.log(Z)


def zcap(T, mode="sector"):
    """Trace the two bosonic flavour legs with an all-ones capper (gauge2d.py:1591-1597)."""
    n = T.shape[4]
    cap = dense(np.ones((n, n)), statistics=(0, 0))
    return einsum("IJKLij,ij->IJKL", T, cap, mode=mode)


def logZ(T, boundary_conditions="periodic", mode="sector"):
    """ln of the supertrace 'IJIJ' (gauge2d.py:1599-1615); anti-periodic signs use the popcount of the
    index in T's own encoder, as the reference does."""
    T = T.copy()
    if boundary_conditions == "anti-periodic":
        d = T.shape[1]
        s = (-1) ** gparity(np.arange(d))
        T.data = np.einsum("IJKL,J->IJKL", T.data, s)
    Z = einsum("IJIJ", T, mode=mode)
    return np.log(Z + 0j) if np.isreal(Z) and Z < 0 else np.log(Z)


def trg_lnz(T, dcut, steps, mode="sector", block_fn=None, record=None):
    """ln Z per site after ``steps`` TRG steps with the driver's log-norm rule (example.py:162-196):
    logNorm = 2 logNorm + ln Tnorm, F = (lnZ(T) + logNorm) / 2^(i+1)."""
    logn = 0.0
    F = None
    for i in range(steps):
        T, nrm = trg(T, dcut, mode=mode, block_fn=block_fn)
        logn = 2 * logn + np.log(nrm)
        F = (logZ(T, mode=mode) + logn) / 2 ** (i + 1)
        if record is not None:
            record.append((float(nrm), complex(F)))
    return F, T


# --------------------------------------------------------------------------------------------------
# synthetic inputs (the generator is defined by this project; documented with the workloads)
# --------------------------------------------------------------------------------------------------
def site_tensor(D, seed, dtype=np.float64):
    """Grassmann-even 4-leg site tensor: N(0,1) * exp(-0.5 * sum of leg popcounts), odd total parity
    zeroed, scaled to max|T| = 1; legs in the parity-preserving encoder, statistics (1,1,-1,-1).
    The popcount is that of the canonical occupation of the stored index, p(encoder(i))."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((D, D, D, D))
    i = np.arange(D)
    w = np.exp(-0.5 * popcount(encoder(i)))
    par = i & 1
    mask = ((par[:, None, None, None] ^ par[None, :, None, None] ^ par[None, None, :, None]
             ^ par[None, None, None, :]) == 0)
    T = X * w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None] * w[None, None, None, :]
    T = T * mask
    T = T / np.abs(T).max()
    return dense(T.astype(dtype), encoder="parity-preserving", statistics=(1, 1, -1, -1))
