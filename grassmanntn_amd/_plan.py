"""Host-side planning of Grassmann contractions: subscript parsing, the sign rule, and its
factorisation onto operands and output.

The sign of a Grassmann contraction depends only on the parities of the fermionic indices
(einsum_ds, __init__.py:1633-2126).  It is the product of
  s1  -- the reordering that places every conjugated summed index right after its partner
         (__init__.py:1791-1999),
  s2  -- sigma(i) of each contracted index (__init__.py:2010-2039), and
  s3  -- the reordering of the remaining fermions into the output order (__init__.py:2045-2126).
s1 * s3 = (-1)^q(p) with q a quadratic form over GF(2) in the parities p of the distinct fermionic
characters.  Every monomial of q involves characters that live together on one operand or on the
output, so q splits into q_A + q_B + q_out; each part becomes a sign table that the device applies
while it gathers that tensor (the kernel computes the parities by popcount).
"""
import itertools

HYBRID = "*"
SEPARATORS = ("|", ":", ";", ",", ".")
DIGITS = "0123456789"
FERMI = (1, -1)
BOSE = (0, HYBRID)
CHARS = tuple("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ") + tuple(
    "αβΓγΔδεζηΘθικλμνΞξΠπρΣσςτυΦϕφχΨψΩω")


class GrassmannError(NameError):
    """Errors of the reference's error() family (__init__.py:57-63 raises NameError)."""


def fail(msg):
    raise GrassmannError(msg)


def fresh_char(used):
    for c in CHARS:
        if c not in used:
            return c
    fail("Error[get_char]: Running out of index character!")


def denumerate(string):
    """Multi-character numbered indices ('i12') become single fresh characters (__init__.py:1610-1631)."""
    toks = []
    for ch in string:
        if ch in DIGITS and toks:
            toks[-1] += ch
        else:
            toks.append(ch)
    long_toks = sorted({t for t in toks if len(t) > 1}, key=len, reverse=True)
    used = string
    subs = []
    for t in long_toks:
        c = fresh_char(used)
        used += c
        subs.append((t, c))
    for t, c in subs:
        string = string.replace(t, c)
    return string


def odd_inversions(seq_from, seq_to, odd):
    """Parity (0/1) of the number of pairs of odd elements whose order differs between the two
    sequences (elements present once in seq_from; duplicates dropped as relative_sign does,
    __init__.py:1580-1590)."""
    src = [c for c in seq_from if seq_from.count(c) == 1]
    pos = {}
    k = 0
    for c in seq_to:
        if c in src and c not in pos:
            pos[c] = k
            k += 1
    o = [pos[c] for c in src if odd.get(c, 0)]
    inv = 0
    for a in range(len(o)):
        for b in range(a + 1, len(o)):
            inv ^= o[a] > o[b]
    return inv


class Contraction:
    """Parsed einsum with its sign rule.

    operands: list of (index string, statistics tuple, shape tuple)
    """

    def __init__(self, subscripts, operands):
        sub = denumerate(subscripts.replace(" ", ""))
        if sub.count("->") > 1:
            fail("Error[einsum]: Only one arrow is allowed in the input string!")
        self.has_output = "->" in sub
        inp, out = sub.split("->") if self.has_output else (sub, "")
        if "," in out:
            fail("Error[einsum]: Output string must not contain commas (',') !")
        self.inputs = inp.split(",")
        self.output = out
        if len(self.inputs) != len(operands):
            fail("Error[einsum]: %d operand strings for %d operands" % (len(self.inputs), len(operands)))
        stats, dims = [], {}
        for s, (st, shp) in zip(self.inputs, operands):
            if len(s) != len(st):
                fail("Error[einsum]: subscript '%s' does not match an operand of rank %d" % (s, len(st)))
            for c, t, d in zip(s, st, shp):
                if t == HYBRID:
                    raise NotImplementedError("einsum over hybrid ('*') indices is not supported")
                if c in dims and dims[c] != d:
                    fail("Error[einsum]: inconsistent dimensions for index '%s'" % c)
                dims[c] = d
            stats += list(st)
        self.dims = dims
        summand = inp.replace(",", "")
        self.stat_of = {}
        for c, t in zip(summand, stats):
            self.stat_of.setdefault(c, []).append(t)
        for c in out:
            if c not in dims:
                fail("Error[einsum]: output index '%s' is not in the input" % c)
        # statistics checks (__init__.py:1691-1741)
        for i, c in enumerate(summand):
            if stats[i] in BOSE:
                continue
            lf = inp.count(c)
            if self.has_output:
                rf = out.count(c)
                if lf > 2 or rf > 1 or (lf + rf) % 2 == 1:
                    fail("Error[einsum]: Inconsistent index statistics.")
            elif lf > 2:
                fail("Error[einsum]: Inconsistent index statistics.")
        for c, st in self.stat_of.items():
            fs = [t for t in st if t in FERMI]
            if len(fs) == 2 and sorted(fs) != [-1, 1]:
                fail("Error[einsum]: The contracted indices have inconsistent statistics!")
        self.fermionic = [c for c in dict.fromkeys(summand) if any(t in FERMI for t in self.stat_of[c])]
        # fermion string with each conjugated summed index renamed (step 1 of einsum_ds)
        used = summand
        seq, partner = [], {}
        for c, t in zip(summand, stats):
            if t == 0:
                continue
            if t == -1 and summand.count(c) > 1:
                n = fresh_char(used)
                used += n
                partner[c] = n
                seq.append(n)
            else:
                seq.append(c)
        self.pairs = [c for c in dict.fromkeys(summand) if c in partner]
        self._partner = partner
        self._orig = {v: k for k, v in partner.items()}
        moved = [c for c in seq if c not in self._orig]
        for c in self.pairs:
            moved.insert(moved.index(c) + 1, partner[c])
        self._seq, self._moved = seq, moved
        self._rest = [c for c in moved if c not in partner and c not in self._orig]
        self._fout = [c for c in out if self.stat_of[c][0] != 0]
        self.out_stats = tuple(self.stat_of[c][0] for c in out)

    def sign_bit(self, par):
        """0/1 exponent of s1*s3 for parities `par` (char -> 0/1) of the fermionic characters."""
        odd = {c: par.get(self._orig.get(c, c), 0) for c in self._seq}
        b = odd_inversions(self._seq, self._moved, odd)
        if self.has_output:
            b ^= odd_inversions(self._rest, self._fout, {c: par.get(c, 0) for c in self._rest})
        return b

    def quadratic_form(self):
        """(linear, quadratic) coefficients over GF(2) of the sign exponent, keyed by characters
        (computed once per contraction)."""
        qf = getattr(self, "_qf", None)
        if qf is not None:
            return qf
        self._qf = self._quadratic_form()
        return self._qf

    def _quadratic_form(self):
        ch = self.fermionic
        z = {c: 0 for c in ch}
        if self.sign_bit(z):
            raise AssertionError("sign rule is not +1 at all-even parities")
        lin = {}
        for c in ch:
            lin[c] = self.sign_bit(dict(z, **{c: 1}))
        quad = {}
        for a, b in itertools.combinations(ch, 2):
            v = self.sign_bit(dict(z, **{a: 1, b: 1})) ^ lin[a] ^ lin[b]
            if v:
                quad[(a, b)] = 1
        if len(ch) <= 10:  # confirm the form reproduces the rule on every parity pattern
            for bits in itertools.product((0, 1), repeat=len(ch)):
                par = dict(zip(ch, bits))
                q = 0
                for c in ch:
                    q ^= lin[c] & par[c]
                for (a, b) in quad:
                    q ^= par[a] & par[b]
                if q != self.sign_bit(par):
                    raise AssertionError("sign rule is not a quadratic form")
        return lin, quad

    def split_form(self, groups):
        """Distribute the quadratic form over `groups` (list of character sets, e.g. [A, B, out]);
        each monomial goes to the first group that holds all of its characters."""
        lin, quad = self.quadratic_form()
        parts = [({}, {}) for _ in groups]
        for c, v in lin.items():
            if not v:
                continue
            for gi, g in enumerate(groups):
                if c in g:
                    parts[gi][0][c] = 1
                    break
            else:
                raise AssertionError("character %s belongs to no group" % c)
        for (a, b) in quad:
            for gi, g in enumerate(groups):
                if a in g and b in g:
                    parts[gi][1][(a, b)] = 1
                    break
            else:
                raise AssertionError("monomial %s%s spans groups" % (a, b))
        return parts


_PLANS = {}


def contraction(subscripts, operands):
    """Contraction(subscripts, operands), shared between calls with the same subscripts, statistics and
    shapes (a TRG step repeats the same handful of contractions every step)."""
    key = (subscripts, tuple((tuple(st), tuple(shp)) for st, shp in operands))
    plan = _PLANS.get(key)
    if plan is None:
        if len(_PLANS) > 4096:
            _PLANS.clear()
        plan = _PLANS[key] = Contraction(subscripts, operands)
    return plan


_TABLES = {}


def sign_table(chars, lin, quad, zero_if_odd=False):
    """2-bit table over the parity bits of `chars` (bit i <-> chars[i]): 0 -> +1, 1 -> -1, 2 -> 0.
    zero_if_odd marks patterns of odd total parity as zero entries."""
    key = (tuple(chars), tuple(sorted(lin.items())), tuple(sorted(quad.items())), bool(zero_if_odd))
    codes = _TABLES.get(key)
    if codes is None:
        if len(_TABLES) > 4096:
            _TABLES.clear()
        codes = _TABLES[key] = _sign_table(chars, lin, quad, zero_if_odd)
    return list(codes)


def _sign_table(chars, lin, quad, zero_if_odd):
    n = len(chars)
    if n > 10:
        raise NotImplementedError("more than 10 fermionic indices on one tensor")
    codes = []
    for m in range(1 << n):
        par = {c: (m >> i) & 1 for i, c in enumerate(chars)}
        if zero_if_odd and (bin(m).count("1") & 1):
            codes.append(2)
            continue
        q = 0
        for c, v in lin.items():
            q ^= v & par.get(c, 0)
        for (a, b) in quad:
            q ^= par.get(a, 0) & par.get(b, 0)
        codes.append(q)
    return codes


def parse_partition(string, name):
    """'(ab)(cd)' or 'ab|cd' -> 'ab|cd' (svd / eig / hconjugate argument, __init__.py:4046-4099)."""
    string = denumerate(string.replace(" ", ""))
    if string.count("(") == string.count(")") and string.count("(") > 0:
        string = string.replace(")(", "|")
        if string.count("(") > 1 or string.count(")") < 1:
            fail("Error[%s]: Parentheses don't match" % name)
        string = string.replace(")", "").replace("(", "")
    if sum(string.count(p) for p in SEPARATORS) != 1:
        fail("Error[%s]: The input string must contain one and only one partition ( '|', ':', ';', ',', or '.' )" % name)
    left, right, seen = "", "", False
    for ch in string:
        if ch in SEPARATORS:
            seen = True
        elif seen:
            right += ch
        else:
            left += ch
    return left, right


def group_info(string, stats, shape):
    """Groups of a join/split string, with bosons moved to the left inside each group
    (__init__.py:3172-3243).  Returns (groups, sorted_groups), each a list of
    (chars, stats, dims, positions)."""
    s = string
    if "(" not in s and ")" not in s:
        for sep in SEPARATORS:
            s = s.replace(sep, ")(")
        s = "(" + s + ")"
    elif any(sep in s for sep in SEPARATORS):
        fail("Error[get_grouping_info]: Do not mix the string format.")
    groups, inside, loc = [], False, 0
    for ch in s:
        if ch == "(" and not inside:
            inside = True
            groups.append([[], [], [], []])
            continue
        if ch == ")" and inside:
            inside = False
            continue
        if ch in "()":
            fail("Error[get_grouping_info]: No nested parenthesis allowed!")
        if loc >= len(stats):
            fail("Error[join_legs]: The number of indices is not consistent with the object's shape.")
        if not inside:
            groups.append([[ch], [stats[loc]], [shape[loc]], [loc]])
        else:
            g = groups[-1]
            g[0].append(ch)
            g[1].append(stats[loc])
            g[2].append(shape[loc])
            g[3].append(loc)
        loc += 1
    if loc != len(stats):
        fail("Error[join_legs]: The number of indices is not consistent with the object's shape.")
    srt = []
    for chars, st, dm, pos in groups:
        items = list(zip(chars, st, dm, pos))
        bos = [x for x in items if x[1] in BOSE][::-1]
        fer = [x for x in items if x[1] not in BOSE]
        items = bos + fer
        srt.append([[x[0] for x in items], [x[1] for x in items], [x[2] for x in items], [x[3] for x in items]])
    return groups, srt
