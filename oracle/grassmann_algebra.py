"""Brute-force exact Grassmann algebra -- TEST INFRASTRUCTURE ONLY (checker for the oracle's signs).

Only ``tests/`` may import this module.  It evaluates a Grassmann contraction from first principles,
independently of the string algorithm that einsum uses, so that a wrong sign rule in the oracle fails:

  * a leg of dimension 2^b carries b Grassmann generators; the canonical index I = sum_k 2^k i_k
    stands for the ordered monomial  psi_(0)^{i_0} psi_(1)^{i_1} ... psi_(b-1)^{i_(b-1)}
    (docs/_sources/grassmanntn/densesparse.rst.txt:61-100); a -1 leg uses barred generators;
  * a tensor is the polynomial sum_I T_I (leg monomials in leg order); operands multiply in the
    order they are written;
  * a contracted pair (psi, psibar) is integrated with  int dpsibar dpsi exp(-psibar psi)  per bit,
    i.e. <1> = 1, <psi psibar> = 1, <psi> = <psibar> = 0
    (docs/_sources/grassmanntn/einsum.rst.txt:27-107);
  * the result's coefficients are read off after reordering each monomial into the output's leg
    order.

Only fermionic legs are supported, in the standard format.  Cost grows as the product of the
operands' sizes, so keep legs to 1-2 bits.
"""

import itertools
import numpy as np


def _bits(d):
    b = int(round(np.log2(d)))
    if 2 ** b != d:
        raise ValueError("fermionic leg dimension must be a power of two")
    return b


def _monomial(chars, stats, idx, shape):
    gens = []
    for c, s, i, d in zip(chars, stats, idx, shape):
        bar = 0 if s == 1 else 1
        for k in range(_bits(d)):
            if (i >> k) & 1:
                gens.append((c, k, bar))
    return gens


def _integrate(gens, pairs):
    """Integrate every contracted character; returns (sign, remaining) or (0, None)."""
    sign = 1
    g = list(gens)
    for c, nbits in pairs:
        for k in range(nbits):
            a, b = (c, k, 0), (c, k, 1)
            ha, hb = a in g, b in g
            if ha != hb:
                return 0, None
            if ha:
                i = g.index(a)
                sign *= -1 if i % 2 else 1
                g.pop(i)
                j = g.index(b)
                sign *= -1 if j % 2 else 1
                g.pop(j)
    return sign, g


def _reorder_sign(gens, key):
    order = sorted(range(len(gens)), key=lambda t: key(gens[t]))
    inv = 0
    for x in range(len(order)):
        for y in range(x + 1, len(order)):
            if order[x] > order[y]:
                inv += 1
    return -1 if inv % 2 else 1


def contract(subscripts, tensors, stats):
    """Evaluate a fermionic einsum from first principles.

    subscripts: e.g. 'ilk,jl->ijk' (no spaces, single-character indices, output required unless it
    is a full contraction); tensors: numpy arrays in the canonical encoder and standard format;
    stats: per-operand statistics tuples (+1 / -1 only).
    """
    if "->" in subscripts:
        inp, out = subscripts.split("->")
    else:
        inp, out = subscripts, ""
    ops = inp.split(",")
    dim = {}
    stat_of = {}
    for s, T, st in zip(ops, tensors, stats):
        for c, d, t in zip(s, T.shape, st):
            dim[c] = d
            stat_of.setdefault(c, []).append(t)
    summed = [c for c in dim if c not in out]
    pairs = [(c, _bits(dim[c])) for c in summed]
    out_stats = [stat_of[c][0] for c in out]
    out_pos = {c: k for k, c in enumerate(out)}

    def key(g):
        return (out_pos[g[0]], g[1])

    terms = []
    for s, T, st in zip(ops, tensors, stats):
        lst = []
        for idx in itertools.product(*[range(d) for d in T.shape]):
            v = T[idx]
            if v != 0:
                lst.append((v, _monomial(s, st, idx, T.shape)))
        terms.append(lst)

    res = np.zeros(tuple(dim[c] for c in out), dtype=np.result_type(*tensors, np.float64))
    for combo in itertools.product(*terms):
        coef = 1.0
        gens = []
        for v, g in combo:
            coef = coef * v
            gens = gens + g
        sgn, rest = _integrate(gens, pairs)
        if sgn == 0:
            continue
        sg2 = _reorder_sign(rest, key)
        idx = [0] * len(out)
        for (c, k, bar) in rest:
            idx[out_pos[c]] |= 1 << k
        res[tuple(idx)] += sgn * sg2 * coef
    if not out:
        return res[()]
    return res, tuple(out_stats)
