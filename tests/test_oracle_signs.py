"""CPU tests that pin the oracle's index maps and sign rules (no GPU)."""
import hashlib
import itertools
import json
import os

import numpy as np
import pytest

from oracle import gtn_oracle as O
from oracle import grassmann_algebra as G
from conftest import random_even

HERE = os.path.dirname(os.path.abspath(__file__))


def _fp(values):
    return hashlib.sha256(",".join(str(int(v)) for v in values).encode()).hexdigest()


def test_param_closed_forms_match_reference_tables():
    """gparity/encoder/sgn closed forms reproduce all 65,536 entries of param.py:4-6."""
    with open(os.path.join(HERE, "golden", "param_tables.json")) as f:
        ref = json.load(f)["tables"]
    i = np.arange(65536)
    got = {"arr_gparity": O.gparity(i), "arr_encoder": O.encoder(i), "arr_sgn": O.sgn(i)}
    for name, vals in got.items():
        assert len(vals) == ref[name]["length"]
        assert _fp(vals) == ref[name]["sha256"], name


def _encoder_from_doc(I, n):
    """Parity-preserving encoder as written in the docs (densesparse.rst.txt:72-86), bit by bit."""
    bits = [(I >> k) & 1 for k in range(n)]
    if sum(bits[1:]) % 2 == 0:
        return sum(b << k for k, b in enumerate(bits))
    return 1 - bits[0] + sum(b << k for k, b in enumerate(bits) if k >= 1)


def test_encoder_doc_formula_and_involution():
    for n in (1, 2, 3, 6, 10, 17):
        I = np.arange(min(2 ** n, 4096)) if n < 17 else np.array([65536, 70001, 131071, 99999])
        e = O.encoder(I)
        assert all(int(e[k]) == _encoder_from_doc(int(I[k]), n) for k in range(len(I)))
        assert np.array_equal(O.encoder(e), I)
        assert np.array_equal(e & 1, O.gparity(I) & 1)


def test_sgn_is_reordering_sign():
    """sigma_I is the sign of psi^I psibar^I -> prod (psi psibar) (the S2 factor, einsum.rst.txt:83-95)."""
    for I in range(64):
        bits = [k for k in range(6) if (I >> k) & 1]
        seq = [(k, 0) for k in bits] + [(k, 1) for k in bits]
        tgt = sum([[(k, 0), (k, 1)] for k in bits], [])
        assert O.perm_sign(seq, tgt, {x: 1 for x in seq}) == int(O.sgn(I))


CASES = [
    ("ilk,jl->ijk", [(1, 1, -1), (-1, -1)], [(2, 4, 2), (2, 4)]),
    ("ij,jk->ik", [(1, 1), (-1, -1)], [(4, 2), (2, 4)]),
    ("ijk->kji", [(1, -1, 1)], [(2, 4, 2)]),
    ("ijkl->jkli", [(1, 1, -1, -1)], [(2, 2, 4, 2)]),
    ("ijkl->klij", [(1, 1, -1, -1)], [(2, 4, 2, 2)]),
    ("ijij", [(1, 1, -1, -1)], [(2, 4, 2, 4)]),
    ("ijkl,klij", [(1, 1, -1, -1), (1, 1, -1, -1)], [(2, 2, 2, 2), (2, 2, 2, 2)]),
    ("abc,cd->bda", [(1, -1, 1), (-1, 1)], [(2, 4, 2), (2, 2)]),
    ("kwz,lxw->lxzk", [(-1, -1, 1), (-1, 1, 1)], [(2, 2, 2), (2, 2, 2)]),
    ("yxi,zyj->jzxi", [(1, -1, 1), (-1, -1, 1)], [(2, 2, 2), (2, 2, 2)]),
    ("lxzk,jzxi->ijkl", [(-1, 1, 1, -1), (1, -1, -1, 1)], [(2, 2, 2, 2), (2, 2, 2, 2)]),
    ("abx,xc->abc", [(1, -1, 1), (-1, 1)], [(2, 2, 4), (4, 2)]),
    ("ajk,jib->aibk", [(1, 1, -1), (-1, -1, 1)], [(2, 2, 2), (2, 2, 2)]),
    ("ab,bc,cd->ad", [(1, 1), (-1, 1), (-1, 1)], [(2, 2), (2, 2), (2, 2)]),
    ("aIb,cId->dcab", [(1, -1, -1), (1, 1, 1)], [(2, 2, 4), (2, 2, 2)]),
]


@pytest.mark.parametrize("sub,stats,shapes", CASES)
@pytest.mark.parametrize("mode", ["literal", "sector"])
def test_einsum_matches_grassmann_algebra(sub, stats, shapes, mode):
    """Odd and even random operands: the oracle's sign rule equals first-principles integration."""
    rng = np.random.default_rng(abs(hash(sub)) % 2 ** 32)
    Ts = [rng.standard_normal(s) for s in shapes]
    ref = G.contract(sub, Ts, stats)
    r = O.einsum(sub, *[O.dense(T, statistics=st) for T, st in zip(Ts, stats)], mode=mode)
    if isinstance(ref, tuple):
        assert r.statistics == ref[1]
        np.testing.assert_allclose(r.data, ref[0], rtol=0, atol=1e-13)
    else:
        assert abs(r - ref) < 1e-13


def test_einsum_doc_closed_form():
    """'ilk,jl->ijk': s_JKL = (-)^{p(L)(p(J)+p(K))+p(J)p(K)} sigma_L (einsum.rst.txt:51-60)."""
    rng = np.random.default_rng(7)
    A = rng.standard_normal((4, 8, 4))
    B = rng.standard_normal((8, 8))
    p = lambda x: O.parity_of_canonical(np.arange(x))
    pJ, pK, pL = p(8)[:, None, None], p(4)[None, :, None], p(8)[None, None, :]
    s = (-1) ** (pL * (pJ + pK) + pJ * pK) * O.sgn(np.arange(8))[None, None, :]
    ref = np.einsum("ilk,jl,jkl->ijk", A, B, s)
    r = O.einsum("ilk,jl->ijk", O.dense(A, statistics=(1, 1, -1)), O.dense(B, statistics=(-1, -1)))
    np.testing.assert_allclose(r.data, ref, atol=1e-13)


def test_literal_vs_sector_with_bosons_and_encoders(rng):
    A = random_even(rng, (4, 3, 8, 4), (1, 0, -1, 1), "parity-preserving")
    B = random_even(rng, (4, 3, 2, 8), (-1, 0, 1, 1))
    a = O.einsum("iIaj,jIbk->akbi", A, B, mode="literal")
    b = O.einsum("iIaj,jIbk->akbi", A, B, mode="sector")
    assert a.encoder == "parity-preserving" and a.statistics == (-1, 1, 1, 1)
    np.testing.assert_allclose(a.data, b.data, atol=1e-12)


def test_pairwise_associativity(rng):
    A = O.dense(rng.standard_normal((2, 4, 2)), statistics=(1, -1, 1))
    B = O.dense(rng.standard_normal((4, 2, 2)), statistics=(1, -1, 1))
    C = O.dense(rng.standard_normal((2, 2, 4)), statistics=(-1, -1, -1))
    full = O.einsum("abc,bde,cef->adf", A, B, C, mode="literal")
    AB = O.einsum("abc,bde->acde", A, B, mode="literal")
    np.testing.assert_allclose(O.einsum("acde,cef->adf", AB, C).data, full.data, atol=1e-13)


def test_sign_tensor_windows():
    """The S1/S3 windows of the T' contraction (Appendix-A derivation of einsum_ds)."""
    plan = O.EinsumPlan("lxzk,jzxi->ijkl", [-1, 1, 1, -1, 1, -1, -1, 1], [4] * 8)
    assert plan.s1_chars == ["z", "k", "j", "x"]
    assert plan.s3_chars == ["l", "k", "j", "i"]
    S1, _ = plan.sign_tensor(1)
    p = O.parity_of_canonical(np.arange(4))
    pz, pk, pj, px = np.ix_(p, p, p, p)
    np.testing.assert_array_equal(S1, (-1) ** ((px + pz) * (pk + pj)))


def test_errors_match_reference():
    A = O.dense(np.zeros((2, 2)), statistics=(1, 1))
    with pytest.raises(NameError):
        O.einsum("ij,jk->ik", A, A)
    with pytest.raises(NameError):
        O.dense(np.zeros((3, 2)), statistics=(1, 1))
    with pytest.raises(NameError):
        O.svd(O.dense(np.ones((2, 2)), statistics=(1, -1)), "ab")
