"""Pinning of the block-format oracle (oracle/gtn_oracle_block.py) against the dense oracle, the param
tables and the linear algebra the block operations must satisfy."""
import numpy as np
import pytest

from oracle import gtn_oracle as O
from oracle import gtn_oracle_block as OB
from conftest import random_even
from cases import EINSUMS

def test_cells_and_sign_vectors():
    """Each cell is the parity-preserving slice [p::2] and each sign vector is sgn(encoder(2i+p)) (the
    closed forms are checked against the digests of the reference's literal tables, param.py:4-6, in
    test_oracle_signs)."""
    rng = np.random.default_rng(0)
    T = random_even(rng, (8, 4, 2, 16), (1, -1, 0, 1), "parity-preserving")
    B = OB.block.from_dense(T)
    for a, d in enumerate(T.shape):
        if T.statistics[a] == 0:
            continue
        for p in (0, 1):
            i = np.arange(p, d, 2)
            assert np.array_equal(B.sgn[p][a], O.sgn(O.encoder(i)))
    for par, c in B.cells.items():
        sl = (slice(par[0], None, 2), slice(par[1], None, 2), slice(None), slice(par[2], None, 2))
        assert np.array_equal(c, T.data[sl])


def test_sign_vectors_closed_form():
    for p in (0, 1):
        i = 2 * np.arange(64) + p
        v = OB._sign_vector(p, 64)
        pc = np.array([bin(int(O.encoder(np.array([x]))[0])).count("1") for x in i])
        assert np.array_equal(v, np.where((pc % 4) < 2, 1, -1))


@pytest.mark.parametrize("enc", ["canonical", "parity-preserving"])
def test_round_trip(enc):
    rng = np.random.default_rng(1)
    T = random_even(rng, (8, 4, 3, 16), (1, -1, 0, -1), enc)
    B = OB.block.from_dense(T)
    assert B.even_shape == (4, 2, 3, 8) and B.odd_shape == (4, 2, 3, 8)
    assert B.effective_shape == (8, 4, 3, 16)
    R = B.todense(enc)
    assert R.encoder == enc
    assert np.array_equal(R.data, T.data)
    assert abs(B.norm - T.norm) < 1e-12 * T.norm


def test_leg_of_dimension_one():
    T = O.dense(np.arange(4.0).reshape(1, 4) * [[1, 0, 0, 1]], statistics=(1, 1))
    B = OB.block.from_dense(T)
    assert B.even_shape == (1, 2) and B.odd_shape == (0, 2)


@pytest.mark.parametrize("op", ["switch_format", "switch_parity"])
def test_switches_match_dense(op):
    rng = np.random.default_rng(2)
    T = random_even(rng, (8, 4, 3, 16), (1, -1, 0, -1), "parity-preserving")
    B = OB.block.from_dense(T)
    R = getattr(B, op)().todense("parity-preserving")
    ref = getattr(T, op)()
    assert R.format == ref.format
    assert np.array_equal(R.data, ref.data)


@pytest.mark.parametrize("sub,stats,shapes", EINSUMS)
def test_einsum_block_matches_dense(sub, stats, shapes):
    """einsum_block's per-parity loop with the operands' sign vectors against einsum_ds semantics."""
    rng = np.random.default_rng(abs(hash(sub)) % 2 ** 32)
    ops = [random_even(rng, sh, st, "parity-preserving") for st, sh in zip(stats, shapes)]
    ref = O.einsum(sub, *ops)
    got = OB.einsum_block(sub, *[OB.block.from_dense(T) for T in ops])
    if isinstance(ref, O.dense):
        assert got.statistics == ref.statistics
        np.testing.assert_allclose(got.todense("parity-preserving").data, ref.force_encoder("parity-preserving").data,
                                   rtol=0, atol=1e-12 * max(1.0, np.abs(ref.data).max()))
    else:
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))


def test_einsum_block_planted_sign_error_is_caught():
    """A dropped sigma vector (planted in a copy of the oracle's inputs) changes the result."""
    rng = np.random.default_rng(7)
    A = random_even(rng, (4, 8, 4), (1, 1, -1), "parity-preserving")
    Bt = random_even(rng, (8, 8), (-1, -1), "parity-preserving")
    bA, bB = OB.block.from_dense(A), OB.block.from_dense(Bt)
    good = OB.einsum_block("ilk,jl->ijk", bA, bB).todense().data
    bad = bB.copy()
    bad.sgn = [[np.abs(v) for v in s] for s in bad.sgn]
    bA2 = bA.copy()
    bA2.sgn = [[np.abs(v) for v in s] for s in bA2.sgn]
    wrong = OB.einsum_block("ilk,jl->ijk", bA2, bad).todense().data
    assert np.abs(good - wrong).max() > 1e-3


@pytest.mark.parametrize("cut", [None, 7, 8, 25])
def test_svd_block_truncation_and_reconstruction(cut):
    rng = np.random.default_rng(3)
    T = random_even(rng, (8, 8, 8, 8), (1, 1, -1, -1), "parity-preserving")
    B = OB.block.from_dense(T)
    U, S, V = OB.svd_block(B, "ab|cd", cut)
    ME, MO = O.parity_blocks(O.join_legs(T, "(ab)(cd)", "matrix", intermediate_stat=(-1, 1)).data)
    sE = np.linalg.svd(ME, compute_uv=False)
    sO = np.linalg.svd(MO, compute_uv=False)
    kE = 32 if cut is None else min(32, -(-cut // 2))
    kO = 32 if cut is None else min(32, cut // 2)
    n = max(kE, kO)
    assert U.even_shape[-1] == n and U.odd_shape[-1] == n
    assert S.even_shape == (n, n) and V.even_shape[0] == n
    nfull = 2 ** int(1 + np.ceil(np.log2(n)))
    assert U.shape[-1] == nfull and S.shape == (nfull, nfull)
    Sm = S.force_format("matrix")
    dE = np.diag(Sm.cells[(0, 0)])
    dO = np.diag(Sm.cells[(1, 1)])
    np.testing.assert_allclose(dE[:kE], sE[:kE], rtol=1e-12)
    np.testing.assert_allclose(dO[:kO], sO[:kO], rtol=1e-12)
    assert np.all(dE[kE:] == 0) and np.all(dO[kO:] == 0)
    R = OB.einsum_block("abx,xy,ycd->abcd", U, S, V)
    if cut is None:
        np.testing.assert_allclose(R.todense().data, T.force_encoder("canonical").data, atol=1e-12)
    else:
        # the truncation error is the discarded singular values
        err = np.linalg.norm(R.todense().data - T.force_encoder("canonical").data)
        np.testing.assert_allclose(err, np.sqrt(np.sum(sE[kE:] ** 2) + np.sum(sO[kO:] ** 2)), rtol=1e-9)


def test_eig_and_hconjugate_block():
    rng = np.random.default_rng(4)
    A = OB.block.from_dense(random_even(rng, (4, 4, 4), (1, 1, -1), "parity-preserving"))
    cA = OB.hconjugate_block(A, "i|jk")
    assert cA.statistics == (-1, 1, -1)
    back = OB.hconjugate_block(cA, "ij|k")
    np.testing.assert_allclose(back.todense().data, A.todense().data, atol=1e-14)
    H = OB.einsum_block("jki,iJK->jkJK", cA, A)
    U, L, V = OB.eig_block(H, "jk|JK")
    R = OB.einsum_block("jka,ab,bJK->jkJK", U, L, V)
    np.testing.assert_allclose(R.todense().data, H.todense().data, atol=1e-12)
    Vc = OB.hconjugate_block(U, "jk|a")
    np.testing.assert_allclose(Vc.todense().data, V.todense().data, atol=1e-12)


def test_power_block_is_a_real_square_root():
    rng = np.random.default_rng(5)
    T = random_even(rng, (8, 8, 8, 8), (1, 1, -1, -1), "parity-preserving")
    _, S, _ = OB.svd_block(OB.block.from_dense(T), "ab|cd", 8)
    r = OB.sqrt_block(S)
    S2 = OB.einsum_block("ab,bc->ac", r, r)
    np.testing.assert_allclose(S2.todense().data, S.todense().data, atol=1e-12)


@pytest.mark.parametrize("dcut", [8, 16])
def test_block_trg_equals_dense_trg_at_power_of_two_bond(dcut):
    """docs/_sources/schwinger.rst.txt:920-922: dense and block TRG give the same norm."""
    T = O.site_tensor(8, 2)
    Td, nd = O.trg(T, dcut)
    Tb, nb = OB.trg_block(OB.block.from_dense(T), dcut)
    assert abs(nd - nb) <= 1e-12 * nd
    assert abs(O.logZ(Td) - OB.logZ_block(Tb)) < 1e-12


def test_block_trg_odd_bond_shapes():
    """Odd bond: ceil/floor halves, both parities padded to the larger count (__init__.py:5076-5100)."""
    T = O.site_tensor(8, 2)
    Tb, nb = OB.trg_block(OB.block.from_dense(T), 7)
    assert Tb.even_shape == (4, 4, 4, 4) and Tb.odd_shape == (4, 4, 4, 4)
    assert Tb.shape == (8, 8, 8, 8)
    # the same numbers as keeping 4 even / 3 odd values in the dense pipeline
    Td, nd = O.trg(T, 8, block_fn=_keep(4, 3))
    assert abs(nd - nb) <= 1e-12 * nd


def _keep(kE, kO):
    state = {"n": 0}

    def fn(M, cut):
        k = kE if state["n"] % 2 == 0 else kO
        state["n"] += 1
        return O.sorted_svd(M, k)
    return fn
