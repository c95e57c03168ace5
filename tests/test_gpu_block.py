"""GPU parity tests of the block format (einsum_block, svd_block / eig_block, hconjugate_block, block
TRG) against the block oracle."""
import numpy as np
import pytest

from oracle import gtn_oracle as O
from oracle import gtn_oracle_block as OB
from conftest import random_even
from cases import EINSUMS

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")


def to_dev_block(T):
    return gtn.dense(T.data, encoder=T.encoder, format=T.format, statistics=T.statistics).toblock()


def cells_equal(got, ref, atol):
    """Same extents, statistics, physical shape and cells."""
    assert got.statistics == ref.statistics
    assert got.even_shape == ref.even_shape and got.odd_shape == ref.odd_shape
    assert tuple(got.shape) == tuple(ref.shape)
    cells = got.data
    for par, c in ref.cells.items():
        np.testing.assert_allclose(cells[par], c, rtol=0, atol=atol)


def test_block_conversion_and_switches():
    rng = np.random.default_rng(2)
    T = random_even(rng, (8, 4, 3, 16), (1, -1, 0, -1), "parity-preserving")
    ref = OB.block.from_dense(T)
    got = to_dev_block(T)
    cells_equal(got, ref, 0.0)
    for op in ("switch_format", "switch_parity"):
        cells_equal(getattr(got, op)(), getattr(ref, op)(), 0.0)
    assert np.array_equal(got.todense("canonical").data.numpy(), ref.todense("canonical").data)
    for p in (0, 1):
        for a in range(4):
            assert np.array_equal(got.sgn[p][a], ref.sgn[p][a])


@pytest.mark.parametrize("sub,stats,shapes", EINSUMS)
def test_einsum_block_matches_oracle(sub, stats, shapes):
    rng = np.random.default_rng(abs(hash(sub)) % 2 ** 32)
    ops = [random_even(rng, sh, st, "parity-preserving") for st, sh in zip(stats, shapes)]
    ref = OB.einsum_block(sub, *[OB.block.from_dense(T) for T in ops])
    got = gtn.einsum(sub, *[to_dev_block(T) for T in ops])
    if isinstance(ref, OB.block):
        cells_equal(got, ref, 1e-12 * max(1.0, max(np.abs(c).max() for c in ref.cells.values())))
    else:
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))


@pytest.mark.parametrize("cut", [None, 7, 8, 25])
def test_svd_block_matches_oracle(cut):
    rng = np.random.default_rng(3)
    T = random_even(rng, (8, 8, 8, 8), (1, 1, -1, -1), "parity-preserving")
    U, S, V = to_dev_block(T).svd("ab|cd", cut)
    Uo, So, Vo = OB.svd_block(OB.block.from_dense(T), "ab|cd", cut)
    for g, r in ((U, Uo), (S, So), (V, Vo)):
        assert g.even_shape == r.even_shape and g.odd_shape == r.odd_shape and tuple(g.shape) == tuple(r.shape)
    cells_equal(S, So, 1e-10)
    R = gtn.einsum("abx,xy,ycd->abcd", U, S, V)
    Ro = OB.einsum_block("abx,xy,ycd->abcd", Uo, So, Vo)
    cells_equal(R, Ro, 1e-9)


def test_eig_and_hconjugate_block():
    rng = np.random.default_rng(4)
    A = random_even(rng, (4, 4, 4), (1, 1, -1), "parity-preserving")
    dA, oA = to_dev_block(A), OB.block.from_dense(A)
    cA = dA.hconjugate("i|jk")
    cells_equal(cA, OB.hconjugate_block(oA, "i|jk"), 0.0)
    H = gtn.einsum("jki,iJK->jkJK", cA, dA)
    U, L, V = H.eig("jk|JK")
    R = gtn.einsum("jka,ab,bJK->jkJK", U, L, V)
    np.testing.assert_allclose(R.todense().data.numpy(), H.todense().data.numpy(), atol=1e-11)


@pytest.mark.parametrize("dcut", [7, 8])
def test_block_trg_matches_oracle(dcut):
    T = O.site_tensor(8, 2)
    ref, nref = OB.trg_block(OB.block.from_dense(T), dcut)
    got, nrm = gtn.trg(to_dev_block(T), dcut)
    assert got.even_shape == ref.even_shape and got.odd_shape == ref.odd_shape
    assert tuple(got.shape) == tuple(ref.shape)
    assert abs(nrm - nref) <= 1e-10 * nref
    assert abs(gtn.logZ(got) - OB.logZ_block(ref)) < 1e-10
