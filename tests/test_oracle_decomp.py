"""CPU tests that pin the oracle's join/split, svd, eig, hconjugate and TRG (no GPU)."""
import numpy as np
import pytest

from oracle import gtn_oracle as O
from conftest import random_even


@pytest.mark.parametrize("enc", ["canonical", "parity-preserving"])
def test_join_split_roundtrip(rng, enc):
    A = random_even(rng, (4, 2, 3, 8), (1, -1, 0, -1), enc)
    M = O.join_legs(A, "(ij)(kl)", intermediate_stat=(1, -1))
    assert M.shape == (8, 24) and M.statistics == (1, "*") and M.encoder == "parity-preserving"
    B = O.split_legs(M, "(ij)(kl)", (1, -1, 0, -1), (4, 2, 3, 8), intermediate_stat=(1, -1))
    B = B.force_encoder(enc)
    np.testing.assert_array_equal(A.data, B.data)


def test_join_is_c_order_reshape_with_grouping_sign(rng):
    """Fermionic join = grouping sign (-1)^p(a) on +1 legs of a -1 group, then C-order reshape
    (first leg = high bits), then the encoder on the fused index (__init__.py:2947-3068)."""
    A = random_even(rng, (4, 8, 2, 4), (1, 1, -1, -1))
    M = O.join_legs(A, "(ij)(kl)", intermediate_stat=(-1, 1))
    p = O.parity_of_canonical
    s = (1 - 2 * p(np.arange(4)))[:, None, None, None] * (1 - 2 * p(np.arange(8)))[None, :, None, None]
    ref = (A.data * s).reshape(32, 8)
    ref = ref[O.encoder(np.arange(32))][:, O.encoder(np.arange(8))]
    np.testing.assert_array_equal(M.data, ref)


def test_switch_encoder_and_format_are_involutions(rng):
    A = random_even(rng, (4, 8, 3), (1, -1, 0))
    np.testing.assert_array_equal(A.switch_encoder().switch_encoder().data, A.data)
    np.testing.assert_array_equal(A.switch_format().switch_format().data, A.data)
    np.testing.assert_array_equal(A.switch_parity().switch_parity().data, A.data)


def test_svd_reconstruction_and_blocks(rng):
    A = random_even(rng, (4, 4, 4), (1, 1, -1))
    U, S, V = A.svd("i|jk")
    USV = O.einsum("ia,ab,bjk->ijk", U, S, V)
    assert (A - USV).norm < 1e-13
    # singular values equal numpy's on the two parity blocks of the joined matrix
    M = O.join_legs(A, "(i)(jk)", "matrix", intermediate_stat=(-1, 1)).force_encoder("parity-preserving")
    ME, MO = O.parity_blocks(M.data)
    ref = np.sort(np.concatenate([np.linalg.svd(ME, compute_uv=False), np.linalg.svd(MO, compute_uv=False)]))
    got = np.sort(np.abs(np.diag(S.force_format("matrix").data)))
    got = got[got > 1e-12]
    np.testing.assert_allclose(got, ref[ref > 1e-12], rtol=1e-12)


def test_svd_truncation_and_padding(rng):
    """int(chi/2) per parity block, then both blocks padded to the next power of two."""
    A = random_even(rng, (8, 8, 8, 8), (1, -1, -1, 1))
    for chi, bond in ((24, 32), (16, 16), (10, 16), (6, 8), (4, 4)):
        U, S, V = A.svd("ab|cd", chi)
        assert U.shape == (8, 8, bond) and S.shape == (bond, bond) and V.shape == (bond, 8, 8)
        nz = np.count_nonzero(np.abs(np.diag(S.data)) > 0)
        assert nz == 2 * (chi // 2)


def test_hconjugate_and_eig_doc_identities(rng):
    A = random_even(rng, (4, 4, 4), (1, 1, -1))
    cA = A.hconjugate("i|jk")
    assert cA.statistics == (-1, 1, -1)
    H = O.einsum("ijk,jkI->iI", A, cA)
    assert (H - H.hconjugate("i|I")).norm < 1e-14
    H2 = O.einsum("jki,iJK->jkJK", A.hconjugate("i|jk"), A)
    U, S, V = H2.eig("jk|JK")
    assert (U - V.hconjugate("a|JK")).norm < 1e-13
    assert (H2 - O.einsum("jka,ab,bJK->jkJK", U, S, V)).norm < 1e-12


def test_trg_error_identity_at_full_bond(rng):
    """At a bond that truncates nothing, Z(T') equals Z(T,T) (gauge2d.py:1737-1740)."""
    T = O.site_tensor(4, seed=5)
    _, nrm, err = O.trg(T, dcut=16, error_test=True)
    assert err < 1e-12 and nrm > 0


@pytest.mark.parametrize("direction", ["y", "x"])
def test_atrg_error_identity_at_full_bond(direction):
    """At bonds that truncate nothing, the ATRG step keeps the two-site partition function:
    Z(T') = einsum('IJIK,iKiJ', T1, T2) (the reference's error_test, gauge2d.py:1855-1857); atrg2dx
    is atrg2dy between the x/y leg swaps (gauge2d.py:1871-1885).  Every one of the step's eleven
    contractions and three decompositions enters this identity."""
    f = O.atrg2dy if direction == "y" else O.atrg2dx
    T1, T2 = O.site_tensor(4, seed=5), O.site_tensor(4, seed=7)
    for mode in ("sector", "literal"):
        T, nrm, err = f(T1, T2, dcut=256, error_test=True, mode=mode)
        assert err < 1e-12 and nrm > 0, (mode, err)
    assert T.statistics == (1, 1, -1, -1)
    assert T.shape == ((16, 4, 16, 4) if direction == "y" else (4, 16, 4, 16))


def test_atrg_identity_detects_mistakes(monkeypatch):
    """The full-bond identity is not satisfied by a step with a dropped square root (U S and S V
    instead of U sqrt(S), sqrt(S) V: the reference's block-power defect) or with one sign of the
    fermionic reordering flipped (an odd-odd exchange counted as even)."""
    T1, T2 = O.site_tensor(4, seed=5), O.site_tensor(4, seed=7)
    with monkeypatch.context() as m:
        m.setattr(O, "sqrt", lambda S: S)
        _, _, err = O.atrg2dy(T1, T2, dcut=256, error_test=True)
        assert err > 1e-3
    real_perm_sign = O.perm_sign

    def wrong_perm_sign(src, dst, odd):
        s = real_perm_sign(src, dst, odd)
        n_odd = sum(1 for c in set(src) if odd.get(c, 0) % 2 == 1)
        return -s if n_odd >= 2 else s

    with monkeypatch.context() as m:
        m.setattr(O, "perm_sign", wrong_perm_sign)
        _, _, err = O.atrg2dy(T1, T2, dcut=256, error_test=True, mode="literal")
        assert err > 1e-6


def test_trg_literal_equals_sector():
    T = O.site_tensor(4, seed=6)
    a, na = O.trg(T, dcut=4, mode="literal")
    b, nb = O.trg(T, dcut=4, mode="sector")
    assert abs(na - nb) < 1e-12 * na
    np.testing.assert_allclose(a.data, b.data, atol=1e-12)


def test_site_tensor_generator():
    T = O.site_tensor(16, seed=0)
    assert T.shape == (16,) * 4 and T.encoder == "parity-preserving" and T.statistics == (1, 1, -1, -1)
    assert abs(np.abs(T.data).max() - 1.0) < 1e-15
    assert O.is_grassmann_even(T)


def _site6(rng, dims):
    """A Grassmann-even site tensor (i1 i2 i3 i4 m n) with fermionic x/y legs and bosonic z legs."""
    return random_even(rng, dims, (1, 1, -1, -1, 0, 0), "parity-preserving")


def test_hotrg3dz_error_identity_at_full_bond(rng):
    """At bonds that truncate nothing the coarse-grained tensor keeps the traced partition function:
    Z(T1, T2) = Z(T') per z configuration (the reference's error_test, gauge2d.py:2133-2136).  Any
    wrong sign among the step's 40 contractions and transpositions breaks this identity."""
    T1, T2 = _site6(rng, (4, 4, 4, 4, 2, 2)), _site6(rng, (4, 4, 4, 4, 2, 2))
    T, nrm, err = O.hotrg3dz(T1, T2, dcut=64, error_test=True)
    assert err < 1e-12 and nrm > 0
    assert T.shape == (16, 16, 16, 16, 2, 2) and T.statistics == (1, 1, -1, -1, 0, 0)
    for bc in ("periodic", "anti-periodic"):
        assert np.isfinite(O.logZhotrg3dz(T, T, bc))


def test_hotrg3dz_literal_equals_sector(rng):
    """The literal (per-element sign tensor) einsum and the per-parity-sector einsum give the same
    truncated HOTRG step."""
    T1, T2 = _site6(rng, (4, 4, 4, 4, 2, 2)), _site6(rng, (4, 4, 4, 4, 2, 2))
    a, na = O.hotrg3dz(T1, T2, dcut=6, mode="literal")
    b, nb = O.hotrg3dz(T1, T2, dcut=6, mode="sector")
    assert abs(na - nb) < 1e-12 * na
    assert abs(O.logZhotrg3dz(a, a) - O.logZhotrg3dz(b, b)) < 1e-10


def test_pinv_moore_penrose(rng):
    """pinv (__init__.py:5961-5983) satisfies A P A = A and P A P = P under the Grassmann einsum."""
    A = random_even(rng, (4, 4, 8, 8), (1, 1, -1, -1), "parity-preserving")
    P = O.pinv(A, "ab|cd")
    assert P.shape == (8, 8, 4, 4)
    scale = np.abs(A.data).max()
    assert np.abs(O.einsum("abcd,cdef,efgh->abgh", A, P, A).data - A.data).max() < 1e-12 * scale
    assert np.abs(O.einsum("cdab,abef,efgh->cdgh", P, A, P).data - P.data).max() < 1e-12 * np.abs(P.data).max()


def test_inv_nz_and_exp_identities(rng):
    """inv_nz is an involution on the entries above rcond, and inv_nz(exp(T)) = exp(-T) there (the
    masks agree: exp(x) != 0 exactly where |x| > rcond); both keep the format and encoder."""
    A = random_even(rng, (4, 8, 4), (1, -1, 1), "parity-preserving")
    kept = np.where(np.abs(A.data) > 1e-10, A.data, 0.0)
    twice = O.inv_nz(O.inv_nz(A))
    assert (twice.format, twice.encoder) == (A.format, A.encoder)
    assert np.abs(twice.data - kept).max() < 1e-13
    lhs = O.inv_nz(O.exp(A)).data
    rhs = O.exp(A * -1.0).data
    assert np.abs(lhs - rhs).max() < 1e-13
