"""The sparse (COO) type (__init__.py:1196-1478): construction, COO accessors, arithmetic and its
operations through the dense device path, on the CPU through the numpy kernel emulator
(tests/kernel_emulator.py)."""
import numpy as np

from conftest import random_even
from kernel_emulator import emulate
from oracle import gtn_oracle as O


def test_sparse_container_and_round_trips():
    rng = np.random.default_rng(41)
    A = random_even(rng, (4, 8, 4, 8), (1, 1, -1, -1), "parity-preserving")
    with emulate():
        import grassmanntn_amd as gtn
        S = gtn.sparse(A.data, encoder="parity-preserving", statistics=A.statistics)
        assert S.nnz == int(np.count_nonzero(A.data)) and S.shape == A.shape
        assert all(A.data[c] == v for c, v in zip(S.coords, S.value))
        assert np.array_equal(S.toarray(), A.data)
        D = gtn.dense(S)
        assert (D.statistics, D.encoder, D.format) == (S.statistics, S.encoder, S.format)
        assert np.array_equal(D.data.numpy(), A.data)
        R = gtn.sparse(D)
        assert np.array_equal(R.toarray(), A.data)
        assert np.array_equal((S * 2.0 - S).toarray(), A.data)
        assert abs(S.norm - np.linalg.norm(A.data)) < 1e-12
        Z = (S * 0.0).remove_zeros()
        assert Z.nnz == 0
        assert np.array_equal(S.switch_encoder().toarray(), A.switch_encoder().data)


def test_sparse_operations_follow_dense():
    """einsum, svd, hconjugate and one TRG step give the dense results, typed sparse where the
    reference returns the first operand's type; Tnorm of a TRG step is the same for both types, as
    the reference's documentation shows for its Schwinger tensor (docs/_sources/schwinger.rst.txt:920)."""
    T0 = O.site_tensor(8, 4)
    with emulate():
        import grassmanntn_amd as gtn
        D = gtn.dense(T0.data, encoder=T0.encoder, statistics=T0.statistics)
        S = gtn.sparse(D)
        e_s = gtn.einsum("ijkl->jkli", S)
        e_d = gtn.einsum("ijkl->jkli", D)
        assert isinstance(e_s, gtn.sparse) and np.array_equal(e_s.toarray(), e_d.data.numpy())
        U, s, V = S.svd("ab|cd", 8)
        Ud, sd, Vd = D.svd("ab|cd", 8)
        assert isinstance(U, gtn.sparse) and np.array_equal(s.toarray(), sd.data.numpy())
        h = S.hconjugate("ab|cd")
        assert isinstance(h, gtn.sparse) and np.array_equal(h.toarray(), D.hconjugate("ab|cd").data.numpy())
        Ts, ns = gtn.trg(S, 8)
        Td, nd = gtn.trg(D, 8)
        assert isinstance(Ts, gtn.sparse) and abs(ns - nd) <= 1e-12 * nd
        assert abs(gtn.logZ(Ts) - gtn.logZ(Td)) < 1e-10
        # sparse input to logZ is densified and taken to the canonical encoder (gauge2d.py:1599-1605)
        ref = gtn.logZ(gtn.dense(Ts).switch_encoder(), "anti-periodic")
        assert gtn.logZ(Ts, "anti-periodic") == ref
