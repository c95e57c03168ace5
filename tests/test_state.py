"""On-disk coarse-graining state (save_state / load_state): round trips of dense and block tensors
and a resumed TRG run, on the CPU through the numpy kernel emulator (tests/kernel_emulator.py)."""
import math

import numpy as np
import pytest

from conftest import random_even
from kernel_emulator import emulate
from oracle import gtn_oracle as O


def test_state_round_trip_dense_and_block(tmp_path):
    rng = np.random.default_rng(21)
    A = random_even(rng, (4, 8, 3, 8), (1, -1, 0, -1), "parity-preserving")
    with emulate():
        import grassmanntn_amd as gtn
        for fmt in ("standard", "matrix"):
            T = gtn.dense(A.data, encoder="parity-preserving", format=fmt, statistics=A.statistics)
            gtn.save_state(tmp_path / "d.npz", T, log_norm=1.25, step=7)
            R, ln, st = gtn.load_state(tmp_path / "d.npz")
            assert (ln, st) == (1.25, 7)
            assert (R.statistics, R.encoder, R.format) == (T.statistics, T.encoder, T.format)
            assert np.array_equal(R.data.numpy(), T.data.numpy())
        B = gtn.dense(A.data, encoder="parity-preserving", statistics=A.statistics).toblock()
        gtn.save_state(tmp_path / "b.npz", B, log_norm=-3.0, step=2)
        R, ln, st = gtn.load_state(tmp_path / "b.npz")
        assert isinstance(R, gtn.block) and (ln, st) == (-3.0, 2)
        assert (R.even_shape, R.odd_shape, tuple(R.shape), R.statistics) == \
            (B.even_shape, B.odd_shape, tuple(B.shape), B.statistics)
        cb, cr = B.data, R.data
        for par in np.ndindex(*cb.shape):
            assert np.array_equal(cr[par], cb[par])
    with np.load(tmp_path / "b.npz", allow_pickle=False) as z:  # plain arrays only
        assert str(z["kind"]) == "block"


def test_trg_resumes_from_a_checkpoint(tmp_path):
    """Two TRG steps in one go and one step, a checkpoint, a reload and one more step give the same
    tensor, ln-norm and ln Z per site."""
    T0 = O.site_tensor(8, 3)
    with emulate():
        import grassmanntn_amd as gtn
        T = gtn.dense(T0.data, encoder=T0.encoder, statistics=T0.statistics)
        logn = 0.0
        for i in range(2):
            T, n = gtn.trg(T, 8)
            logn = 2 * logn + math.log(n)
        F = (gtn.logZ(T) + logn) / 4
        S = gtn.dense(T0.data, encoder=T0.encoder, statistics=T0.statistics)
        S, n = gtn.trg(S, 8)
        gtn.save_state(tmp_path / "s.npz", S, log_norm=math.log(n), step=1)
        S, logn2, step = gtn.load_state(tmp_path / "s.npz")
        S, n = gtn.trg(S, 8)
        logn2 = 2 * logn2 + math.log(n)
        F2 = (gtn.logZ(S) + logn2) / 2 ** (step + 1)
        assert np.array_equal(S.data.numpy(), T.data.numpy())
        assert logn2 == logn and F2 == F
