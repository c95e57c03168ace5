"""The TRG step's speculative SVD rank counts (the step assumes int(dcut/2) values per parity block
and checks the read-back counts afterwards), on the CPU through the numpy kernel emulator
(tests/kernel_emulator.py): a full-rank step keeps its result, a step with deficient blocks is
recomputed and matches the oracle's gauge2d.trg (gauge2d.py:1647-1755)."""
import numpy as np

from conftest import diagonal_site
from kernel_emulator import emulate
from oracle import gtn_oracle as O


def _spy(core):
    seen = []
    orig = core._Pending.verify

    def verify(self):
        r = orig(self)
        seen.append(r)
        return r

    core._Pending.verify = verify
    return seen, orig


def test_full_rank_step_keeps_speculative_result():
    T0 = O.site_tensor(8, 2)
    with emulate():
        import grassmanntn_amd as gtn
        from grassmanntn_amd import core
        seen, orig = _spy(core)
        try:
            T = gtn.dense(T0.data, encoder=T0.encoder, statistics=T0.statistics)
            got, n = gtn.trg(T, 8)
            ref, nref = gtn.trg(T, 8, _speculate=False)
        finally:
            core._Pending.verify = orig
        assert seen == [True]
        assert n == nref and np.array_equal(got.data.numpy(), ref.data.numpy())


def test_deficient_blocks_recompute_and_match_oracle():
    T0 = diagonal_site(8, 7)
    ref, nref = O.trg(T0, 8)
    with emulate():
        import grassmanntn_amd as gtn
        from grassmanntn_amd import core
        seen, orig = _spy(core)
        try:
            got, n = gtn.trg(gtn.dense(T0.data, encoder=T0.encoder, statistics=T0.statistics), 8)
        finally:
            core._Pending.verify = orig
        assert seen == [False]
        assert tuple(got.shape) == tuple(ref.shape)
        assert abs(n - nref) <= 1e-12 * nref
        np.testing.assert_allclose(got.data.numpy(), ref.data, rtol=0, atol=1e-12 * np.abs(ref.data).max())
