"""GPU parity at the working size: leg dim 64, chi 64.

* TRG for 20 steps (the full run of the workload) against the oracle's committed golden
  (tests/golden/trg_golden.json "d64_chi64", written by tests/golden/make_trg_golden.py): per-step
  Tnorm and ln Z per site, in FP64 and with the contractions in FP32 (SVD, normalisation and the
  log-norm sum in FP64) -- the ln Z rule of example.py:186-193 over the steps of gauge2d.py:1647-1755;
* the standalone contraction 'ijkl,klmn->ijmn' of two seeded Grassmann-even 64^4 tensors against the
  oracle (FP64, 1e-12 relative), and its S1 sign tensor plus the T' contraction's S1 / S3 (64^4
  entries each) bit-exact.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import gtn_oracle as O
from conftest import device_sign_tensor

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")

HERE = os.path.dirname(os.path.abspath(__file__))


def golden(name="d64_chi64"):
    return json.load(open(os.path.join(HERE, "golden", "trg_golden.json")))[name]


def to_dev(T):
    return gtn.dense(T.data, encoder=T.encoder, format=T.format, statistics=T.statistics)


def run_trg(T0, chi, steps, schedule, diag_scale=False):
    """steps TRG steps from the oracle tensor T0; schedule(i) -> (contraction, svd) precision.
    Returns per step (Tnorm, ln Z per site, sum of |T[I,J,I,J]| or None) and the number of
    recomputed steps."""
    from grassmanntn_amd.gauge2d import trg_counters
    rec0 = trg_counters.recomputed
    T = to_dev(T0)
    logn, out = 0.0, []
    for i in range(steps):
        with gtn.precision(*schedule(i)):
            T, nrm = gtn.trg(T, chi)
        logn = 2 * logn + math.log(nrm)
        scale = None
        if diag_scale:
            h = T.data.numpy()
            d = h.shape[0] * h.shape[1]
            scale = float(np.abs(h.reshape(d, d)[np.arange(d), np.arange(d)]).sum())
        with gtn.precision("f64", "f64"):
            out.append((nrm, (gtn.logZ(T) + logn) / 2 ** (i + 1), scale))
    return out, trg_counters.recomputed - rec0


def test_c1_trg_fp64_matches_golden():
    """FP64 contractions and SVD over the 20 steps: ln Z per site after 20 steps within 1e-10 of the
    oracle; over the first 6 steps every norm within 1e-9 relative and every intermediate supertrace
    Z_n = einsum('IJIJ', T_n) within 1e-9 of the scale of its terms, sum |T_n[I,J,I,J]|.

    Why these forms: an intermediate Z_n can be much smaller than that scale (|Z_2| = 3e-4 here), so
    ln Z_n magnifies the entry-level differences of two FP64 SVDs (LAPACK's and the device's); and
    the coarse-graining flow amplifies those differences about twofold per step (measured: 2e-12 in
    the step-1 norm, 6e-10 at step 6, 1.8e-9 at step 7, 4e-8 at step 12), up to 1e-3 at step 20, where the kept and dropped singular
    values at the truncation edge are nearly degenerate and either basis of that pair is an exact
    SVD.  ln Z per site weighs step n by 2^-n, so the differences of late steps do not reach it."""
    g = golden()
    assert (g["leg_dim"], g["chi"], g["steps"]) == (64, 64, 20)
    res, recomputed = run_trg(O.site_tensor(64, g["seed"]), 64, 20, lambda i: ("f64", "f64"), diag_scale=True)
    logn = 0.0
    worst = []
    for i, (nrm, lnz, scale) in enumerate(res):
        ref = complex(g["lnz_per_site_real"][i], g["lnz_per_site_imag"][i])
        logn = 2 * logn + math.log(g["tnorm"][i])
        z_ref = np.exp(ref * 2 ** (i + 1) - logn)   # the oracle's supertrace of its T_n
        dz = abs(z_ref) * abs(lnz - ref) * 2 ** (i + 1)
        print("step %2d  dTnorm/Tnorm %.2e  dlnZ/site %.2e  |Z_n| %.2e  |dZ_n|/sum|diag| %.2e"
              % (i + 1, abs(nrm / g["tnorm"][i] - 1), abs(lnz - ref), abs(z_ref), dz / scale))
        worst.append((abs(nrm / g["tnorm"][i] - 1), dz / scale))
    assert max(w[0] for w in worst[:6]) < 1e-9
    assert max(w[1] for w in worst[:6]) < 1e-9
    ref = complex(g["lnz_per_site_real"][-1], g["lnz_per_site_imag"][-1])
    assert abs(res[-1][1] - ref) < 1e-10, (res[-1][1], ref)
    from grassmanntn_amd.gauge2d import trg_counters
    print("recomputed steps: %d %s" % (recomputed, trg_counters.misses))


@pytest.mark.parametrize("prec", ["f32", "bf16x6"])
def test_c1_trg_mixed_matches_golden(prec):
    """The mixed schedule (gtn.mixed_schedule): contractions FP64 in the first five steps (their
    normalisation errors weigh 2^-1 .. 2^-5 in ln Z per site), then FP32 MFMA (or split BF16, six
    products, FP32 accumulation); the parity-block SVD, normalisation and log-norm sum FP64 at every
    step: |d ln Z/site| <= 1e-8 after 20 steps (measured 2.0e-9 f32, 1.2e-9 bf16x6)."""
    g = golden()
    res, _ = run_trg(O.site_tensor(64, g["seed"]), 64, 20, lambda i: gtn.mixed_schedule(i, prec))
    ref = complex(g["lnz_per_site_real"][-1], g["lnz_per_site_imag"][-1])
    err = abs(res[-1][1] - ref)
    print("C1 mixed (5 x f64, then %s contractions) |dlnZ/site| = %.3e" % (prec, err))
    assert err <= 1e-8, (prec, err)


@pytest.mark.parametrize("prec", ["f32", "bf16x6"])
def test_c1_trg_reduced_from_step_one(prec):
    """Every step's contractions in reduced precision (SVD in FP64): measured 1.1e-7 (f32) and 7.4e-8
    (bf16x6) -- the FP32 rounding of the first steps' products enters ln Z per site with weight 1/2,
    1/4, ..., which is why the mixed schedule keeps three FP64 steps (bound: a factor ~3)."""
    g = golden()
    res, _ = run_trg(O.site_tensor(64, g["seed"]), 64, 20, lambda i: (prec, "f64"))
    ref = complex(g["lnz_per_site_real"][-1], g["lnz_per_site_imag"][-1])
    err = abs(res[-1][1] - ref)
    print("C1 %s contractions from step 1, f64 SVD: |dlnZ/site| = %.3e" % (prec, err))
    assert err <= 3.5e-7, (prec, err)


def test_c1_trg_split_bf16_range_finder_matches_fp64():
    """Every contraction FP64, the range finder in split BF16 with its last three iterations FP64
    (gtn.precision('f64', 'bf16x3:3')) against the golden after 20 steps: the FP64 iterations damp
    the BF16 rounding left in the subspace by (sigma_{p+1} / sigma_k)^2 each.  Measured over seeds
    1-4 (batched): 2.6e-10 with three FP64 iterations, 1.8e-9 with two, 2.4e-11 with six; the
    all-FP64 range finder is at ~5e-11.  Bound 1e-9."""
    g = golden()
    ref = complex(g["lnz_per_site_real"][-1], g["lnz_per_site_imag"][-1])
    res, _ = run_trg(O.site_tensor(64, g["seed"]), 64, 20, lambda i: ("f64", "bf16x3:3"))
    err = abs(res[-1][1] - ref)
    print("C1 f64 contractions, split-BF16 range finder: |dlnZ/site| = %.3e" % err)
    assert err <= 1e-9, err


def test_c1_trg_int8_range_finder_matches_golden():
    """The range finder's products in FP64 on the int8 matrix cores (eight 7-bit slices per operand
    row, exact integer products, gtn.precision('f64', 'f64_int8')), every contraction FP64: ln Z per
    site after 20 steps within the FP64 path's bound of the golden."""
    g = golden()
    ref = complex(g["lnz_per_site_real"][-1], g["lnz_per_site_imag"][-1])
    res, _ = run_trg(O.site_tensor(64, g["seed"]), 64, 20, lambda i: ("f64", "f64_int8"))
    err = abs(res[-1][1] - ref)
    print("C1 f64 contractions, FP64-by-int8 range finder: |dlnZ/site| = %.3e" % err)
    assert err <= 2e-10, err


def test_c1_trg_gram_range_finder_matches_golden():
    """The range finder iterating on the FP64 Gram matrix M M^T (gtn.precision('f64', 'f64_gram'): one
    FP64 product per iteration), every contraction FP64: ln Z per site after 20 steps within the
    FP64 path's bound of the golden."""
    g = golden()
    ref = complex(g["lnz_per_site_real"][-1], g["lnz_per_site_imag"][-1])
    res, _ = run_trg(O.site_tensor(64, g["seed"]), 64, 20, lambda i: ("f64", "f64_gram"))
    err = abs(res[-1][1] - ref)
    print("C1 f64 contractions, Gram-iteration range finder: |dlnZ/site| = %.3e" % err)
    assert err <= 2e-10, err


def test_c1_einsum_ijkl_klmn():
    """'ijkl,klmn->ijmn' on two seeded 64^4 Grassmann-even tensors (a 4096 x 4096 x 4096 contraction
    whose nonzero parity blocks are two 2048^3 GEMMs): FP64 result within 1e-12 of the oracle's
    largest entry."""
    A, B = O.site_tensor(64, 2), O.site_tensor(64, 3)
    ref = O.einsum("ijkl,klmn->ijmn", A, B)
    out = gtn.einsum("ijkl,klmn->ijmn", to_dev(A), to_dev(B))
    assert out.statistics == ref.statistics and out.encoder == ref.encoder
    got = out.data.numpy()
    scale = np.abs(ref.data).max()
    err = np.abs(got - ref.data).max() / scale
    print("ijkl,klmn->ijmn at 64^4: max |d| / max |ref| = %.2e" % err)
    assert err <= 1e-12


@pytest.mark.parametrize("sub,stats", [
    ("ijkl,klmn->ijmn", [1, 1, -1, -1, 1, 1, -1, -1]),
    ("lxzk,jzxi->ijkl", [-1, 1, 1, -1, 1, -1, -1, 1]),
    ("kwz,lxw->lxzk", [-1, -1, 1, -1, 1, 1]),
    ("ijkl->jkli", [1, 1, -1, -1]),
])
def test_c1_sign_tensors_bit_exact(sub, stats):
    """The reference's S1 / S3 sign tensors at leg dim 64 (up to 64^4 = 16.8 M entries), generated by
    the device popcount kernel, equal the oracle's entry for entry."""
    plan = O.EinsumPlan(sub, stats, [64] * len(stats))
    n = 0
    for which in (1, 3):
        ref, chars = plan.sign_tensor(which)
        if ref is None:
            continue
        got = device_sign_tensor(plan, which, chars, ref.shape)
        assert np.array_equal(got, ref.astype(np.int8)), (sub, which)
        n += 1
    assert n >= 1
