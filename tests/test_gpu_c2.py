"""GPU parity at chi 96 (leg dim 64): the dense truncation keeps int(96/2) = 48 values per parity
block padded to 64 (the power-of-two rule of __init__.py:4005-4015), so the bond is 128 from the
first step on and the second step decomposes 16384 x 16384 matrices (two 8192 x 8192 parity blocks
each).  Two steps against the oracle golden (tests/golden/trg_golden.json "d64_chi96"): FP64, and
contractions in FP32 or split BF16 (six products, FP32 accumulation) with the SVD, normalisation and
log-norm sum in FP64; per step the norm's relative error (the Frobenius-norm error of the step) and
ln Z per site."""
import json
import math
import os

import numpy as np
import pytest

from oracle import gtn_oracle as O

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")

HERE = os.path.dirname(os.path.abspath(__file__))


def _run(prec, steps, g):
    import time
    T0 = O.site_tensor(64, g["seed"])
    T = gtn.dense(T0.data, encoder=T0.encoder, statistics=T0.statistics)
    logn, logn_ref, out = 0.0, 0.0, []
    for i in range(steps):
        t0 = time.time()
        with gtn.precision(prec, "f64"):
            T, nrm = gtn.trg(T, 96)
        assert tuple(T.shape) == (128, 128, 128, 128)
        logn = 2 * logn + math.log(nrm)
        logn_ref = 2 * logn_ref + math.log(g["tnorm"][i])
        with gtn.precision("f64", "f64"):
            z = gtn.einsum("IJIJ", T)
        h = T.data.numpy().reshape(128 * 128, 128 * 128)
        scale = float(np.abs(np.diagonal(h)).sum())
        del h
        ref = complex(g["lnz_per_site_real"][i], g["lnz_per_site_imag"][i])
        z_ref = np.exp(ref * 2 ** (i + 1) - logn_ref)
        lnz = (np.log(complex(z)) + logn) / 2 ** (i + 1)
        dn, dz = abs(nrm / g["tnorm"][i] - 1), abs(z - z_ref) / scale
        print("chi 96 %s step %d (%.1f s): |dTnorm|/Tnorm %.2e  |dZ_n|/sum|diag| %.2e  |dlnZ/site| %.2e"
              % (prec, i + 1, time.time() - t0, dn, dz, abs(lnz - ref)), flush=True)
        out.append((dn, dz, abs(lnz - ref)))
    return out


def test_chi96_fp64_three_steps_match_golden():
    """FP64, all three golden steps: every norm within 1e-9 relative, every supertrace within 1e-9 of
    its term scale (an intermediate Z_n can be small -- Z_1 < 0 here -- so its logarithm is
    ill-conditioned) and ln Z per site after the third step within 1e-10."""
    g = json.load(open(os.path.join(HERE, "golden", "trg_golden.json")))["d64_chi96"]
    res = _run("f64", 3, g)
    assert max(r[0] for r in res) < 1e-9 and max(r[1] for r in res) < 1e-9
    assert res[-1][2] < 1e-10


@pytest.mark.parametrize("prec", ["f32", "bf16x6"])
def test_chi96_reduced_precision_two_steps(prec):
    """Reduced-precision contractions (SVD, normalisation and log-norm sum FP64), two steps: the
    norms' relative errors and the supertraces against their term scale stay below 1e-6."""
    g = json.load(open(os.path.join(HERE, "golden", "trg_golden.json")))["d64_chi96"]
    res = _run(prec, 2, g)
    assert max(r[0] for r in res) < 1e-6 and max(r[1] for r in res) < 1e-6
