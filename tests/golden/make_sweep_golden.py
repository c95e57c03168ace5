"""Write tests/golden/sweep_golden.json: ln Z per site from the FP64 oracle for a stratified subset of
the coupling sweep -- 512 points T(beta, m) = A0 + beta A1 + m A2 (generators: seeded site tensors of
leg dim 32, seeds 10, 11, 12; beta ~ U[0.1, 2], m ~ U[0, 1], chi from {16, 24, 32, 48, 64}, all drawn
from seed 100), 16 TRG steps -- namely the first point of every chi class (about 20 minutes on 8
cores).  Run in the build container:  python tests/golden/make_sweep_golden.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import gtn_oracle as O  # noqa: E402


def points(n=512, seed=100, chis=(16, 24, 32, 48, 64)):
    rng = np.random.default_rng(seed)
    b = rng.uniform(0.1, 2.0, n)
    m = rng.uniform(0.0, 1.0, n)
    chi = rng.choice(np.asarray(chis), n)
    return b, m, chi


def main():
    D, steps = 32, 16
    A = [O.site_tensor(D, s).data for s in (10, 11, 12)]
    beta, m, chi = points()
    out = {"leg_dim": D, "steps": steps, "gen_seeds": [10, 11, 12], "point_seed": 100, "n_points": 512,
           "points": []}
    for c in sorted(set(int(x) for x in chi)):
        i = int(np.nonzero(chi == c)[0][0])
        T = O.dense(A[0] + beta[i] * A[1] + m[i] * A[2], encoder="parity-preserving", statistics=(1, 1, -1, -1))
        t0 = time.time()
        F, _ = O.trg_lnz(T, c, steps)
        out["points"].append({"index": i, "beta": float(beta[i]), "m": float(m[i]), "chi": c,
                              "lnz_per_site": [F.real, F.imag], "oracle_seconds": time.time() - t0})
        print(i, c, F, "%.1fs" % (time.time() - t0), flush=True)
        with open(os.path.join(HERE, "sweep_golden.json"), "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
