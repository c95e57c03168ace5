"""Shared inputs of the CPU sweep / surrogate tests: a few coupling points of leg-dim-8 generators
(seeds 1, 2, 3) at bond 8 and their oracle ln Z per site after 2 TRG steps."""
import numpy as np

from oracle import gtn_oracle as O


def small_sweep_labels(n=5, seed=9):
    rng = np.random.default_rng(seed)
    beta, m = rng.uniform(0.1, 2.0, n), rng.uniform(0.0, 1.0, n)
    A = [O.site_tensor(8, s).data for s in (1, 2, 3)]
    ref = []
    for i in range(n):
        T = O.dense(A[0] + beta[i] * A[1] + m[i] * A[2], encoder="parity-preserving", statistics=(1, 1, -1, -1))
        ref.append(O.trg_lnz(T, 8, 2)[0])
    return beta, m, np.array(ref)
