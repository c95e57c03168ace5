"""Coupling sweeps: many (beta, m) points through the batched TRG, grouped by bond dimension.

The synthetic workload of this module: three seeded Grassmann-even generators A0, A1, A2 (site
tensors of leg dim D, statistics (1, 1, -1, -1), parity-preserving legs) define the site tensor of a
coupling point as T(beta, m) = A0 + beta A1 + m A2.  Points draw beta ~ U[beta_lo, beta_hi],
m ~ U[m_lo, m_hi] and a bond dimension chi from a list.  Every point is coarse-grained by TRG
(gauge2d.py:1647-1755) and reported as ln Z per site (example.py:186-193).  Points of equal chi share
their step shapes, so they run together as batches of BatchedTRG (grassmanntn_amd.batch): one
recorded graph per (step shape, arithmetic), replayed for every step of every batch.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._device import DeviceArray
from .batch import BatchedTRG, _bond
from .core import _gemm, _gemm_desc


def _popcount(x):
    x = np.asarray(x, dtype=np.int64)
    c = np.zeros_like(x)
    while np.any(x):
        c += x & 1
        x = x >> 1
    return c


def site_array(D, seed):
    """The synthetic site tensor (host float64 array (D, D, D, D)): N(0,1) entries from
    numpy.random.default_rng(seed) times exp(-0.5 * sum of the legs' occupation counts) (the
    canonical occupation of the stored parity-preserving index), odd total parity zeroed, scaled to
    max |T| = 1."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((D, D, D, D))
    i = np.arange(D)
    enc = i ^ (_popcount(i >> 1) & 1)
    w = np.exp(-0.5 * _popcount(enc))
    par = i & 1
    mask = (par[:, None, None, None] ^ par[None, :, None, None] ^ par[None, None, :, None]
            ^ par[None, None, None, :]) == 0
    T = X * w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None] * w[None, None, None, :] * mask
    return T / np.abs(T).max()


def coupling_points(n, seed, chis=(16, 24, 32, 48, 64), beta=(0.1, 2.0), m=(0.0, 1.0)):
    """n points from numpy.random.default_rng(seed): arrays beta (U[beta]), m (U[m]) and chi
    (uniform over `chis`), drawn in that order."""
    rng = np.random.default_rng(seed)
    b = rng.uniform(beta[0], beta[1], n)
    mm = rng.uniform(m[0], m[1], n)
    chi = rng.choice(np.asarray(chis), n)
    return b, mm, chi


class CouplingSweep:
    """ln Z per site of coupling points T(beta, m) = A0 + beta A1 + m A2.

    D: leg dim of the generators (seeds `gen_seeds`); steps: TRG steps per point; schedule: the
    per-step arithmetic (see core.mixed_schedule); max_batch(chi) -> largest batch of points of bond
    chi run together (default: about 24 GiB of step buffers per batch).  prepare(beta, m, chi)
    records the graphs every batch of those points needs (run() records missing ones itself);
    run(beta, m, chi) -> ln Z per site of every point."""

    def __init__(self, D=32, steps=16, gen_seeds=(10, 11, 12), schedule=None, max_batch=None, graphs=True):
        self.D, self.steps, self.schedule, self.graphs = D, steps, schedule, graphs
        self.A = DeviceArray.from_numpy(np.stack([site_array(D, s).reshape(-1) for s in gen_seeds]))
        self.max_batch = max_batch or self._default_batch
        self.runs = {}

    def _default_batch(self, chi):
        """Points per batch at bond chi: a third of the device's memory (at least 24 GiB) over an
        upper estimate of one point's step buffers, eager and recorded (24 arrays of d^4 FP64 values
        at bond d; measured on MI355X: 2.1 GB per point at leg dim 32 -> bond 64, 0.12 GB at bond 32),
        at most 512."""
        d = _bond(min(self.D * self.D // 2, int(chi / 2)))
        per_point = 24 * 8 * max(d, self.D) ** 4
        if not hasattr(self, "_budget"):
            self._budget = max(24 << 30, L.mem_info()["device_total"] // 3)
        return int(max(1, min(512, self._budget // per_point)))

    def _sites(self, coef):
        """(P, D^4) device array of A0 + beta A1 + m A2 for coef rows (1, beta, m): one GEMM."""
        P = coef.shape[0]
        n = self.A.shape[1]
        cf = DeviceArray.from_numpy(np.ascontiguousarray(coef, dtype=np.float64))
        out = DeviceArray.empty((P, n))
        d = _gemm_desc(cf.data_ptr(), 3, 1, self.A.data_ptr(), n, 1, out.data_ptr(), n, 1, P, n, 3)
        _gemm([d], L.GTN_F64, "f64")
        return out

    def batches(self, beta, m, chi):
        """[(chi, batch size, indices of the points of one batch)]: points grouped by chi, spread
        evenly over the fewest batches of at most max_batch(chi) points; the batch size is that
        count rounded up to a multiple of max(8, q / 16), q the next power of two (or the maximum):
        at most 16 sizes per octave -- few recorded graphs serve any number of points -- and at most
        1/16 of a batch is padding."""
        out = []
        for c in sorted(set(int(x) for x in chi)):
            idx = np.nonzero(np.asarray(chi) == c)[0]
            B = self.max_batch(c)
            nb = -(-len(idx) // B)          # batches needed, then points spread evenly over them so
            per = -(-len(idx) // nb)        # that they share one size (and one set of graphs)
            P = min(B, self.batch_size(per))
            for s in range(0, len(idx), per):
                out.append((c, P, idx[s:s + per]))
        return out

    @staticmethod
    def batch_size(n):
        """Slots of a batch holding n points: n rounded up to a multiple of max(8, q / 16) with q the
        smallest power of two >= n (powers of two up to 8 stay as they are)."""
        q = 1 << int(np.ceil(np.log2(max(1, n))))
        if q <= 8:
            return q
        g = max(8, q // 16)
        return -(-n // g) * g

    def _run_for(self, chi, P, beta=None, m=None, idx=None):
        key = (int(chi), int(P))
        if key not in self.runs:
            run = BatchedTRG(P, self.D, int(chi), self.steps, schedule=self.schedule, graphs=self.graphs)
            if self.graphs:
                run.prepare(self._sites(self._coef(beta, m, self._pad(idx, P))))
            self.runs[key] = run
        return self.runs[key]

    def prepare(self, beta, m, chi):
        """Warm-up and graph capture of every (chi, batch size) the points need."""
        beta, m = np.asarray(beta), np.asarray(m)
        for c, P, idx in self.batches(beta, m, chi):
            self._run_for(c, P, beta, m, idx)

    @staticmethod
    def _pad(idx, P):
        """idx padded to P entries by repeating its last point (the extra results are dropped)."""
        return np.concatenate([idx, np.full(P - len(idx), idx[-1])]) if len(idx) < P else idx

    @staticmethod
    def _coef(beta, m, idx):
        return np.stack([np.ones(len(idx)), beta[idx], m[idx]], axis=1)

    def run(self, beta, m, chi, n_sv=0, progress=None):
        """ln Z per site (complex where a final supertrace is negative) of every point; with n_sv > 0
        also the n_sv leading singular values of every point's final step (T1's blocks, decreasing,
        divided by the largest; zero-padded), as (lnz, sv).  progress(message), when given, is called
        after every batch."""
        beta, m, chi = np.asarray(beta), np.asarray(m), np.asarray(chi)
        out = np.empty(len(beta), dtype=np.complex128)
        sv = np.zeros((len(beta), n_sv))
        for c, P, idx in self.batches(beta, m, chi):
            run = self._run_for(c, P, beta, m, idx)
            full = self._pad(idx, run.P)
            run.start(self._sites(self._coef(beta, m, full)))
            run.advance()
            out[idx] = run.result()[:len(idx)]
            if n_sv:
                k = min(n_sv, run.final_sv.shape[1])
                sv[idx, :k] = run.final_sv[:len(idx), :k]
            if progress:
                progress("chi %d: %d points done" % (c, len(idx)))
        return (out, sv) if n_sv else out

    def close(self):
        for r in self.runs.values():
            r.close()
        self.runs = {}
