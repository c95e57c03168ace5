"""Learned surrogate of the coupling sweep: an MLP ensemble that maps (beta, m) to ln Z per site and
the leading normalised singular values of the final TRG step, with TRG as the fallback.

Network: layers 2 -> 256 -> 256 -> 256 -> 17, GELU between layers, a linear output.  Weights are
rounded to BF16 in every product and the products accumulate in FP32 (the grouped MFMA GEMM in
GTN_PREC_BF16_SPLIT_A: the activation operand keeps hi + lo BF16 terms); FP32 master weights are
updated by AdamW with a cosine learning-rate schedule.  Every member of the ensemble (different
initialisation and the same data) runs in the same launches (the GEMMs batch over members), and a
full-batch training step -- forward, weighted MSE, backward, AdamW -- is recorded once as a HIP
graph and replayed.  Inputs and targets are standardised with the training set's mean and spread.

Surrogate.query(beta, m): the ensemble mean of ln Z per site where the ensemble's spread is at most
`tol`, and the full batched TRG (the method itself, which also judges the surrogate's error)
everywhere else.
"""
import ctypes as C
import time

import numpy as np

from . import _lib as L
from ._device import DeviceArray
from .batch import Graph, _new_pool
from .core import _gemm, _gemm_desc

PREC = "bf16_split_a"


class MLPEnsemble:
    def __init__(self, dims=(2, 256, 256, 256, 17), members=5, seed=0):
        self.dims, self.E = tuple(int(d) for d in dims), int(members)
        self.L = len(self.dims) - 1
        self.offsets = []
        off = 0
        for l in range(self.L):
            a, b = self.dims[l], self.dims[l + 1]
            self.offsets.append((off, off + self.E * a * b))
            off += self.E * a * b + self.E * b
        self.count = off
        host = np.zeros(self.count, dtype=np.float32)
        for e in range(self.E):
            rng = np.random.default_rng(seed * 1000003 + e)
            for l in range(self.L):
                a, b = self.dims[l], self.dims[l + 1]
                W = rng.standard_normal((a, b)) * np.sqrt(1.0 / a)
                ow, _ = self.offsets[l]
                host[ow + e * a * b: ow + (e + 1) * a * b] = W.reshape(-1)
        self.params = DeviceArray.from_numpy(host)
        self.grads = DeviceArray.zeros(self.count, np.float32)
        self.m = DeviceArray.zeros(self.count, np.float32)
        self.v = DeviceArray.zeros(self.count, np.float32)
        self.step = DeviceArray.zeros(1)
        self._n = None
        self.graph = None

    # ---- parameter views ----------------------------------------------------------------------
    def W(self, l, buf=None):
        buf = buf or self.params
        a, b = self.dims[l], self.dims[l + 1]
        ow, _ = self.offsets[l]
        return buf.flat(ow, self.E * a * b, (self.E, a, b))

    def b(self, l, buf=None):
        buf = buf or self.params
        _, ob = self.offsets[l]
        return buf.flat(ob, self.E * self.dims[l + 1], (self.E, self.dims[l + 1]))

    def host_params(self):
        h = self.params.numpy()
        Ws = [self.W(l, DeviceArray.from_numpy(h)).numpy() for l in range(self.L)]
        bs = [self.b(l, DeviceArray.from_numpy(h)).numpy() for l in range(self.L)]
        return [w.astype(np.float64) for w in Ws], [x.astype(np.float64) for x in bs]

    # ---- buffers for n rows -------------------------------------------------------------------
    def _buffers(self, n):
        if self._n == n:
            return
        self._n = n
        # A[l]: layer l's output (GEMM result, then bias and GELU in place); P[l]: its pre-activation
        self.A = [DeviceArray.empty((self.E, n, self.dims[l + 1]), np.float32) for l in range(self.L)]
        self.P = [DeviceArray.empty((self.E, n, self.dims[l + 1]), np.float32) for l in range(self.L - 1)]
        self.dZ = [DeviceArray.empty((self.E, n, self.dims[l + 1]), np.float32) for l in range(self.L)]
        self.loss = DeviceArray.zeros(self.E)
        cmax = max(self.dims[1:])
        self.colpart = DeviceArray.empty(L.lib().gtn_mlp_colsum_workspace(self.E, cmax) // 8)

    def _mm(self, A, sam, sak, sab, B, sbk, sbn, sbb, Cm, scm, M, N, K, scb):
        d = _gemm_desc(A.data_ptr(), sam, sak, B.data_ptr(), sbk, sbn, Cm.data_ptr(), scm, 1, M, N, K,
                       batch=self.E, sab=sab, sbb=sbb, scb=scb)
        _gemm([d], L.GTN_F32, PREC)

    def forward(self, X):
        """X: device (n, d_0) FP32 shared by the members -> output (E, n, d_L) (a buffer of this
        ensemble, overwritten by the next call)."""
        n = X.shape[0]
        self._buffers(n)
        lib = L.lib()
        H, hshared = X, True
        for l in range(self.L):
            a, b = self.dims[l], self.dims[l + 1]
            A = self.A[l]
            self._mm(H, a, 1, 0 if hshared else n * a, self.W(l), b, 1, a * b, A, b, n, b, a, n * b)
            last = l == self.L - 1
            L.check(lib.gtn_mlp_bias_act(L.ptr(A), self.E, n, b, L.ptr(self.b(l)),
                                         L.GTN_ACT_NONE if last else L.GTN_ACT_GELU,
                                         None if last else L.ptr(self.P[l]), L.stream_ptr()), "gtn_mlp_bias_act")
            H, hshared = A, False
        return H

    def backward(self, X, dout):
        """Gradients of the summed member losses into self.grads, given dL/dout (E, n, d_L)."""
        n = X.shape[0]
        lib = L.lib()
        dZ = dout
        for l in range(self.L - 1, -1, -1):
            a, b = self.dims[l], self.dims[l + 1]
            Hin, shared = (X, True) if l == 0 else (self.A[l - 1], False)
            # dW = H^T dZ (E, a, b); db = column sums of dZ
            self._mm(Hin, 1, a, 0 if shared else n * a, dZ, b, 1, n * b, self.W(l, self.grads), b, a, b, n, a * b)
            L.check(lib.gtn_mlp_colsum(L.ptr(dZ), self.E, n, b, L.ptr(self.b(l, self.grads)), L.ptr(self.colpart),
                                       L.stream_ptr()), "gtn_mlp_colsum")
            if l > 0:
                dH = self.dZ[l - 1]
                # dH = dZ W^T (E, n, a), then through GELU' of the layer below's pre-activation
                self._mm(dZ, b, 1, n * b, self.W(l), 1, b, a * b, dH, a, n, a, b, n * a)
                L.check(lib.gtn_mlp_act_backward(L.ptr(dH), L.ptr(self.P[l - 1]), self.E * n * a, L.GTN_ACT_GELU,
                                                 L.stream_ptr()), "gtn_mlp_act_backward")
                dZ = dH

    def loss_grad(self, out, T, w):
        """dL/dout into the last dZ buffer, per-member losses into self.loss; T (n, d_L) device FP32."""
        n = T.shape[0]
        dout = self.dZ[self.L - 1]
        L.check(L.lib().gtn_mlp_mse(L.ptr(out), L.ptr(T), L.ptr(w), self.E, n, self.dims[-1], L.ptr(dout),
                                    L.ptr(self.loss), L.stream_ptr()), "gtn_mlp_mse")
        return dout

    def adamw(self, opts):
        L.check(L.lib().gtn_adamw(L.ptr(self.params), L.ptr(self.grads), L.ptr(self.m), L.ptr(self.v), self.count,
                                  C.byref(opts), L.ptr(self.step), L.stream_ptr()), "gtn_adamw")

    def train_step(self, X, T, w, opts):
        out = self.forward(X)
        dout = self.loss_grad(out, T, w)
        self.backward(X, dout)
        self.adamw(opts)

    def fit(self, X, T, w, steps, lr=2e-3, lr_min=2e-5, weight_decay=1e-5, graphs=True):
        """Full-batch AdamW on device inputs X (n, d_0) and targets T (n, d_L); returns the per-member
        losses after training (host)."""
        opts = L.AdamwOpts()
        opts.lr_max, opts.lr_min, opts.beta1, opts.beta2, opts.eps = lr, lr_min, 0.9, 0.999, 1e-8
        opts.weight_decay, opts.total_steps = weight_decay, steps
        if not graphs:
            for _ in range(steps):
                self.train_step(X, T, w, opts)
            return self.loss.numpy()
        self._buffers(X.shape[0])
        self.train_step(X, T, w, opts)  # warm-up (fills plan caches), counts as step 1
        g = Graph(lambda: self.train_step(X, T, w, opts), _new_pool())
        for _ in range(steps - 1):
            g.launch()
        g.destroy()
        return self.loss.numpy()

    def predict(self, X):
        """Ensemble mean and spread (device arrays (n, d_L)) for device inputs X (n, d_0)."""
        out = self.forward(X)
        n, d = X.shape[0], self.dims[-1]
        mean = DeviceArray.empty((n, d), np.float32)
        sd = DeviceArray.empty((n, d), np.float32)
        L.check(L.lib().gtn_ensemble_stats(L.ptr(out), self.E, n * d, L.ptr(mean), L.ptr(sd), L.stream_ptr()),
                "gtn_ensemble_stats")
        return mean, sd


class Surrogate:
    """ln Z per site of coupling points from a trained ensemble, with the batched TRG as fallback.

    sweep: a grassmanntn_amd.sweep.CouplingSweep that labels points (the method itself); chi: the
    bond of the labels; tol: largest ensemble spread of ln Z per site accepted without fallback."""

    def __init__(self, sweep, chi=32, members=5, seed=0, tol=1e-4, dims=(2, 256, 256, 256, 17)):
        self.sweep, self.chi, self.tol = sweep, int(chi), float(tol)
        self.ens = MLPEnsemble(dims, members, seed)

    # ---- data ---------------------------------------------------------------------------------
    def labels(self, beta, m):
        """TRG labels (n, 17): ln Z per site (real part) and the 16 leading singular values of the
        final step, each divided by the largest."""
        chi = np.full(len(beta), self.chi)
        lnz, sv = self.sweep.run(beta, m, chi, n_sv=16)
        return np.concatenate([lnz.real[:, None], sv], axis=1)

    def _x(self, beta, m):
        x = np.stack([(np.asarray(beta) - self.xmu[0]) / self.xsd[0], (np.asarray(m) - self.xmu[1]) / self.xsd[1]], 1)
        return DeviceArray.from_numpy(x.astype(np.float32))

    def train(self, beta, m, Y, steps=4000, lnz_weight=8.0, **kw):
        """Fit the ensemble on labelled points (Y from labels()); returns the final member losses."""
        beta, m = np.asarray(beta), np.asarray(m)
        self.xmu = np.array([beta.mean(), m.mean()])
        self.xsd = np.array([beta.std(), m.std()])
        self.ymu, self.ysd = Y.mean(axis=0), Y.std(axis=0) + 1e-12
        T = DeviceArray.from_numpy(((Y - self.ymu) / self.ysd).astype(np.float32))
        w = np.ones(Y.shape[1], dtype=np.float32)
        w[0] = lnz_weight
        w /= w.sum()
        return self.ens.fit(self._x(beta, m), T, DeviceArray.from_numpy(w), steps, **kw)

    def predict(self, beta, m):
        """Host arrays: ensemble mean (n, 17) and spread (n, 17) in label units."""
        mean, sd = self.ens.predict(self._x(beta, m))
        return mean.numpy() * self.ysd + self.ymu, sd.numpy() * self.ysd

    def query(self, beta, m):
        """ln Z per site of every point: the ensemble mean where its spread is <= tol, else TRG.
        Returns (values, fallback mask, seconds in the surrogate, seconds in the fallback)."""
        from ._device import synchronize
        t0 = time.perf_counter()
        mean, sd = self.predict(beta, m)
        lnz = mean[:, 0].astype(np.float64)
        fb = sd[:, 0] > self.tol
        t1 = time.perf_counter()
        if np.any(fb):
            idx = np.nonzero(fb)[0]
            lnz[idx] = self.sweep.run(np.asarray(beta)[idx], np.asarray(m)[idx], np.full(len(idx), self.chi)).real
        synchronize()
        return lnz, fb, t1 - t0, time.perf_counter() - t1
