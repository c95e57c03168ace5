"""CPU oracle of the surrogate MLP ensemble -- TEST INFRASTRUCTURE ONLY (used by tests/ as the
checker of the device's MLP kernels; never imported by the product).

An ensemble of E members, each layers x -> GELU(x W_l + b_l) with a linear last layer, as in
grassmanntn_amd.surrogate.  Two arithmetic modes:
  * "f64": everything in FP64 (pinned by finite differences of the loss, tests/test_surrogate_oracle);
  * "device": the device's arithmetic -- in every product the first operand is split into two BF16
    terms (hi = bf16(x), lo = bf16(x - hi)) and the second rounded to one BF16 term, products
    accumulated exactly here (FP32 on the device), activations stored in FP32.  BF16 rounding is
    FP64 -> FP32 -> BF16, each round-to-nearest-even, as the device converts.
GELU(x) = 0.5 x (1 + erf(x / sqrt 2)).  Loss per member: sum_{r,j} w_j (out_rj - t_rj)^2 / n.
This restates no reference code: the reference has no learned component.
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round to BF16 (through FP32, round-to-nearest-even at both steps); returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(A, B, mode):
    if mode == "f64":
        return A @ B
    hi = bf16(A)
    lo = bf16(A - hi)
    return (hi + lo) @ bf16(B)


def _store(x, mode):
    return x.astype(np.float32).astype(np.float64) if mode == "device" else x


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(Ws, bs, X, mode="f64"):
    """Ws[l]: (E, d_l, d_{l+1}), bs[l]: (E, d_{l+1}); X (n, d_0) shared by the members.  Returns the
    output (E, n, d_L) and the cache (inputs and pre-activations of every layer)."""
    E = Ws[0].shape[0]
    H = np.broadcast_to(X, (E,) + X.shape)
    cache = []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Z = np.stack([_mm(H[e], W[e], mode) for e in range(E)]) + b[:, None, :]
        Z = _store(Z, mode)
        cache.append((H, Z))
        H = _store(gelu(Z), mode) if l < len(Ws) - 1 else Z
    return H, cache


def loss_and_dout(out, T, w):
    n = T.shape[0]
    d = out - T[None]
    return (w * d * d).sum(axis=(1, 2)) / n, 2.0 * w * d / n


def backward(Ws, cache, dout, mode="f64"):
    """Gradients (dWs, dbs) of sum over members of their losses, given dL/dout."""
    E = Ws[0].shape[0]
    dWs, dbs = [None] * len(Ws), [None] * len(Ws)
    dZ = _store(dout, mode)
    for l in range(len(Ws) - 1, -1, -1):
        H, Z = cache[l]
        dWs[l] = np.stack([_mm(H[e].T, dZ[e], mode) for e in range(E)])
        dbs[l] = dZ.sum(axis=1)
        if l > 0:
            dH = np.stack([_mm(dZ[e], Ws[l][e].T, mode) for e in range(E)])
            dZ = _store(_store(dH, mode) * gelu_grad(cache[l - 1][1]), mode)
    return dWs, dbs


def adamw(w, g, m, v, t, lr, beta1, beta2, eps, wd):
    """One AdamW step (decoupled weight decay) at step t (1-based); returns (w, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return w * (1 - lr * wd) - lr * mh / (np.sqrt(vh) + eps), m, v
