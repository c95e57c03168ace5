import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def random_even(rng, shape, stats, encoder="canonical"):
    """Random Grassmann tensor with the odd-total-parity entries removed (what gtn.random does with
    trimming, __init__.py:6021-6045), built from a seeded generator."""
    from oracle import gtn_oracle as O
    X = rng.standard_normal(shape)
    par = np.zeros(shape, dtype=np.int64)
    for ax, (d, s) in enumerate(zip(shape, stats)):
        if s in (1, -1):
            p = O.parity_of_canonical(np.arange(d))
            sh = [1] * len(shape)
            sh[ax] = d
            par = par + p.reshape(sh)
    X = X * (par % 2 == 0)
    T = O.dense(X, statistics=stats)
    if encoder == "parity-preserving":
        T = T.switch_encoder()
    return T


def diagonal_site(D, seed, terms=3):
    """T = sum_n c_n e(a_n, b_n, a_n, b_n) over `terms` distinct pairs of even indices, an oracle dense
    tensor with statistics (1, 1, -1, -1): both TRG matrices ('jk|li' and 'kl|ij') are diagonal, with
    `terms` nonzero elements in their even parity block and none in the odd one, so each SVD keeps
    `terms` values exactly (no rounding noise to resolve)."""
    from oracle import gtn_oracle as O
    rng = np.random.default_rng(seed)
    pairs = rng.choice((D // 2) ** 2, terms, replace=False)
    T = np.zeros((D,) * 4)
    for n, q in enumerate(pairs):
        a, b = 2 * (q // (D // 2)), 2 * (q % (D // 2))
        T[a, b, a, b] = 1.0 - 0.25 * n
    return O.dense(T, encoder="parity-preserving", statistics=(1, 1, -1, -1))


def device_sign_tensor(plan, which, chars, shape):
    """The sign tensor S1 (which=1) or S3 (which=3) of an oracle EinsumPlan over the axes `chars`,
    generated on the device by gtn_sign_tensor from the plan's per-parity-pattern signs (a 1-bit
    table, bit j of the pattern = parity of axis j); returned as a host int8 array."""
    import ctypes as C
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray
    n = len(chars)
    table = [0] * 32
    for m in range(1 << n):
        par = {c: 0 for c in plan.chars}
        for j, c in enumerate(chars):
            par[c] = (m >> j) & 1
        s = plan.s1(par) if which == 1 else plan.s3(par)
        if s < 0:
            table[m >> 5] |= 1 << (m & 31)
    out = DeviceArray.empty(tuple(shape), np.int8)
    dims = (C.c_int64 * n)(*shape)
    tab = (C.c_uint32 * 32)(*table)
    L.check(L.lib().gtn_sign_tensor(n, dims, tab, L.ptr(out), L.stream_ptr()), "gtn_sign_tensor")
    return out.numpy()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
