"""GPU checks of the surrogate's kernels and of the surrogate with its TRG fallback.

* forward and backward of a 5-member ensemble of the full network (2 -> 256 -> 256 -> 256 -> 17,
  GELU) on the device against the MLP oracle's emulation of the device arithmetic (BF16 weights,
  split-BF16 activations; the oracle accumulates exactly, the device in FP32);
* one AdamW step of the device kernel against the oracle's formula;
* a reduced coupling sweep (leg-dim-16 generators, bond 16, 8 steps): an ensemble trained on 256
  TRG-labelled points predicts 128 other points; where its spread exceeds the tolerance the TRG value
  is returned (exactly the TRG's), elsewhere the surrogate's error against TRG is measured.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")


def test_ensemble_forward_backward_match_oracle():
    from grassmanntn_amd._device import DeviceArray
    from grassmanntn_amd.surrogate import MLPEnsemble
    rng = np.random.default_rng(11)
    n = 512
    ens = MLPEnsemble(members=5, seed=2)
    X = rng.uniform(-1.7, 1.7, (n, 2)).astype(np.float32)
    T = rng.standard_normal((n, 17)).astype(np.float32)
    w = np.full(17, 1.0 / 17, dtype=np.float32)
    Xd = DeviceArray.from_numpy(X)
    out = ens.forward(Xd).numpy()
    dout = ens.loss_grad(ens.forward(Xd), DeviceArray.from_numpy(T), DeviceArray.from_numpy(w))
    ens.backward(Xd, dout)
    g = ens.grads.numpy()
    Ws, bs = ens.host_params()
    ref, cache = M.forward(Ws, bs, X.astype(np.float64), mode="device")
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print("forward max |d| / max |ref| = %.2e" % rel)
    assert rel < 1e-5
    _, dref = M.loss_and_dout(ref, T.astype(np.float64), w.astype(np.float64))
    dWs, dbs = M.backward(Ws, cache, dref, mode="device")
    gd = DeviceArray.from_numpy(g)
    for l in range(4):
        gw, gb = ens.W(l, gd).numpy(), ens.b(l, gd).numpy()
        ew = np.abs(gw - dWs[l]).max() / np.abs(dWs[l]).max()
        eb = np.abs(gb - dbs[l]).max() / np.abs(dbs[l]).max()
        print("layer %d: dW %.2e, db %.2e (relative to the largest entry)" % (l, ew, eb))
        assert ew < 1e-4 and eb < 1e-4


def test_adamw_kernel_matches_oracle():
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray
    import ctypes as C
    rng = np.random.default_rng(5)
    n = 10000
    w, g = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    m, v = (0.1 * rng.standard_normal(n)).astype(np.float32), np.abs(rng.standard_normal(n)).astype(np.float32)
    dw, dg, dm, dv = (DeviceArray.from_numpy(a) for a in (w, g, m, v))
    step = DeviceArray.from_numpy(np.array([4.0]))
    o = L.AdamwOpts()
    o.lr_max, o.lr_min, o.beta1, o.beta2, o.eps, o.weight_decay, o.total_steps = 1e-3, 1e-5, 0.9, 0.999, 1e-8, 1e-2, 11
    L.check(L.lib().gtn_adamw(L.ptr(dw), L.ptr(dg), L.ptr(dm), L.ptr(dv), n, C.byref(o), L.ptr(step), L.stream_ptr()), "adamw")
    lr = 1e-5 + 0.5 * (1e-3 - 1e-5) * (1 + np.cos(np.pi * 4 / 10))
    rw, rm, rv = M.adamw(w.astype(np.float64), g.astype(np.float64), m.astype(np.float64), v.astype(np.float64),
                         5, lr, 0.9, 0.999, 1e-8, 1e-2)
    assert step.numpy()[0] == 5.0
    np.testing.assert_allclose(dw.numpy(), rw, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(dm.numpy(), rm, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(dv.numpy(), rv, rtol=1e-6, atol=1e-7)


def test_surrogate_with_fallback_on_trg_labels():
    from grassmanntn_amd.sweep import CouplingSweep
    from grassmanntn_amd.surrogate import Surrogate
    rng = np.random.default_rng(21)
    bt, mt = rng.uniform(0.1, 2.0, 256), rng.uniform(0.0, 1.0, 256)
    bq, mq = rng.uniform(0.1, 2.0, 128), rng.uniform(0.0, 1.0, 128)
    sw = CouplingSweep(D=16, steps=8, gen_seeds=(10, 11, 12), schedule=lambda i: gtn.mixed_schedule(i, "f32"))
    sur = Surrogate(sw, chi=16, members=5, seed=0, tol=1e-4)
    Y = sur.labels(bt, mt)
    losses = sur.train(bt, mt, Y, steps=1500)
    print("final member losses:", losses)
    truth = sw.run(bq, mq, np.full(128, 16)).real
    lnz, fb, t_nn, t_fb = sur.query(bq, mq)
    mean, sd = sur.predict(bq, mq)
    err = np.abs(mean[:, 0] - truth)
    print("surrogate |d lnZ/site| vs TRG: median %.2e, max %.2e; fallback %d of 128" % (np.median(err), err.max(), fb.sum()))
    assert np.all(lnz[fb] == truth[fb])
    assert np.median(err) < 1e-2 * np.std(truth)
    sw.close()
