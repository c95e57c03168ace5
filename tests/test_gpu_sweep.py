"""GPU parity of the coupling sweep (grassmanntn_amd.sweep) against the oracle: the stratified subset
of the 512-point sweep committed in tests/golden/sweep_golden.json (the first point of every chi
class in {16, 24, 32, 48, 64}: leg-dim-32 generators, 16 TRG steps), run through the batched,
graph-replayed TRG with the mixed schedule (five FP64 steps, then FP32 contractions; SVD,
normalisation and log-norm sum in FP64): ln Z per site within 1e-8 of the oracle for every class."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sweep_subset_matches_oracle():
    from grassmanntn_amd.sweep import CouplingSweep, coupling_points
    g = json.load(open(os.path.join(HERE, "golden", "sweep_golden.json")))
    beta, m, chi = coupling_points(g["n_points"], g["point_seed"])
    idx = np.array([p["index"] for p in g["points"]])
    ref = np.array([complex(*p["lnz_per_site"]) for p in g["points"]])
    sw = CouplingSweep(D=g["leg_dim"], steps=g["steps"], gen_seeds=tuple(g["gen_seeds"]),
                       schedule=lambda i: gtn.mixed_schedule(i, "f32"))
    got = sw.run(beta[idx], m[idx], chi[idx])
    sw.close()
    err = np.abs(got - ref)
    for p, e in zip(g["points"], err):
        print("chi %d point %d |dlnZ/site| %.2e" % (p["chi"], p["index"], e))
    assert err.max() <= 1e-8


def test_memory_report_and_default_batch():
    """gtn_mem_info counts a fresh allocation as live and reports the device's memory; the sweep's
    default batch at bond 64 (leg-dim-32 generators, chi 48 and 64) is a third of that memory over
    24 x 64^4 FP64 values per point."""
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray
    from grassmanntn_amd.sweep import CouplingSweep
    m0 = L.mem_info()
    x = DeviceArray.empty((8 << 20,))  # 64 MiB
    m1 = L.mem_info()
    assert m1["live"] - m0["live"] >= x.nbytes
    assert 0 < m1["device_free"] <= m1["device_total"]
    del x
    sw = CouplingSweep(D=32, steps=2)
    want = max(1, min(512, max(24 << 30, m1["device_total"] // 3) // (24 * 8 * 64 ** 4)))
    assert sw.max_batch(64) == sw.max_batch(48) == want
