"""Write tests/golden/trg_golden.json: per-step Tnorm and ln Z per site of the FP64 oracle for the
synthetic TRG workloads (Grassmann-even site tensor, statistics (1,1,-1,-1), parity-preserving legs):
  d16_chi16  : leg dim 16, seed 0, chi 16, 10 steps
  d64_chi64  : leg dim 64, seed 1, chi 64, 20 steps   (about 12 minutes on 8 cores)
  d64_chi64_s<n> : the same workload from seeds 2, 3, 4 (members of the batched chi-64 runs)
  d64_chi96  : leg dim 64, seed 1, chi 96 (bond 128 after the power-of-two padding), 3 steps
Run in the build container:  python tests/golden/make_trg_golden.py [name prefix|all]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import gtn_oracle as O  # noqa: E402

CASES = {"d16_chi16": (16, 0, 16, 10), "d64_chi64": (64, 1, 64, 20),
         "d64_chi64_s2": (64, 2, 64, 20), "d64_chi64_s3": (64, 3, 64, 20), "d64_chi64_s4": (64, 4, 64, 20),
         "d64_chi96": (64, 1, 96, 3)}


def run(name):
    D, seed, chi, steps = CASES[name]
    T = O.site_tensor(D, seed)
    rec = []
    t0 = time.time()
    F, _ = O.trg_lnz(T, chi, steps, record=rec)
    return {"leg_dim": D, "seed": seed, "chi": chi, "steps": steps,
            "tnorm": [r[0] for r in rec],
            "lnz_per_site_real": [r[1].real for r in rec], "lnz_per_site_imag": [r[1].imag for r in rec],
            "oracle_seconds": time.time() - t0}


def main(which):
    path = os.path.join(HERE, "trg_golden.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    names = list(CASES) if which == "all" else [n for n in CASES if n == which or n.startswith(which + "_s")
                                                or (which in ("d16", "d64") and n.startswith(which))]
    for n in names:
        res = run(n)
        out = json.load(open(path)) if os.path.exists(path) else {}  # other runs may have written meanwhile
        out[n] = res
        print(n, out[n]["lnz_per_site_real"][-1], "%.1fs" % out[n]["oracle_seconds"], flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "all")
