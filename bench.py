#!/usr/bin/env python
"""TRG coarse-graining throughput on one MI355X: mixed precision vs FP64.

Workload (leg dim 64, chi 64, 20 steps): a seeded Grassmann-even 4-leg site tensor (entries N(0,1)
times exp(-0.5 * sum of leg popcounts), odd total parity zeroed, max|T| = 1, parity-preserving legs,
statistics (1,1,-1,-1)) coarse-grained by Levin-Nave TRG for 20 steps.  A step is one gtn.trg call
(two signed permutations, two parity-block SVDs, six contractions, the norm).  `--warmup W` runs W
untimed steps first (from the same tensor, results discarded); the timed region is then `--steps K`
steps from the initial tensor, so that ln Z per site after K <= 20 steps is always checked against the
FP64 CPU oracle's golden values.

Prints one JSON line: value = steps/s of the mixed-precision path (the first five steps in FP64, then
contractions in FP32 MFMA with FP32 accumulation; the parity-block SVD -- range finder included -- the
normalisation and the log-norm sum in FP64 at every step), with the FP64 path's steps/s, ln Z per site of both against the oracle, the number of
steps recomputed after an assumed rank count failed, the roofline of the T' contraction GEMM (time from
HIP events around each launch; HBM bytes from the committed rocprofv3 counter summary) and the CPU
baseline (the numpy oracle timed on the host).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

D, CHI, SEED = 64, 64, 1
MFMA_PEAK_TF = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0, "bf16x6": 2500.0}
SPLIT_PRODUCTS = {"bf16x3": 3, "bf16x6": 6}


def site_tensor(d, seed):
    """The synthetic workload's site tensor on the host (numpy)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((d, d, d, d))
    i = np.arange(d)
    pc = lambda x: np.array([bin(int(v)).count("1") for v in x])
    enc = i ^ (pc(i >> 1) & 1)
    w = np.exp(-0.5 * pc(enc))
    par = i & 1
    mask = (par[:, None, None, None] ^ par[None, :, None, None] ^ par[None, None, :, None] ^ par[None, None, None, :]) == 0
    T = X * w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None] * w[None, None, None, :] * mask
    return T / np.abs(T).max()


VERBOSE = False


def einsum_microbench(gtn, prec, reps=10):
    """The standalone contraction 'ijkl,klmn->ijmn' of two seeded Grassmann-even 64^4 FP64 site
    tensors (statistics (1,1,-1,-1), parity-preserving): a 4096 x 4096 x 4096 contraction after the
    sign application, whose two nonzero parity blocks are 2048^3 GEMMs.  Mean over `reps` calls of
    the whole einsum (operand gathers with the signs, GEMMs, output gather), timed with HIP events."""
    from grassmanntn_amd._device import Event, synchronize
    st = (1, 1, -1, -1)
    A = gtn.dense(site_tensor(D, SEED + 1), encoder="parity-preserving", statistics=st)
    B = gtn.dense(site_tensor(D, SEED + 2), encoder="parity-preserving", statistics=st)
    with gtn.precision(prec, "f64"):
        for _ in range(2):
            C = gtn.einsum("ijkl,klmn->ijmn", A, B)
        synchronize()
        e0, e1 = Event(), Event()
        e0.record()
        for _ in range(reps):
            C = gtn.einsum("ijkl,klmn->ijmn", A, B)
        e1.record()
        synchronize()
    ms = e0.elapsed_ms(e1) / reps
    flop = 2 * 2 * float(D ** 2 // 2) ** 3  # two parity blocks, 2 M N K each
    del C
    return {"precision": prec, "ms": ms, "flop": flop, "tflops": flop / ms / 1e9 if ms > 0 else None, "reps": reps}


def run_path(gtn, prec, svd_prec, steps, warmup, T0, fp64_steps=0):
    """`warmup` untimed TRG steps from T0 (discarded), then `steps` timed steps from T0; step i runs its
    contractions in FP64 when i < fp64_steps, else in `prec` (SVD range finder in `svd_prec`)."""
    from grassmanntn_amd._device import Event, synchronize
    from grassmanntn_amd.core import GemmProfile
    from grassmanntn_amd.gauge2d import trg_counters

    def step(T, i):
        p = ("f64", "f64") if i < fp64_steps else (prec, svd_prec)
        if VERBOSE:
            print("step %d (%s contractions, %s range finder)" % (i, p[0], p[1]), file=sys.stderr, flush=True)
        with gtn.precision(*p):
            out = gtn.trg(T, CHI)
        if VERBOSE:
            print("  Tnorm %.17g" % out[1], file=sys.stderr, flush=True)
        return out

    T = gtn.dense(T0, encoder="parity-preserving", statistics=(1, 1, -1, -1))
    for i in range(warmup):
        T, _ = step(T, i)
    T = gtn.dense(T0, encoder="parity-preserving", statistics=(1, 1, -1, -1))
    logn = 0.0
    synchronize()
    rec0 = trg_counters.recomputed
    e0, e1 = Event(), Event()
    with GemmProfile() as prof:
        t0 = time.perf_counter()
        e0.record()
        for i in range(steps):
            T, nrm = step(T, i)
            logn = 2 * logn + math.log(nrm)
        e1.record()
        synchronize()
        wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_ms(e1)
    with gtn.precision("f64", "f64"):
        F = (gtn.logZ(T) + logn) / 2 ** steps
    # T' contraction: two parity-block GEMMs of (chi^2/2) x (D^2/2) x (chi^2/2)
    key = [k for k in prof.ms if k[0] == CHI * CHI // 2 and k[2] == D * D // 2 and k[3] == 2 and k[5] == prec]
    tp = prof.ms[key[0]] if key else []
    return {"wall_s": wall, "gpu_ms": gpu_ms, "lnz": complex(F), "tprime_ms": tp,
            "recomputed_steps": trg_counters.recomputed - rec0, "recompute_misses": list(trg_counters.misses)}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE passes scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "tprime_gemm_d64_chi64_hbm_bytes.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("machine-readable: "):
            d = json.loads(line[len("machine-readable: "):])
            if d.get("kernel") == kernel:
                return d["traffic_bytes"]
    return None


def cpu_baseline(steps=2):
    """`steps` oracle TRG steps at leg dim 64, chi 64 on the host (numpy + BLAS on the process's
    threads), each step coarse-graining the previous one's normalised output; plus the oracle's literal
    mode (whole sign tensors S1 / S3 materialised, as the reference builds them) over the 10 steps of
    the leg-dim-16, chi-16 workload."""
    from oracle import gtn_oracle as O
    T = O.site_tensor(D, SEED)
    t0 = time.perf_counter()
    for _ in range(steps):
        T, _n = O.trg(T, CHI)
    dt = time.perf_counter() - t0
    T = O.site_tensor(16, 0)
    t1 = time.perf_counter()
    for _ in range(10):
        T, _n = O.trg(T, 16, mode="literal")
    dt_lit = time.perf_counter() - t1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    try:
        blas = np.show_config(mode="dicts")["Build Dependencies"]["blas"]
        blas = "%s %s" % (blas.get("name"), blas.get("version"))
    except Exception:  # numpy builds without the dict form of show_config
        blas = "unknown"
    return {"value": steps / dt, "unit": "TRG steps/s", "cores": threads, "kind": "port",
            "affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(), "blas": blas,
            "sample": "%d TRG steps of the oracle (numpy FP64, per-parity-sector BLAS contractions, LAPACK "
                      "block SVD) at leg dim 64, chi 64; %.1f s" % (steps, dt),
            "literal_mode_d16_chi16": {"value": 10 / dt_lit, "unit": "TRG steps/s",
                                       "sample": "10 TRG steps at leg dim 16, chi 16 with the sign tensors "
                                                 "materialised; %.2f s" % dt_lit}}


def copy_calibration(mib, reps=5):
    """Run the library's 2-D copy kernel over `mib` MiB of FP64 `reps` times (reads and writes `mib` MiB
    per launch): the known-size reference for scaling HBM byte counters."""
    from grassmanntn_amd._device import DeviceArray, copy2d, synchronize
    n = mib * (1 << 20) // 8
    cols = 1 << 12
    rows = n // cols
    src = DeviceArray.zeros((rows, cols))
    dst = DeviceArray.empty((rows, cols))
    for _ in range(reps):
        copy2d(dst, cols, src, cols, rows, cols)
    synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": rows * cols * 8,
                                           "bytes_written_per_launch": rows * cols * 8, "launches": reps}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run the copy kernel over this many MiB (HBM counter calibration)")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="f32", help="contraction arithmetic of the mixed path")
    ap.add_argument("--svd-precision", default="f64", help="range-finder arithmetic of the mixed path")
    ap.add_argument("--fp64-steps", type=int, default=5,
                    help="leading TRG steps of the mixed path run in FP64 (their normalisation errors "
                         "weigh 2^-n in ln Z per site)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-einsum", action="store_true", help="skip the standalone 'ijkl,klmn->ijmn' "
                    "contraction of two 64^4 tensors")
    ap.add_argument("--verbose", action="store_true", help="per-step progress and norms on stderr")
    ap.add_argument("--svd-iters", type=int, default=0, help="subspace iterations of the block SVD "
                    "(0: the library default)")
    ap.add_argument("--mixed-svd-iters", type=int, default=0, help="subspace iterations of the block "
                    "SVD in the mixed path (0: as the FP64 path)")
    ap.add_argument("--svd-orth-every", type=int, default=0, help="re-orthonormalise the range-finder "
                    "basis every this many iterations (0: the library default)")
    ap.add_argument("--svd-oversample", type=int, default=-1, help="range-finder columns beyond the kept "
                    "rank (-1: the library default)")
    args = ap.parse_args()
    global VERBOSE
    VERBOSE = args.verbose
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build runs the TRG hot path on a single GPU "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import grassmanntn_amd as gtn
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
        return
    if args.svd_iters:
        gtn.settings.svd_iters = args.svd_iters
    if args.svd_orth_every:
        gtn.settings.svd_orth_every = args.svd_orth_every
    if args.svd_oversample >= 0:
        gtn.settings.svd_oversample = args.svd_oversample
    T0 = site_tensor(D, SEED)
    fp64_iters = gtn.settings.svd_iters
    mixed_iters = args.mixed_svd_iters or fp64_iters
    gtn.settings.svd_iters = mixed_iters
    mixed = run_path(gtn, args.precision, args.svd_precision, args.steps, args.warmup, T0, args.fp64_steps)
    gtn.settings.svd_iters = fp64_iters
    fp64 = None if args.no_fp64 else run_path(gtn, "f64", "f64", args.steps, args.warmup, T0)
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "trg_golden.json")
    n_total = args.steps
    if os.path.exists(gpath):
        g = json.load(open(gpath)).get("d64_chi64")
        if g and n_total <= len(g["lnz_per_site_real"]):
            golden = complex(g["lnz_per_site_real"][n_total - 1], g["lnz_per_site_imag"][n_total - 1])
    sps = args.steps / (mixed["gpu_ms"] / 1000.0)
    tp = mixed["tprime_ms"]
    flop = 2 * 2.0 * (CHI * CHI // 2) * (D * D // 2) * (CHI * CHI // 2)
    avg_ms = sum(tp) / len(tp) if tp else float("nan")
    achieved = flop / (avg_ms / 1000.0) / 1e12 if tp else None
    peak = MFMA_PEAK_TF[args.precision]
    out = {
        "metric": "TRG steps/s at chi=64 (mixed precision), leg dim 64",
        "value": sps,
        "unit": "TRG steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["gpu_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "%s contractions, %s range finder, f64 SVD/normalisation (first %d steps all f64)" % (
            args.precision, args.svd_precision, args.fp64_steps),
        "data": "synthetic",
        "config": {"workload": "TRG, leg dim 64, chi 64, %d steps from the seed-%d site tensor" % (n_total, SEED),
                   "leg_dim": D, "chi": CHI, "steps_total": n_total, "svd_iters": mixed_iters,
                   "svd_iters_fp64_path": fp64_iters, "svd_oversample": gtn.settings.svd_oversample,
                   "svd_orth_every": gtn.settings.svd_orth_every},
        "lnz_per_site_mixed": [mixed["lnz"].real, mixed["lnz"].imag],
        "recomputed_steps_mixed": mixed["recomputed_steps"],
        "recompute_misses_mixed": mixed["recompute_misses"],
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": (achieved / peak) if achieved else None,
                     "traffic": measured_traffic("tprime_gemm_" + args.precision),
                     "kernel": "T' contraction, 2 parity-block GEMMs of 2048^3, %s" % args.precision,
                     "flop_per_launch": flop, "avg_launch_ms": avg_ms, "launches": len(tp)},
    }
    terms = SPLIT_PRODUCTS.get(args.precision)
    if terms and achieved:
        # split-BF16 forms issue `terms` BF16 matrix-core products per FP64-equivalent product
        out["roofline"]["matrix_core_products_per_flop"] = terms
        out["roofline"]["matrix_core_tflops"] = terms * achieved
        out["roofline"]["matrix_core_frac"] = terms * achieved / peak
    if fp64 is not None:
        out["fp64_steps_per_s"] = args.steps / (fp64["gpu_ms"] / 1000.0)
        out["speedup_mixed_vs_fp64"] = out["value"] / out["fp64_steps_per_s"]
        out["lnz_per_site_fp64"] = [fp64["lnz"].real, fp64["lnz"].imag]
        out["recomputed_steps_fp64"] = fp64["recomputed_steps"]
        out["dlnz_mixed_vs_fp64_gpu"] = abs(mixed["lnz"] - fp64["lnz"])
        if fp64["tprime_ms"]:
            ms64 = sum(fp64["tprime_ms"]) / len(fp64["tprime_ms"])
            a64 = flop / (ms64 / 1000.0) / 1e12
            out["fp64_tprime_tflops"] = a64
            out["roofline_fp64_path"] = {"bound": "mfma", "achieved": a64, "peak": MFMA_PEAK_TF["f64"],
                                         "unit": "TFLOP/s", "frac": a64 / MFMA_PEAK_TF["f64"],
                                         "traffic": measured_traffic("tprime_gemm_f64"),
                                         "kernel": "T' contraction, 2 parity-block GEMMs of 2048^3, f64",
                                         "avg_launch_ms": ms64, "launches": len(fp64["tprime_ms"])}
    if golden is not None:
        out["oracle_check"] = "ln Z per site after %d steps vs tests/golden/trg_golden.json d64_chi64" % n_total
        out["dlnz_mixed_vs_oracle"] = abs(mixed["lnz"] - golden)
        if fp64 is not None:
            out["dlnz_fp64_vs_oracle"] = abs(fp64["lnz"] - golden)
    else:
        out["oracle_check"] = "skipped: the golden covers 20 steps, %d were timed" % n_total
    if not args.no_einsum:
        out["einsum_ijkl_klmn_d64"] = [einsum_microbench(gtn, "f64"), einsum_microbench(gtn, args.precision)]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
