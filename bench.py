#!/usr/bin/env python
"""TRG coarse-graining throughput on one MI355X: mixed precision vs FP64, batched coupling sweeps
and the learned surrogate.

Headline (value): TRG steps/s at leg dim 64, chi 64 of the mixed-precision batched path -- P seeded
Grassmann-even site tensors (entries N(0,1) times exp(-0.5 * sum of leg popcounts), odd total parity
zeroed, max|T| = 1, parity-preserving legs, statistics (1,1,-1,-1); seeds 1..P) coarse-grained
together for --steps (20) steps, every step one replay of a recorded HIP graph; arithmetic
(gtn.mixed_schedule): contractions FP64 in the first five steps, then FP32 MFMA with FP32
accumulation; the parity-block SVD (range finder iterating on the FP64 Gram matrix, Rayleigh-Ritz,
vectors), the normalisation and the log-norm sum FP64 at every step.  Steps/s counts (point, step)
pairs.  Beside it: the same batch with other range-finder arithmetic (split BF16 with FP64 closing
iterations; alternating FP64 products; FP64 on int8 slices), each with its speed-up and ln Z error,
the plain FP64 GPU path (one tensor, the eager gtn.trg loop, the same FP64 SVD)
and the FP64 batched path (same batch, all FP64), ln Z per site of every point with an oracle golden
against that golden, the roofline of the step's dominant kernel (the range finder's FP64 GEMM, timed
with HIP events on the same shape), the coupling sweep (512 points, leg-dim-32 generators, chi from
{16, 24, 32, 48, 64}, 16 steps: ln Z points/s), the learned surrogate (a 5-member MLP ensemble
trained on 8192 TRG-labelled points at chi 32, queried on 4096 points with TRG fallback where the
ensemble spread exceeds 1e-4: ln Z points/s and its error against TRG), the standalone
'ijkl,klmn->ijmn' contraction of two 64^4 tensors and the CPU baseline (the numpy oracle).

`--warmup W` replays W untimed steps (and the plain path runs W untimed eager steps) before timing;
every timed run starts from the initial tensors, so that ln Z after the 20 steps of the workload is
always checked against the oracle.
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
GOLDEN_SEEDS = {1: "d64_chi64", 2: "d64_chi64_s2", 3: "d64_chi64_s3", 4: "d64_chi64_s4"}


def site_tensor(d, seed):
    """The synthetic workload's site tensor on the host (numpy)."""
    from grassmanntn_amd.sweep import site_array
    return site_array(d, seed)


def golden_lnz(n_steps):
    """{seed: ln Z per site after n_steps} from the committed oracle goldens (leg dim 64, chi 64)."""
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "trg_golden.json")))
    out = {}
    for seed, name in GOLDEN_SEEDS.items():
        e = g.get(name)
        if e and n_steps <= len(e["lnz_per_site_real"]):
            out[seed] = complex(e["lnz_per_site_real"][n_steps - 1], e["lnz_per_site_imag"][n_steps - 1])
    return out


VERBOSE = False
_T0 = time.perf_counter()


def progress(msg):
    """A progress line on stderr (the run's stages take seconds to a minute each)."""
    print("[bench %6.1fs] %s" % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


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
        p = ("f64", svd_prec) if i < fp64_steps else (prec, svd_prec)
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
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summaries
    (profiles/*_hbm_bytes.txt: FETCH_SIZE and WRITE_SIZE passes scaled by a copy kernel of known
    size), or None."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith("_hbm_bytes.txt"):
            continue
        for line in open(os.path.join(pdir, name)):
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


def batched_path(gtn, P, steps, warmup, schedule, T0s):
    """P points of leg dim 64 at chi 64 through the graph-replayed batched TRG: graphs recorded
    (untimed), `warmup` untimed replays, then the timed run of `steps` steps from T0s (device-resident
    before timing).  Returns timings and ln Z per site of every point."""
    from grassmanntn_amd._device import DeviceArray, Event, synchronize
    from grassmanntn_amd.batch import BatchedTRG
    t0 = time.perf_counter()
    run = BatchedTRG(P, D, CHI, steps, schedule=schedule)
    X0 = DeviceArray.from_numpy(T0s.reshape(P, -1))
    run.prepare(X0)
    run.start(X0)
    for i in range(warmup):
        run.graphs[run._key(min(i, steps - 1))].launch()
    synchronize()
    prep = time.perf_counter() - t0
    e0, e1 = Event(), Event()
    w0 = time.perf_counter()
    e0.record()
    run.start(X0)
    run.advance()
    e1.record()
    synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_ms(e1)
    lnz = run.result()
    run.close()
    return {"gpu_ms": ms, "wall_s": wall, "prepare_s": prep, "lnz": lnz}


def rf_gemm_roofline(P, reps=20):
    """The batched step's dominant kernel: the range finder's FP64 GEMM of its parity blocks (4P blocks
    of 2048 x 2048 times 2048 x 128, one launch), run `reps` times on the same shape and timed with
    HIP events on the library's stream.  Algorithmic work 2 M N K per block."""
    import ctypes as C
    from grassmanntn_amd import _lib as L
    from grassmanntn_amd._device import DeviceArray, Event, synchronize
    from grassmanntn_amd.core import _gemm_desc
    nb, m, k, p = 4 * P, D * D // 2, D * D // 2, 128
    # operands of N(0, 1) entries (one seeded block repeated): the matrix cores' clock under load
    # depends on the data they multiply, and zero operands would flatter the kernel
    rng = np.random.default_rng(SEED)
    a1 = DeviceArray.from_numpy(rng.standard_normal((m, k)))
    b1 = DeviceArray.from_numpy(rng.standard_normal((k, p)))
    A = DeviceArray.empty((nb, m, k))
    B = DeviceArray.empty((nb, k, p))
    for i in range(nb):
        A.flat(i * m * k, m * k, shape=(m, k)).copy_from(a1)
        B.flat(i * k * p, k * p, shape=(k, p)).copy_from(b1)
    Cm = DeviceArray.empty((nb, m, p))
    d = _gemm_desc(A.data_ptr(), k, 1, B.data_ptr(), p, 1, Cm.data_ptr(), p, 1, m, p, k, batch=nb,
                   sab=m * k, sbb=k * p, scb=m * p)
    arr = (L.GemmDesc * 1)(d)
    lib = L.lib()
    for _ in range(2):
        L.check(lib.gtn_gemm_grouped(arr, 1, L.GTN_F64, L.PREC["f64"], L.stream_ptr()), "gemm")
    synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(reps):
        L.check(lib.gtn_gemm_grouped(arr, 1, L.GTN_F64, L.PREC["f64"], L.stream_ptr()), "gemm")
    e1.record()
    synchronize()
    ms = e0.elapsed_ms(e1) / reps
    flop = 2.0 * m * p * k * nb
    ach = flop / (ms / 1000.0) / 1e12
    # the launch configuration the autotuner settled on for this signature (first calls above)
    tile = 128 if nb * (m // 128) >= 256 else 64
    rep = (L.TuneEntry * 256)()
    for e in rep[:min(lib.gtn_autotune_report(rep, 256), 256)]:
        if (e.M, e.N, e.K, e.batch, e.n_groups, e.precision) == (m, p, k, nb, 1, L.PREC["f64"]):
            tile = e.tile
    return {"bound": "mfma", "achieved": ach, "peak": MFMA_PEAK_TF["f64"], "unit": "TFLOP/s",
            "frac": ach / MFMA_PEAK_TF["f64"], "traffic": measured_traffic("rf_gemm_f64"),
            "kernel": "range-finder GEMM k_gemm_fp<double,double,%d,%d,16>: %d parity blocks of 2048 x 2048 "
                      "times 2048 x 128 (N(0,1) entries) per launch, FP64" % (tile, tile, nb),
            "flop_per_launch": flop, "avg_launch_ms": ms, "launches": reps}


def coupling_sweep(gtn):
    """The 512-point coupling sweep (leg-dim-32 generators, seeds 10/11/12; beta, m, chi from seed 100;
    16 steps, mixed schedule): ln Z points/s, and the points of the committed oracle subset."""
    from grassmanntn_amd._device import synchronize
    from grassmanntn_amd.sweep import CouplingSweep, coupling_points
    beta, m, chi = coupling_points(512, 100)
    sw = CouplingSweep(D=32, steps=16, gen_seeds=(10, 11, 12), schedule=lambda i: gtn.mixed_schedule(i, "f32"))
    t0 = time.perf_counter()
    sw.prepare(beta, m, chi)
    synchronize()
    prep = time.perf_counter() - t0
    progress("coupling sweep: graphs recorded (%.1f s)" % prep)
    t0 = time.perf_counter()
    lnz = sw.run(beta, m, chi, progress=progress)
    synchronize()
    wall = time.perf_counter() - t0
    sw.close()
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "sweep_golden.json")))
    errs = {str(p["chi"]): abs(lnz[p["index"]] - complex(*p["lnz_per_site"])) for p in g["points"]}
    counts = {str(c): int(np.sum(chi == c)) for c in sorted(set(int(x) for x in chi))}
    return {"points": 512, "steps": 16, "wall_s": wall, "prepare_s": prep, "lnz_points_per_s": 512 / wall,
            "point_steps_per_s": 512 * 16 / wall, "points_per_chi": counts,
            "dlnz_vs_oracle_by_chi": errs, "dlnz_vs_oracle_max": max(errs.values())}


def surrogate_workload(gtn, n_train=8192, n_query=4096, tol=1e-4, train_steps=12000):
    """The learned surrogate: TRG labels at chi 32 for n_train points (seed 101), a 5-member ensemble
    trained on them, then n_query points (seed 102) answered by the surrogate with TRG fallback where
    the ensemble spread exceeds tol, against TRG on the same points."""
    from grassmanntn_amd._device import synchronize
    from grassmanntn_amd.sweep import CouplingSweep
    from grassmanntn_amd.surrogate import Surrogate
    rt, rq = np.random.default_rng(101), np.random.default_rng(102)
    bt, mt = rt.uniform(0.1, 2.0, n_train), rt.uniform(0.0, 1.0, n_train)
    bq, mq = rq.uniform(0.1, 2.0, n_query), rq.uniform(0.0, 1.0, n_query)
    sw = CouplingSweep(D=32, steps=16, gen_seeds=(10, 11, 12), schedule=lambda i: gtn.mixed_schedule(i, "f32"))
    sur = Surrogate(sw, chi=32, members=5, seed=0, tol=tol)
    sw.prepare(bt[:1024], mt[:1024], np.full(1024, 32))
    progress("surrogate: TRG labels of %d points" % n_train)
    t0 = time.perf_counter()
    Y = sur.labels(bt, mt)
    synchronize()
    t_label = time.perf_counter() - t0
    progress("surrogate: training (%d steps)" % train_steps)
    t0 = time.perf_counter()
    losses = sur.train(bt, mt, Y, steps=train_steps, lr=3e-3, lr_min=3e-6, lnz_weight=256.0)
    synchronize()
    t_train = time.perf_counter() - t0
    progress("surrogate: TRG on %d queries" % n_query)
    t0 = time.perf_counter()
    truth = sw.run(bq, mq, np.full(n_query, 32)).real
    synchronize()
    t_trg = time.perf_counter() - t0
    progress("surrogate: queries with fallback")
    t0 = time.perf_counter()
    lnz, fb, t_nn, t_fb = sur.query(bq, mq)
    t_q = time.perf_counter() - t0
    mean, sd = sur.predict(bq, mq)
    sw.close()
    err = np.abs(lnz - truth)
    err_nn = np.abs(mean[:, 0] - truth)
    acc = ~fb
    table = {}
    for t in (1e-4, 3e-4, 1e-3, 3e-3):
        ok = sd[:, 0] <= t
        table[str(t)] = {"accepted_frac": float(ok.mean()),
                         "max_err_accepted": float(err_nn[ok].max()) if ok.any() else None,
                         "frac_accepted_err_over_tol": float((err_nn[ok] > t).mean()) if ok.any() else None}
    return {"train_points": n_train, "queries": n_query, "chi": 32, "tol": tol,
            "label_s": t_label, "label_points_per_s": n_train / t_label, "train_s": t_train,
            "train_steps": train_steps, "final_member_losses": [float(x) for x in losses],
            "trg_query_s": t_trg, "trg_lnz_points_per_s": n_query / t_trg,
            "surrogate_query_s": t_q, "surrogate_lnz_points_per_s": n_query / t_q,
            "surrogate_gain_vs_trg": (n_query / t_q) / (n_query / t_trg),
            "inference_only_points_per_s": n_query / t_nn if t_nn > 0 else None,
            "fallback_frac": float(fb.mean()), "max_err_with_fallback": float(err.max()),
            "accepted_max_err": float(err[acc].max()) if acc.any() else None,
            "surrogate_err_median": float(np.median(err_nn)), "surrogate_err_p90": float(np.percentile(err_nn, 90)),
            "tolerance_table": table}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run the copy kernel over this many MiB (HBM counter calibration)")
    ap.add_argument("--rf-gemm-only", action="store_true",
                    help="only time the range finder's FP64 GEMM shape (the roofline kernel; counter runs)")
    ap.add_argument("--plain-only", action="store_true",
                    help="only run the plain FP64 GPU path (profiles of the baseline)")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="f32", help="contraction arithmetic of the mixed path")
    ap.add_argument("--svd-precision", default=None, help="range-finder arithmetic of the mixed path "
                    "(default: the library's mixed schedule, FP64)")
    ap.add_argument("--fp64-steps", type=int, default=5,
                    help="leading TRG steps of the mixed path run in FP64 (their normalisation errors "
                         "weigh 2^-n in ln Z per site)")
    ap.add_argument("--batch", type=int, default=64, help="points coarse-grained together in the headline path")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 512-point coupling sweep")
    ap.add_argument("--no-c4", action="store_true", help="skip the learned surrogate")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rf-variants", action="store_true", help="skip the mixed path with other range-finder "
                    "arithmetic")
    ap.add_argument("--no-einsum", action="store_true", help="skip the standalone 'ijkl,klmn->ijmn' "
                    "contraction of two 64^4 tensors")
    ap.add_argument("--verbose", action="store_true", help="per-step progress and norms on stderr")
    ap.add_argument("--svd-iters", type=int, default=0, help="subspace iterations of the block SVD "
                    "(0: the library default)")
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
    if args.rf_gemm_only:
        print(json.dumps({"roofline": rf_gemm_roofline(args.batch)}))
        return
    if args.plain_only:
        plain = run_path(gtn, "f64", gtn.settings.svd_precision, args.steps, args.warmup, site_tensor(D, SEED))
        print(json.dumps({"fp64_steps_per_s": args.steps / (plain["gpu_ms"] / 1000.0), "lnz": str(plain["lnz"])}))
        return
    if args.svd_iters:
        gtn.settings.svd_iters = args.svd_iters
    if args.svd_orth_every:
        gtn.settings.svd_orth_every = args.svd_orth_every
    if args.svd_oversample >= 0:
        gtn.settings.svd_oversample = args.svd_oversample
    P, K = args.batch, args.steps
    prec = args.precision
    svd_prec = args.svd_precision or gtn.settings.svd_precision
    mixed_sched = lambda i: gtn.mixed_schedule(i, prec, svd=svd_prec, fp64_steps=args.fp64_steps)
    T0s = np.stack([site_tensor(D, SEED + s) for s in range(P)])
    golden = golden_lnz(K)
    progress("mixed batched path: %d points, %d steps" % (P, K))
    mixed = batched_path(gtn, P, K, args.warmup, mixed_sched, T0s)
    progress("mixed batched path: %.1f ms" % mixed["gpu_ms"])
    value = P * K / (mixed["gpu_ms"] / 1000.0)
    out = {
        "metric": "TRG steps/s at chi=64 and ln Z points/s, mixed vs FP64 GPU at |dlnZ|<=1e-8",
        "value": value,
        "unit": "TRG steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": mixed["gpu_ms"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 contractions in steps 1-%d, then %s; SVD (range finder %s), normalisation and log-norm "
                 "sum f64" % (args.fp64_steps, prec, mixed_sched(0)[1]),
        "data": "synthetic",
        "config": {"workload": "TRG at leg dim 64, chi 64: %d seeded site tensors (seeds %d..%d) coarse-grained "
                               "together for %d steps, one HIP-graph replay per step" % (P, SEED, SEED + P - 1, K),
                   "leg_dim": D, "chi": CHI, "points": P, "steps": K, "svd_iters": gtn.settings.svd_iters,
                   "svd_oversample": gtn.settings.svd_oversample, "svd_orth_every": gtn.settings.svd_orth_every},
        "ms_per_point_step_mixed": mixed["gpu_ms"] / (P * K),
        "graph_prepare_s": mixed["prepare_s"],
    }
    if golden:
        errs = {str(s): abs(mixed["lnz"][s - SEED] - g) for s, g in golden.items() if s - SEED < P}
        out["oracle_check"] = "ln Z per site after %d steps of the points with seeds %s vs tests/golden/trg_golden.json" % (
            K, sorted(int(x) for x in errs))
        out["dlnz_mixed_vs_oracle_by_seed"] = errs
        out["dlnz_mixed_vs_oracle"] = max(errs.values())
    else:
        out["oracle_check"] = "skipped: the goldens cover 20 steps, %d were timed" % K
    if not args.no_rf_variants:
        # the same batch and contraction schedule with other range-finder arithmetic
        out["range_finder_variants"] = {}
        for spec, what in (("bf16x3:3", "split BF16 (three BF16 products, FP32 accumulation) except the last "
                                        "three iterations, FP64; Rayleigh-Ritz FP64"),
                           ("f64", "FP64, alternating products M^T Y and M Z per iteration"),
                           ("f64_int8", "FP64 products on the int8 matrix cores (eight 7-bit slices per operand "
                                        "row, exact integer products)")):
            if spec == svd_prec:
                continue
            progress("mixed batched path, range finder %s" % spec)
            sched = lambda i, spec=spec: gtn.mixed_schedule(i, prec, svd=spec, fp64_steps=args.fp64_steps)
            var = batched_path(gtn, P, K, args.warmup, sched, T0s)
            v = {"range_finder": what, "steps_per_s": P * K / (var["gpu_ms"] / 1000.0),
                 "ms_per_point_step": var["gpu_ms"] / (P * K)}
            if golden:
                e = {str(s): abs(var["lnz"][s - SEED] - g) for s, g in golden.items() if s - SEED < P}
                v["dlnz_vs_oracle_by_seed"] = e
                v["dlnz_vs_oracle"] = max(e.values())
            out["range_finder_variants"][spec] = v
    if not args.no_fp64:
        progress("FP64 batched path")
        f64b = batched_path(gtn, P, K, args.warmup, lambda i: ("f64", gtn.settings.svd_precision), T0s)
        out["fp64_batched_steps_per_s"] = P * K / (f64b["gpu_ms"] / 1000.0)
        out["speedup_mixed_vs_fp64_batched"] = value / out["fp64_batched_steps_per_s"]
        if golden:
            out["dlnz_fp64_batched_vs_oracle"] = max(abs(f64b["lnz"][s - SEED] - g) for s, g in golden.items()
                                                     if s - SEED < P)
        progress("plain FP64 path")
        plain = run_path(gtn, "f64", gtn.settings.svd_precision, K, args.warmup, T0s[0])
        out["fp64_steps_per_s"] = K / (plain["gpu_ms"] / 1000.0)
        out["speedup_mixed_vs_fp64"] = value / out["fp64_steps_per_s"]
        for v in out.get("range_finder_variants", {}).values():
            v["speedup_vs_fp64"] = v["steps_per_s"] / out["fp64_steps_per_s"]
        out["fp64_path"] = "plain FP64 GPU path: one tensor (seed %d), the eager gtn.trg loop, %d steps" % (SEED, K)
        out["recomputed_steps_fp64"] = plain["recomputed_steps"]
        if SEED in golden:
            out["dlnz_fp64_vs_oracle"] = abs(plain["lnz"] - golden[SEED])
        if plain["tprime_ms"]:
            flop = 2 * 2.0 * (CHI * CHI // 2) * (D * D // 2) * (CHI * CHI // 2)
            ms64 = sum(plain["tprime_ms"]) / len(plain["tprime_ms"])
            a64 = flop / (ms64 / 1000.0) / 1e12
            out["roofline_tprime_fp64"] = {"bound": "mfma", "achieved": a64, "peak": MFMA_PEAK_TF["f64"],
                                           "unit": "TFLOP/s", "frac": a64 / MFMA_PEAK_TF["f64"],
                                           "traffic": measured_traffic("tprime_gemm_f64"),
                                           "kernel": "T' contraction, 2 parity-block GEMMs of 2048^3, f64 (plain path)",
                                           "avg_launch_ms": ms64, "launches": len(plain["tprime_ms"])}
    progress("range-finder GEMM roofline")
    out["roofline"] = rf_gemm_roofline(P)
    if not args.no_c3:
        progress("coupling sweep: 512 points")
        out["coupling_sweep"] = coupling_sweep(gtn)
        out["lnz_points_per_s_trg_sweep"] = out["coupling_sweep"]["lnz_points_per_s"]
    if not args.no_c4:
        progress("surrogate: labels, training, queries")
        sur = surrogate_workload(gtn)
        out["surrogate"] = sur
        out["lnz_points_per_s"] = sur["surrogate_lnz_points_per_s"]
        out["lnz_points_per_s_trg"] = sur["trg_lnz_points_per_s"]
    if not args.no_einsum:
        progress("einsum 'ijkl,klmn->ijmn' at 64^4")
        out["einsum_ijkl_klmn_d64"] = [einsum_microbench(gtn, "f64"), einsum_microbench(gtn, prec)]
    if not args.no_cpu_baseline:
        progress("CPU baseline (numpy oracle)")
        out["cpu_baseline"] = cpu_baseline()
    progress("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
