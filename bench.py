"""Benchmark: Scott-Q field lines/s on one MI355X.

Workload (one "step" = one full Q map): random multipolar potential field of
40 point sources (default_rng(0)), 256^3 float32 grid; Scott method, RKF45 at
an absolute per-step tolerance of 1e-6 cells (initial step 0.25 cell); seeds
on the z = 0 plane at 1024 x 1024 pixels (linspace(0, 255, 1024) per axis).
Inputs (field and seeds) are resident in HBM before timing starts; the timed
region is the trace kernel including the per-line Q assembly, outputs written
to HBM.

Reported beside it, on the same GPU: the FP64 one-thread-per-line kernel
(``fp64_per_line``), and on the host CPU the FP64 oracle over a bounded sample
of the same seeds (``cpu_baseline``).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP32_PEAK = 157.3e12       # FLOP/s, MI355X FP32 vector
N_SIMDS = 1024             # 256 CUs x 4 SIMD16
SCLK_PEAK_MHZ = 2400.0     # MI355X peak engine clock
BYTES_PER_EVAL_F32 = 96    # 8 corners x 3 components x 4 B per RHS evaluation
BYTES_PER_EVAL_F64 = 192
FLOPS_PER_EVAL = 350       # algorithmic flops of one Scott stage evaluation (analytic grad b)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (plane 20, volume 2)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (plane 2, volume 1)")
    ap.add_argument("--rates", default="5,20,50,100,200",
                    help="serve: Poisson arrival rates (requests/s) to sweep")
    ap.add_argument("--duration", type=float, default=2.0,
                    help="serve: seconds of arrivals per rate")
    ap.add_argument("--workload", default="plane",
                    choices=["plane", "volume", "batch", "serve", "stretch", "neighbour", "twist", "c1"],
                    help="plane: the 1024^2 Scott Q map (default); volume: 3-D Q on every node "
                         "of the 512x256x256 rope-arcade field; batch: 256 uneven plane requests "
                         "over 3 resident fields in one graph-captured run; serve: Poisson "
                         "arrivals of 128^2..2048^2 maps over 16 resident snapshots; stretch: "
                         "1024^2 plain RKF45 lines on a stretched 256^3 grid; neighbour: the "
                         "1024^2 neighbour-difference Q map (plain RKF45, QCalcPlane, pseudo lines); "
                         "twist: the default map with the twist number (J = index-space curl B)")
    ap.add_argument("--pixels", type=int, default=1024)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch knobs instead of autotuning them")
    ap.add_argument("--per-line-fp32", action="store_true",
                    help="also time the FP32 one-thread-per-line kernel (always on for volume)")
    ap.add_argument("--precision", default="fp32c", choices=["fp32c", "mixed", "fp64"])
    ap.add_argument("--schedule", default=None, choices=["persistent", "per_line"],
                    help="persistent (default) or per_line; the stretch workload defaults to per_line, "
                         "which is the faster schedule for its short lines")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--headline-only", action="store_true",
                    help="plane workload: skip the C1/C3/C4/C5 blocks of the default record")
    ap.add_argument("--no-mixed", action="store_true",
                    help="plane workload: skip the mixed-precision block")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 1 GiB device-to-device copy (known bytes) for counter scaling")
    return ap.parse_args()


PROFILES = {"plane": "profiles/scott_rkf45_multipole256_1024sq.md",
            "volume": "profiles/scott_rkf45_rope512x256x256_volume.md",
            "stretch": "profiles/plain_rkf45_stretched_multipole256_1024sq.md",
            "neighbour": "profiles/neighbour_qmap_plain_rkf45_multipole256_1024sq.md"}
TRAFFIC_HOW = "rocprofv3 FETCH_SIZE x2.00 + WRITE_SIZE x1.00, copy-calibrated"


def hbm_roofline(workload, kernel, kernel_ms, evals, bpe, traffic_applies=True):
    """HBM roofline of a tracing kernel: achieved = measured HBM bytes per
    launch (the workload's rocprofv3 profile) over the live kernel time;
    beside it the no-reuse stencil model (bytes per stage evaluation x
    evaluations), which is a model of the cache-less traffic, not a bound."""
    t = kernel_ms * 1e-3
    traffic = measured_traffic(workload) if traffic_applies else None
    return {"bound": "hbm", "kernel": kernel, "kernel_ms": kernel_ms, "unit": "GB/s",
            "achieved": traffic / t / 1e9 if traffic else None, "peak": HBM_PEAK / 1e9,
            "frac": traffic / t / HBM_PEAK if traffic else None, "traffic": traffic,
            "traffic_source": "%s (%s)" % (PROFILES.get(workload), TRAFFIC_HOW) if traffic else None,
            "stencil_model": {"bytes_per_eval": bpe, "evals_per_launch": evals,
                              "achieved_GBs": evals * bpe / t / 1e9, "frac": evals * bpe / t / HBM_PEAK}}


def measured_traffic(workload):
    """HBM bytes per launch of the workload's dominant kernel as measured
    with rocprofv3 (FETCH_SIZE + WRITE_SIZE, copy-calibrated) and recorded in
    the profile summary for the workload; None if absent."""
    if workload not in PROFILES:
        return None
    try:
        with open(os.path.join(ROOT, PROFILES[workload])) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(field_np, seeds_at, nlines, sample, log):
    """The FP64 C oracle on a uniform sample of the same seeds, on all the
    threads the process may use, plus a single-thread figure on a smaller
    sample.  Returns (record, sample indices, oracle results)."""
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    idx = np.sort(rng.choice(nlines, size=min(sample, nlines), replace=False))
    seeds = seeds_at(idx)
    nt = O.num_threads()
    t0 = time.perf_counter()
    ref = O.trace_scott(field_np, seeds, method="rkf45", s0=0.25, tol=1e-6, nthreads=nt)
    dt = time.perf_counter() - t0
    n1 = min(len(idx), 4096 if nlines < 1 << 24 else 1024)
    t1 = time.perf_counter()
    O.trace_scott(field_np, seeds[:n1], method="rkf45", s0=0.25, tol=1e-6, nthreads=1)
    dt1 = time.perf_counter() - t1
    log(f"cpu baseline: {len(idx)} lines in {dt:.2f} s on {nt} threads; {n1} lines in "
        f"{dt1:.2f} s on 1 thread")
    rec = {"value": len(idx) / dt, "unit": "lines/s", "cores": nt, "kind": "port",
           "sample": f"{len(idx)} seeds drawn uniformly (default_rng(1)) from the same "
                     f"{nlines} seeds, FP64 C oracle, OpenMP dynamic schedule",
           "cpu_model": cpu_model(), "single_thread_lines_per_s": n1 / dt1,
           "oracle_evals_per_line": float(ref["evals"].mean())}
    return rec, idx, ref


def error_vs_oracle(out, idx, ref, floors=None, controller="fp32"):
    """Per-line error of the GPU map against the FP64 oracle on the sampled
    lines: class mismatches, |dlog10 Q| distribution, endpoint error in
    cells; with ``floors`` (class_floors) the class-mismatch budget the path
    is held to: the oracle's one-ulp spread + 1e-4 of the lines for a path
    on the FP64 controller, the oracle's float model's + 1e-4 for the FP32
    error norm."""
    g = {k: out[k][idx].cpu().numpy() for k in ("flag_start", "flag_end", "q_0", "end_x", "end_y",
                                                 "end_z")}
    cls = (g["flag_start"] == ref["flag_start"]) & (g["flag_end"] == ref["flag_end"])
    a, b = g["q_0"].astype(np.float64), ref["q0"]
    d = np.full(a.shape, np.inf)
    ok = np.isfinite(a) & np.isfinite(b) & (a > 0) & (b > 0)
    d[ok] = np.abs(np.log10(a[ok]) - np.log10(b[ok]))
    d[(np.isnan(a) & np.isnan(b)) | ((a == 0) & (b == 0))] = 0.0
    dm = d[cls]
    fin = dm[np.isfinite(dm)]
    end = np.column_stack([g["end_x"], g["end_y"], g["end_z"]]).astype(np.float64)
    e = np.abs(end - ref["end"]).max(axis=1)[cls]
    e = e[np.isfinite(e)]
    rec = {"lines": int(len(idx)), "class_match": float(cls.mean()), "class_mismatches": int((~cls).sum()),
           "frac_dlog10q_le_1e-3": float(np.mean(dm <= 1e-3)),
           "dlog10q_median": float(np.median(fin)), "dlog10q_p999": float(np.quantile(fin, 0.999)),
           "endpoint_err_cells_median": float(np.median(e)),
           "endpoint_err_cells_p99": float(np.quantile(e, 0.99))}
    if floors:
        budget = floors["budget_fp64_controller" if controller == "fp64" else "budget_fp32_error_norm"]
        rec["class_mismatch_budget"] = budget
        rec["within_budget"] = bool(rec["class_mismatches"] <= budget)
        rec["budget_is"] = ("the oracle's own class mismatches under one-float32-ulp seed moves + 1e-4 N"
                            if controller == "fp64" else
                            "the class mismatches of the oracle's float model (the reference's own float "
                            "arithmetic) + 1e-4 N")
    return rec


def bench_batch(args, fq, torch, log):
    """256 uneven Q-map requests (arbitrary planes of 64^2..1024^2 pixels, RK4
    or RKF45 at tol 1e-7..1e-5) over three resident fields -- the quadrupole
    64^3, the random multipole 256^3 (seed 0) and a 320x320x256 multipole
    (seed 7) -- traced as one batch captured in a HIP graph.  One step = one
    replay.  Beside it: the same requests traced one call at a time (FP32 and
    FP64 per-line) and the FP64 oracle on a sample of the lines."""
    from fastqsl_amd import tune
    from fastqsl_amd.batch import ScottBatch
    t0 = time.perf_counter()
    arrs = [fq.synth.quadrupole_64(dtype=np.float64, device="cuda"),
            fq.synth.multipole_256(seed=0, device="cuda"),
            fq.synth.multipole(shape=(320, 320, 256), seed=7, device="cuda")]
    fields = [fq.Field(*a) for a in arrs]
    specs = fq.synth.batch_requests([f.shape for f in fields], count=256, seed=0)
    reqs = []
    for r in specs:
        x, y, z = fq.synth.plane_seeds_abc_torch(r["A"], r["B"], r["C"], r["n"], r["n"])
        reqs.append(dict(field=fields[r["field"]], x=x, y=y, z=z, method=r["method"],
                         tol=r["tol"] or 1e-6, step=r["step"]))
    sizes = np.array([r["n"] ** 2 for r in specs])
    nlines = int(sizes.sum())
    log(f"batch: {len(reqs)} requests, {nlines} lines, fields ready in "
        f"{time.perf_counter() - t0:.1f} s")
    outputs = ("q_0", "q_perp", "flag_start", "flag_end", "evals", "end_x", "end_y", "end_z")
    tuning = None
    if not args.no_autotune:
        t1 = time.perf_counter()
        cfg = tune.autotune_batch(reqs, precision=args.precision)
        tuning = {"cfg": cfg, "seconds": time.perf_counter() - t1,
                  "how": "coordinate descent over the leading requests (<= 2^23 lines), one "
                         "plan per point"}
        log(f"autotuned in {tuning['seconds']:.1f} s: {cfg}")
    b = ScottBatch(reqs, precision=args.precision, outputs=outputs)
    g = b.graph()
    for _ in range(args.warmup):
        g.replay()
    torch.cuda.synchronize()
    step_ms = []
    for _ in range(args.steps):
        ts = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    done = b.times_ms()
    ms_step = float(np.mean(step_ms))
    value = nlines / (ms_step * 1e-3)
    evals = int(sum(int(r["evals"].to(torch.int64).sum().item()) for r in b.results))
    log(f"batch {args.precision} graph: {value:.4g} lines/s, {ms_step:.1f} ms per replay, "
        f"{evals / nlines:.0f} evals/line")
    bpe = BYTES_PER_EVAL_F64 if args.precision == "fp64" else BYTES_PER_EVAL_F32
    extra = {}

    def one_by_one(precision, schedule):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for r in reqs:
            fq.trace_scott(r["field"], r["x"], r["y"], r["z"], method=r["method"], step=r["step"],
                           tol=r["tol"], precision=precision, schedule=schedule,
                           outputs=("q_0", "flag_start", "flag_end"))
        torch.cuda.synchronize()
        return time.perf_counter() - t1

    if not args.no_fp64_compare:
        dt = one_by_one("fp32c", "persistent")
        extra["one_call_per_request_fp32c"] = {"lines_per_s": nlines / dt, "ms": dt * 1e3}
        dt = one_by_one("fp64", "per_line")
        extra["fp64_per_line"] = {"lines_per_s": nlines / dt, "ms": dt * 1e3,
                                  "how": "one FP64 one-thread-per-line launch per request"}
        extra["speedup_vs_fp64_per_line"] = value / (nlines / dt)
        log(f"one call per request: fp32c {extra['one_call_per_request_fp32c']['ms']:.0f} ms, "
            f"fp64 per-line {dt * 1e3:.0f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        rng = np.random.default_rng(1)
        sample = min(args.cpu_sample, 1 << 14)
        pick = np.sort(rng.choice(nlines, size=sample, replace=False))
        starts = np.concatenate([[0], np.cumsum(sizes)])
        which = np.searchsorted(starts, pick, side="right") - 1
        fnp = [tuple(np.asarray(c, np.float64) for c in a) for a in arrs]
        nt = O.num_threads()
        t_cpu = 0.0
        cls_ok, dq = [], []
        for k in np.unique(which):
            li = pick[which == k] - starts[k]
            r, sp = reqs[k], specs[k]
            seeds = torch.stack([r["x"], r["y"], r["z"]], 1)[torch.as_tensor(li, device="cuda")].cpu().numpy()
            t1 = time.perf_counter()
            ref = O.trace_scott(fnp[sp["field"]], seeds, method=r["method"], s0=r["step"],
                                tol=r["tol"], nthreads=nt)
            t_cpu += time.perf_counter() - t1
            res = b.results[k]
            gq = res["q_0"][torch.as_tensor(li, device="cuda")].cpu().numpy().astype(np.float64)
            fs = res["flag_start"][torch.as_tensor(li, device="cuda")].cpu().numpy()
            fe = res["flag_end"][torch.as_tensor(li, device="cuda")].cpu().numpy()
            c = (fs == ref["flag_start"]) & (fe == ref["flag_end"])
            d = np.full(len(li), np.inf)
            ok = np.isfinite(gq) & np.isfinite(ref["q0"]) & (gq > 0) & (ref["q0"] > 0)
            d[ok] = np.abs(np.log10(gq[ok]) - np.log10(ref["q0"][ok]))
            d[(np.isnan(gq) & np.isnan(ref["q0"])) | ((gq == 0) & (ref["q0"] == 0))] = 0.0
            cls_ok.append(c)
            dq.append(np.where(c, d, np.inf))
        cls_ok = np.concatenate(cls_ok)
        dq = np.concatenate(dq)
        fin = dq[np.isfinite(dq)]
        cpu = {"value": sample / t_cpu, "unit": "lines/s", "cores": nt, "kind": "port",
               "sample": f"{sample} lines drawn uniformly (default_rng(1)) from the batch, each "
                         "with its request's method/step/tolerance, FP64 C oracle",
               "cpu_model": cpu_model()}
        extra["error_vs_oracle"] = {"lines": sample, "class_match": float(cls_ok.mean()),
                                    "frac_dlog10q_le_1e-3": float(np.mean(dq <= 1e-3)),
                                    "dlog10q_median": float(np.median(fin)),
                                    "dlog10q_p999": float(np.quantile(fin, 0.999))}
        log(f"cpu baseline: {sample} lines in {t_cpu:.2f} s on {nt} threads")
    nz = sizes > 0
    rec = {
        "metric": "Scott-Q field lines/s",
        "value": value,
        "unit": "lines/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "latency_ms": {"p50": float(np.percentile(done[nz], 50)),
                       "p99": float(np.percentile(done[nz], 99)),
                       "what": "per request, from the start of the batch run to its last line "
                               "finishing (device clock), last replay"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32c" else "f64",
        "data": "synthetic",
        "config": {"workload": "256 uneven requests (planes 64^2..1024^2 px, RK4/RKF45 tol "
                               "1e-7..1e-5) over quadrupole 64^3, multipole 256^3 (seed 0), "
                               "multipole 320x320x256 (seed 7); one graph-captured batch",
                   "lines_per_step": nlines, "precision": args.precision,
                   "schedule": "batch", "parallelism": "single GPU"},
        "roofline": dict(hbm_roofline("batch", "persistent batch groups (one replay)", ms_step, evals, bpe),
                         note="HBM traffic of a replay not profiled; time per replay (host-synchronised), "
                              "all groups"),
        "cpu_baseline": cpu,
        "tuning": tuning,
        **extra,
    }
    b.close()
    for f in fields:
        f.close()
    return rec


def bench_serve(args, fq, torch, log):
    """Serving: Poisson arrivals of Q-map requests (z = 0 windows of 128^2..
    2048^2 pixels, RKF45 tol 1e-6) against 16 resident snapshots of the
    random-multipole field at 320x320x256 (each with a 1-5 % Fourier-mode
    potential perturbation); continuous batching on two streams with HBM
    admission, FP32-compensated.  Beside it: FP64 one-request-at-a-time with
    the one-thread-per-line kernel (the plain baseline) at the low rates."""
    from fastqsl_amd import serve
    t0 = time.perf_counter()
    shape = (320, 320, 256)
    fields = [fq.Field(*fq.synth.snapshot(i, shape=shape, device="cuda")) for i in range(16)]
    hbm = sum(f.nbytes for f in fields)
    log(f"16 snapshots resident ({hbm / 1e9:.1f} GB) in {time.perf_counter() - t0:.1f} s")
    rates = [float(r) for r in args.rates.split(",")]
    sweep = []
    # warm up the kernels and the allocator
    serve.QServer(fields).run(serve.poisson_requests(100, 0.05, 16, shape, seed=99, min_count=4))
    for rate in rates:
        reqs = serve.poisson_requests(rate, args.duration, 16, shape, seed=int(rate))
        st = serve.QServer(fields, precision=args.precision, streams=2).run(reqs)
        row = {"rate": rate, "requests": st["requests"], "lines": st["lines"],
               "lines_per_s": st["lines_per_s"], "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
               "batches": st["batches"], "mean_batch": st["mean_batch"]}
        sweep.append(row)
        log(f"rate {rate:g}/s: {st['requests']} req, {st['lines_per_s']:.3g} lines/s, "
            f"p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, {st['batches']} batches")
    base = []
    if not args.no_fp64_compare:
        for rate in [r for r in rates if r <= 20]:
            reqs = serve.poisson_requests(rate, args.duration, 16, shape, seed=int(rate))
            st = serve.QServer(fields, precision="fp64", mode="serial", schedule="per_line").run(reqs)
            base.append({"rate": rate, "requests": st["requests"], "lines_per_s": st["lines_per_s"],
                         "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"]})
            log(f"fp64 plain serial, rate {rate:g}/s: {st['lines_per_s']:.3g} lines/s, "
                f"p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms")
    top = sweep[-1]
    rec = {
        "metric": "Scott-Q field lines/s",
        "value": top["lines_per_s"],
        "unit": "lines/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 1,
        "ms_per_step": args.duration * 1e3,
        "latency_ms": {"p50": top["p50_ms"], "p99": top["p99_ms"],
                       "what": f"per request, arrival to results resident, at {top['rate']:g} req/s"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32c" else "f64",
        "data": "synthetic",
        "config": {"workload": "Poisson arrivals of z=0 Q maps (128^2..2048^2 px, RKF45 tol 1e-6) "
                               "over 16 resident 320x320x256 snapshots of the random multipole "
                               "(1-5 % Fourier-mode perturbations); continuous batching, 2 streams",
                   "precision": args.precision, "schedule": "serve", "parallelism": "single GPU",
                   "duration_s_per_rate": args.duration},
        "roofline": None,
        "cpu_baseline": None,
        "sweep": sweep,
        "fp64_plain_serial": base,
    }
    for f in fields:
        f.close()
    return rec


def bench_neighbour(args, fq, torch, log):
    """The neighbour-difference Q map of demo/Q_RKF45CuPy.ipynb cells [6]-[11]
    on the random multipole (default_rng(0)) 256^3 float32: plain RKF45
    (TraceBlineAdaptive.cu) with s_len 1/8 and tol_coef 0.5 from 1024^2 seeds
    on z = 1e-4 (a quarter-cell stride), QCalcPlane (FastQSL.py:65-117), four
    pseudo lines per B_flag pixel traced with tol_coef 0.1 (CookPseudoLine,
    FastQSL.py:4-57), Qcut and the merge.  One step = one full map; inputs
    resident; the K timed maps run as one job over four streams
    (qmap.neighbour_qmaps), and the line also carries the one-map-at-a-time
    rate.  The CPU leg runs the same pipeline (FP64 C oracle lines, the
    NumPy float32 restatements of the map steps) on the central block of the
    seed grid."""
    from fastqsl_amd import qmap
    dev = torch.device("cuda", 0)
    steps = args.steps or 10
    warm = args.warmup if args.warmup is not None else 2
    if args.calibrate:
        calibration_copy(torch, dev, log)
    bx, by, bz = fq.synth.multipole_256(seed=0, device="cuda")
    field = fq.Field(bx, by, bz, layouts=("soa64", "cell32"))
    n = args.pixels
    g = np.linspace(0, 255, n, dtype=np.float32)
    stride = 255.0 / (n - 1)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.full(X.size, 1e-4, np.float32)])
    st = torch.as_tensor(seeds, device=dev)
    shape = (n, n)
    kw = dict(method="rkf45", step=0.125, tol_coef=0.5, pseudo_tol_scale=0.2,
              precision=args.precision, schedule=args.schedule)

    def qmap_once():
        return qmap.neighbour_qmap(field, st, shape, stride, **kw)

    for _ in range(warm):
        res = qmap_once()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        res = qmap_once()
    torch.cuda.synchronize()
    wall1 = time.perf_counter() - t1
    npseudo = int(res["n_pseudo"])
    # the job: K maps over four streams, so a map's pseudo-line trace (a few
    # long lines) runs beside the other maps' main traces
    nstreams = 4
    jobs = [(st, shape, stride)] * steps
    for _ in range(warm):
        qmap.neighbour_qmaps(field, jobs[:nstreams], streams=nstreams, **kw)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    piped = qmap.neighbour_qmaps(field, jobs, streams=nstreams, **kw)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    value = n * n * steps / wall
    same = all(bool(torch.equal(p["q_all"].nan_to_num(-1.0), res["q_all"].nan_to_num(-1.0))) for p in piped)
    del piped
    # the dominant kernel: the main trace in the pipeline's seed order (16 x 16
    # pixel blocks), timed alone with HIP events on its stream
    perm = qmap.block_order(shape, (16, 16))
    x, y, z = (st[:, i][perm].contiguous() for i in range(3))
    tr, kms = None, []
    for _ in range(max(3, steps // 2)):
        tr = fq.trace_plain(field, x, y, z, method="rkf45", step=0.125, tol_coef=0.5,
                            precision=args.precision, schedule=args.schedule, out=tr)
        kms.append(fq.last_kernel_ms())
    evals = int(tr["evals"].to(torch.int64).sum().item())
    k = float(np.mean(kms))
    bpe = BYTES_PER_EVAL_F64 if args.precision == "fp64" else BYTES_PER_EVAL_F32
    log(f"neighbour {args.precision}/{args.schedule}: {value:.4g} pixels/s over {nstreams} streams "
        f"({wall / steps * 1e3:.2f} ms per map; one at a time {wall1 / steps * 1e3:.2f} ms), {npseudo} pseudo "
        f"pixels; main trace {k:.2f} ms, {evals / (n * n):.1f} evals/line; maps identical: {same}")
    extra = {"one_map_at_a_time": {"pixels_per_s": n * n * steps / wall1, "ms_per_map": wall1 / steps * 1e3},
             "streams": nstreams, "maps_identical_to_one_at_a_time": same}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        from oracle import qplane
        side = min(n, int(round(np.sqrt(args.cpu_sample << 2))))
        r0 = (n - side) // 2
        blk = seeds.reshape(n, n, 3)[r0:r0 + side, r0:r0 + side].reshape(-1, 3)
        f64 = tuple(np.asarray(b, np.float64) for b in (bx, by, bz))
        nt = O.num_threads()
        t3 = time.perf_counter()
        o = O.trace_plain(f64, blk, method="rkf45", s0=0.125, tol_coef=0.5, nthreads=nt)
        gr = lambda a: np.asarray(a, np.float32).reshape(side, side)
        ge = lambda a, i: gr(a[:, i])
        qin = qplane.qcalc_plane(ge(o["end"], 0), ge(o["end"], 1), ge(o["end"], 2), gr(o["flag_end"]),
                                 ge(o["start"], 0), ge(o["start"], 1), ge(o["start"], 2), gr(o["flag_start"]),
                                 ge(o["b_start"], 0), ge(o["b_start"], 1), ge(o["b_start"], 2),
                                 ge(o["b_end"], 0), ge(o["b_end"], 1), ge(o["b_end"], 2),
                                 ge(o["b_this"], 0), ge(o["b_this"], 1), ge(o["b_this"], 2),
                                 gr(o["b_flag"]), stride)
        ps, b0s, b0e = qplane.pseudo_seeds(o["start"][:, 0], o["start"][:, 1], o["start"][:, 2],
                                           o["flag_start"], o["b_start"][:, 2], o["b_end"][:, 2],
                                           o["b_flag"], stride)
        qc = np.zeros(0, np.float32)
        if len(ps):
            op = O.trace_plain(f64, ps, method="rkf45", s0=0.125, tol_coef=0.1, nthreads=nt)
            qc = qplane.qcut(op["start"], op["flag_start"], op["end"], op["flag_end"], b0s, b0e, stride)
        qref = qplane.merge(qin, qc, gr(o["b_flag"]))
        dt = time.perf_counter() - t3
        cpu = {"value": side * side / dt, "unit": "pixels/s", "cores": nt, "kind": "port",
               "sample": f"the central {side}x{side} block of the same seed grid as its own map "
                         f"({side * side} lines + {len(ps)} pseudo lines), FP64 C oracle lines "
                         "(OpenMP dynamic schedule) and the NumPy float32 map steps",
               "cpu_model": cpu_model()}
        qg = res["q_all"][r0 + 1:r0 + side - 1, r0 + 1:r0 + side - 1].cpu().numpy()
        fin = np.isfinite(qg) & np.isfinite(qref) & (qg > 0) & (qref > 0)
        dl = np.abs(np.log10(qg[fin]) - np.log10(qref[fin]))
        extra["error_vs_oracle"] = {
            "pixels": int(qref.size), "nan_match": float(np.mean(np.isnan(qg) == np.isnan(qref))),
            "dlog10q_median": float(np.median(dl)), "frac_dlog10q_le_1e-2": float(np.mean(dl <= 1e-2)),
            "note": "plain RKF45 step sequences fork on round-off; the map differences amplify endpoint "
                    "differences near separatrices"}
        log(f"cpu baseline: {side}x{side} block in {dt:.2f} s on {nt} threads")
    rec = {
        "metric": "Q-map pixels/s (neighbour difference, plain RKF45)",
        "value": value, "unit": "pixels/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32c" else "f64", "data": "synthetic",
        "config": {"workload": "random multipole (40 sources, default_rng(0)) 256^3 float32, plain RKF45 "
                               "s_len 1/8 tol_coef 0.5, z=1e-4 plane %dx%d, QCalcPlane + pseudo lines "
                               "(tol_coef 0.1) + Qcut; K maps over %d streams" % (n, n, nstreams),
                   "pixels_per_step": n * n, "pseudo_pixels": npseudo, "precision": args.precision,
                   "schedule": args.schedule, "parallelism": "single GPU"},
        "roofline": hbm_roofline("neighbour", "k_plain_%s (main trace)" % args.schedule, k, evals, bpe,
                                 (args.precision, args.schedule, n) == ("fp32c", "persistent", 1024)),
        "cpu_baseline": cpu,
        **extra,
    }
    field.close()
    return rec


def bench_twist(args, fq, torch, log):
    """Scott Q with the twist number (TraceBlineScott.cu:423-457) on the
    random multipole (default_rng(0)) 256^3 float32, RKF45 tol 1e-6, 1024^2
    seeds on z = 0; J from the twist notebooks' index-space curl
    (postproc.curl_index, demo/Q_RK4.ipynb cell [11]) as a second resident
    field.  One step = one full map with twist; the same map without twist
    is timed beside it, and the FP64 oracle (with J) runs on a sample of the
    seeds."""
    from fastqsl_amd import postproc
    from fastqsl_amd.trace import SCOTT_OUTPUTS
    dev = torch.device("cuda", 0)
    steps = args.steps or 10
    warm = args.warmup if args.warmup is not None else 2
    if args.calibrate:
        calibration_copy(torch, dev, log)
    bx, by, bz = fq.synth.multipole_256(seed=0, device="cuda")
    jx, jy, jz = postproc.curl_index(bx, by, bz)
    field = fq.Field(bx, by, bz, layouts=("soa64", "cell32"), current=(jx, jy, jz))
    n = args.pixels
    g = np.linspace(0, 255, n, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size, np.float32)])
    sx, sy, sz = (torch.as_tensor(np.ascontiguousarray(seeds[:, i]), device=dev) for i in range(3))
    nlines = seeds.shape[0]
    persistent = args.schedule == "persistent"
    outs = {}

    def run(twist):
        o = SCOTT_OUTPUTS + (("twist",) if twist else ())
        outs[twist] = fq.trace_scott(field, sx, sy, sz, method="rkf45", step=0.25, tol=1e-6,
                                     precision=args.precision, schedule=args.schedule, outputs=o,
                                     out=outs.get(twist))
        return outs[twist]

    def timed(twist):
        for _ in range(warm):
            run(twist)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            run(twist)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        kms = []
        for _ in range(max(3, steps // 2)):
            run(twist)
            kms.append(fq._lib.last_step_ms() if persistent else fq.last_kernel_ms())
        return wall, float(np.mean(kms))

    wall, k = timed(True)
    wall0, k0 = timed(False)
    out = outs[True]
    evals = int(out["evals"].to(torch.int64).sum().item())
    value = nlines * steps / wall
    bpe = BYTES_PER_EVAL_F64 if args.precision == "fp64" else BYTES_PER_EVAL_F32
    log(f"twist {args.precision}/{args.schedule}: {value:.4g} lines/s ({wall / steps * 1e3:.2f} ms per map; "
        f"without twist {wall0 / steps * 1e3:.2f} ms), stepping kernel {k:.2f} ms (without twist {k0:.2f})")
    extra = {"without_twist": {"lines_per_s": nlines * steps / wall0, "ms_per_map": wall0 / steps * 1e3,
                               "kernel_ms": k0}}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        rng = np.random.default_rng(1)
        idx = np.sort(rng.choice(nlines, size=min(args.cpu_sample, nlines), replace=False))
        f64 = tuple(np.asarray(b.cpu().numpy() if hasattr(b, "cpu") else b, np.float64) for b in (bx, by, bz))
        j64 = tuple(np.asarray(b.cpu().numpy() if hasattr(b, "cpu") else b, np.float64) for b in (jx, jy, jz))
        nt = O.num_threads()
        t3 = time.perf_counter()
        ref = O.trace_scott(f64, seeds[idx], method="rkf45", s0=0.25, tol=1e-6, nthreads=nt, current=j64)
        dt = time.perf_counter() - t3
        cpu = {"value": len(idx) / dt, "unit": "lines/s", "cores": nt, "kind": "port",
               "sample": f"{len(idx)} seeds drawn uniformly (default_rng(1)) from the same {nlines} seeds, "
                         "FP64 C oracle with the same J, OpenMP dynamic schedule", "cpu_model": cpu_model()}
        tw = out["twist"][idx].cpu().numpy()
        cls = (out["flag_start"][idx].cpu().numpy() == ref["flag_start"]) & \
              (out["flag_end"][idx].cpu().numpy() == ref["flag_end"])
        d = np.abs(tw - ref["twist"])[cls]
        scale = np.abs(ref["twist"])[cls]
        extra["error_vs_oracle"] = {
            "lines": int(len(idx)), "class_match": float(cls.mean()),
            "twist_abs_err_median": float(np.median(d)),
            "frac_twist_rel_err_le_1e-3": float(np.mean(d <= 1e-3 * np.maximum(scale, 1e-3))),
            "note": "RKF45 step sequences fork on round-off; twist integrates J.B/|B|^2/4pi along them"}
        log(f"cpu baseline: {len(idx)} lines in {dt:.2f} s on {nt} threads")
    rec = {
        "metric": "Scott-Q field lines/s with twist",
        "value": value, "unit": "lines/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32c" else "f64", "data": "synthetic",
        "config": {"workload": "random multipole (40 sources, default_rng(0)) 256^3 float32 with J = index-space "
                               "curl B (current-free up to the discretisation, so twists are small), Scott Q + "
                               "twist, RKF45 tol 1e-6, z=0 plane %dx%d" % (n, n),
                   "lines_per_step": nlines, "precision": args.precision, "schedule": args.schedule,
                   "parallelism": "single GPU"},
        "roofline": dict(hbm_roofline("twist", ("stepping kernel k_scott_persistent_w2 (twist)" if persistent
                                                 else "k_scott_per_line (twist)"), k, evals, bpe),
                         note="HBM traffic not profiled; stage evaluations only (the two J and B "
                              "interpolations per accepted step for the twist are not counted)"),
        "cpu_baseline": cpu,
        **extra,
    }
    field.close()
    return rec


def calibration_copy(torch, dev, log):
    """A 1 GiB device-to-device copy of known bytes, for scaling the HBM
    byte counters of a profiled run."""
    src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log("calibration copy: 1073741824 B read + 1073741824 B written")
    del src, dst


def bench_stretch(args, fq, torch, log):
    """Plain RKF45 tracing on a stretched grid (TraceBlineAdaptiveStretch.cu):
    the random multipole (default_rng(0)) scaled by 10 and sampled on a 256^3
    float32 grid with x and y clustered at the centre (sinh, strength 1.5;
    cells 0.055-0.13) and z spacing growing with height (geometric, strength
    3; cells 0.013-0.24), box [-10,10]^2 x [0,20]; 1024^2 seeds on z = 1e-3
    over [-9.5, 9.5]^2, s_len 0.01, tol_coef 1e-2, the reference's bisection
    cell rule (flag_uni false).  FP32c (one thread per line by default:
    these lines are short, ~200 evaluations, and adjacent seeds in one
    wavefront share cache lines) against FP64 one thread per line on the
    same GPU, and the FP64 oracle on the host."""
    dev = torch.device("cuda", 0)
    steps = args.steps or 5
    warm = args.warmup if args.warmup is not None else 1
    t0 = time.perf_counter()
    (bx, by, bz), axes = fq.synth.stretched_multipole((256, 256, 256), seed=0, device="cuda", scale=10.0)
    field = fq.Field(bx, by, bz, layouts=("soa64", "cell32"), axes=axes, lookup="bisect")
    n = args.pixels
    g = np.linspace(-9.5, 9.5, n, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.full(X.size, 1e-3, np.float32)])
    sx, sy, sz = (torch.as_tensor(np.ascontiguousarray(seeds[:, i]), device=dev) for i in range(3))
    nlines = seeds.shape[0]
    log(f"stretched field resident ({field.nbytes / 1e9:.2f} GB) in {time.perf_counter() - t0:.1f} s; "
        f"{nlines} lines per step")
    kw = dict(method="rkf45", step=0.01, tol_coef=1e-2)
    outs = {}
    if args.calibrate:
        calibration_copy(torch, dev, log)

    def run(precision, schedule):
        key = (precision, schedule)
        outs[key] = fq.trace_plain(field, sx, sy, sz, precision=precision, schedule=schedule,
                                   out=outs.get(key), **kw)
        return outs[key]

    for _ in range(warm):
        run(args.precision, args.schedule)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        run(args.precision, args.schedule)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    kms = []
    for _ in range(steps):
        run(args.precision, args.schedule)
        kms.append(fq.last_kernel_ms())
    out = outs[(args.precision, args.schedule)]
    evals = int(out["evals"].to(torch.int64).sum().item())
    value = nlines * steps / wall
    k = float(np.mean(kms))
    bpe = BYTES_PER_EVAL_F64 if args.precision == "fp64" else BYTES_PER_EVAL_F32
    log(f"stretch {args.precision}/{args.schedule}: {value:.4g} lines/s, kernel {k:.2f} ms, "
        f"{evals / nlines:.1f} evals/line")
    extra = {}
    if not args.no_fp64_compare:
        run("fp64", "per_line")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        run("fp64", "per_line")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t2
        extra["fp64_per_line"] = {"lines_per_s": nlines / dt, "ms_per_step": dt * 1e3}
        extra["speedup_vs_fp64_per_line"] = value / (nlines / dt)
        log(f"stretch fp64/per_line: {nlines / dt:.4g} lines/s ({dt * 1e3:.1f} ms)")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        rng = np.random.default_rng(1)
        idx = np.sort(rng.choice(nlines, size=min(args.cpu_sample << 2, nlines), replace=False))
        f64 = tuple(np.asarray(b, np.float64) for b in (bx, by, bz))
        ax64 = [np.asarray(a, np.float64) for a in axes]
        nt = O.num_threads()
        t3 = time.perf_counter()
        ref = O.trace_stretch(f64, ax64, seeds[idx], s0=0.01, tol_coef=1e-2, rule="bisect", nthreads=nt)
        dt = time.perf_counter() - t3
        cpu = {"value": len(idx) / dt, "unit": "lines/s", "cores": nt, "kind": "port",
               "sample": f"{len(idx)} seeds drawn uniformly (default_rng(1)) from the same {nlines} "
                         "seeds, FP64 C oracle (stretched tracer), OpenMP dynamic schedule",
               "cpu_model": cpu_model(), "oracle_evals_per_line": float(ref["evals"].mean())}
        g_ = {key: out[key][idx].cpu().numpy() for key in ("flag_start", "flag_end", "end_x", "end_y",
                                                           "end_z")}
        cls = (g_["flag_start"] == ref["flag_start"]) & (g_["flag_end"] == ref["flag_end"])
        end = np.column_stack([g_["end_x"], g_["end_y"], g_["end_z"]]).astype(np.float64)
        e = np.abs(end - ref["end"]).max(axis=1)[cls]
        extra["error_vs_oracle"] = {"lines": int(len(idx)), "class_match": float(cls.mean()),
                                    "endpoint_err_median": float(np.median(e)),
                                    "frac_endpoint_err_le_1e-4": float(np.mean(e <= 1e-4))}
        log(f"cpu baseline: {len(idx)} lines in {dt:.2f} s on {nt} threads")
    flags = torch.stack([out["flag_start"], out["flag_end"]]).cpu().numpy()
    classes, counts = np.unique(flags, return_counts=True)
    rec = {
        "metric": "field lines/s (plain RKF45, stretched grid)",
        "value": value, "unit": "lines/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32c" else "f64", "data": "synthetic",
        "config": {"workload": "random multipole (default_rng(0), scaled x10) on a stretched 256^3 "
                               "float32 grid over [-10,10]^2x[0,20] (sinh x/y, geometric z), plain "
                               "RKF45 tol_coef 1e-2 s_len 0.01, bisection cell rule, z=1e-3 plane "
                               "%dx%d" % (n, n),
                   "lines_per_step": nlines, "precision": args.precision, "schedule": args.schedule,
                   "parallelism": "single GPU"},
        "roofline": hbm_roofline("stretch", "k_plain_%s (stretched)" % args.schedule, k, evals, bpe,
                                 (args.precision, args.schedule, n) == ("fp32c", "per_line", 1024)),
        "cpu_baseline": cpu,
        "classes": {str(int(c)): int(v) for c, v in zip(classes, counts)},
        **extra,
    }
    field.close()
    return rec


def profile_values(workload):
    """key = value lines of the workload's profile summary (HBM traffic,
    VALU instructions and engine clock of its dominant kernel, as measured
    with rocprofv3 on the current build)."""
    vals = {}
    try:
        with open(os.path.join(ROOT, PROFILES[workload])) as fh:
            for line in fh:
                if "=" in line and not line.startswith((" ", "|", "#")):
                    k, v = line.split("=", 1)
                    try:
                        vals[k.strip()] = float(v.split()[0])
                    except (ValueError, IndexError):
                        pass
    except (OSError, KeyError):
        pass
    return vals


def valu_roofline(workload, kernel_ms, evals, bpe, oracle_evals=None, kernel="", flops_per_eval=FLOPS_PER_EVAL):
    """Roofline of a stepping kernel bound by VALU issue.  achieved / peak is
    the issue rate of VALU wave instructions (4 cycles per wave64 instruction
    on a SIMD16) against the chip's (1,024 SIMDs at the engine clock); the
    instruction count per launch comes from the workload's rocprofv3 profile
    (SQ_INSTS_VALU), the kernel time is measured live.  Beside it: HBM
    traffic (profile) and algorithmic flops against their peaks, and the
    no-reuse stencil model (96 B per stage evaluation)."""
    pv = profile_values(workload)
    t = kernel_ms * 1e-3
    insts = pv.get("valu_insts_per_launch")
    clk = pv.get("sclk_mhz", SCLK_PEAK_MHZ) * 1e6
    peak = N_SIMDS * clk / 4.0
    traffic = pv.get("traffic_bytes_per_launch")
    rec = {"bound": "valu", "kernel": kernel, "kernel_ms": kernel_ms,
           "unit": "G VALU wave-instructions/s",
           "achieved": insts / t / 1e9 if insts else None, "peak": peak / 1e9,
           "frac": insts / t / peak if insts else None,
           "valu_insts_per_launch": insts, "sclk_mhz": clk / 1e6, "simds": N_SIMDS,
           "traffic": traffic,
           "hbm": {"achieved_GBs": traffic / t / 1e9 if traffic else None, "peak_GBs": HBM_PEAK / 1e9,
                   "frac": traffic / t / HBM_PEAK if traffic else None},
           "flops": {"per_eval": flops_per_eval, "achieved_TFLOPs": evals * flops_per_eval / t / 1e12,
                     "peak_TFLOPs": FP32_PEAK / 1e12, "frac": evals * flops_per_eval / t / FP32_PEAK},
           "stencil_model": {"bytes_per_eval": bpe, "evals_per_launch": evals,
                             "achieved_GBs": evals * bpe / t / 1e9, "frac": evals * bpe / t / HBM_PEAK},
           "source": PROFILES.get(workload)}
    if oracle_evals is not None:
        rec["stencil_model"]["frac_at_oracle_evals"] = oracle_evals * bpe / t / HBM_PEAK
    return rec


def class_floors(field_np, seeds_sample, ref, hi, log):
    """The oracle's own spread on the sample: from seeds moved by one
    float32 ulp (the floor of an FP32 path) and in its float model (the
    reference's own float arithmetic)."""
    from oracle import oracle as O
    from tests import parity
    t0 = time.perf_counter()
    ulp = O.trace_scott(field_np, parity.ulp_moved(seeds_sample, hi, np.random.default_rng(5)), method="rkf45",
                        s0=0.25, tol=1e-6)
    f32 = O.trace_scott_f32(field_np, seeds_sample, s0=0.25, tol=1e-6)
    log(f"oracle floors in {time.perf_counter() - t0:.1f} s")
    cm = lambda a: int((~parity.class_match(a["flag_start"], a["flag_end"], ref["flag_start"],
                                            ref["flag_end"])).sum())
    fr = lambda a: float(np.mean(parity.log_q_diff(a["q0"], ref["q0"]) <= 1e-3))
    n = len(seeds_sample)
    return {"lines": n, "one_ulp_oracle": {"class_mismatches": cm(ulp), "frac_dlog10q_le_1e-3": fr(ulp)},
            "float_model": {"class_mismatches": cm(f32), "frac_dlog10q_le_1e-3": fr(f32),
                            "evals_per_line": float(f32["evals"].mean())},
            "budget_fp64_controller": cm(ulp) + 1e-4 * n, "budget_fp32_error_norm": cm(f32) + 1e-4 * n}


def bench_trace(args, fq, torch, log, volume):
    """The plane map (C2, the headline) or the 3-D cube (C3)."""
    dev = torch.device("cuda", 0)
    if args.steps is None:
        args.steps = 2 if volume else 20
    if args.warmup is None:
        args.warmup = 1 if volume else 2
    t0 = time.perf_counter()
    if volume:
        bx, by, bz = fq.synth.rope_arcade(device="cuda")
        from fastqsl_amd import qmap
        shape = bx.shape
        sx, sy, sz, _ = qmap.cube_seeds(*(np.arange(m, dtype=np.float32) for m in shape))
        # traced in 4^3 bricks of voxels, as qmap.scott_qcube does
        brick = qmap.block_order(shape, (4, 4, 4))
        sx, sy, sz = sx[brick].contiguous(), sy[brick].contiguous(), sz[brick].contiguous()
        brick_host = brick.cpu().numpy()
        nx_, ny_, nz_ = shape

        def seeds_at(idx):   # line idx traces voxel brick[idx] (cube order: x slowest, z fastest)
            v = brick_host[idx].astype(np.int64)
            return np.column_stack([v // (ny_ * nz_), (v // nz_) % ny_, v % nz_]).astype(np.float32)
        nlines = int(sx.numel())
        workload = ("Gold-Hoyle rope (1.5 turns, r 0.2, axis z 0.4) + 2 bipoles, 512x256x256 "
                    "float32, 3-D Scott Q on every grid node (traced in 4^3 bricks), RKF45 tol 1e-6")
        hi = (shape[0] - 1, shape[1] - 1)
    else:
        bx, by, bz = fq.synth.multipole_256(seed=0, device="cuda")
        n = args.pixels
        g = np.linspace(0, 255, n, dtype=np.float32)
        X, Y = np.meshgrid(g, g)
        seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size, np.float32)])
        sx, sy, sz = (torch.as_tensor(np.ascontiguousarray(seeds[:, i]), device=dev) for i in range(3))
        seeds_at = lambda idx: seeds[idx]
        nlines = seeds.shape[0]
        workload = ("random multipole (40 sources, default_rng(0)) 256^3 float32, "
                    "Scott Q, RKF45 tol 1e-6, z=0 plane %dx%d" % (n, n))
        hi = (255, 255)
    log(f"field generated in {time.perf_counter() - t0:.1f} s; {nlines} lines per step")

    layouts = ("soa64", "cell32")
    field = fq.Field(bx, by, bz, layouts=layouts)
    log(f"field resident: {field.nbytes / 1e9:.2f} GB ({', '.join(layouts)})")

    outs = {}
    kw = dict(method="rkf45", step=0.25, tol=1e-6)

    def run(precision, schedule):
        # output buffers are reused across steps, one set per path
        key = (precision, schedule)
        outs[key] = fq.trace_scott(field, sx, sy, sz, precision=precision, schedule=schedule,
                                   out=outs.get(key), **kw)
        return outs[key]

    tuning = None
    if not args.no_autotune and args.schedule == "persistent":
        from fastqsl_amd import tune
        t1 = time.perf_counter()
        cfg = tune.autotune(field, sx, sy, sz, precision=args.precision, use_cache=False, **kw)
        tuning = {"cfg": cfg, "seconds": time.perf_counter() - t1,
                  "how": "coordinate descent on this workload (a contiguous 2^22-line slice "
                         "when larger), median of 3 HIP-event timings per point"}
        log(f"autotuned in {tuning['seconds']:.1f} s: {cfg}")

    if args.calibrate:
        calibration_copy(torch, dev, log)

    def timed(precision, schedule, steps, warmup):
        for _ in range(warmup):
            run(precision, schedule)
        torch.cuda.synchronize()
        # throughput: K maps back to back on the stream, synchronised on both sides
        t0 = time.perf_counter()
        for _ in range(steps):
            run(precision, schedule)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # latency and kernel times: the same K maps, each waited for
        kernel_ms, trace_ms, step_ms = [], [], []
        for _ in range(steps):
            ts = time.perf_counter()
            run(precision, schedule)
            trace_ms.append(fq.last_kernel_ms())         # waits for the map's end event
            # the stepping kernel alone (persistent schedule); the whole trace otherwise
            kernel_ms.append(fq._lib.last_step_ms() if schedule == "persistent" else trace_ms[-1])
            step_ms.append((time.perf_counter() - ts) * 1e3)
        out = outs[(precision, schedule)]
        evals = int(out["evals"].to(torch.int64).sum().item())
        return dict(wall=wall, kms=float(np.mean(kernel_ms)), trace_ms=float(np.mean(trace_ms)),
                    step_ms=step_ms, evals=evals, value=nlines * steps / wall, ms_step=wall / steps * 1e3)

    head = timed(args.precision, args.schedule, args.steps, args.warmup)
    out = outs[(args.precision, args.schedule)]
    flags = torch.stack([out["flag_start"], out["flag_end"]]).cpu().numpy()
    value, kms, evals = head["value"], head["kms"], head["evals"]
    bpe = BYTES_PER_EVAL_F64 if args.precision == "fp64" else BYTES_PER_EVAL_F32
    log(f"{args.precision}/{args.schedule}: {value:.4g} lines/s, {head['ms_step']:.2f} ms/step, "
        f"trace {head['trace_ms']:.2f} ms, stepping kernel {kms:.2f} ms, {evals / nlines:.1f} evals/line")

    extra = {}
    if args.precision == "fp32c" and args.schedule == "persistent" and not args.no_mixed:
        m = timed("mixed", "persistent", max(3, args.steps // 4), 1)
        extra["mixed"] = {"lines_per_s": m["value"], "ms_per_step": m["ms_step"], "kernel_ms": m["kms"],
                          "evals_per_line": m["evals"] / nlines,
                          "what": "FP64 position path and error norm, FP32 Jacobian: the FP64 "
                                  "controller's step sequences (reference-parity precision)"}
        log(f"mixed/persistent: {m['value']:.4g} lines/s, {m['ms_step']:.2f} ms/step, stepping kernel "
            f"{m['kms']:.2f} ms, {m['evals'] / nlines:.1f} evals/line")
    if (volume or args.per_line_fp32) and (args.precision, args.schedule) == ("fp32c", "persistent"):
        run("fp32c", "per_line")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        run("fp32c", "per_line")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        extra["fp32c_per_line"] = {"lines_per_s": nlines / dt, "ms_per_step": dt * 1e3,
                                   "kernel_ms": fq.last_kernel_ms()}
        extra["persistent_vs_per_line_fp32c"] = value / (nlines / dt)
        log(f"fp32c/per_line: {nlines / dt:.4g} lines/s ({dt * 1e3:.1f} ms)")
    if not args.no_fp64_compare:
        run("fp64", "per_line")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        run("fp64", "per_line")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        extra["fp64_per_line"] = {"lines_per_s": nlines / dt, "ms_per_step": dt * 1e3,
                                  "kernel_ms": fq.last_kernel_ms()}
        extra["speedup_vs_fp64_per_line"] = value / (nlines / dt)
        if "mixed" in extra:
            extra["mixed"]["speedup_vs_fp64_per_line"] = extra["mixed"]["lines_per_s"] / (nlines / dt)
        log(f"fp64/per_line: {nlines / dt:.4g} lines/s ({dt * 1e3:.1f} ms)")

    cpu = None
    oracle_evals = None
    if not args.no_cpu_baseline:
        field_np = tuple(np.asfortranarray(np.asarray(b.cpu().numpy() if hasattr(b, "cpu") else b, np.float64))
                         for b in (bx, by, bz))
        cpu, idx, ref = cpu_baseline(field_np, seeds_at, nlines,
                                     args.cpu_sample if not volume else min(args.cpu_sample, 1 << 15),
                                     log)
        oracle_evals = cpu["oracle_evals_per_line"] * nlines
        floors = class_floors(field_np, seeds_at(idx), ref, hi, log) if not volume else None
        extra["error_vs_oracle"] = error_vs_oracle(out, idx, ref, floors, "fp32")
        if ("mixed", "persistent") in outs:
            extra["mixed"]["error_vs_oracle"] = error_vs_oracle(outs[("mixed", "persistent")], idx, ref, floors,
                                                                "fp64")
        if ("fp64", "per_line") in outs:
            extra["fp64_per_line"]["error_vs_oracle"] = error_vs_oracle(outs[("fp64", "per_line")], idx, ref,
                                                                        floors, "fp64")
        if floors:
            extra["oracle_floors"] = floors

    classes, counts = np.unique(flags, return_counts=True)
    prof = "volume" if volume else "plane"
    waves = fq._lib.tune_get("scott_waves") if args.precision == "fp32c" else 0
    roof = valu_roofline(prof, kms, evals, bpe, oracle_evals,
                         kernel=("stepping kernel k_scott_persistent%s" % ("_w%d" % waves if waves else "")
                                 if args.schedule == "persistent" else "k_scott_per_line"))
    if not (args.precision == "fp32c" and args.schedule == "persistent" and (volume or args.pixels == 1024)):
        roof.update(achieved=None, frac=None, valu_insts_per_launch=None, traffic=None,
                    note="the profile covers the default FP32c persistent workload only")
    rec = {
        "metric": "Scott-Q field lines/s",
        "value": value,
        "unit": "lines/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_step"],
        "latency_ms": {"p50": float(np.percentile(head["step_ms"], 50)),
                       "p99": float(np.percentile(head["step_ms"], 99)),
                       "what": "one full Q map, launch to results resident, each map waited for "
                               "(a second pass of the same K maps)"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "lines_per_step": nlines, "precision": args.precision,
                   "schedule": args.schedule, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "tuning": tuning,
        "classes": {str(int(c)): int(k) for c, k in zip(classes, counts)},
        **extra,
    }
    field.close()
    return rec


def bench_c1(args, fq, torch, log):
    """The quadrupole workload: 64^3 float64 field, Scott Q with RK4 at a fixed
    0.25-cell step, 256^2 seeds on z = 0; FP64 on the GPU (persistent and one
    thread per line), the FP32c and mixed paths beside it, the FP64 oracle on
    all 65,536 lines on the host.  Parity is counted off the field's two
    mirror planes x = y and x + y = 63 (512 pixels whose lines run into the
    central null in exact arithmetic)."""
    from tests import parity
    field_np = fq.synth.quadrupole_64(dtype=np.float64, device="cpu")
    g = np.linspace(0, 63, 256, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)]).astype(np.float32)
    sx, sy, sz = (torch.as_tensor(np.ascontiguousarray(seeds[:, i]), device="cuda") for i in range(3))
    n = seeds.shape[0]
    row, col = np.divmod(np.arange(n), 256)
    off = (row != col) & (row + col != 255)
    res = {}
    with fq.Field(*field_np) as f:
        for prec, sched in (("fp64", "persistent"), ("fp64", "per_line"), ("fp32c", "persistent"),
                            ("mixed", "persistent")):
            r = None
            for _ in range(2):
                r = fq.trace_scott(f, sx, sy, sz, method="rk4", step=0.25, precision=prec, schedule=sched)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 5
            for _ in range(reps):
                r = fq.trace_scott(f, sx, sy, sz, method="rk4", step=0.25, precision=prec, schedule=sched,
                                   out=r)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / reps
            res[f"{prec}_{sched}"] = {"lines_per_s": n / dt, "ms_per_map": dt * 1e3,
                                      "_r": {k: r[k].cpu().numpy() for k in ("flag_start", "flag_end", "q_0")}}
    from oracle import oracle as O
    nt = O.num_threads()
    t0 = time.perf_counter()
    ref = O.trace_scott(field_np, seeds, method="rk4", s0=0.25, nthreads=nt)
    dt = time.perf_counter() - t0
    for k, v in res.items():
        r = v.pop("_r")
        cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
        d = parity.log_q_diff(r["q_0"], ref["q0"])
        v["error_vs_oracle"] = {"lines_off_mirror_planes": int(off.sum()), "class_match": float(cls[off].mean()),
                                "frac_dlog10q_le_1e-3": float(np.mean(d[off] <= 1e-3)),
                                "dlog10q_max": float(d[off][np.isfinite(d[off])].max()),
                                "class_match_all_pixels": float(cls.mean())}
    log(f"C1: fp64 persistent {res['fp64_persistent']['ms_per_map']:.2f} ms, fp32c "
        f"{res['fp32c_persistent']['ms_per_map']:.2f} ms; oracle {dt:.2f} s on {nt} threads")
    return {"workload": "quadrupole (q = +-1 alternating at (+-0.3, +-0.3, -0.15)) 64^3 float64, Scott Q, RK4 "
                        "step 0.25, z=0 plane 256x256",
            "lines_per_step": n, "value": res["fp64_persistent"]["lines_per_s"], "unit": "lines/s",
            "paths": res, "cpu_baseline": {"value": n / dt, "unit": "lines/s", "cores": nt, "kind": "port",
                                           "sample": "all 65,536 lines, FP64 C oracle"}}


def summary_of(rec, keys=("value", "unit", "ms_per_step", "latency_ms", "speedup_vs_fp64_per_line")):
    out = {k: rec[k] for k in keys if k in rec}
    out["workload"] = rec.get("config", {}).get("workload")
    for k in ("fp64_per_line", "fp32c_per_line", "persistent_vs_per_line_fp32c", "one_call_per_request_fp32c",
              "error_vs_oracle", "cpu_baseline", "sweep", "fp64_plain_serial", "mixed", "tuning",
              "one_map_at_a_time", "streams", "maps_identical_to_one_at_a_time", "without_twist"):
        if rec.get(k) is not None:
            out[k] = rec[k]
    if isinstance(rec.get("roofline"), dict):
        out["roofline"] = {k: rec["roofline"].get(k) for k in ("bound", "frac", "kernel_ms", "stencil_model",
                                                               "hbm", "achieved", "peak", "unit")
                           if k in rec["roofline"]}
    return out


def other_configs(args, fq, torch, log):
    """C1, C3, C4 and C5 at their stated sizes, each shortened to a few
    timed repetitions, for the default run's record."""
    import copy
    out = {}
    t0 = time.perf_counter()
    out["C1_quadrupole_rk4"] = bench_c1(args, fq, torch, log)
    a = copy.copy(args)
    # default knobs: autotuning on the cube (10 s) picks them with the longest-first order
    a.workload, a.steps, a.warmup, a.no_autotune, a.cpu_sample = "volume", 1, 1, True, 1 << 14
    a.no_mixed = True
    out["C3_volume"] = summary_of(bench_trace(a, fq, torch, log, True))
    torch.cuda.empty_cache()
    a = copy.copy(args)
    a.workload, a.steps, a.warmup, a.no_autotune = "batch", 2, 1, True
    out["C4_batch"] = summary_of(bench_batch(a, fq, torch, log))
    torch.cuda.empty_cache()
    a = copy.copy(args)
    a.workload, a.rates, a.duration = "serve", "20", 2.0
    out["C5_serve"] = summary_of(bench_serve(a, fq, torch, log))
    torch.cuda.empty_cache()
    log(f"other configs in {time.perf_counter() - t0:.1f} s")
    return out


def other_workloads(args, fq, torch, log):
    """The workloads either side of the Scott path -- the neighbour-difference
    map, the stretched-grid tracer and the twist number -- at their bench
    sizes with their default repetitions, for the default run's record."""
    import copy
    out = {}
    t0 = time.perf_counter()
    for key, fn, workload in (("neighbour_qmap", bench_neighbour, "neighbour"),
                              ("stretched_grid", bench_stretch, "stretch"),
                              ("twist", bench_twist, "twist")):
        a = copy.copy(args)
        a.workload, a.steps, a.warmup = workload, None, None
        a.schedule = "per_line" if workload == "stretch" else "persistent"
        rec = fn(a, fq, torch, log)
        out[key] = dict(metric=rec["metric"], **summary_of(rec))
        torch.cuda.empty_cache()
    log(f"other workloads in {time.perf_counter() - t0:.1f} s")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(lines are independent, so a multi-GPU run would shard the seeds over one process per GPU)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import fastqsl_amd as fq

    log = lambda m: print(m, file=sys.stderr, flush=True)
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)

    if args.schedule is None:
        args.schedule = "per_line" if args.workload == "stretch" else "persistent"
    if args.workload == "serve":
        rec = bench_serve(args, fq, torch, log)
    elif args.workload == "stretch":
        rec = bench_stretch(args, fq, torch, log)
    elif args.workload == "neighbour":
        rec = bench_neighbour(args, fq, torch, log)
    elif args.workload == "twist":
        rec = bench_twist(args, fq, torch, log)
    elif args.workload == "batch":
        if args.steps is None:
            args.steps = 3
        if args.warmup is None:
            args.warmup = 1
        rec = bench_batch(args, fq, torch, log)
    elif args.workload == "c1":
        rec = bench_c1(args, fq, torch, log)
    else:
        rec = bench_trace(args, fq, torch, log, args.workload == "volume")
        if args.workload == "plane" and not args.headline_only:
            torch.cuda.empty_cache()
            rec["configs"] = other_configs(args, fq, torch, log)
            rec["other_workloads"] = other_workloads(args, fq, torch, log)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
