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
BYTES_PER_EVAL_F32 = 96    # 8 corners x 3 components x 4 B per RHS evaluation
BYTES_PER_EVAL_F64 = 192


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pixels", type=int, default=1024)
    ap.add_argument("--precision", default="fp32c", choices=["fp32c", "fp64"])
    ap.add_argument("--schedule", default="persistent", choices=["persistent", "per_line"])
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 1 GiB device-to-device copy (known bytes) for counter scaling")
    return ap.parse_args()


PROFILE = os.path.join(ROOT, "profiles", "scott_rkf45_multipole256_1024sq.md")


def measured_traffic():
    """HBM bytes per launch of the stepping kernel as measured with rocprofv3
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated) and recorded in the profile
    summary for this workload; None if absent."""
    try:
        with open(PROFILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def cpu_baseline(field_np, seeds, sample, log):
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    idx = rng.choice(len(seeds), size=min(sample, len(seeds)), replace=False)
    nt = O.num_threads()
    t0 = time.perf_counter()
    O.trace_scott(field_np, seeds[idx], method="rkf45", s0=0.25, tol=1e-6, nthreads=nt)
    dt = time.perf_counter() - t0
    log(f"cpu baseline: {len(idx)} lines in {dt:.2f} s on {nt} threads")
    return {"value": len(idx) / dt, "unit": "lines/s", "cores": nt, "kind": "port",
            "sample": f"{len(idx)} seeds drawn uniformly (default_rng(1)) from the same "
                      f"{len(seeds)}-seed plane, FP64 C oracle, OpenMP dynamic schedule"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(lines are independent, so a multi-GPU run shards seeds; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import fastqsl_amd as fq

    log = lambda m: print(m, file=sys.stderr, flush=True)
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)

    t0 = time.perf_counter()
    bx, by, bz = fq.synth.multipole_256(seed=0, device="cuda")
    log(f"field generated in {time.perf_counter() - t0:.1f} s")
    n = args.pixels
    g = np.linspace(0, 255, n, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size, np.float32)])
    sx, sy, sz = (torch.as_tensor(np.ascontiguousarray(seeds[:, i]), device=dev) for i in range(3))
    nlines = seeds.shape[0]

    layouts = ("soa64", "cell32")
    field = fq.Field(bx, by, bz, layouts=layouts)
    log(f"field resident: {field.nbytes / 1e9:.2f} GB ({', '.join(layouts)})")

    out = None
    kw = dict(method="rkf45", step=0.25, tol=1e-6)

    def run(precision, schedule):
        nonlocal out
        out = fq.trace_scott(field, sx, sy, sz, precision=precision, schedule=schedule,
                             out=out, **kw)
        return out

    if args.calibrate:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 1073741824 B read + 1073741824 B written")
        del src, dst

    for _ in range(args.warmup):
        run(args.precision, args.schedule)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(args.precision, args.schedule)
        kernel_ms.append(fq.last_kernel_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    evals = int(out["evals"].to(torch.int64).sum().item())
    flags = torch.stack([out["flag_start"], out["flag_end"]]).cpu().numpy()
    ms_step = wall / args.steps * 1e3
    value = nlines * args.steps / wall
    kms = float(np.mean(kernel_ms))
    bpe = BYTES_PER_EVAL_F32 if args.precision == "fp32c" else BYTES_PER_EVAL_F64
    achieved = evals * bpe / (kms * 1e-3)
    log(f"{args.precision}/{args.schedule}: {value:.4g} lines/s, {ms_step:.2f} ms/step, "
        f"kernel {kms:.2f} ms, {evals / nlines:.1f} evals/line")

    extra = {}
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
        log(f"fp64/per_line: {nlines / dt:.4g} lines/s ({dt * 1e3:.1f} ms)")

    cpu = None
    if not args.no_cpu_baseline:
        field_np = tuple(np.asarray(b, dtype=np.float64) for b in (bx, by, bz))
        cpu = cpu_baseline(field_np, seeds, args.cpu_sample, log)

    classes, counts = np.unique(flags, return_counts=True)
    rec = {
        "metric": "Scott-Q field lines/s",
        "value": value,
        "unit": "lines/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32c" else "f64",
        "data": "synthetic",
        "config": {"workload": "random multipole (40 sources, default_rng(0)) 256^3 float32, "
                               "Scott Q, RKF45 tol 1e-6, z=0 plane %dx%d" % (n, n),
                   "lines_per_step": nlines, "precision": args.precision,
                   "schedule": args.schedule, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": (measured_traffic() if args.precision == "fp32c" and n == 1024 else None),
                     "traffic_source": "profiles/scott_rkf45_multipole256_1024sq.md (rocprofv3 "
                                       "FETCH_SIZE x2.00 + WRITE_SIZE x1.00, copy-calibrated)",
                     "kernel_ms": kms, "evals_per_launch": evals, "bytes_per_eval": bpe},
        "cpu_baseline": cpu,
        "classes": {str(int(c)): int(k) for c, k in zip(classes, counts)},
        **extra,
    }
    print(json.dumps(rec), flush=True)
    field.close()


if __name__ == "__main__":
    main()
