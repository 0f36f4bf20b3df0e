"""Autotuner for the persistent tracers' launch parameters.

The knobs (``_lib.TUNE_KNOBS``) change speed only -- a line's result does not
depend on the lane, wave, queue or launch that traces it -- so they can be
picked per workload by measurement:

* ``scott_waves``  waves per SIMD the FP32c Scott stepping kernel is built for
                   (3 = 168 VGPRs per lane, 2 = 256; 0 = the compiler's choice);
* ``refill_min``   idle lanes a wavefront gathers before it claims new lines
                   (the queue chunk: one atomic per chunk);
* ``queues``       one seed queue, or one per XCD (8), the launch's blocks
                   spread over them;
* ``grid_pct``     persistent grid size in % of one full occupancy wave;
* ``order``        queue order: 2 = longest line first (by a coarse length
                   estimate), 0 = seed order, 1 = longest first only up to 64
                   lines per persistent lane;
* ``order_buckets``length classes of that order (fewer keep more locality).

``autotune`` runs coordinate descent over these on the actual request, or on
a contiguous slice of it when it is larger than ``sample`` lines (contiguous,
because neighbouring seeds share cache lines of the field: a random subset
has a different L2 behaviour and picks different knobs), keeps the fastest
by the HIP-event time of the whole trace (prep, stepping, Q assembly),
applies it and caches it per (device, field shape, method, precision,
tolerance decade) in a JSON file.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import _lib

SPACE = {
    "order": (1, 2, 0),
    "order_buckets": (64, 16, 256),
    "scott_waves": (3, 2, 0),
    "refill_min": (8, 4, 16, 32, 1),
    "queues": (1, 8),
    "grid_pct": (100, 75, 50),
}
DEFAULTS = {"order": 1, "order_buckets": 64, "scott_waves": 3, "refill_min": 8, "queues": 1,
            "grid_pct": 100}


def cache_path() -> str:
    return os.environ.get("FQSL_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "fastqsl_amd", "tune.json"))


def cache_key(device: str, shape, method: str, precision: str, tol: float) -> str:
    dec = "none" if not tol or tol <= 0 else "1e" + str(int(math.floor(math.log10(tol) + 1e-9)))
    return f"{device}|{'x'.join(str(int(s)) for s in shape)}|{method}|{precision}|tol{dec}"


def load_cache(path=None) -> dict:
    try:
        with open(path or cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def save_cache(data: dict, path=None) -> None:
    p = path or cache_path()
    os.makedirs(os.path.dirname(p), exist_ok=True)
    tmp = p + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, p)


def current() -> dict:
    """The Scott tracer's knobs (the ones the autotuner searches)."""
    return {k: _lib.tune_get(k) for k in SPACE}


def apply(cfg: dict) -> None:
    for k, v in cfg.items():
        _lib.tune_set(k, int(v))


def coordinate_descent(measure, space=SPACE, start=DEFAULTS, rounds=2, log=None):
    """Minimise measure(cfg) -> ms over `space`, one knob at a time, starting
    from `start`; stops after a round without improvement.  Returns
    (best cfg, best ms, {cfg as tuple: ms} of every point measured)."""
    seen = {}

    def m(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            seen[key] = float(measure(dict(cfg)))
            if log:
                log(f"tune {dict(cfg)}: {seen[key]:.3f} ms")
        return seen[key]

    best = dict(start)
    best_ms = m(best)
    for _ in range(rounds):
        improved = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best, **{knob: v})
                ms = m(cand)
                if ms < best_ms * 0.98:          # keep a change only for a clear gain
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    return best, best_ms, seen


def autotune(field, x, y, z, method="rkf45", step=0.25, tol=1e-6, precision="fp32c",
             sample=1 << 22, repeats=3, use_cache=True, log=None):
    """Pick the knobs for tracing seeds (x, y, z) on `field` with the persistent
    Scott tracer; applies them and returns the chosen configuration."""
    import torch
    from .trace import trace_scott
    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    key = cache_key(dev, field.shape, method, precision, tol)
    if use_cache:
        hit = load_cache().get(key)
        if hit:
            apply(hit["cfg"])
            return dict(hit["cfg"])
    xs, ys, zs = (torch.as_tensor(np.asarray(a, np.float32)) if not isinstance(a, torch.Tensor) else a
                  for a in (x, y, z))
    n = int(xs.numel())
    if n > sample:
        lo = (n - sample) // 2
        xs, ys, zs = (a.reshape(-1)[lo:lo + sample] for a in (xs, ys, zs))
    before = current()
    out = {}

    def measure(cfg):
        apply(cfg)
        ms = []
        for r in range(repeats + 1):
            res = trace_scott(field, xs, ys, zs, method=method, step=step, tol=tol,
                              precision=precision, schedule="persistent",
                              outputs=("q_0", "flag_start", "flag_end"), out=out.get("o"))
            out["o"] = res
            t = _lib.last_kernel_ms()
            if r:
                ms.append(t)
        return float(np.median(ms))

    try:
        best, best_ms, seen = coordinate_descent(measure, log=log)
    except Exception:
        apply(before)
        raise
    apply(best)
    if use_cache:
        data = load_cache()
        data[key] = {"cfg": best, "ms": best_ms, "lines": int(xs.numel()),
                     "points": len(seen)}
        save_cache(data)
    return best


BATCH_SPACE = {
    "order": (1, 2, 0),
    "order_buckets": (64, 16, 256),
    "scott_waves": (3, 2, 0),
    "refill_min": (8, 4, 16, 32),
    "grid_pct": (100, 75),
}


def autotune_batch(requests, precision="fp32c", outputs=("q_0", "flag_start", "flag_end"),
                   max_lines=1 << 23, log=None):
    """Pick the knobs for a batch of requests (``batch.ScottBatch``): the
    stepping kernel's waves per SIMD and refill chunk are fixed when a batch is
    planned, the grid size when it runs.  Measures on the leading requests up
    to `max_lines` lines (each point plans a fresh batch and times its second
    run), applies the best configuration and returns it."""
    from .batch import ScottBatch
    sub, lines = [], 0
    for r in requests:
        n = int(r["x"].numel()) if hasattr(r["x"], "numel") else len(r["x"])
        if sub and lines + n > max_lines:
            break
        sub.append(r)
        lines += n

    def measure(cfg):
        apply(cfg)
        with ScottBatch(sub, precision=precision, outputs=outputs) as b:
            b.run()
            _lib.last_kernel_ms()
            b.run()
            return _lib.last_kernel_ms()

    before = current()
    try:
        best, best_ms, _ = coordinate_descent(measure, space=BATCH_SPACE,
                                              start={k: DEFAULTS[k] for k in BATCH_SPACE}, rounds=1,
                                              log=log)
    except Exception:
        apply(before)
        raise
    apply(best)
    return best
