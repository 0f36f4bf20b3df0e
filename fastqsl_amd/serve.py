"""A continuous-batching Q-map server over resident field snapshots.

Requests (a snapshot, a seed plane of n x n pixels, method, step, tolerance)
arrive over time.  The server keeps every snapshot resident and runs this
loop on the host:

* requests that have arrived wait in FIFO order;
* whenever one of ``streams`` device streams is free, the waiting requests
  are admitted in order while the HBM they need (seeds, outputs and the
  batch's per-line scratch) fits the budget left by the batches still on the
  device, up to ``max_batch_lines`` lines, and traced as one batch
  (``batch.ScottBatch``: one persistent launch per snapshot and method) on
  that stream -- so a batch starts while earlier ones are still running and
  their work overlaps on the device;
* a batch's memory returns to the budget when its completion event has
  fired.

Completion times come from device events against an event recorded at the
start, so a request's latency is (device completion) - (arrival), both on the
run's clock.  ``mode="serial"`` is the plain baseline: one request at a time,
in arrival order, with ``trace_scott`` (e.g. the FP64 one-thread-per-line
kernel).

Each request's result does not depend on the batch it was traced in (the
per-line code is the same and independent of lane and launch).
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field as dc_field

import numpy as np

from .batch import ScottBatch
from .trace import trace_scott

# bytes per line on the device: seeds (12) + outputs kept (q_0, q_perp, two
# classes, evals: 20) + batch scratch (seed frame 8 words + finished-line record)
LINE_BYTES = {"fp32c": 12 + 20 + 8 * 4 + 96, "fp64": 12 + 20 + 8 * 8 + 176}
OUTPUTS = ("q_0", "q_perp", "flag_start", "flag_end", "evals")


@dataclass
class QRequest:
    rid: int
    t_arrival: float              # seconds from the start of the run
    snapshot: int
    x0: float                     # window origin on the z = z0 plane (cells)
    y0: float
    side: float                   # window side (cells)
    n: int                        # pixels per side
    z0: float = 0.0
    method: str = "rkf45"
    step: float = 0.25
    tol: float = 1e-6
    result: dict = dc_field(default=None, repr=False)
    t_done: float = math.nan

    @property
    def lines(self) -> int:
        return self.n * self.n

    def seeds(self, device="cuda"):
        import torch
        g = torch.linspace(0.0, 1.0, self.n, dtype=torch.float64, device=device)
        Y, X = torch.meshgrid(self.y0 + self.side * g, self.x0 + self.side * g, indexing="ij")
        x = X.reshape(-1).to(torch.float32)
        y = Y.reshape(-1).to(torch.float32)
        return x, y, torch.full_like(x, self.z0)


def poisson_requests(rate, duration, nsnap, shape, seed=0, side=(128, 2048), min_count=1,
                     **kw) -> list:
    """Arrivals of a Poisson process of `rate` requests/s over `duration` s
    (at least `min_count`).  Each request: snapshot uniform, n log-uniform in
    `side`, a square window on the z = 0 plane with side uniform in
    [min(64, m/2), m] cells (m = min(nx, ny) - 1) and origin uniform inside
    the field.  Draw
    order per request (default_rng(seed)): gap, snapshot, n, window side,
    x0, y0."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    nx, ny = shape[0], shape[1]
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration and len(out) >= min_count:
            break
        snap = int(rng.integers(nsnap))
        n = int(round(math.exp(rng.uniform(math.log(side[0]), math.log(side[1])))))
        m = min(nx, ny) - 1
        w = float(rng.uniform(min(64, m / 2), m))
        x0 = float(rng.uniform(0, nx - 1 - w))
        y0 = float(rng.uniform(0, ny - 1 - w))
        out.append(QRequest(len(out), t, snap, x0, y0, w, n, **kw))
    return out


class QServer:
    def __init__(self, fields, precision="fp32c", mode="batch", hbm_budget=48 << 30, streams=2,
                 max_batch_lines=1 << 24, schedule="per_line", keep_results=False):
        import torch
        self.torch = torch
        self.fields = fields
        self.precision = precision
        self.mode = mode
        self.budget = int(hbm_budget)
        self.streams = [torch.cuda.Stream() for _ in range(streams if mode == "batch" else 1)]
        self.max_batch_lines = int(max_batch_lines)
        self.schedule = schedule
        self.keep = keep_results

    def bytes_for(self, req: QRequest) -> int:
        return req.lines * LINE_BYTES[self.precision]

    def run(self, requests, log=None) -> dict:
        """Serve `requests` (sorted by arrival) in real time; returns a dict
        with per-request latencies (ms) and totals."""
        torch = self.torch
        reqs = sorted(requests, key=lambda r: r.t_arrival)
        pending = collections.deque(reqs)
        waiting = collections.deque()
        inflight = []          # (stream index, done event, bytes, requests, batch, tensors)
        free = list(range(len(self.streams)))
        used = 0
        nbatches = 0
        batch_sizes = []
        start_ev = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start_ev.record()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        while pending or waiting or inflight:
            now = clock()
            while pending and pending[0].t_arrival <= now:
                waiting.append(pending.popleft())
            # retire finished batches
            still = []
            for item in inflight:
                si, ev, nbytes, group, b, keep = item
                if ev.query():
                    t_done = start_ev.elapsed_time(ev) / 1e3
                    for r in group:
                        r.t_done = t_done
                    used -= nbytes
                    free.append(si)
                    if b is not None:
                        b.close()
                else:
                    still.append(item)
            inflight = still
            if waiting and free:
                group, nbytes, lines = [], 0, 0
                while waiting:
                    r = waiting[0]
                    rb = self.bytes_for(r)
                    fits = used + nbytes + rb <= self.budget and lines + r.lines <= self.max_batch_lines
                    if group and not fits:
                        break
                    if not group and not fits and inflight:
                        break            # wait for memory to free up
                    group.append(waiting.popleft())
                    nbytes += rb
                    lines += r.lines
                    if self.mode != "batch":
                        break
                if group:
                    si = free.pop(0)
                    s = self.streams[si]
                    with torch.cuda.stream(s):
                        b = self._launch(group, s)
                        ev = torch.cuda.Event(enable_timing=True)
                        ev.record(s)
                    used += nbytes
                    nbatches += 1
                    batch_sizes.append(len(group))
                    inflight.append((si, ev, nbytes, group, b, None))
                    continue
            if not (pending and pending[0].t_arrival <= clock()):
                time.sleep(1e-4)
        torch.cuda.synchronize()
        wall = clock()
        lat = np.array([(r.t_done - r.t_arrival) * 1e3 for r in reqs])
        lines = sum(r.lines for r in reqs)
        span = max(r.t_done for r in reqs) - reqs[0].t_arrival if reqs else 0.0
        return {"requests": len(reqs), "lines": lines, "batches": nbatches,
                "mean_batch": float(np.mean(batch_sizes)) if batch_sizes else 0.0,
                "latency_ms": lat, "p50_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
                "p99_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0,
                "lines_per_s": lines / span if span > 0 else 0.0, "wall_s": wall}

    def _launch(self, group, stream):
        if self.mode == "batch":
            specs = []
            for r in group:
                x, y, z = r.seeds()
                specs.append(dict(field=self.fields[r.snapshot], x=x, y=y, z=z, method=r.method,
                                  step=r.step, tol=r.tol))
            b = ScottBatch(specs, precision=self.precision, outputs=OUTPUTS, stream=stream)
            b.run(stream)
            for r, res in zip(group, b.results):
                r.result = res if self.keep else None
            return b
        r = group[0]
        x, y, z = r.seeds()
        res = trace_scott(self.fields[r.snapshot], x, y, z, method=r.method, step=r.step, tol=r.tol,
                          precision=self.precision, schedule=self.schedule, outputs=OUTPUTS,
                          stream=stream)
        r.result = res if self.keep else None
        return None


class NativeServer:
    """The library's own request server (``fqsl_server_*``, ``fqsl_submit``,
    ``fqsl_poll``, ``fqsl_wait``): a consumer thread in the library traces
    the requests that have arrived as one batch whenever the device is free.
    ``submit`` returns an id; the request's seeds and outputs are kept alive
    here until ``wait`` returns them."""

    def __init__(self, precision="fp32c", device=None, max_lines=0):
        import ctypes
        import torch
        from . import _lib
        self._lib = _lib
        dev = torch.cuda.current_device() if device is None else int(device)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().fqsl_server_create(dev, _lib.PRECISIONS[precision], int(max_lines),
                                                 ctypes.byref(h)))
        self._h = h
        self._live = {}

    def submit(self, field, x, y, z, method="rkf45", step=0.25, tol=1e-6, outputs=OUTPUTS, stream=None):
        import ctypes
        import torch
        from .trace import SCOTT_OUTPUTS, _dtype_of, _seed_tensors, _stream
        _lib = self._lib
        xs, ys, zs = _seed_tensors(x, y, z)
        n = xs.numel()
        res = {k: torch.empty(n, dtype=_dtype_of(k), device=xs.device) for k in outputs}
        o = _lib.fqsl_scott_out(**{k: (res[k].data_ptr() if k in res else None) for k in SCOTT_OUTPUTS})
        rq = _lib.fqsl_scott_request(field=field.handle, method=_lib.METHODS[method], reserved=0,
                                     step=float(step), tol=float(tol), n=n, x=xs.data_ptr(),
                                     y=ys.data_ptr(), z=zs.data_ptr(), out=o)
        rid = ctypes.c_uint64()
        _lib.check(_lib.lib().fqsl_submit(self._h, ctypes.byref(rq), _stream(stream), ctypes.byref(rid)))
        self._live[rid.value] = (res, (xs, ys, zs), field)
        return rid.value

    def poll(self, rid) -> bool:
        import ctypes
        done = ctypes.c_int()
        self._lib.check(self._lib.lib().fqsl_poll(self._h, int(rid), ctypes.byref(done)))
        return bool(done.value)

    def wait(self, rid) -> dict:
        rc = self._lib.lib().fqsl_wait(self._h, int(rid))
        res = self._live.pop(int(rid), (None,))[0]
        self._lib.check(rc)
        return res

    def close(self):
        if self._h:
            self._lib.check(self._lib.lib().fqsl_server_destroy(self._h))
            self._h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

