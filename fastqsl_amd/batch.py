"""Batches of Scott Q-map requests over resident fields.

A ``ScottBatch`` plans many requests -- each with its own field, seeds,
method, step and tolerance -- into one device workspace
(``fqsl_scott_batch_create``) and traces them all with one persistent launch
per (field, method) group (``fqsl_scott_batch_run``).  A run only launches
kernels, so it can be captured in a HIP graph and replayed
(``ScottBatch.graph``).  Per-request completion times come from the device
clock (``ScottBatch.times_ms``).

The reference has no batch call (its notebooks launch TraceAllBlineScott per
map, demo/Q_RKF45CuPy3DScott.ipynb cell [7]); each request's results are those
of ``trace_scott`` on its own seeds, bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .trace import SCOTT_OUTPUTS, _dtype_of, _seed_tensors, _stream


class ScottBatch:
    """requests: sequence of dicts with keys field (Field), x, y, z (seed
    arrays or tensors), and optionally method ("rkf45"), step (0.25),
    tol (1e-6)."""

    def __init__(self, requests, precision="fp32c", outputs=("q_0", "q_perp", "flag_start",
                                                              "flag_end", "evals"), stream=None):
        import torch
        self.precision = precision
        self.requests = []
        self.results = []
        table = (_lib.fqsl_scott_request * len(requests))()
        for k, rq in enumerate(requests):
            xs, ys, zs = _seed_tensors(rq["x"], rq["y"], rq["z"])
            n = xs.numel()
            res = {}
            for name in outputs:
                if name not in SCOTT_OUTPUTS:
                    raise ValueError(f"unknown output {name}")
                res[name] = torch.empty(n, dtype=_dtype_of(name), device=xs.device)
            o = _lib.fqsl_scott_out(**{kk: (res[kk].data_ptr() if kk in res else None)
                                       for kk in SCOTT_OUTPUTS})
            method = rq.get("method", "rkf45")
            table[k] = _lib.fqsl_scott_request(
                field=rq["field"].handle, method=_lib.METHODS[method], reserved=0,
                step=float(rq.get("step", 0.25)), tol=float(rq.get("tol", 1e-6)), n=n,
                x=xs.data_ptr(), y=ys.data_ptr(), z=zs.data_ptr(), out=o)
            self.requests.append(dict(rq, seeds=(xs, ys, zs), n=n, method=method))
            self.results.append(res)
        self._table = table
        L = _lib.lib()
        prec = _lib.PRECISIONS[precision]
        nbytes = int(L.fqsl_scott_batch_workspace_bytes(table, len(requests), prec))
        # torch's caching allocator returns blocks aligned to 512 bytes
        self.workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
        self.workspace_bytes = nbytes
        h = ctypes.c_void_p()
        _lib.check(L.fqsl_scott_batch_create(table, len(requests), prec, self.workspace.data_ptr(),
                                             nbytes, _stream(stream), ctypes.byref(h)))
        self._h = h
        self.lines = sum(r["n"] for r in self.requests)

    def run(self, stream=None):
        _lib.check(_lib.lib().fqsl_scott_batch_run(self._h, _stream(stream)))

    def graph(self):
        """A torch CUDA graph (HIP graph) holding one run; ``g.replay()``
        re-traces every request."""
        import torch
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.run(s)                        # warm-up outside capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.run(s)
        torch.cuda.synchronize()
        return g

    def times_ms(self, stream=None) -> np.ndarray:
        """Per request: ms from the start of the last run until its last line
        finished stepping (device clock)."""
        out = (ctypes.c_double * len(self.requests))()
        _lib.check(_lib.lib().fqsl_scott_batch_times(self._h, out, _stream(stream)))
        return np.frombuffer(out, dtype=np.float64).copy()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().fqsl_scott_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
