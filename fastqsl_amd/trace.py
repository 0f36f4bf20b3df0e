"""Field-line tracing on the GPU: Scott-Jacobian Q and position-only traces.

Arrays live on the current CUDA (HIP) device as torch tensors; every call is
enqueued on torch's current stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .field import Field

SCOTT_OUTPUTS = ("start_x", "start_y", "start_z", "end_x", "end_y", "end_z", "flag_start",
                 "flag_end", "b_this_x", "b_this_y", "b_this_z", "q_0", "q_perp", "line_len",
                 "evals")
PLAIN_OUTPUTS = ("start_x", "start_y", "start_z", "end_x", "end_y", "end_z", "flag_start",
                 "flag_end", "b_this_x", "b_this_y", "b_this_z", "b_flag", "b_start_x",
                 "b_start_y", "b_start_z", "b_end_x", "b_end_y", "b_end_z", "line_len", "evals")
# requested explicitly (``outputs=... + ("twist",)``); needs Field(..., current=...)
OPTIONAL_OUTPUTS = ("twist",)


def _torch():
    import torch
    return torch


def _dtype_of(name):
    torch = _torch()
    if name.startswith("flag") or name == "b_flag":
        return torch.int32
    if name in ("line_len", "twist"):
        return torch.float64
    if name == "evals":
        return torch.int32  # uint32 on the device; counts stay far below 2^31
    return torch.float32


def _seed_tensors(x, y, z):
    torch = _torch()
    out = []
    for a in (x, y, z):
        if isinstance(a, torch.Tensor):
            t = a.to(device="cuda", dtype=torch.float32).contiguous()
        else:
            t = torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float32)), device="cuda")
        out.append(t.reshape(-1))
    if not (out[0].numel() == out[1].numel() == out[2].numel()):
        raise ValueError("seed arrays differ in length")
    return out


def _stream(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def trace_scott(field: Field, x, y, z, method="rkf45", step=0.25, tol=1e-6, precision="fp32c",
                schedule="persistent", outputs=SCOTT_OUTPUTS, out=None, stream=None):
    """Trace seeds (x, y, z) in grid-index coordinates with the Scott method
    (TraceAllBlineScott, fastqslpy/cu/TraceBlineScott.cu:521-595) and return a
    dict of device tensors: footpoints, classes, q_0, q_perp, B at the seed,
    line length and the number of right-hand-side evaluations per line.

    ``tol`` is the absolute per-step tolerance in cells (RKF45); ``step`` is
    the initial (RKF45) or fixed (RK4) step in cells."""
    torch = _torch()
    xs, ys, zs = _seed_tensors(x, y, z)
    n = xs.numel()
    res = dict(out) if out is not None else {}
    for name in outputs:
        if name not in SCOTT_OUTPUTS + OPTIONAL_OUTPUTS:
            raise ValueError(f"unknown output {name}")
        if name not in res:
            res[name] = torch.empty(n, dtype=_dtype_of(name), device=xs.device)
    o = _lib.fqsl_scott_out(**{k: (res[k].data_ptr() if k in res and k in outputs else None)
                               for k in SCOTT_OUTPUTS + OPTIONAL_OUTPUTS})
    p = _lib.fqsl_scott_params(method=_lib.METHODS[method], precision=_lib.PRECISIONS[precision],
                               schedule=_lib.SCHEDULES[schedule], reserved=0, step=float(step),
                               tol=float(tol))
    _lib.check(_lib.lib().fqsl_trace_scott(field.handle, ctypes.byref(p), n, xs.data_ptr(),
                                           ys.data_ptr(), zs.data_ptr(), ctypes.byref(o),
                                           _stream(stream)))
    res["_seeds"] = (xs, ys, zs)  # keep the inputs alive until the stream has run
    return res


def trace_plain(field: Field, x, y, z, ncross=(0.0, 0.0, 1.0), method="rkf45", step=0.25,
                tol_coef=1.0, precision="fp32c", schedule="persistent", outputs=PLAIN_OUTPUTS,
                out=None, stream=None, n_device=None):
    """Position-only tracing, forward then backward (TraceAllBline,
    TraceBlineAdaptive.cu:429-510 for RKF45, TraceBline.cu:383-464 for RK4).

    ``n_device``: a CUDA int64 tensor whose first element, written earlier on
    the stream, is the number of seeds to trace (at most the seed arrays'
    length); it is read on the device, so nothing waits for it on the host.
    Outputs past that count are left unwritten."""
    torch = _torch()
    xs, ys, zs = _seed_tensors(x, y, z)
    n = xs.numel()
    res = dict(out) if out is not None else {}
    for name in outputs:
        if name not in PLAIN_OUTPUTS + OPTIONAL_OUTPUTS:
            raise ValueError(f"unknown output {name}")
        if name not in res:
            res[name] = torch.empty(n, dtype=_dtype_of(name), device=xs.device)
    o = _lib.fqsl_plain_out(**{k: (res[k].data_ptr() if k in res and k in outputs else None)
                               for k in PLAIN_OUTPUTS + OPTIONAL_OUTPUTS})
    nc = (ctypes.c_float * 3)(*[float(v) for v in ncross])
    p = _lib.fqsl_plain_params(method=_lib.METHODS[method], precision=_lib.PRECISIONS[precision],
                               schedule=_lib.SCHEDULES[schedule], reserved=0, step=float(step),
                               tol_coef=float(tol_coef), ncross=nc, reserved2=0.0)
    if n_device is None:
        _lib.check(_lib.lib().fqsl_trace_plain(field.handle, ctypes.byref(p), n, xs.data_ptr(),
                                               ys.data_ptr(), zs.data_ptr(), ctypes.byref(o),
                                               _stream(stream)))
    else:
        if not (isinstance(n_device, torch.Tensor) and n_device.is_cuda
                and n_device.dtype == torch.int64 and n_device.numel() >= 1):
            raise ValueError("n_device must be a CUDA int64 tensor")
        _lib.check(_lib.lib().fqsl_trace_plain_n(field.handle, ctypes.byref(p), n,
                                                 n_device.data_ptr(), xs.data_ptr(), ys.data_ptr(),
                                                 zs.data_ptr(), ctypes.byref(o), _stream(stream)))
    res["_seeds"] = (xs, ys, zs)
    return res
