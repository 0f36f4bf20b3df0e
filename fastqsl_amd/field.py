"""A magnetic field held resident on the GPU (``fqsl_field`` handle)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _host_xfast(b, dtype):
    """(nx, ny, nz) array -> contiguous x-fastest buffer of `dtype`."""
    a = np.asarray(b)
    if a.ndim != 3:
        raise ValueError("field components must be 3-D arrays of shape (nx, ny, nz)")
    return np.ascontiguousarray(a.transpose(2, 1, 0), dtype=dtype)


class Field:
    """Bx, By, Bz of shape (nx, ny, nz) (x fastest in memory, the layout the
    reference gets from ``np.transpose(sav, (2, 1, 0))``), uploaded to the
    current device with the requested layouts:

    * ``soa64`` three double arrays (FP64 tracer),
    * ``cell32`` 128-byte cell stencils (FP32 tracer),
    * ``soa32`` three float arrays.

    Accepts NumPy arrays or torch tensors (CPU or GPU, shape (nx, ny, nz)).
    """

    def __init__(self, bx, by, bz, layouts=("soa64", "cell32"), dtype=None):
        L = _lib.lib()
        mask = 0
        for name in layouts:
            mask |= {"soa32": _lib.LAYOUT_SOA32, "soa64": _lib.LAYOUT_SOA64,
                     "cell32": _lib.LAYOUT_CELL32}[name]
        self.layouts = tuple(layouts)
        self._h = ctypes.c_void_p()
        self._keep = None
        torch = _maybe_torch(bx)
        if torch is not None and bx.is_cuda:
            comps = [t.permute(2, 1, 0).contiguous() for t in (bx, by, bz)]
            if dtype is not None:
                comps = [c.to(getattr(torch, np.dtype(dtype).name)) for c in comps]
            if comps[0].dtype not in (torch.float32, torch.float64):
                comps = [c.float() for c in comps]
            dt = _lib.DTYPE_F64 if comps[0].dtype == torch.float64 else _lib.DTYPE_F32
            shape = tuple(bx.shape)
            ptrs = [c.data_ptr() for c in comps]
            on_dev = 1
            self._keep = comps
        else:
            if torch is not None:
                bx, by, bz = (t.detach().cpu().numpy() for t in (bx, by, bz))
            src_dt = np.dtype(dtype) if dtype is not None else np.asarray(bx).dtype
            if src_dt not in (np.float32, np.float64):
                src_dt = np.dtype(np.float32)
            comps = [_host_xfast(b, src_dt) for b in (bx, by, bz)]
            dt = _lib.DTYPE_F64 if src_dt == np.float64 else _lib.DTYPE_F32
            shape = np.asarray(bx).shape
            ptrs = [c.ctypes.data for c in comps]
            on_dev = 0
        self.shape = tuple(int(s) for s in shape)
        _lib.check(L.fqsl_field_create(ptrs[0], ptrs[1], ptrs[2], *self.shape, dt, on_dev, mask,
                                       ctypes.byref(self._h)))
        self._keep = None

    @property
    def handle(self):
        if not self._h:
            raise _lib.FqslError("field has been closed")
        return self._h

    @property
    def nbytes(self) -> int:
        return int(_lib.lib().fqsl_field_bytes(self.handle))

    def close(self):
        if self._h:
            _lib.lib().fqsl_field_destroy(self._h)
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


def _maybe_torch(x):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return None
    return torch if isinstance(x, torch.Tensor) else None
