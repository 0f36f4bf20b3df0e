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


def _device_xfast(comps, dtype=None):
    """Three (nx, ny, nz) arrays -> (pointers, dtype code, on-device flag,
    buffers to keep alive until the library has copied them)."""
    torch = _maybe_torch(comps[0])
    if torch is not None and comps[0].is_cuda:
        cs = [t.permute(2, 1, 0).contiguous() for t in comps]
        if dtype is not None:
            cs = [c.to(getattr(torch, np.dtype(dtype).name)) for c in cs]
        if cs[0].dtype not in (torch.float32, torch.float64):
            cs = [c.float() for c in cs]
        dt = _lib.DTYPE_F64 if cs[0].dtype == torch.float64 else _lib.DTYPE_F32
        return [c.data_ptr() for c in cs], dt, 1, cs
    if torch is not None:
        comps = [t.detach().cpu().numpy() for t in comps]
    src_dt = np.dtype(dtype) if dtype is not None else np.asarray(comps[0]).dtype
    if src_dt not in (np.float32, np.float64):
        src_dt = np.dtype(np.float32)
    cs = [_host_xfast(b, src_dt) for b in comps]
    dt = _lib.DTYPE_F64 if src_dt == np.float64 else _lib.DTYPE_F32
    return [c.ctypes.data for c in cs], dt, 0, cs


class Field:
    """Bx, By, Bz of shape (nx, ny, nz) (x fastest in memory, the layout the
    reference gets from ``np.transpose(sav, (2, 1, 0))``), uploaded to the
    current device with the requested layouts:

    * ``soa64`` three double arrays (FP64 tracer),
    * ``cell32`` 128-byte cells of trilinear coefficients (FP32 tracer),
    * ``soa32`` three float arrays.

    ``current`` = (Jx, Jy, Jz) on the same grid (e.g. from
    ``FastQSL.curl_of_B_grid``) enables the twist output of the tracers.
    ``axes`` = (xa, ya, za) makes the grid stretched (node (i, j, k) at
    (xa[i], ya[j], za[k])): ``trace_plain`` then runs the stretched RKF45
    tracer of TraceBlineAdaptiveStretch.cu in those coordinates, with the
    cell rule ``lookup`` ("bisect": GetIndexBiSearch, the reference with
    flag_uni false; "uniform": GetIndexUniStretch, flag_uni true; "exact":
    the cell that holds the point, landing on the box faces).
    Accepts NumPy arrays or torch tensors (CPU or GPU, shape (nx, ny, nz)).
    """

    def __init__(self, bx, by, bz, layouts=("soa64", "cell32"), dtype=None, current=None, axes=None,
                 lookup="bisect"):
        L = _lib.lib()
        mask = 0
        for name in layouts:
            mask |= {"soa32": _lib.LAYOUT_SOA32, "soa64": _lib.LAYOUT_SOA64,
                     "cell32": _lib.LAYOUT_CELL32}[name]
        self.layouts = tuple(layouts)
        self._h = ctypes.c_void_p()
        self.has_current = False
        self.shape = tuple(int(v) for v in np.shape(bx))
        ptrs, dt, on_dev, keep = _device_xfast((bx, by, bz), dtype)
        _lib.check(L.fqsl_field_create(ptrs[0], ptrs[1], ptrs[2], *self.shape, dt, on_dev, mask,
                                       ctypes.byref(self._h)))
        del keep
        self.axes = None
        self.lookup = None
        if current is not None:
            self.set_current(*current)
        if axes is not None:
            self.set_axes(*axes, lookup=lookup)

    def set_current(self, jx, jy, jz):
        """Attach the current density J (same grid as B) for the twist number."""
        if tuple(int(v) for v in np.shape(jx)) != self.shape:
            raise ValueError(f"current of shape {np.shape(jx)} on a field of shape {self.shape}")
        ptrs, dt, on_dev, keep = _device_xfast((jx, jy, jz))
        _lib.check(_lib.lib().fqsl_field_set_current(self.handle, ptrs[0], ptrs[1], ptrs[2], dt, on_dev))
        del keep
        self.has_current = True

    def set_axes(self, xa, ya, za, lookup="bisect"):
        """Node coordinates of a stretched grid (strictly increasing, one per
        node); ``set_axes(None, None, None)`` returns to grid-index
        coordinates."""
        L = _lib.lib()
        if xa is None and ya is None and za is None:
            _lib.check(L.fqsl_field_set_axes(self.handle, None, None, None, _lib.DTYPE_F64, 0, 0))
            self.axes = None
            self.lookup = None
            return
        if lookup not in _lib.LOOKUPS:
            raise ValueError(f"lookup must be one of {sorted(_lib.LOOKUPS)}")
        arrs = []
        for k, a in enumerate((xa, ya, za)):
            torch = _maybe_torch(a)
            a = a.detach().cpu().numpy() if torch is not None else np.asarray(a)
            if a.ndim != 1 or a.shape[0] != self.shape[k]:
                raise ValueError(f"axis {k} has shape {a.shape}; the field has {self.shape[k]} nodes")
            arrs.append(np.ascontiguousarray(a, dtype=np.float64 if a.dtype == np.float64 else np.float32))
        if len({a.dtype for a in arrs}) != 1:
            arrs = [a.astype(np.float64) for a in arrs]
        dt = _lib.DTYPE_F64 if arrs[0].dtype == np.float64 else _lib.DTYPE_F32
        _lib.check(L.fqsl_field_set_axes(self.handle, arrs[0].ctypes.data, arrs[1].ctypes.data,
                                         arrs[2].ctypes.data, dt, 0, _lib.LOOKUPS[lookup]))
        self.axes = tuple(a.copy() for a in arrs)
        self.lookup = lookup

    @property
    def stretched(self) -> bool:
        return bool(_lib.lib().fqsl_field_stretched(self.handle))

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
