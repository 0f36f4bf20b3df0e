"""The steps either side of the Q map that the reference's notebooks run inline.

* ``slog10q``     -- log10(Q) * sign(Bn) [* (flag <= 6)], the signed map the
  notebooks plot (demo/Q_RKF45CuPy.ipynb cells [13]-[14]).
* ``scott_fixup`` -- the mixed method of demo/Q_RKF45MixCuPyDiffScott.ipynb
  cell [9]: pixels of a neighbour-difference map with B_flag set or Q = NaN
  are traced again with the Scott method and take its q_perp.
* ``curl_index``  -- curl B in grid-index units with the one-sided z = 0 rows,
  as demo/Q_RK4.ipynb cell [11] builds the current density for the twist.

All run as HIP kernels of ``libfqsl.so``; arrays may be CUDA tensors (kept on
the device) or NumPy arrays (copied in and out).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .field import Field
from .trace import trace_scott


def _torch():
    import torch
    return torch


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _dev(a, dtype):
    """CUDA tensor of `dtype` (contiguous, flattened view kept by the caller)."""
    torch = _torch()
    t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(a))
    return t.to(device="cuda", dtype=getattr(torch, np.dtype(dtype).name)).contiguous()


def slog10q(q, bn, flag=None):
    """log10(q) * sign(bn), times (flag <= 6) when ``flag`` is given (the
    exit class of the start footpoint in the notebooks).  Same shape as q;
    NumPy in -> NumPy out."""
    torch = _torch()
    is_np = not isinstance(q, torch.Tensor)
    qt, bt = _dev(q, np.float32), _dev(bn, np.float32)
    if qt.shape != bt.shape:
        raise ValueError(f"q {tuple(qt.shape)} and bn {tuple(bt.shape)} differ")
    ft = None
    if flag is not None:
        ft = _dev(flag, np.int32)
        if ft.shape != qt.shape:
            raise ValueError("flag must have q's shape")
    out = torch.empty_like(qt)
    _lib.check(_lib.lib().fqsl_slog10q(qt.numel(), qt.data_ptr(), bt.data_ptr(),
                                       ft.data_ptr() if ft is not None else None, out.data_ptr(), _stream()))
    return out.cpu().numpy() if is_np else out


def bad_pixels(b_flag, q, x, y, z):
    """Seeds and flat pixel indices of the pixels with b_flag + isnan(q) > 0.5
    (pixel order, as cupy.where gives them).  Returns (x, y, z, pixel) CUDA
    tensors."""
    torch = _torch()
    bf = _dev(b_flag, np.int32).reshape(-1)
    qt = _dev(q, np.float32).reshape(-1)
    xs, ys, zs = (_dev(a, np.float32).reshape(-1) for a in (x, y, z))
    n = bf.numel()
    if not (qt.numel() == xs.numel() == ys.numel() == zs.numel() == n):
        raise ValueError("b_flag, q and the seeds must have one value per pixel")
    L = _lib.lib()
    cnt = ctypes.c_uint64()
    null = ctypes.c_void_p()
    _lib.check(L.fqsl_bad_pixels(n, bf.data_ptr(), qt.data_ptr(), xs.data_ptr(), ys.data_ptr(), zs.data_ptr(),
                                 0, null, null, null, null, ctypes.byref(cnt), _stream()))
    m = int(cnt.value)
    out = [torch.empty(m, dtype=torch.float32, device="cuda") for _ in range(3)]
    pix = torch.empty(m, dtype=torch.int64, device="cuda")
    if m:
        _lib.check(L.fqsl_bad_pixels(n, bf.data_ptr(), qt.data_ptr(), xs.data_ptr(), ys.data_ptr(),
                                     zs.data_ptr(), m, out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                                     pix.data_ptr(), ctypes.byref(cnt), _stream()))
    return out[0], out[1], out[2], pix


def scott_fixup(field: Field, q_all, b_flag, x, y, z, method="rkf45", step=0.125, tol=1e-6,
                precision="fp32c", schedule="persistent", use="q_perp"):
    """The mixed method (demo/Q_RKF45MixCuPyDiffScott.ipynb cell [9]): trace
    the pixels with B_flag set or Q = NaN again with the Scott method and
    write its ``use`` (q_perp, as the notebook; or q_0) into ``q_all`` in
    place.  ``q_all`` is the full neighbour-difference map (a contiguous CUDA
    float32 tensor, one value per seed); x, y, z the seeds of every pixel.
    Returns the number of pixels replaced."""
    torch = _torch()
    if not (isinstance(q_all, torch.Tensor) and q_all.is_cuda and q_all.dtype == torch.float32
            and q_all.is_contiguous()):
        raise ValueError("q_all must be a contiguous CUDA float32 tensor")
    if use not in ("q_perp", "q_0"):
        raise ValueError("use must be 'q_perp' or 'q_0'")
    bx, by, bz, pix = bad_pixels(b_flag, q_all, x, y, z)
    m = int(pix.numel())
    if m == 0:
        return 0
    r = trace_scott(field, bx, by, bz, method=method, step=step, tol=tol, precision=precision,
                    schedule=schedule, outputs=(use,))
    _lib.check(_lib.lib().fqsl_scatter_f32(m, pix.data_ptr(), r[use].data_ptr(), q_all.data_ptr(), _stream()))
    return m


def curl_index(Bx, By, Bz):
    """curl B of three (nx, ny, nz) arrays in grid-index units, the twist
    notebooks' array code (demo/Q_RK4.ipynb cell [11]); float32 or float64,
    result in the same dtype and layout (NumPy in -> NumPy out)."""
    torch = _torch()
    is_np = not isinstance(Bx, torch.Tensor)
    dt = np.float64 if (np.asarray(Bx).dtype == np.float64 if is_np else Bx.dtype == torch.float64) \
        else np.float32
    shape = tuple(int(v) for v in (np.shape(Bx) if is_np else Bx.shape))
    comps = []
    for b in (Bx, By, Bz):
        if is_np:
            t = torch.as_tensor(np.ascontiguousarray(np.asarray(b, dt).transpose(2, 1, 0))).cuda()
        else:
            t = b.to(device="cuda", dtype=getattr(torch, np.dtype(dt).name)).permute(2, 1, 0).contiguous()
        comps.append(t)
    out = [torch.empty_like(comps[0]) for _ in range(3)]
    code = _lib.DTYPE_F64 if dt == np.float64 else _lib.DTYPE_F32
    _lib.check(_lib.lib().fqsl_curl_index(comps[0].data_ptr(), comps[1].data_ptr(), comps[2].data_ptr(),
                                          *shape, code, out[0].data_ptr(), out[1].data_ptr(),
                                          out[2].data_ptr(), _stream()))
    res = tuple(o.permute(2, 1, 0) for o in out)
    if is_np:
        return tuple(np.asarray(o.cpu().numpy()) for o in res)
    return res
