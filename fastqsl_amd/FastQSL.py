"""Drop-in for ``fastqslpy.FastQSL`` (fastqslpy/FastQSL.py) plus the Qcut
assembly the demo notebooks do inline (demo/Q_RKF45CuPy.ipynb cells [10]-[11])
and the current density for the twist number (curl_of_B_grid, diff_3point,
copy_boundary, FastQSL.py:120-163).

``QCalcPlane``, ``CookPseudoLine`` and ``ResReshape`` keep the reference's
positional signatures and return values.  Inputs may be torch CUDA tensors or
NumPy arrays; results come back in the same kind (CUDA tensors stay on the
device).  The arithmetic runs in the HIP kernels of libfqsl.so
(fqsl_qcalc_plane, fqsl_pseudo_seeds, fqsl_qcut_scatter) and follows the
reference's float32 expression trees operation by operation.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


class _Dev:
    """Device views of a mix of NumPy arrays and CUDA tensors."""

    def __init__(self):
        self.any_numpy = False
        self.keep = []

    def __call__(self, a, dtype):
        torch = _torch()
        tdt = {np.float32: torch.float32, np.int32: torch.int32}[dtype]
        if isinstance(a, torch.Tensor):
            t = a.to(device="cuda", dtype=tdt).contiguous()
        else:
            self.any_numpy = True
            t = torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=dtype)), device="cuda")
        self.keep.append(t)
        return t

    def out(self, t):
        return t.cpu().numpy() if self.any_numpy else t


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def ResReshape(shape, *args):
    """Reshape flat per-line outputs to the seed-grid shape (FastQSL.py:59-63)."""
    return [a.reshape(shape) for a in args]


def QCalcPlane(x_end_arr, y_end_arr, z_end_arr, flag_end_arr,
               x_start_arr, y_start_arr, z_start_arr, flag_start_arr,
               Bx_in_arr, By_in_arr, Bz_in_arr,
               Bx_out_arr, By_out_arr, Bz_out_arr,
               Bx_0_arr, By_0_arr, Bz_0_arr,
               B_flag_arr, stride_step, CrossBC_dir=(0.0, 0.0, 1.0)):
    """Neighbour-difference Q on a (ny, nx) seed grid (FastQSL.py:65-117);
    returns (ny-2, nx-2) float32 with Q < 1 clamped to 1 and NaN where a
    neighbour has another start or end class."""
    torch = _torch()
    d = _Dev()
    ny, nx = np.shape(flag_end_arr)
    f32, i32 = np.float32, np.int32
    xe, ye, ze = d(x_end_arr, f32), d(y_end_arr, f32), d(z_end_arr, f32)
    fe = d(flag_end_arr, i32)
    xs, ys, zs = d(x_start_arr, f32), d(y_start_arr, f32), d(z_start_arr, f32)
    fs = d(flag_start_arr, i32)
    bzin, bzout = d(Bz_in_arr, f32), d(Bz_out_arr, f32)
    b0x, b0y, b0z = d(Bx_0_arr, f32), d(By_0_arr, f32), d(Bz_0_arr, f32)
    cd = CrossBC_dir.detach().cpu().numpy() if isinstance(CrossBC_dir, torch.Tensor) else CrossBC_dir
    cdh = (ctypes.c_float * 3)(*[float(v) for v in np.asarray(cd, np.float32).reshape(3)])
    q = torch.empty((ny - 2, nx - 2), dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().fqsl_qcalc_plane(
        int(ny), int(nx), xe.data_ptr(), ye.data_ptr(), ze.data_ptr(), fe.data_ptr(),
        xs.data_ptr(), ys.data_ptr(), zs.data_ptr(), fs.data_ptr(), bzin.data_ptr(),
        bzout.data_ptr(), b0x.data_ptr(), b0y.data_ptr(), b0z.data_ptr(), float(stride_step),
        cdh, q.data_ptr(), _stream()))
    return d.out(q)


def cook_pseudo_lines(x_start, y_start, z_start, flag_start, bz_start, bz_end, b_flag, stride):
    """Device-side CookPseudoLine core: four seeds per b_flag == 1 pixel (in
    pixel order) around its start footpoint.  Returns a dict of CUDA tensors:
    seed_x/y/z (4m), bz0_start, bz0_end (m), pixel (m, flat pixel index)."""
    torch = _torch()
    d = _Dev()
    f32, i32 = np.float32, np.int32
    bf = d(b_flag, i32).reshape(-1)
    xs, ys, zs = (d(a, f32).reshape(-1) for a in (x_start, y_start, z_start))
    fs = d(flag_start, i32).reshape(-1)
    bzs, bze = d(bz_start, f32).reshape(-1), d(bz_end, f32).reshape(-1)
    n = bf.numel()
    L = _lib.lib()
    cnt = ctypes.c_uint64()
    null = ctypes.c_void_p()
    _lib.check(L.fqsl_pseudo_seeds(n, bf.data_ptr(), xs.data_ptr(), ys.data_ptr(), zs.data_ptr(),
                                   fs.data_ptr(), bzs.data_ptr(), bze.data_ptr(),
                                   ctypes.c_float(float(stride)), 0, null, null, null, null, null,
                                   null, ctypes.byref(cnt), _stream()))
    m = int(cnt.value)
    out = dict(seed_x=torch.empty(4 * m, dtype=torch.float32, device="cuda"),
               seed_y=torch.empty(4 * m, dtype=torch.float32, device="cuda"),
               seed_z=torch.empty(4 * m, dtype=torch.float32, device="cuda"),
               bz0_start=torch.empty(m, dtype=torch.float32, device="cuda"),
               bz0_end=torch.empty(m, dtype=torch.float32, device="cuda"),
               pixel=torch.empty(m, dtype=torch.int64, device="cuda"))
    if m:
        _lib.check(L.fqsl_pseudo_seeds(n, bf.data_ptr(), xs.data_ptr(), ys.data_ptr(), zs.data_ptr(),
                                       fs.data_ptr(), bzs.data_ptr(), bze.data_ptr(),
                                       ctypes.c_float(float(stride)), m, out["seed_x"].data_ptr(),
                                       out["seed_y"].data_ptr(), out["seed_z"].data_ptr(),
                                       out["bz0_start"].data_ptr(), out["bz0_end"].data_ptr(),
                                       out["pixel"].data_ptr(), ctypes.byref(cnt), _stream()))
    return out


def cook_pseudo_lines_async(x_start, y_start, z_start, flag_start, bz_start, bz_end, b_flag, stride):
    """``cook_pseudo_lines`` without waiting for the count: the arrays are
    sized for every pixel (the most there can be) and ``counts`` is a CUDA
    int64 tensor [pixels, 4 x pixels] written on the stream; only its leading
    entries of each array are set.  Feed ``counts[1:]`` to
    ``trace_plain(..., n_device=...)`` and ``counts[:1]`` to
    ``Qcut(..., m_device=...)`` so the pseudo-line pipeline never waits on the
    host."""
    torch = _torch()
    d = _Dev()
    f32, i32 = np.float32, np.int32
    bf = d(b_flag, i32).reshape(-1)
    xs, ys, zs = (d(a, f32).reshape(-1) for a in (x_start, y_start, z_start))
    fs = d(flag_start, i32).reshape(-1)
    bzs, bze = d(bz_start, f32).reshape(-1), d(bz_end, f32).reshape(-1)
    n = bf.numel()
    L = _lib.lib()
    wsb = int(L.fqsl_pseudo_workspace_bytes(n))
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    out = dict(seed_x=torch.empty(4 * n, dtype=torch.float32, device="cuda"),
               seed_y=torch.empty(4 * n, dtype=torch.float32, device="cuda"),
               seed_z=torch.empty(4 * n, dtype=torch.float32, device="cuda"),
               bz0_start=torch.empty(n, dtype=torch.float32, device="cuda"),
               bz0_end=torch.empty(n, dtype=torch.float32, device="cuda"),
               pixel=torch.empty(n, dtype=torch.int64, device="cuda"),
               counts=torch.empty(2, dtype=torch.int64, device="cuda"))
    _lib.check(L.fqsl_pseudo_seeds_async(n, bf.data_ptr(), xs.data_ptr(), ys.data_ptr(), zs.data_ptr(),
                                         fs.data_ptr(), bzs.data_ptr(), bze.data_ptr(),
                                         ctypes.c_float(float(stride)), n, out["seed_x"].data_ptr(),
                                         out["seed_y"].data_ptr(), out["seed_z"].data_ptr(),
                                         out["bz0_start"].data_ptr(), out["bz0_end"].data_ptr(),
                                         out["pixel"].data_ptr(), out["counts"].data_ptr(),
                                         ws.data_ptr(), wsb, _stream()))
    out["_keep"] = (bf, xs, ys, zs, fs, bzs, bze, ws)  # inputs and scratch live until the stream ran
    return out


def CookPseudoLine(x_end_arr, y_end_arr, z_end_arr, flag_end_arr,
                   x_start_arr, y_start_arr, z_start_arr, flag_start_arr,
                   Bx_in_arr, By_in_arr, Bz_in_arr,
                   Bx_out_arr, By_out_arr, Bz_out_arr,
                   Bz_0_arr, B_flag_arr, stride_this):
    """FastQSL.py:4-57: seeds for the pseudo lines and zeroed output arrays
    for the second TraceAllBline launch, in the reference's 25-tuple order."""
    torch = _torch()
    c = cook_pseudo_lines(x_start_arr, y_start_arr, z_start_arr, flag_start_arr, Bz_in_arr,
                          Bz_out_arr, B_flag_arr, stride_this)
    n4 = c["seed_x"].numel()
    z = lambda dt=torch.float32: torch.zeros(n4, dtype=dt, device="cuda")
    (cut_start_x, cut_start_y, cut_start_z, cut_end_x, cut_end_y, cut_end_z, Bx_start_cut,
     By_start_cut, Bz_start_cut, Bx_end_cut, By_end_cut, Bz_end_cut, Bx_inp_cut, By_inp_cut,
     Bz_inp_cut) = [z() for _ in range(15)]
    line_len_cut = z(torch.float64)
    N_cut = torch.tensor([n4], dtype=torch.int64, device="cuda")  # unsigned long long on the device
    B_flag_cut, flag_cut_start, flag_cut_end = z(torch.int32), z(torch.int32), z(torch.int32)
    res = (c["seed_x"], c["seed_y"], c["seed_z"],
           cut_start_x, cut_start_y, cut_start_z, flag_cut_start,
           cut_end_x, cut_end_y, cut_end_z, flag_cut_end,
           Bx_inp_cut, By_inp_cut, Bz_inp_cut, B_flag_cut,
           Bx_start_cut, By_start_cut, Bz_start_cut,
           Bx_end_cut, By_end_cut, Bz_end_cut, N_cut, line_len_cut,
           c["bz0_start"], c["bz0_end"])
    if not isinstance(B_flag_arr, torch.Tensor):
        res = tuple(t.cpu().numpy() for t in res)
        res = res[:21] + (res[21].astype(np.uint64),) + res[22:]
    return res


def Qcut(cut_start_x, cut_start_y, cut_start_z, flag_cut_start,
         cut_end_x, cut_end_y, cut_end_z, flag_cut_end, Bz0_start, Bz0_end, stride_this,
         Q_all=None, B_flag_arr=None, pixel=None, m_device=None):
    """Q of the pseudo-line pixels from their four traces (demo/Q_RKF45CuPy.ipynb
    cell [10]).  With ``Q_all`` and the pixel indices (or ``B_flag_arr``) the
    values are scattered into ``Q_all`` in place (cell [11],
    ``Q_all[B_flag_arr==1] = Qcut``); otherwise a length-m array is returned.
    ``m_device`` (CUDA int64 tensor, with ``Q_all`` and ``pixel``): only the
    first of its value pixels are assembled, the count read on the device."""
    torch = _torch()
    d = _Dev()
    f32, i32 = np.float32, np.int32
    sx, sy, sz = (d(a, f32).reshape(-1) for a in (cut_start_x, cut_start_y, cut_start_z))
    ex, ey, ez = (d(a, f32).reshape(-1) for a in (cut_end_x, cut_end_y, cut_end_z))
    fs, fe = d(flag_cut_start, i32).reshape(-1), d(flag_cut_end, i32).reshape(-1)
    bzs, bze = d(Bz0_start, f32).reshape(-1), d(Bz0_end, f32).reshape(-1)
    m = bzs.numel()
    if Q_all is None:
        target = torch.empty(m, dtype=torch.float32, device="cuda")
        pix = torch.arange(m, dtype=torch.int64, device="cuda")
    else:
        if not (isinstance(Q_all, torch.Tensor) and Q_all.is_cuda and Q_all.is_contiguous()):
            raise ValueError("Q_all must be a contiguous CUDA float32 tensor")
        target = Q_all
        if pixel is None:
            bf = d(B_flag_arr, i32).reshape(-1)
            pix = torch.nonzero(bf == 1).reshape(-1)
        else:
            pix = pixel.to(device="cuda", dtype=torch.int64).contiguous()
        if pix.numel() != m:
            raise ValueError(f"{pix.numel()} flagged pixels but {m} pseudo-line groups")
    if m_device is not None:
        if Q_all is None or pixel is None:
            raise ValueError("m_device needs Q_all and pixel")
        if not (isinstance(m_device, torch.Tensor) and m_device.is_cuda and m_device.dtype == torch.int64):
            raise ValueError("m_device must be a CUDA int64 tensor")
        _lib.check(_lib.lib().fqsl_qcut_scatter_n(
            m, m_device.data_ptr(), sx.data_ptr(), sy.data_ptr(), sz.data_ptr(), fs.data_ptr(),
            ex.data_ptr(), ey.data_ptr(), ez.data_ptr(), fe.data_ptr(), bzs.data_ptr(), bze.data_ptr(),
            ctypes.c_float(float(stride_this)), pix.data_ptr(), target.data_ptr(), _stream()))
        return Q_all
    _lib.check(_lib.lib().fqsl_qcut_scatter(
        m, sx.data_ptr(), sy.data_ptr(), sz.data_ptr(), fs.data_ptr(), ex.data_ptr(), ey.data_ptr(),
        ez.data_ptr(), fe.data_ptr(), bzs.data_ptr(), bze.data_ptr(), ctypes.c_float(float(stride_this)),
        pix.data_ptr(), target.data_ptr(), _stream()))
    if Q_all is None:
        return d.out(target)
    return Q_all


# ----------------------------------------------------------- current density
def _xfast_dev(a, dtype=None):
    """(nx, ny, nz) array or tensor -> x-fastest CUDA tensor, its dtype code,
    and whether the caller passed NumPy."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        t = a.to(device="cuda")
        is_np = False
    else:
        t = torch.as_tensor(np.asarray(a), device="cuda")
        is_np = True
    if dtype is not None:
        t = t.to(dtype)
    elif t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    t = t.permute(2, 1, 0).contiguous()
    code = _lib.DTYPE_F64 if t.dtype == torch.float64 else _lib.DTYPE_F32
    return t, code, is_np


def _axis(a):
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=torch.float64).reshape(-1).contiguous()
    return torch.as_tensor(np.asarray(a, np.float64).reshape(-1), device="cuda")


def _back(t, is_np):
    """x-fastest (nz, ny, nx) device tensor -> (nx, ny, nz) view, NumPy if asked."""
    v = t.permute(2, 1, 0)
    return v.cpu().numpy() if is_np else v


def curl_of_B_grid(Bx, By, Bz, arr_x, arr_y, arr_z):
    """curl B on the node grid of axes arr_x/y/z (FastQSL.py:120-131): float64
    (nx, ny, nz) arrays, three-point differences on non-uniform spacing inside,
    boundary nodes copied from the nearest interior node.  Axes are used in
    float64."""
    torch = _torch()
    bx, code, is_np = _xfast_dev(Bx)
    by, _, _ = _xfast_dev(By, bx.dtype)
    bz, _, _ = _xfast_dev(Bz, bx.dtype)
    nx, ny, nz = (int(v) for v in np.shape(Bx))
    ax, ay, az = _axis(arr_x), _axis(arr_y), _axis(arr_z)
    if (ax.numel(), ay.numel(), az.numel()) != (nx, ny, nz):
        raise ValueError("axis lengths must match the field shape")
    out = [torch.empty((nz, ny, nx), dtype=torch.float64, device="cuda") for _ in range(3)]
    _lib.check(_lib.lib().fqsl_curl(bx.data_ptr(), by.data_ptr(), bz.data_ptr(), nx, ny, nz, code,
                                    ax.data_ptr(), ay.data_ptr(), az.data_ptr(), out[0].data_ptr(),
                                    out[1].data_ptr(), out[2].data_ptr(), _stream()))
    return [_back(o, is_np) for o in out]


def diff_3point(Cube, arr_x, axis):
    """Three-point derivative along `axis` at the interior nodes
    (FastQSL.py:143-163): a float64 (nx-2, ny-2, nz-2) array."""
    torch = _torch()
    c, code, is_np = _xfast_dev(Cube)
    nx, ny, nz = (int(v) for v in np.shape(Cube))
    ax = _axis(arr_x)
    if ax.numel() != (nx, ny, nz)[axis]:
        raise ValueError("axis length must match the cube along that axis")
    out = torch.empty((nz - 2, ny - 2, nx - 2), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().fqsl_diff_3point(c.data_ptr(), nx, ny, nz, code, ax.data_ptr(), int(axis),
                                           out.data_ptr(), _stream()))
    return _back(out, is_np)


def copy_boundary(arr):
    """Boundary nodes take the value of the nearest interior node
    (FastQSL.py:133-141); in place for a CUDA float64 tensor (returned), a
    copy for anything else."""
    torch = _torch()
    inplace = isinstance(arr, torch.Tensor) and arr.is_cuda and arr.dtype == torch.float64
    t, _, is_np = _xfast_dev(arr, torch.float64)
    nx, ny, nz = (int(v) for v in np.shape(arr))
    _lib.check(_lib.lib().fqsl_copy_boundary(t.data_ptr(), nx, ny, nz, _stream()))
    res = _back(t, is_np)
    if inplace:
        arr.copy_(res)
        return arr
    return res
