"""NumPy float32 restatement of the neighbour-difference Q assembly.

TEST INFRASTRUCTURE ONLY (the checker for the HIP Q-assembly kernels; only
``tests/`` and ``bench.py``'s ``cpu_baseline`` leg import it).

Each function follows the reference's CuPy expression tree operation by
operation, in float32, so that a kernel that evaluates the same tree without
contraction matches it bit for bit:

* ``qcalc_plane``    FastQSL.QCalcPlane           fastqslpy/FastQSL.py:65-117
* ``pseudo_seeds``   FastQSL.CookPseudoLine       fastqslpy/FastQSL.py:4-57
* ``qcut``           Qcut assembly                demo/Q_RKF45CuPy.ipynb cell [10]
* ``merge``          Q_all / Q_final merge        demo/Q_RKF45CuPy.ipynb cell [11]
* ``slog10q``        signed log map               demo/Q_RKF45CuPy.ipynb cells [13]-[14]
* ``bad_pixels``     the mixed method's selection demo/Q_RKF45MixCuPyDiffScott.ipynb cell [9]
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def _footpoint_xy(x, y, z, flag):
    """Exit-face -> 2-D coordinates (FastQSL.py:76-84): z-face (x,y),
    y-face (z,x), x-face (y,z); any other class leaves (0,0)."""
    X = np.zeros(flag.shape, f32)
    Y = np.zeros(flag.shape, f32)
    dim = np.floor_divide(flag - 1, 2)
    iz, iy, ix = dim == 2, dim == 1, dim == 0
    X[iz] = x[iz]; Y[iz] = y[iz]
    X[iy] = z[iy]; Y[iy] = x[iy]
    X[ix] = y[ix]; Y[ix] = z[ix]
    return X, Y


def qcalc_plane(x_end, y_end, z_end, flag_end, x_start, y_start, z_start, flag_start,
                bx_in, by_in, bz_in, bx_out, by_out, bz_out, bx_0, by_0, bz_0, b_flag,
                stride_step, cross_dir=(0.0, 0.0, 1.0)):
    """QCalcPlane (FastQSL.py:65-117) on (ny, nx) arrays; returns (ny-2, nx-2)."""
    a32 = lambda v: np.asarray(v, dtype=f32)
    x_end, y_end, z_end = a32(x_end), a32(y_end), a32(z_end)
    x_start, y_start, z_start = a32(x_start), a32(y_start), a32(z_start)
    bz_in, bz_out = a32(bz_in), a32(bz_out)
    bx_0, by_0, bz_0 = a32(bx_0), a32(by_0), a32(bz_0)
    fs = np.asarray(flag_start, np.int32)
    fe = np.asarray(flag_end, np.int32)
    cd = np.asarray(cross_dir, dtype=f32)

    X1, Y1 = _footpoint_xy(x_start, y_start, z_start, fs)
    X2, Y2 = _footpoint_xy(x_end, y_end, z_end, fe)

    c = (slice(1, -1), slice(1, -1))
    nan_arr = np.zeros(fs.shape, f32)
    nan_arr[c] = ((np.abs(fs[2:, 1:-1] - fs[c]) + np.abs(fs[0:-2, 1:-1] - fs[c])
                   + np.abs(fs[1:-1, 2:] - fs[c]) + np.abs(fs[1:-1, 0:-2] - fs[c]))
                  + (np.abs(fe[2:, 1:-1] - fe[c]) + np.abs(fe[0:-2, 1:-1] - fe[c])
                     + np.abs(fe[1:-1, 2:] - fe[c]) + np.abs(fe[1:-1, 0:-2] - fe[c])))

    dx2xc = X2[2:, 1:-1] - X2[0:-2, 1:-1]; dx2yc = X2[1:-1, 2:] - X2[1:-1, 0:-2]
    dy2xc = Y2[2:, 1:-1] - Y2[0:-2, 1:-1]; dy2yc = Y2[1:-1, 2:] - Y2[1:-1, 0:-2]
    dx1xc = X1[2:, 1:-1] - X1[0:-2, 1:-1]; dx1yc = X1[1:-1, 2:] - X1[1:-1, 0:-2]
    dy1xc = Y1[2:, 1:-1] - Y1[0:-2, 1:-1]; dy1yc = Y1[1:-1, 2:] - Y1[1:-1, 0:-2]

    a = dx2xc * dy1yc - dx2yc * dy1xc
    b = dx2yc * dx1xc - dx2xc * dx1yc
    cc = dy2xc * dy1yc - dy2yc * dy1xc
    d = dy2yc * dx1xc - dy2xc * dx1yc

    b_norm = bx_0 * cd[0] + by_0 * cd[1] + bz_0 * cd[2]
    bn = np.abs(b_norm[c])
    scale = f32((1.0 / (2.0 * float(stride_step))) ** 4.0)
    bnr = np.abs(bz_in[c]) / (bn * bn) * np.abs(bz_out[c]) * scale
    Q = (a * a + b * b + cc * cc + d * d) * bnr
    Q[Q < 1.0] = 1.0
    Q[nan_arr[c] > 0] = np.nan
    return Q


def pseudo_seeds(x_start, y_start, z_start, flag_start, bz_start, bz_end, b_flag, stride):
    """CookPseudoLine (FastQSL.py:4-57): four seeds per B_flag==1 pixel at its
    start footpoint +/- stride along the two in-face axes, in the order
    (+a, +b, -a, -b).  Returns (seeds (4m, 3) float32, Bz0_start, Bz0_end)."""
    sel = np.asarray(b_flag).reshape(-1) == 1
    xs = np.asarray(x_start, f32).reshape(-1)[sel]
    ys = np.asarray(y_start, f32).reshape(-1)[sel]
    zs = np.asarray(z_start, f32).reshape(-1)[sel]
    fl = np.asarray(flag_start, np.int32).reshape(-1)[sel]
    st = f32(stride)
    m = xs.shape[0]
    cx = np.repeat(xs, 4)
    cy = np.repeat(ys, 4)
    cz = np.repeat(zs, 4)
    dim = np.floor_divide(fl - 1, 2)
    iz, iy, ix = np.where(dim == 2)[0], np.where(dim == 1)[0], np.where(dim == 0)[0]
    cx[iz * 4 + 0] = xs[iz] + st; cx[iz * 4 + 2] = xs[iz] - st
    cy[iz * 4 + 1] = ys[iz] + st; cy[iz * 4 + 3] = ys[iz] - st
    cx[iy * 4 + 0] = xs[iy] + st; cx[iy * 4 + 2] = xs[iy] - st
    cz[iy * 4 + 1] = zs[iy] + st; cz[iy * 4 + 3] = zs[iy] - st
    cy[ix * 4 + 0] = ys[ix] + st; cy[ix * 4 + 2] = ys[ix] - st
    cz[ix * 4 + 1] = zs[ix] + st; cz[ix * 4 + 3] = zs[ix] - st
    seeds = np.stack([cx, cy, cz], axis=1) if m else np.zeros((0, 3), f32)
    bz0s = np.asarray(bz_start, f32).reshape(-1)[sel]
    bz0e = np.asarray(bz_end, f32).reshape(-1)[sel]
    return seeds, bz0s, bz0e


def qcut(start, flag_start, end, flag_end, bz0_start, bz0_end, stride):
    """Qcut assembly (demo/Q_RKF45CuPy.ipynb cell [10]) from the traces of the
    4m pseudo seeds; returns m values."""
    start = np.asarray(start, f32); end = np.asarray(end, f32)
    X1, Y1 = _footpoint_xy(start[:, 0], start[:, 1], start[:, 2], np.asarray(flag_start, np.int32))
    X2, Y2 = _footpoint_xy(end[:, 0], end[:, 1], end[:, 2], np.asarray(flag_end, np.int32))
    dx2xc = X2[0::4] - X2[2::4]; dx2yc = X2[1::4] - X2[3::4]
    dy2xc = Y2[0::4] - Y2[2::4]; dy2yc = Y2[1::4] - Y2[3::4]
    dx1xc = X1[0::4] - X1[2::4]; dx1yc = X1[1::4] - X1[3::4]
    dy1xc = Y1[0::4] - Y1[2::4]; dy1yc = Y1[1::4] - Y1[3::4]
    a = dx2xc * dy1yc - dx2yc * dy1xc
    b = dx2yc * dx1xc - dx2xc * dx1yc
    c = dy2xc * dy1yc - dy2yc * dy1xc
    d = dy2yc * dx1xc - dy2xc * dx1yc
    st = f32(stride)
    scale = f32((f32(1) / st / f32(2)) ** f32(4))
    bnr = np.abs(np.asarray(bz0_end, f32)) / np.abs(np.asarray(bz0_start, f32)) * scale
    Q = (a * a + b * b + c * c + d * d) * bnr
    Q[Q < 1.0] = 1.0
    return Q


def merge(Q_inner, qcut_vals, b_flag):
    """Q_all[1:-1,1:-1] = Q; Q_all[B_flag==1] = Qcut; Q_final = Q_all[1:-1,1:-1]
    (demo/Q_RKF45CuPy.ipynb cell [11])."""
    bf = np.asarray(b_flag)
    Q_all = np.zeros(bf.shape, f32)
    Q_all[1:-1, 1:-1] = Q_inner
    Q_all[bf == 1] = qcut_vals
    return Q_all[1:-1, 1:-1]


def slog10q(q, bn, flag=None):
    """np.log10(Q) * np.sign(Bn) [* (flag <= 6)], float32, left to right."""
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.log10(np.asarray(q, f32)) * np.sign(np.asarray(bn, f32))
        if flag is not None:
            r = r * (np.asarray(flag) <= 6)
    return r


def bad_pixels(b_flag, q):
    """Flat indices where B_flag + isnan(Q) > 0.5, ascending."""
    return np.where((np.asarray(b_flag).reshape(-1) + np.isnan(np.asarray(q).reshape(-1))) > 0.5)[0]
