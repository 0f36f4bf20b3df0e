"""NumPy restatement of the current-density helpers of fastqslpy.FastQSL.

TEST INFRASTRUCTURE ONLY (the checker for the HIP curl kernel and the source
of J for the twist tests; the product never imports it).

* ``diff_3point``     FastQSL.diff_3point      fastqslpy/FastQSL.py:143-163
* ``copy_boundary``   FastQSL.copy_boundary    fastqslpy/FastQSL.py:133-141
* ``curl_of_B_grid``  FastQSL.curl_of_B_grid   fastqslpy/FastQSL.py:120-131
* ``curl_index``      the notebooks' index-space curl with one-sided z = 0 rows
                                               demo/Q_RK4.ipynb cell [11]

The three-point derivative on a non-uniform axis is the weighted mean of the
two one-sided differences, each weighted by the other side's spacing; arrays
are indexed [x, y, z] as the reference's (its fields carry x fastest in
memory).  The arithmetic follows the reference's float64 expression order.
"""
from __future__ import annotations

import numpy as np


def diff_3point(cube, arr, axis):
    cube = np.asarray(cube)
    dx = np.diff(np.asarray(arr, np.float64))
    shp = [1, 1, 1]
    shp[axis] = dx.size
    d = dx.reshape(shp)
    sl = lambda a, b: tuple(slice(a, b) if ax == axis else slice(1, -1) for ax in range(3))
    dl = tuple(slice(0, -1) if ax == axis else slice(None) for ax in range(3))
    dr = tuple(slice(1, None) if ax == axis else slice(None) for ax in range(3))
    c0, cp, cm = cube[sl(1, -1)], cube[sl(2, None)], cube[sl(0, -2)]
    return ((cp - c0) / d[dr] * d[dl] / (d[dr] + d[dl])
            + (c0 - cm) / d[dl] * d[dr] / (d[dr] + d[dl]))


def copy_boundary(arr):
    arr[:, :, 0] = arr[:, :, 1]
    arr[:, 0, :] = arr[:, 1, :]
    arr[0, :, :] = arr[1, :, :]
    arr[:, :, -1] = arr[:, :, -2]
    arr[:, -1, :] = arr[:, -2, :]
    arr[-1, :, :] = arr[-2, :, :]
    return arr


def curl_of_B_grid(bx, by, bz, arr_x, arr_y, arr_z):
    """curl B on the (nx, ny, nz) node grid with axes arr_x/y/z; float64."""
    cx = np.zeros(np.shape(bx))
    cy = np.zeros(np.shape(by))
    cz = np.zeros(np.shape(bz))
    c = (slice(1, -1),) * 3
    cx[c] = diff_3point(bz, arr_y, 1) - diff_3point(by, arr_z, 2)
    cy[c] = diff_3point(bx, arr_z, 2) - diff_3point(bz, arr_x, 0)
    cz[c] = diff_3point(by, arr_x, 0) - diff_3point(bx, arr_y, 1)
    return copy_boundary(cx), copy_boundary(cy), copy_boundary(cz)


def curl_index(bx, by, bz):
    """curl B in grid-index units as demo/Q_RK4.ipynb cell [11] writes it
    (arrays [x, y, z], B's dtype, the notebook's slices and operation
    order): central differences / 2 on the sliced index ranges, then at z = 0
    the one-sided (-3 c0 + 4 c1 - c2) / 2 along z for the x and y
    components; zero elsewhere."""
    bx, by, bz = (np.asarray(b) for b in (bx, by, bz))
    cx, cy, cz = np.zeros_like(bx), np.zeros_like(by), np.zeros_like(bz)
    cx[:, 1:-1, 1:-1] = ((bz[:, 2:, 1:-1] - bz[:, 0:-2, 1:-1]) / 2.
                         - (by[:, 1:-1, 2:] - by[:, 1:-1, 0:-2]) / 2)
    cy[1:-1, :, 1:-1] = ((bx[1:-1, :, 2:] - bx[1:-1, :, 0:-2]) / 2.
                         - (bz[2:, :, 1:-1] - bz[0:-2, :, 1:-1]) / 2)
    cz[1:-1, 1:-1, :] = ((by[2:, 1:-1, :] - by[0:-2, 1:-1, :]) / 2.
                         - (bx[1:-1, 2:, :] - bx[1:-1, 0:-2, :]) / 2)
    cx[1:-1, 1:-1, 0] = ((bz[1:-1, 2:, 0] - bz[1:-1, 0:-2, 0]) / 2.
                         - (-3. * by[1:-1, 1:-1, 0] + 4. * by[1:-1, 1:-1, 1] - by[1:-1, 1:-1, 2]) / 2)
    cy[1:-1, 1:-1, 0] = ((-3. * bx[1:-1, 1:-1, 0] + 4. * bx[1:-1, 1:-1, 1] - bx[1:-1, 1:-1, 2]) / 2.
                         - (bz[2:, 1:-1, 0] - bz[0:-2, 1:-1, 0]) / 2)
    return cx, cy, cz
