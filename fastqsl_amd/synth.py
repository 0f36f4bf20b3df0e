"""Synthetic magnetic fields and seed grids for the benchmark workloads.

All fields are sampled at grid nodes and returned as three arrays of shape
(nx, ny, nz) -- the layout the reference gets from
``np.transpose(readsav(...), (2, 1, 0))`` (demo/Q_RKF45CuPy.ipynb cell [2]) --
stored x-fastest.  Tracing happens in grid-index coordinates (cells).

Workloads:

* ``quadrupole_64``     4 alternating point sources, 64^3 float64
* ``multipole_256``     40 random point sources (default_rng(seed)), 256^3 float32
* ``rope_arcade``       Gold-Hoyle flux rope along x plus two bipoles, 512x256x256
* ``multipole``         the random-multipole generator at any shape/seed
* ``stretched_multipole`` the same sources on a stretched (non-uniform) grid,
                        traced in the axes' coordinates

Generation runs in torch (on the GPU when one is present) in float64 and is
cast to the requested dtype at the end; the result is returned as NumPy.
"""
from __future__ import annotations

import math

import numpy as np


def _torch():
    import torch
    return torch


def _axes(shape, box, device, torch):
    (x0, x1), (y0, y1), (z0, z1) = box
    nx, ny, nz = shape
    xs = torch.linspace(x0, x1, nx, dtype=torch.float64, device=device)
    ys = torch.linspace(y0, y1, ny, dtype=torch.float64, device=device)
    zs = torch.linspace(z0, z1, nz, dtype=torch.float64, device=device)
    return xs, ys, zs


def _to_numpy(bx, by, bz, dtype):
    # torch tensors are (nz, ny, nx) C-order (x fastest); hand back (nx, ny, nz) views
    out = []
    for b in (bx, by, bz):
        a = b.to("cpu").numpy().astype(dtype, copy=False)
        out.append(a.transpose(2, 1, 0))
    return tuple(out)


def _default_device(device):
    torch = _torch()
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    return device


def point_sources(shape, charges, positions, box=((-1, 1), (-1, 1), (0, 2)), dtype=np.float32,
                  device=None, zchunk=16, axes=None):
    """B = sum_i q_i (r - r_i) / |r - r_i|^3 sampled at the grid nodes (evenly
    spaced over ``box``, or at the node coordinates ``axes`` = (xa, ya, za))."""
    torch = _torch()
    device = _default_device(device)
    if axes is None:
        xs, ys, zs = _axes(shape, box, device, torch)
    else:
        xs, ys, zs = (torch.as_tensor(np.asarray(a, dtype=np.float64), device=device) for a in axes)
        shape = (xs.shape[0], ys.shape[0], zs.shape[0])
    nx, ny, nz = shape
    bx = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    by = torch.zeros_like(bx)
    bz = torch.zeros_like(bx)
    q = torch.as_tensor(np.asarray(charges, dtype=np.float64), device=device)
    P = torch.as_tensor(np.asarray(positions, dtype=np.float64).reshape(-1, 3), device=device)
    Y, X = torch.meshgrid(ys, xs, indexing="ij")  # (ny, nx)
    for k0 in range(0, nz, zchunk):
        k1 = min(nz, k0 + zchunk)
        Z = zs[k0:k1].view(-1, 1, 1)
        for i in range(q.shape[0]):
            dx = X - P[i, 0]
            dy = Y - P[i, 1]
            dz = Z - P[i, 2]
            r2 = dx * dx + dy * dy + dz * dz
            w = q[i] / (r2 * torch.sqrt(r2))
            bx[k0:k1] += w * dx
            by[k0:k1] += w * dy
            bz[k0:k1] += w * dz
    return _to_numpy(bx, by, bz, dtype)


def quadrupole_64(dtype=np.float64, device=None, n=64):
    """Four sub-surface point sources q=(+1,-1,+1,-1) at (+.3,+.3), (-.3,+.3),
    (-.3,-.3), (+.3,-.3), depth 0.15, in [-1,1]^2 x [0,2] on an n^3 grid."""
    charges = [1.0, -1.0, 1.0, -1.0]
    pos = [[0.3, 0.3, -0.15], [-0.3, 0.3, -0.15], [-0.3, -0.3, -0.15], [0.3, -0.3, -0.15]]
    return point_sources((n, n, n), charges, pos, dtype=dtype, device=device)


def multipole_sources(seed=0, count=40):
    """Draw order: x ~ U[-1,1], y ~ U[-1,1], depth ~ U[0.05,0.3], q ~ N(0,1),
    each as one vector of ``count`` draws from ``default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, count)
    y = rng.uniform(-1.0, 1.0, count)
    depth = rng.uniform(0.05, 0.3, count)
    q = rng.normal(0.0, 1.0, count)
    return q, np.stack([x, y, -depth], axis=1)


def multipole(shape=(256, 256, 256), seed=0, count=40, dtype=np.float32, device=None):
    """Random multipolar potential field in [-1,1]^2 x [0,2]."""
    q, pos = multipole_sources(seed, count)
    return point_sources(shape, q, pos, dtype=dtype, device=device)


def multipole_256(seed=0, dtype=np.float32, device=None):
    return multipole((256, 256, 256), seed=seed, dtype=dtype, device=device)


def rope_arcade(shape=(512, 256, 256), dtype=np.float32, device=None, turns=1.5, radius=0.2,
                axis_z=0.4, b0=1.0, bipole_q=1.0):
    """Uniform-twist Gold-Hoyle rope along x (B_x = B0/(1+T^2 r^2),
    B_phi = B0 T r/(1+T^2 r^2), T = 2 pi turns / L, cut at r <= radius, axis at
    (y=0, z=axis_z)) plus two point-source bipoles: +q at (-0.5, +-0.5, -0.1)
    and -q at (+0.5, +-0.5, -0.1).  Box x in [-2,2], y in [-1,1], z in [0,2]."""
    torch = _torch()
    device = _default_device(device)
    box = ((-2, 2), (-1, 1), (0, 2))
    charges = [bipole_q, -bipole_q, bipole_q, -bipole_q]
    pos = [[-0.5, 0.5, -0.1], [0.5, 0.5, -0.1], [-0.5, -0.5, -0.1], [0.5, -0.5, -0.1]]
    bx, by, bz = point_sources(shape, charges, pos, box=box, dtype=np.float64, device=device)
    xs, ys, zs = _axes(shape, box, device, torch)
    L = box[0][1] - box[0][0]
    T = 2 * math.pi * turns / L
    Z, Y = torch.meshgrid(zs, ys, indexing="ij")  # (nz, ny)
    dy = Y
    dz = Z - axis_z
    r = torch.sqrt(dy * dy + dz * dz)
    inside = (r <= radius).to(torch.float64)
    den = 1.0 + T * T * r * r
    bxr = b0 / den * inside
    bphi = b0 * T * r / den * inside
    # phi direction about +x: (0, -dz, dy)/r
    rs = torch.where(r > 0, r, torch.ones_like(r))
    byr = -bphi * dz / rs
    bzr = bphi * dy / rs
    nx = shape[0]
    add = [t.unsqueeze(2).expand(-1, -1, nx).to("cpu").numpy().transpose(2, 1, 0)
           for t in (bxr, byr, bzr)]
    return tuple((b + a).astype(dtype) for b, a in zip((bx, by, bz), add))


def stretched_axis(n, lo, hi, kind="tanh", strength=2.0, dtype=np.float32):
    """n strictly increasing node coordinates from lo to hi.  ``sinh``: nodes
    cluster at the centre (spacing ratio about cosh(strength) from the ends to
    the centre); ``tanh``: nodes cluster at both ends; ``geometric``: spacing
    grows by a constant factor (1 + strength / n) from lo upward, as a
    solar-atmosphere z axis does; ``uniform``: evenly spaced."""
    t = np.linspace(0.0, 1.0, n)
    if kind == "sinh":
        u = 0.5 + 0.5 * np.sinh(strength * (2 * t - 1)) / np.sinh(strength)
    elif kind == "tanh":
        u = 0.5 + 0.5 * np.tanh(strength * (2 * t - 1)) / np.tanh(strength)
    elif kind == "geometric":
        r = 1.0 + strength / n
        w = r ** np.arange(n - 1)
        u = np.r_[0.0, np.cumsum(w)] / np.sum(w)
    elif kind == "uniform":
        u = t
    else:
        raise ValueError(f"unknown axis kind {kind!r}")
    a = (lo + (hi - lo) * u).astype(dtype)
    a[0], a[-1] = dtype(lo), dtype(hi)
    if np.any(np.diff(a.astype(np.float64)) <= 0):
        raise ValueError("axis is not strictly increasing at this precision")
    return a


def stretched_multipole(shape=(256, 256, 256), seed=0, count=40, dtype=np.float32, device=None,
                        box=((-1.0, 1.0), (-1.0, 1.0), (0.0, 2.0)), kinds=("sinh", "sinh", "geometric"),
                        strength=(1.5, 1.5, 3.0), scale=1.0):
    """The random multipole of ``multipole`` sampled on a stretched grid: x
    and y clustered at the centre, z spacing growing with height.  ``scale``
    multiplies the box and the source positions (the field lines scale with
    them).  Returns ((bx, by, bz), (xa, ya, za)) with float32 axes in the
    box's units."""
    axes = tuple(stretched_axis(n, lo * scale, hi * scale, kind, st)
                 for n, (lo, hi), kind, st in zip(shape, box, kinds, strength))
    q, pos = multipole_sources(seed, count)
    field = point_sources(shape, q, pos * scale, dtype=dtype, device=device, axes=axes)
    return field, axes


def hyperbolic(shape, k=0.02, b0=1.0, dtype=np.float64, center=None):
    """Linear X-point field B = (k (x-xc), -k (y-yc), b0) in index units.
    Trilinear interpolation reproduces it exactly; lines obey
    x - xc = (x0 - xc) exp(k z / b0), y - yc = (y0 - yc) exp(-k z / b0)."""
    nx, ny, nz = shape
    xc, yc = center if center is not None else ((nx - 1) / 2.0, (ny - 1) / 2.0)
    i = np.arange(nx, dtype=np.float64).reshape(-1, 1, 1)
    j = np.arange(ny, dtype=np.float64).reshape(1, -1, 1)
    bx = np.broadcast_to(k * (i - xc), shape)
    by = np.broadcast_to(-k * (j - yc), shape)
    bz = np.full(shape, b0)
    return tuple(np.asfortranarray(np.array(b, dtype=dtype)) for b in (bx, by, bz))


def rotation(shape, omega=0.05, b0=1.0, dtype=np.float64, center=None):
    """Linear rotating field B = (-omega (y-yc), omega (x-xc), b0) in index
    units: lines are helices about the axis (xc, yc), and trilinear
    interpolation reproduces the field exactly, with no kinks at cell faces,
    so an adaptive controller's error estimate varies smoothly along a line."""
    nx, ny, nz = shape
    xc, yc = center if center is not None else ((nx - 1) / 2.0, (ny - 1) / 2.0)
    i = np.arange(nx, dtype=np.float64).reshape(-1, 1, 1)
    j = np.arange(ny, dtype=np.float64).reshape(1, -1, 1)
    bx = np.broadcast_to(-omega * (j - yc), shape)
    by = np.broadcast_to(omega * (i - xc), shape)
    bz = np.full(shape, b0)
    return tuple(np.asfortranarray(np.array(b, dtype=dtype)) for b in (bx, by, bz))


def uniform(shape, b, dtype=np.float64):
    return tuple(np.asfortranarray(np.full(shape, float(c), dtype=dtype)) for c in b)


# ------------------------------------------------------------------ seeds
def plane_seeds_z(nx_pix, ny_pix, x_range, y_range, z=0.0, dtype=np.float32):
    """Seed grid on a constant-z plane: ``linspace`` x ``meshgrid`` ('xy'
    indexing, rows = y), flattened row-major (demo/Q_RKF45CuPy.ipynb cell [6]).
    Returns (seeds (n,3), (ny_pix, nx_pix))."""
    xi = np.linspace(x_range[0], x_range[1], nx_pix, dtype=dtype)
    yi = np.linspace(y_range[0], y_range[1], ny_pix, dtype=dtype)
    X, Y = np.meshgrid(xi, yi)
    Z = np.full_like(X, z)
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1), X.shape


def plane_seeds_abc(A, B, C, n_u, n_v, dtype=np.float32):
    """Seed grid spanning the parallelogram A + u (B-A) + v (C-A), u, v in
    [0,1] (demo/Q_RKF45CuPyPlane.ipynb cell [4]); rows follow C-A.  Returns
    (seeds, (n_v, n_u), unit normal (B-A)x(C-A)/|.|)."""
    A = np.asarray(A, np.float64); B = np.asarray(B, np.float64); C = np.asarray(C, np.float64)
    u = np.linspace(0.0, 1.0, n_u)
    v = np.linspace(0.0, 1.0, n_v)
    U, V = np.meshgrid(u, v)
    P = A[None, None, :] + (B - A)[None, None, :] * U[..., None] + (C - A)[None, None, :] * V[..., None]
    nrm = np.cross(B - A, C - A)
    nrm = nrm / np.linalg.norm(nrm)
    return P.reshape(-1, 3).astype(dtype), U.shape, nrm.astype(dtype)


def volume_seeds(shape, dtype=np.float32):
    """One seed per grid node of an (nx, ny, nz) cube, x fastest."""
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(dtype)


def batch_requests(shapes, count=256, seed=0, side=(64, 1024), tol=(1e-7, 1e-5), step=0.25):
    """Uneven Q-map requests over resident fields of the given shapes: each
    picks a field uniformly, an arbitrarily oriented plane through three
    points A, B, C drawn uniformly in that field's box (grid-index
    coordinates), a side of n x n pixels with n log-uniform in `side`, RK4 or
    RKF45 with equal odds, and for RKF45 a tolerance log-uniform in `tol`.
    Draw order per request (default_rng(seed)): field, A, B, C, n, method,
    tol.  Returns a list of dicts (field index, A, B, C, n, method, tol, step)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        fi = int(rng.integers(len(shapes)))
        hi = np.asarray(shapes[fi], np.float64) - 1
        A, B, C = (rng.uniform(0, 1, 3) * hi for _ in range(3))
        n = int(round(math.exp(rng.uniform(math.log(side[0]), math.log(side[1])))))
        method = "rk4" if rng.uniform() < 0.5 else "rkf45"
        t = float(math.exp(rng.uniform(math.log(tol[0]), math.log(tol[1]))))
        out.append(dict(field=fi, A=A, B=B, C=C, n=n, method=method,
                        tol=t if method == "rkf45" else 0.0, step=step))
    return out


def plane_seeds_abc_torch(A, B, C, n_u, n_v, device="cuda"):
    """plane_seeds_abc on the device: x, y, z float32 tensors of n_v * n_u
    seeds (rows follow C - A), computed in float64 as the host version."""
    torch = _torch()
    A, B, C = (torch.as_tensor(np.asarray(v, np.float64), device=device) for v in (A, B, C))
    u = torch.linspace(0.0, 1.0, n_u, dtype=torch.float64, device=device)
    v = torch.linspace(0.0, 1.0, n_v, dtype=torch.float64, device=device)
    V, U = torch.meshgrid(v, u, indexing="ij")
    P = A + (B - A) * U[..., None] + (C - A) * V[..., None]
    P = P.reshape(-1, 3).to(torch.float32)
    return P[:, 0].contiguous(), P[:, 1].contiguous(), P[:, 2].contiguous()


def fourier_perturbation(shape, seed, box=((-1, 1), (-1, 1), (0, 2)), modes=8, kmax=4,
                         device=None):
    """dB = -grad(dPhi) of a harmonic potential perturbation
    dPhi = sum_m a_m cos(kx_m x + ky_m y + phi_m) exp(-k_m z), k_m = |(kx_m, ky_m)|
    (each mode solves Laplace's equation, so dB is itself a potential field).
    Draw order (default_rng(seed)): per mode, kx and ky as pi * integers in
    [-kmax, kmax] (not both zero), a ~ N(0,1), phi ~ U[0, 2 pi).  Returns
    unnormalised (dBx, dBy, dBz) torch float64 tensors of shape (nz, ny, nx)."""
    torch = _torch()
    device = _default_device(device)
    rng = np.random.default_rng(seed)
    xs, ys, zs = _axes(shape, box, device, torch)
    Z, Y, X = torch.meshgrid(zs, ys, xs, indexing="ij")
    dbx = torch.zeros_like(X)
    dby = torch.zeros_like(X)
    dbz = torch.zeros_like(X)
    for _ in range(modes):
        while True:
            ix, iy = rng.integers(-kmax, kmax + 1, 2)
            if ix or iy:
                break
        a = rng.normal()
        phi = rng.uniform(0, 2 * math.pi)
        kx, ky = math.pi * ix, math.pi * iy
        k = math.hypot(kx, ky)
        arg = kx * X + ky * Y + phi
        decay = torch.exp(-k * Z)
        s = a * torch.sin(arg) * decay
        dbx += kx * s
        dby += ky * s
        dbz += a * k * torch.cos(arg) * decay
    return dbx, dby, dbz


def snapshot(index, shape=(320, 320, 256), base_seed=0, dtype=np.float32, device=None):
    """Snapshot `index` of a field time series: the random multipole of
    `base_seed` plus a Fourier-mode potential perturbation whose RMS |dB| on
    the z = 0 plane is eps times that of the base field, eps ~ U[0.01, 0.05]
    drawn first from default_rng(1000 + index) (the modes then from
    default_rng(index))."""
    torch = _torch()
    device = _default_device(device)
    eps = float(np.random.default_rng(1000 + index).uniform(0.01, 0.05))
    q, pos = multipole_sources(base_seed, 40)
    base = point_sources(shape, q, pos, dtype=np.float64, device=device)
    bt = [torch.as_tensor(np.ascontiguousarray(b.transpose(2, 1, 0)), device=device) for b in base]
    dbx, dby, dbz = fourier_perturbation(shape, index, device=device)
    rms = lambda t: torch.sqrt(torch.mean(t[0] ** 2 + t[1] ** 2 + t[2] ** 2))
    scale = eps * rms([b[0] for b in bt]) / rms([dbx[0], dby[0], dbz[0]])
    out = [b + scale * d for b, d in zip(bt, (dbx, dby, dbz))]
    return _to_numpy(*out, dtype)
