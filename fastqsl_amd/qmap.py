"""Q maps on a plane of seeds: the two pipelines the reference's notebooks run.

* ``scott_qmap``     -- Scott method, one trace per pixel, q_0 / q_perp
  (demo/Q_RKF45CuPyScott.ipynb cells [5]-[7]); ``scott_qmap_plane`` the same
  through the library's C plane entry, seeds formed on the device
  (``neighbour_qmap_plane`` likewise for the neighbour-difference map).
* ``neighbour_qmap`` -- position-only traces, QCalcPlane, pseudo lines for
  pixels where B is nearly tangent to the plane, Qcut and merge
  (demo/Q_RKF45CuPy.ipynb cells [6]-[11]).
* ``mixed_qmap``     -- ``neighbour_qmap`` with the pixels that have B_flag set
  or Q = NaN traced again by the Scott method
  (demo/Q_RKF45MixCuPyDiffScott.ipynb cells [7]-[9]).
* ``neighbour_qmaps`` -- several ``neighbour_qmap`` jobs over HIP streams, so
  one map's pseudo-line trace (a few long lines that leave most of the GPU
  idle) runs beside the other maps' main traces.

Seed grids use the notebooks' conventions (``linspace`` x ``meshgrid``, rows
= the second in-plane axis), see ``synth.plane_seeds_z`` / ``plane_seeds_abc``.
Everything stays on the GPU; results are torch CUDA tensors.
"""
from __future__ import annotations

import collections
import threading

import numpy as np

from . import FastQSL, postproc
from .field import Field
from .trace import trace_plain, trace_scott


def _grid(seeds, shape):
    import torch
    s = seeds if isinstance(seeds, torch.Tensor) else torch.as_tensor(np.asarray(seeds, np.float32))
    s = s.to(device="cuda", dtype=torch.float32).reshape(-1, 3)
    if s.shape[0] != int(np.prod(shape)):
        raise ValueError(f"{s.shape[0]} seeds do not fill a {shape} grid")
    return s[:, 0].contiguous(), s[:, 1].contiguous(), s[:, 2].contiguous()


def scott_qmap(field: Field, seeds, shape, method="rkf45", step=0.25, tol=1e-6,
               precision="fp32c", schedule="persistent", tile=16):
    """Scott Q on a seed grid of `shape` (rows, cols), the pixels traced in
    ``tile`` x ``tile`` blocks (0: row by row; the persistent queue's
    longest-first order then keeps blocks together within a length class,
    ~1 % on the 1024^2 multipole map); returns a dict of (rows, cols)
    tensors: q_0, q_perp, flag_start, flag_end, b_this_z, ..."""
    x, y, z = _grid(seeds, shape)
    perm = block_order(shape, (tile, tile)) if tile and tile > 1 else None
    if perm is not None:
        x, y, z = x[perm], y[perm], z[perm]
    r = trace_scott(field, x, y, z, method=method, step=step, tol=tol, precision=precision,
                    schedule=schedule)
    if perm is not None:
        r = _unpermute(r, perm)
    return {k: v.reshape(shape) for k, v in r.items() if k != "_seeds"}


def scott_qmap_plane(field: Field, plane, n_u, n_v, method="rkf45", step=0.25, tol=1e-6,
                     precision="fp32c", schedule="persistent", outputs=None, stream=None):
    """Scott Q map of a plane through the library's own plane entry
    (``fqsl_qmap_plane``): the seeds are formed on the device as
    ``synth.plane_seeds_z`` (``plane = ("z", x0, x1, y0, y1, z0)``) or
    ``synth.plane_seeds_abc`` (``plane = ("abc", A, B, C)``) form them, bit for
    bit, and traced in row order.  Returns a dict of (n_v, n_u) tensors."""
    import ctypes
    import torch
    from . import _lib
    from .trace import SCOTT_OUTPUTS, _dtype_of, _stream
    pl = _plane_struct(plane, n_u, n_v)
    n = pl.n_u * pl.n_v
    res = {k: torch.empty(n, dtype=_dtype_of(k), device="cuda") for k in (outputs or SCOTT_OUTPUTS)}
    o = _lib.fqsl_scott_out(**{k: (res[k].data_ptr() if k in res else None) for k in SCOTT_OUTPUTS})
    prm = _lib.fqsl_scott_params(method=_lib.METHODS[method], precision=_lib.PRECISIONS[precision],
                                 schedule=_lib.SCHEDULES[schedule], reserved=0, step=float(step), tol=float(tol))
    _lib.check(_lib.lib().fqsl_qmap_plane(field.handle, ctypes.byref(pl), ctypes.byref(prm), ctypes.byref(o),
                                          _stream(stream)))
    return {k: v.reshape(pl.n_v, pl.n_u) for k, v in res.items()}


def _plane_struct(plane, n_u, n_v):
    from . import _lib
    pl = _lib.fqsl_plane()
    pl.n_u, pl.n_v = int(n_u), int(n_v)
    kind = plane[0]
    if kind == "z":
        pl.kind = _lib.PLANE_Z
        vals = [float(v) for v in plane[1:6]]
    elif kind == "abc":
        pl.kind = _lib.PLANE_ABC
        vals = [float(c) for pt in plane[1:4] for c in pt]
    else:
        raise ValueError(f"plane kind {kind!r}: 'z' or 'abc'")
    for k, v in enumerate(vals):
        pl.p[k] = v
    return pl


def neighbour_qmap_plane(field: Field, plane, n_u, n_v, stride, normal=(0.0, 0.0, 1.0), method="rkf45",
                         step=0.125, tol_coef=1.0, pseudo_stride=None, pseudo_tol_scale=0.1, pseudo_normal=None,
                         precision="fp32c", schedule="persistent", stream=None):
    """``neighbour_qmap`` through the library's own plane entry
    (``fqsl_qmap_plane_neighbour``): seeds formed on the device as
    ``scott_qmap_plane`` forms them, every step of the pipeline on one stream.
    Returns q (n_v-2, n_u-2), q_all (n_v, n_u) and n_pseudo (a CUDA int64
    scalar)."""
    import ctypes
    import torch
    from . import _lib
    from .trace import _stream
    pl = _plane_struct(plane, n_u, n_v)
    prm = _lib.fqsl_neighbour_params(
        method=_lib.METHODS[method], precision=_lib.PRECISIONS[precision], schedule=_lib.SCHEDULES[schedule],
        reserved=0, step=float(step), tol_coef=float(tol_coef), pseudo_tol_scale=float(pseudo_tol_scale),
        stride=float(stride), pseudo_stride=float(stride if pseudo_stride is None else pseudo_stride))
    pn = normal if pseudo_normal is None else pseudo_normal
    for k in range(3):
        prm.normal[k] = float(normal[k])
        prm.pseudo_normal[k] = float(pn[k])
    q_all = torch.empty((pl.n_v, pl.n_u), dtype=torch.float32, device="cuda")
    npseudo = torch.zeros((), dtype=torch.int64, device="cuda")
    _lib.check(_lib.lib().fqsl_qmap_plane_neighbour(field.handle, ctypes.byref(pl), ctypes.byref(prm),
                                                    q_all.data_ptr(), npseudo.data_ptr(), _stream(stream)))
    return dict(q=q_all[1:-1, 1:-1], q_all=q_all, n_pseudo=npseudo)


CUBE_OUTPUTS = ("q_0", "q_perp", "flag_start", "flag_end", "evals")


def cube_seeds(x_axis, y_axis, z_axis):
    """Seeds of a 3-D sampling box: every (x, y, z) of the three 1-D axes,
    x slowest and z fastest, so that results reshape to the (nx, ny, nz) cube
    the reference fills slice by slice (``Qube[:, :, iz]`` from
    ``meshgrid(x_i, y_i, indexing='ij')`` at each z,
    demo/Q_RKF45CuPy3DScott.ipynb cells [4]-[7])."""
    import torch
    ax = [a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a, np.float32))
          for a in (x_axis, y_axis, z_axis)]
    ax = [a.to(device="cuda", dtype=torch.float32).reshape(-1) for a in ax]
    X, Y, Z = torch.meshgrid(*ax, indexing="ij")
    return X.reshape(-1), Y.reshape(-1), Z.reshape(-1), tuple(int(a.numel()) for a in ax)


def scott_qcube(field: Field, x_axis, y_axis, z_axis, method="rkf45", step=0.25, tol=1e-6,
                precision="fp32c", schedule="persistent", outputs=CUBE_OUTPUTS, brick=4):
    """3-D Scott Q: one line per sample of the box x_axis x y_axis x z_axis
    (grid-index coordinates), all voxels in one launch, traced in ``brick``^3
    blocks of samples (0: x slowest, z fastest; on the 512x256x256 rope
    arcade 4^3 bricks trace 5-10 % faster).  Returns a dict of (nx, ny, nz)
    tensors (``Qube`` of the reference's 3-D notebooks)."""
    x, y, z, shape = cube_seeds(x_axis, y_axis, z_axis)
    perm = block_order(shape, (brick,) * 3) if brick and brick > 1 else None
    if perm is not None:
        x, y, z = x[perm], y[perm], z[perm]
    r = trace_scott(field, x, y, z, method=method, step=step, tol=tol, precision=precision,
                    schedule=schedule, outputs=outputs)
    if perm is not None:
        r = _unpermute(r, perm)
    return {k: v.reshape(shape) for k, v in r.items() if k != "_seeds"}


def _block_order_compute(shape, block):
    nd = len(shape)
    if all(s % b == 0 for s, b in zip(shape, block)):
        idx = np.arange(int(np.prod(shape)), dtype=np.int32).reshape(
            [v for s, b in zip(shape, block) for v in (s // b, b)])
        return idx.transpose(list(range(0, 2 * nd, 2)) + list(range(1, 2 * nd, 2))).reshape(-1).copy()
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    key = np.zeros(shape, np.int64)
    for g, s, b in zip(grids, shape, block):  # block index
        key = key * (-(-s // b)) + g // b
    for g, b in zip(grids, block):            # position inside the block
        key = key * b + g % b
    return np.argsort(key.reshape(-1), kind="stable").astype(np.int32)


# Seed orders are cached on the host and per device, each cache bounded by
# bytes (least recently used out first; an order larger than the bound is not
# kept), under one lock: the multi-stream jobs call in from several threads.
ORDER_CACHE_BYTES = {"host": 256 << 20, "device": 512 << 20}
_order_lock = threading.Lock()
_order_host = collections.OrderedDict()
_order_dev = collections.OrderedDict()


def _cache_put(cache, key, value, nbytes, budget):
    if nbytes > budget:
        return
    cache[key] = value
    cache.move_to_end(key)
    while sum(v.nbytes if isinstance(v, np.ndarray) else v.numel() * v.element_size()
              for v in cache.values()) > budget:
        cache.popitem(last=False)


def _block_order_host(shape, block):
    """Flat indices (int32) of a C-ordered grid of ``shape`` visited block by
    block (blocks in C order, C order inside a block; partial blocks at the
    far edges)."""
    shape, block = tuple(int(v) for v in shape), tuple(int(v) for v in block)
    if int(np.prod(shape)) >= 2 ** 31:
        raise ValueError(f"grid {shape} has more than 2^31 - 1 seeds")
    key = (shape, block)
    with _order_lock:
        p = _order_host.get(key)
        if p is not None:
            _order_host.move_to_end(key)
            return p
    p = _block_order_compute(shape, block)
    with _order_lock:
        _cache_put(_order_host, key, p, p.nbytes, ORDER_CACHE_BYTES["host"])
    return p


def block_order(shape, block):
    """Device tensor (int32) of ``_block_order_host`` (cached per shape, block
    and device): tracing seeds in this order starts a wavefront's lines close
    together, so they share cached cells; a line's result does not depend on
    the order."""
    import torch
    key = (tuple(int(v) for v in shape), tuple(int(v) for v in block), torch.cuda.current_device())
    with _order_lock:
        t = _order_dev.get(key)
        if t is not None:
            _order_dev.move_to_end(key)
    if t is None:
        t = torch.as_tensor(_block_order_host(key[0], key[1]), device="cuda")
        torch.cuda.current_stream().synchronize()  # complete before another stream or thread reads it
        with _order_lock:
            _cache_put(_order_dev, key, t, t.numel() * t.element_size(), ORDER_CACHE_BYTES["device"])
    t.record_stream(torch.cuda.current_stream())  # an evicted order outlives its last use on this stream
    return t


def _unpermute(r, perm):
    """Per-line outputs traced in ``perm`` order, back in seed order."""
    import torch
    out = {}
    for k, v in r.items():
        if k == "_seeds":
            continue
        u = torch.empty_like(v)
        u[perm] = v
        out[k] = u
    return out


def neighbour_qmap(field: Field, seeds, shape, stride, normal=(0.0, 0.0, 1.0), method="rkf45",
                   step=0.125, tol_coef=1.0, pseudo_stride=None, pseudo_tol_scale=0.1,
                   pseudo_normal=None, precision="fp32c", schedule="persistent", tile=16):
    """Neighbour-difference Q (QCalcPlane) with the pseudo-line fix-up for
    pixels where B is nearly tangent to the plane (B_flag == 1).

    ``stride`` is the pixel spacing in cells (``stride_step`` of the
    notebooks); ``pseudo_stride`` the offset of the four pseudo seeds (default
    ``stride``); the pseudo lines are traced with ``tol_coef *
    pseudo_tol_scale`` and ``pseudo_normal`` (default ``normal``).  The main
    lines are traced in ``tile`` x ``tile`` pixel blocks (0: row by row), so
    a wavefront's lines start close together and share cached cells; results
    do not depend on the order.  Returns a dict with ``q`` = Q_final of
    shape (rows-2, cols-2), ``q_all`` (rows, cols), the per-pixel traces and
    the number of pseudo pixels (``n_pseudo``, a CUDA int64 scalar written on
    the stream: ``int(res["n_pseudo"])`` waits for it)."""
    import torch
    x, y, z = _grid(seeds, shape)
    perm = block_order(shape, (tile, tile)) if tile and tile > 1 else None
    if perm is not None:
        x, y, z = x[perm], y[perm], z[perm]
    r = trace_plain(field, x, y, z, ncross=normal, method=method, step=step, tol_coef=tol_coef,
                    precision=precision, schedule=schedule)
    if perm is not None:  # back to pixel order
        r = _unpermute(r, perm)
    g = {k: v.reshape(shape) for k, v in r.items() if k != "_seeds"}
    q_inner = FastQSL.QCalcPlane(g["end_x"], g["end_y"], g["end_z"], g["flag_end"],
                                 g["start_x"], g["start_y"], g["start_z"], g["flag_start"],
                                 g["b_start_x"], g["b_start_y"], g["b_start_z"],
                                 g["b_end_x"], g["b_end_y"], g["b_end_z"],
                                 g["b_this_x"], g["b_this_y"], g["b_this_z"],
                                 g["b_flag"], stride, normal)
    ps = stride if pseudo_stride is None else pseudo_stride
    # the pseudo-line pipeline stays on the stream: the compaction writes its
    # count on the device and the pseudo trace and Qcut read it there, so the
    # host never waits (and no host thread blocks while other streams run)
    c = FastQSL.cook_pseudo_lines_async(g["start_x"], g["start_y"], g["start_z"], g["flag_start"],
                                        g["b_start_z"], g["b_end_z"], g["b_flag"], ps)
    q_all = torch.zeros(shape, dtype=torch.float32, device="cuda")
    q_all[1:-1, 1:-1] = q_inner
    rc = trace_plain(field, c["seed_x"], c["seed_y"], c["seed_z"],
                     ncross=normal if pseudo_normal is None else pseudo_normal, method=method,
                     step=step, tol_coef=tol_coef * pseudo_tol_scale, precision=precision,
                     schedule=schedule, n_device=c["counts"][1:],
                     outputs=("start_x", "start_y", "start_z", "end_x", "end_y", "end_z",
                              "flag_start", "flag_end"))
    FastQSL.Qcut(rc["start_x"], rc["start_y"], rc["start_z"], rc["flag_start"],
                 rc["end_x"], rc["end_y"], rc["end_z"], rc["flag_end"],
                 c["bz0_start"], c["bz0_end"], ps, Q_all=q_all, pixel=c["pixel"],
                 m_device=c["counts"][:1])
    return dict(q=q_all[1:-1, 1:-1], q_all=q_all, q_inner=q_inner, traces=g,
                n_pseudo=c["counts"][0], _pseudo=(c, rc))


def neighbour_qmaps(field: Field, jobs, streams=4, **kw):
    """``neighbour_qmap`` for each (seeds, shape, stride) of ``jobs``, job j
    issued on HIP stream j mod ``streams`` (default priority) from the
    calling thread.  A map holds no host synchronisation (its pseudo-line
    count stays on the device), so one thread keeps every stream fed, and a
    map's pseudo-line trace runs beside the main traces of the maps on the
    other streams; four streams match the four hardware queues a process
    gets by default.  Keyword arguments go to every ``neighbour_qmap``.
    Returns the results in job order; they are complete on the caller's
    current stream (it waits for every pool stream)."""
    import torch
    jobs = list(jobs)
    nstreams = max(1, min(int(streams), len(jobs)))
    pool = [torch.cuda.Stream() for _ in range(nstreams)]
    base = torch.cuda.current_stream()
    for s in pool:
        s.wait_stream(base)   # inputs written on the caller's stream are visible
    out = []
    for j, (seeds, shape, stride) in enumerate(jobs):
        with torch.cuda.stream(pool[j % nstreams]):
            out.append(neighbour_qmap(field, seeds, shape, stride, **kw))
    for s in pool:
        base.wait_stream(s)

    def keep_for(v):   # results allocated on a pool stream, used on the caller's
        if isinstance(v, torch.Tensor) and v.is_cuda:
            v.record_stream(base)
        elif isinstance(v, dict):
            for u in v.values():
                keep_for(u)
        elif isinstance(v, (tuple, list)):
            for u in v:
                keep_for(u)
    keep_for(out)
    return out


def mixed_qmap(field: Field, seeds, shape, stride, normal=(0.0, 0.0, 1.0), method="rkf45", step=0.125,
               tol_coef=1.0, scott_tol=None, scott_use="q_perp", precision="fp32c", schedule="persistent",
               **neighbour_kw):
    """The mixed method: ``neighbour_qmap``, then every pixel with B_flag set
    or Q = NaN takes the Scott method's q_perp (the notebook's choice;
    ``scott_use="q_0"`` for the other one) from a Scott trace of its seed
    with the same step and a per-step tolerance ``scott_tol`` (default
    1e-3 * tol_coef, the reference kernel's mapping of the coefficient).
    Returns the ``neighbour_qmap`` dict with ``q_all`` / ``q`` updated and
    ``n_scott`` = the number of pixels replaced."""
    res = neighbour_qmap(field, seeds, shape, stride, normal=normal, method=method, step=step,
                         tol_coef=tol_coef, precision=precision, schedule=schedule, **neighbour_kw)
    x, y, z = _grid(seeds, shape)
    tol = 1e-3 * tol_coef if scott_tol is None else scott_tol
    res["n_scott"] = postproc.scott_fixup(field, res["q_all"], res["traces"]["b_flag"], x, y, z, method=method,
                                          step=step, tol=tol, precision=precision, schedule=schedule,
                                          use=scott_use)
    res["q"] = res["q_all"][1:-1, 1:-1]
    return res
