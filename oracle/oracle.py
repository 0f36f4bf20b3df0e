"""ctypes wrapper around the FP64 CPU oracle (``libfqsl_oracle.so``).

TEST INFRASTRUCTURE ONLY: this module is the correctness checker for the HIP
kernels and the timed CPU baseline.  Only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` import it.  The product package
``fastqsl_amd`` never does.

Fields are passed as three arrays of shape (nx, ny, nz) (the reference's
``np.transpose(sav, (2,1,0))`` convention, x fastest in memory; see
TraceBlineScott.cu:36-39).  Seeds are (n, 3) arrays of grid-index coordinates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBNAME = os.path.join(_HERE, "libfqsl_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIBNAME):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIBNAME


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBNAME):
            build()
        L = ctypes.CDLL(_LIBNAME)
        L.orc_trace_scott.restype = ctypes.c_int
        L.orc_trace_scott.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_long, _dp, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      _dp, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _lp, _lp,
                                      _dp, _dp, _dp, _dp]
        L.orc_trace_scott_f32.restype = ctypes.c_int
        L.orc_trace_scott_f32.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_long, _dp, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, _dp, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _lp, _lp]
        L.orc_trace_plain.restype = ctypes.c_int
        L.orc_trace_plain.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_long, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int,
                                      _dp, _dp, _ip, _ip, _dp, _dp, _dp, _ip, _dp, _lp,
                                      _dp, _dp, _dp, _dp]
        L.orc_interp.restype = None
        L.orc_interp.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_long, _dp, _dp, ctypes.c_int]
        L.orc_grad_b.restype = None
        L.orc_grad_b.argtypes = L.orc_interp.argtypes
        L.orc_check_flag.restype = None
        L.orc_check_flag.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                     _dp, _ip]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_scott_step.restype = ctypes.c_double
        L.orc_scott_step.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_stretch_cell.restype = None
        L.orc_stretch_cell.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_long, _dp, _ip, _dp]
        L.orc_interp_stretch.restype = None
        L.orc_interp_stretch.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         _dp, _dp, _dp, ctypes.c_int, ctypes.c_long, _dp, _dp,
                                         ctypes.c_int]
        L.orc_trace_stretch.restype = ctypes.c_int
        L.orc_trace_stretch.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _dp, _dp, _dp, ctypes.c_int,
                                        ctypes.c_long, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int,
                                        _dp, _dp, _ip, _ip, _dp, _dp, _dp, _ip, _dp, _lp,
                                        _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _flat_field(bx, by, bz):
    arrs = []
    for b in (bx, by, bz):
        b = np.asarray(b, dtype=np.float64)
        if b.ndim != 3:
            raise ValueError("field components must be 3-D (nx, ny, nz)")
        arrs.append(np.ascontiguousarray(b.transpose(2, 1, 0)))  # x fastest
    shape = np.asarray(bx).shape
    return arrs, shape


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def _optr(a, t=_dp):
    return None if a is None else _ptr(a, t)


def _seeds(seeds):
    s = np.ascontiguousarray(np.asarray(seeds, dtype=np.float64).reshape(-1, 3))
    return s


def num_threads() -> int:
    return lib().orc_num_threads()


def interp(field, pts, mode=0):
    """mode 0: Interp3dxyzn (raw), 1: normalised, 2: Interp3d clamped rule."""
    (bx, by, bz), shape = _flat_field(*field)
    p = _seeds(pts)
    out = np.empty_like(p)
    lib().orc_interp(_ptr(bx), _ptr(by), _ptr(bz), *shape, p.shape[0], _ptr(p), _ptr(out), mode)
    return out


def grad_b(field, pts, mode=0):
    """Returns (n, 3, 3) with [:, j, c] = d b_c / d x_j.  mode 0: difference
    quotient with h = 1e-3 (grad_unit_vec_B), 1: closed form in the cell."""
    (bx, by, bz), shape = _flat_field(*field)
    p = _seeds(pts)
    out = np.empty((p.shape[0], 3, 3))
    lib().orc_grad_b(_ptr(bx), _ptr(by), _ptr(bz), *shape, p.shape[0], _ptr(p), _ptr(out), mode)
    return out


def check_flag(shape, pts):
    p = _seeds(pts)
    out = np.empty(p.shape[0], dtype=np.int32)
    lib().orc_check_flag(*shape, p.shape[0], _ptr(p), _ptr(out, _ip))
    return out


METHODS = {"rkf45": 0, "rk4": 1}


def scott_step(field, state9, s, method="rkf45", grad_mode=0):
    """One integrator step of the 9-vector (x, u, v); returns (new state, err)."""
    (bx, by, bz), shape = _flat_field(*field)
    y = np.ascontiguousarray(np.asarray(state9, dtype=np.float64).reshape(9))
    out = np.empty(9)
    e = lib().orc_scott_step(_ptr(bx), _ptr(by), _ptr(bz), *shape, _ptr(y), _ptr(out), float(s),
                             METHODS[method], int(grad_mode))
    return out, e


def trace_scott(field, seeds, method="rkf45", s0=0.25, tol=1e-6, grad_mode=0, nthreads=0, current=None):
    """Scott tracer (TraceAllBlineScott, TraceBlineScott.cu:521-595) in FP64.
    With ``current`` = (Jx, Jy, Jz) the result also has the twist number."""
    (bx, by, bz), shape = _flat_field(*field)
    cur = _flat_field(*current)[0] if current is not None else (None, None, None)
    s = _seeds(seeds)
    n = s.shape[0]
    out = dict(start=np.empty((n, 3)), end=np.empty((n, 3)),
               flag_start=np.empty(n, np.int32), flag_end=np.empty(n, np.int32),
               q0=np.empty(n), qperp=np.empty(n), length=np.empty(n),
               b_this=np.empty((n, 3)), evals=np.empty(n, np.int64), steps=np.empty(n, np.int64),
               twist=np.zeros(n))
    rc = lib().orc_trace_scott(_ptr(bx), _ptr(by), _ptr(bz), *shape, n, _ptr(s), METHODS[method],
                               float(s0), float(tol), int(grad_mode), int(nthreads),
                               _ptr(out["start"]), _ptr(out["end"]),
                               _ptr(out["flag_start"], _ip), _ptr(out["flag_end"], _ip),
                               _ptr(out["q0"]), _ptr(out["qperp"]), _ptr(out["length"]),
                               _ptr(out["b_this"]), _ptr(out["evals"], _lp),
                               _ptr(out["steps"], _lp), _optr(cur[0]), _optr(cur[1]), _optr(cur[2]),
                               _ptr(out["twist"]) if current is not None else None)
    if rc != 0:
        raise ValueError("oracle rejected the field shape")
    return out


def trace_scott_f32(field, seeds, s0=0.25, tol=1e-6, nthreads=0):
    """The Scott tracer in the reference kernel's own float arithmetic
    (TraceBlineScott.cu:95-508), RKF45: a model of what the FP32 reference
    produces (its step sequences, evaluation counts and classes), not of its
    bits.  Same outputs as ``trace_scott`` (no twist)."""
    (bx, by, bz), shape = _flat_field(*field)
    s = _seeds(seeds)
    n = s.shape[0]
    out = dict(start=np.empty((n, 3)), end=np.empty((n, 3)),
               flag_start=np.empty(n, np.int32), flag_end=np.empty(n, np.int32),
               q0=np.empty(n), qperp=np.empty(n), length=np.empty(n),
               b_this=np.empty((n, 3)), evals=np.empty(n, np.int64), steps=np.empty(n, np.int64))
    rc = lib().orc_trace_scott_f32(_ptr(bx), _ptr(by), _ptr(bz), *shape, n, _ptr(s), float(s0), float(tol),
                                   int(nthreads), _ptr(out["start"]), _ptr(out["end"]),
                                   _ptr(out["flag_start"], _ip), _ptr(out["flag_end"], _ip),
                                   _ptr(out["q0"]), _ptr(out["qperp"]), _ptr(out["length"]),
                                   _ptr(out["b_this"]), _ptr(out["evals"], _lp), _ptr(out["steps"], _lp))
    if rc != 0:
        raise ValueError("oracle rejected the field shape")
    return out


def trace_plain(field, seeds, ncross=(0.0, 0.0, 1.0), method="rkf45", s0=0.25, tol_coef=1.0,
                nthreads=0, current=None):
    """Position-only tracer (TraceAllBline, TraceBlineAdaptive.cu:429-510 /
    TraceBline.cu:383-464) in FP64; with ``current`` also the twist number."""
    (bx, by, bz), shape = _flat_field(*field)
    cur = _flat_field(*current)[0] if current is not None else (None, None, None)
    s = _seeds(seeds)
    n = s.shape[0]
    nc = np.ascontiguousarray(np.asarray(ncross, dtype=np.float64))
    out = dict(start=np.empty((n, 3)), end=np.empty((n, 3)),
               flag_start=np.empty(n, np.int32), flag_end=np.empty(n, np.int32),
               b_start=np.empty((n, 3)), b_end=np.empty((n, 3)), b_this=np.empty((n, 3)),
               b_flag=np.empty(n, np.int32), length=np.empty(n), evals=np.empty(n, np.int64),
               twist=np.zeros(n))
    rc = lib().orc_trace_plain(_ptr(bx), _ptr(by), _ptr(bz), *shape, n, _ptr(s), _ptr(nc),
                               METHODS[method], float(s0), float(tol_coef), int(nthreads),
                               _ptr(out["start"]), _ptr(out["end"]),
                               _ptr(out["flag_start"], _ip), _ptr(out["flag_end"], _ip),
                               _ptr(out["b_start"]), _ptr(out["b_end"]), _ptr(out["b_this"]),
                               _ptr(out["b_flag"], _ip), _ptr(out["length"]),
                               _ptr(out["evals"], _lp), _optr(cur[0]), _optr(cur[1]), _optr(cur[2]),
                               _ptr(out["twist"]) if current is not None else None)
    if rc != 0:
        raise ValueError("oracle rejected the field shape")
    return out


# cell rules of the stretched tracer: the reference's bisection
# (GetIndexBiSearch), its evenly-spaced formula (GetIndexUniStretch), and the
# cell that holds the point
LOOKUPS = {"bisect": 0, "uniform": 1, "exact": 2}


def _axes(axes):
    return [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in axes]


def stretch_cell(axis, pts, rule="bisect"):
    """(cell, weight) along one stretched axis at each point."""
    a = np.ascontiguousarray(np.asarray(axis, dtype=np.float64))
    p = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1))
    cell = np.empty(p.shape[0], np.int32)
    w = np.empty(p.shape[0])
    lib().orc_stretch_cell(_ptr(a), a.shape[0], LOOKUPS[rule], p.shape[0], _ptr(p), _ptr(cell, _ip),
                           _ptr(w))
    return cell, w


def interp_stretch(field, axes, pts, rule="bisect", normalize=False):
    """Interp3dxyzn of the stretched tracer at points in axis coordinates."""
    (bx, by, bz), shape = _flat_field(*field)
    xa, ya, za = _axes(axes)
    p = _seeds(pts)
    out = np.empty_like(p)
    lib().orc_interp_stretch(_ptr(bx), _ptr(by), _ptr(bz), *shape, _ptr(xa), _ptr(ya), _ptr(za),
                             LOOKUPS[rule], p.shape[0], _ptr(p), _ptr(out), int(bool(normalize)))
    return out


def trace_stretch(field, axes, seeds, ncross=(0.0, 0.0, 1.0), s0=0.01, tol_coef=1.0, rule="bisect",
                  nthreads=0, current=None):
    """Stretched-grid RKF45 tracer (TraceAllBline of
    TraceBlineAdaptiveStretch.cu:422-508) in FP64; seeds and results in the
    axes' coordinates; with ``current`` also the twist number."""
    (bx, by, bz), shape = _flat_field(*field)
    xa, ya, za = _axes(axes)
    if (xa.shape[0], ya.shape[0], za.shape[0]) != tuple(shape):
        raise ValueError("axis lengths must match the field shape")
    cur = _flat_field(*current)[0] if current is not None else (None, None, None)
    s = _seeds(seeds)
    n = s.shape[0]
    nc = np.ascontiguousarray(np.asarray(ncross, dtype=np.float64))
    out = dict(start=np.empty((n, 3)), end=np.empty((n, 3)),
               flag_start=np.empty(n, np.int32), flag_end=np.empty(n, np.int32),
               b_start=np.empty((n, 3)), b_end=np.empty((n, 3)), b_this=np.empty((n, 3)),
               b_flag=np.empty(n, np.int32), length=np.empty(n), evals=np.empty(n, np.int64),
               twist=np.zeros(n))
    rc = lib().orc_trace_stretch(_ptr(bx), _ptr(by), _ptr(bz), *shape, _ptr(xa), _ptr(ya), _ptr(za),
                                 LOOKUPS[rule], n, _ptr(s), _ptr(nc), float(s0), float(tol_coef),
                                 int(nthreads),
                                 _ptr(out["start"]), _ptr(out["end"]),
                                 _ptr(out["flag_start"], _ip), _ptr(out["flag_end"], _ip),
                                 _ptr(out["b_start"]), _ptr(out["b_end"]), _ptr(out["b_this"]),
                                 _ptr(out["b_flag"], _ip), _ptr(out["length"]),
                                 _ptr(out["evals"], _lp), _optr(cur[0]), _optr(cur[1]), _optr(cur[2]),
                                 _ptr(out["twist"]) if current is not None else None)
    if rc != 0:
        raise ValueError("oracle rejected the field shape or lookup rule")
    return out
