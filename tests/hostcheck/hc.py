"""ctypes access to the test-only CPU build of the tracer header."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libhostcheck.so")
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_lib = None


def build():
    src = os.path.join(HERE, "hostcheck.cpp")
    hdrs = [os.path.join(ROOT, "fastqsl_amd", "csrc", h) for h in ("fqsl_tracer.h", "fqsl_stretch.h")]
    if os.path.exists(LIB) and os.path.getmtime(LIB) > max(os.path.getmtime(x) for x in [src] + hdrs):
        return LIB
    subprocess.run(["hipcc", "-x", "hip", "--offload-host-only", "-O2", "-fPIC", "-shared", "-std=c++17",
                    "-I" + os.path.join(ROOT, "fastqsl_amd", "csrc"), src, "-o", LIB], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def _field(f):
    arrs = [np.ascontiguousarray(np.asarray(b, np.float64).transpose(2, 1, 0)) for b in f]
    return arrs, np.asarray(f[0]).shape


def P(a, t=_dp):
    return a.ctypes.data_as(t)


def grad_b(f, pts):
    (bx, by, bz), sh = _field(f)
    p = np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1, 3))
    out = np.empty((p.shape[0], 3, 3))
    lib().hc_grad_b(P(bx), P(by), P(bz), *sh, ctypes.c_long(p.shape[0]), P(p), P(out))
    return out


def trace_scott(f, seeds, precision=0, method=0, s0=0.25, tol=1e-6, current=None):
    (bx, by, bz), sh = _field(f)
    j = _field(current)[0] if current is not None else (None, None, None)
    s = np.ascontiguousarray(np.asarray(seeds, np.float32).reshape(-1, 3))
    n = s.shape[0]
    o = dict(start=np.empty((n, 3)), end=np.empty((n, 3)), flag_start=np.empty(n, np.int32),
             flag_end=np.empty(n, np.int32), q0=np.empty(n), qperp=np.empty(n), length=np.empty(n),
             evals=np.empty(n, np.int64), twist=np.zeros(n))
    L = lib()
    L.hc_trace_scott.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_long, _fp,
                                 _dp, _dp, _ip, _ip, _dp, _dp, _dp, _lp, _dp, _dp, _dp, _dp]
    op = lambda a: None if a is None else P(a)
    L.hc_trace_scott(P(bx), P(by), P(bz), *sh, precision, method, s0, tol, n, P(s, _fp),
                     P(o["start"]), P(o["end"]), P(o["flag_start"], _ip), P(o["flag_end"], _ip),
                     P(o["q0"]), P(o["qperp"]), P(o["length"]), P(o["evals"], _lp),
                     op(j[0]), op(j[1]), op(j[2]), P(o["twist"]) if current is not None else None)
    return o


def trace_plain(f, seeds, ncross=(0, 0, 1), precision=0, method=0, s0=0.25, tol_coef=1.0):
    (bx, by, bz), sh = _field(f)
    s = np.ascontiguousarray(np.asarray(seeds, np.float32).reshape(-1, 3))
    nc = np.ascontiguousarray(np.asarray(ncross, np.float32))
    n = s.shape[0]
    o = dict(start=np.empty((n, 3)), end=np.empty((n, 3)), flag_start=np.empty(n, np.int32),
             flag_end=np.empty(n, np.int32), length=np.empty(n), evals=np.empty(n, np.int64))
    L = lib()
    L.hc_trace_plain.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double, _fp, ctypes.c_long, _fp,
                                 _dp, _dp, _ip, _ip, _dp, _lp]
    L.hc_trace_plain(P(bx), P(by), P(bz), *sh, precision, method, s0, tol_coef, P(nc, _fp), n, P(s, _fp),
                     P(o["start"]), P(o["end"]), P(o["flag_start"], _ip), P(o["flag_end"], _ip),
                     P(o["length"]), P(o["evals"], _lp))
    return o


def pow_minus_tenth(e2):
    e = np.ascontiguousarray(np.asarray(e2, np.float64).reshape(-1))
    out = np.empty_like(e)
    L = lib()
    L.hc_pow_minus_tenth.argtypes = [ctypes.c_long, _dp, _dp]
    L.hc_pow_minus_tenth(e.size, P(e), P(out))
    return out


RULES = {"bisect": 0, "uniform": 1, "exact": 2}


def stretch_cell(axis, pts, rule="bisect", precision=1):
    a = np.ascontiguousarray(np.asarray(axis, np.float64))
    p = np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1))
    cell = np.empty(p.size, np.int32)
    w = np.empty(p.size)
    L = lib()
    L.hc_stretch_cell.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long, _dp, _ip, _dp]
    rc = L.hc_stretch_cell(P(a), a.size, RULES[rule], precision, p.size, P(p), P(cell, _ip), P(w))
    assert rc == 0
    return cell, w


def trace_stretch(f, axes, seeds, ncross=(0, 0, 1), rule="bisect", precision=1, s0=0.01, tol_coef=1.0,
                  current=None):
    (bx, by, bz), sh = _field(f)
    xa, ya, za = [np.ascontiguousarray(np.asarray(a, np.float64)) for a in axes]
    j = _field(current)[0] if current is not None else (None, None, None)
    s = np.ascontiguousarray(np.asarray(seeds, np.float32).reshape(-1, 3))
    nc = np.ascontiguousarray(np.asarray(ncross, np.float32))
    n = s.shape[0]
    o = dict(start=np.empty((n, 3)), end=np.empty((n, 3)), flag_start=np.empty(n, np.int32),
             flag_end=np.empty(n, np.int32), length=np.empty(n), evals=np.empty(n, np.int64),
             twist=np.zeros(n))
    L = lib()
    L.hc_trace_stretch.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _fp,
                                   ctypes.c_long, _fp, _dp, _dp, _ip, _ip, _dp, _lp, _dp, _dp, _dp, _dp]
    op = lambda a: None if a is None else P(a)
    rc = L.hc_trace_stretch(P(bx), P(by), P(bz), *sh, P(xa), P(ya), P(za), RULES[rule], precision, s0,
                            tol_coef, P(nc, _fp), n, P(s, _fp), P(o["start"]), P(o["end"]),
                            P(o["flag_start"], _ip), P(o["flag_end"], _ip), P(o["length"]),
                            P(o["evals"], _lp), op(j[0]), op(j[1]), op(j[2]),
                            P(o["twist"]) if current is not None else None)
    assert rc == 0
    return o
