"""Loader and ctypes prototypes for ``libfqsl.so`` (declared in include/fqsl.h).

The library is built in-tree by ``fastqsl_amd.build`` (hipcc, gfx950).  There
is no fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfqsl.so")

FQSL_OK = 0
LOOKUPS = {"bisect": 0, "uniform": 1, "exact": 2}  # FQSL_LOOKUP_*
DTYPE_F32, DTYPE_F64 = 0, 1
METHODS = {"rkf45": 0, "rk4": 1}
PRECISIONS = {"fp32c": 0, "fp64": 1, "mixed": 2}
SCHEDULES = {"persistent": 0, "per_line": 1}
LAYOUT_SOA32, LAYOUT_SOA64, LAYOUT_CELL32, LAYOUT_ALL = 1, 2, 4, 7

_lib = None

vp = ctypes.c_void_p
u64 = ctypes.c_uint64
i32 = ctypes.c_int


class FqslError(RuntimeError):
    pass


class fqsl_scott_params(ctypes.Structure):
    _fields_ = [("method", i32), ("precision", i32), ("schedule", i32), ("reserved", i32),
                ("step", ctypes.c_double), ("tol", ctypes.c_double)]


class fqsl_scott_out(ctypes.Structure):
    _fields_ = [(n, vp) for n in ("start_x", "start_y", "start_z", "end_x", "end_y", "end_z",
                                  "flag_start", "flag_end", "b_this_x", "b_this_y", "b_this_z",
                                  "q_0", "q_perp", "line_len", "evals", "twist")]


PLANE_Z, PLANE_ABC = 0, 1  # FQSL_PLANE_*


class fqsl_plane(ctypes.Structure):
    _fields_ = [("kind", i32), ("n_u", i32), ("n_v", i32), ("reserved", i32), ("p", ctypes.c_double * 9)]


class fqsl_neighbour_params(ctypes.Structure):
    _fields_ = [("method", i32), ("precision", i32), ("schedule", i32), ("reserved", i32),
                ("step", ctypes.c_double), ("tol_coef", ctypes.c_double), ("pseudo_tol_scale", ctypes.c_double),
                ("stride", ctypes.c_double), ("pseudo_stride", ctypes.c_double),
                ("normal", ctypes.c_float * 3), ("pseudo_normal", ctypes.c_float * 3)]


class fqsl_plain_params(ctypes.Structure):
    _fields_ = [("method", i32), ("precision", i32), ("schedule", i32), ("reserved", i32),
                ("step", ctypes.c_double), ("tol_coef", ctypes.c_double),
                ("ncross", ctypes.c_float * 3), ("reserved2", ctypes.c_float)]


class fqsl_plain_out(ctypes.Structure):
    _fields_ = [(n, vp) for n in ("start_x", "start_y", "start_z", "end_x", "end_y", "end_z",
                                  "flag_start", "flag_end", "b_this_x", "b_this_y", "b_this_z",
                                  "b_flag", "b_start_x", "b_start_y", "b_start_z", "b_end_x",
                                  "b_end_y", "b_end_z", "line_len", "evals", "twist")]


class fqsl_scott_request(ctypes.Structure):
    _fields_ = [("field", vp), ("method", i32), ("reserved", i32), ("step", ctypes.c_double),
                ("tol", ctypes.c_double), ("n", u64), ("x", vp), ("y", vp), ("z", vp),
                ("out", fqsl_scott_out)]


# name -> (restype, argtypes)
PROTOTYPES = {
    "fqsl_last_error": (ctypes.c_char_p, []),
    "fqsl_version": (ctypes.c_char_p, []),
    "fqsl_field_create": (i32, [vp, vp, vp, i32, i32, i32, i32, i32, i32, ctypes.POINTER(vp)]),
    "fqsl_field_destroy": (i32, [vp]),
    "fqsl_field_bytes": (ctypes.c_size_t, [vp]),
    "fqsl_field_shape": (i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]),
    "fqsl_field_set_current": (i32, [vp, vp, vp, vp, i32, i32]),
    "fqsl_field_set_axes": (i32, [vp, vp, vp, vp, i32, i32, i32]),
    "fqsl_field_stretched": (i32, [vp]),
    "fqsl_curl": (i32, [vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]),
    "fqsl_diff_3point": (i32, [vp, i32, i32, i32, i32, vp, i32, vp, vp]),
    "fqsl_copy_boundary": (i32, [vp, i32, i32, i32, vp]),
    "fqsl_trace_scott": (i32, [vp, ctypes.POINTER(fqsl_scott_params), u64, vp, vp, vp,
                               ctypes.POINTER(fqsl_scott_out), vp]),
    "fqsl_trace_scott_noalloc": (i32, [vp, ctypes.POINTER(fqsl_scott_params), u64, vp, vp, vp,
                                       ctypes.POINTER(fqsl_scott_out), vp]),
    "fqsl_trace_plain": (i32, [vp, ctypes.POINTER(fqsl_plain_params), u64, vp, vp, vp,
                               ctypes.POINTER(fqsl_plain_out), vp]),
    "fqsl_qcalc_plane": (i32, [i32, i32] + [vp] * 13 + [ctypes.c_double, vp, vp, vp]),
    "fqsl_pseudo_seeds": (i32, [u64, vp, vp, vp, vp, vp, vp, vp, ctypes.c_float, u64,
                                vp, vp, vp, vp, vp, vp, ctypes.POINTER(u64), vp]),
    "fqsl_qcut_scatter": (i32, [u64] + [vp] * 10 + [ctypes.c_float, vp, vp, vp]),
    "fqsl_trace_plain_n": (i32, [vp, ctypes.POINTER(fqsl_plain_params), u64, vp, vp, vp, vp,
                                 ctypes.POINTER(fqsl_plain_out), vp]),
    "fqsl_pseudo_workspace_bytes": (ctypes.c_size_t, [u64]),
    "fqsl_pseudo_seeds_async": (i32, [u64, vp, vp, vp, vp, vp, vp, vp, ctypes.c_float, u64,
                                      vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp]),
    "fqsl_qcut_scatter_n": (i32, [u64, vp] + [vp] * 10 + [ctypes.c_float, vp, vp, vp]),
    "fqsl_slog10q": (i32, [u64, vp, vp, vp, vp, vp]),
    "fqsl_bad_pixels": (i32, [u64, vp, vp, vp, vp, vp, u64, vp, vp, vp, vp, ctypes.POINTER(u64), vp]),
    "fqsl_scatter_f32": (i32, [u64, vp, vp, vp, vp]),
    "fqsl_curl_index": (i32, [vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]),
    "fqsl_TraceAllBlineScott": (i32, [vp] * 30 + [i32, i32, vp]),
    "fqsl_TraceAllBline": (i32, [i32] + [vp] * 35 + [i32, i32, vp]),
    "fqsl_TraceAllBlineStretch": (i32, [vp] * 39 + [i32, i32, vp]),
    "fqsl_scott_batch_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(fqsl_scott_request), i32, i32]),
    "fqsl_scott_batch_create": (i32, [ctypes.POINTER(fqsl_scott_request), i32, i32, vp, ctypes.c_size_t, vp,
                                      ctypes.POINTER(vp)]),
    "fqsl_scott_batch_run": (i32, [vp, vp]),
    "fqsl_scott_batch_times": (i32, [vp, ctypes.POINTER(ctypes.c_double), vp]),
    "fqsl_scott_batch_destroy": (i32, [vp]),
    "fqsl_qmap_plane": (i32, [vp, ctypes.POINTER(fqsl_plane), ctypes.POINTER(fqsl_scott_params),
                              ctypes.POINTER(fqsl_scott_out), vp]),
    "fqsl_qmap_plane_neighbour": (i32, [vp, ctypes.POINTER(fqsl_plane), ctypes.POINTER(fqsl_neighbour_params),
                                        vp, vp, vp]),
    "fqsl_server_create": (i32, [i32, i32, u64, ctypes.POINTER(vp)]),
    "fqsl_submit": (i32, [vp, ctypes.POINTER(fqsl_scott_request), vp, ctypes.POINTER(u64)]),
    "fqsl_poll": (i32, [vp, u64, ctypes.POINTER(i32)]),
    "fqsl_wait": (i32, [vp, u64]),
    "fqsl_server_destroy": (i32, [vp]),
    "fqsl_last_kernel_ms": (i32, [ctypes.POINTER(ctypes.c_float)]),
    "fqsl_last_step_ms": (i32, [ctypes.POINTER(ctypes.c_float)]),
    "fqsl_tune_set": (i32, [i32, i32]),
    "fqsl_tune_get": (i32, [i32, ctypes.POINTER(i32)]),
}


def lib():
    """The loaded library; raises FqslError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise FqslError(f"{LIBPATH} is missing: build it with fastqsl_amd.build.build() "
                            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != FQSL_OK:
        msg = lib().fqsl_last_error().decode(errors="replace")
        raise FqslError(f"libfqsl error {rc}: {msg}")


TUNE_KNOBS = {"scott_waves": 0, "grid_pct": 1, "queues": 2, "refill_min": 3, "order": 4,
              "order_buckets": 5, "plain_refill": 6, "stretch_lds": 7}


def tune_set(knob: str, value: int) -> None:
    check(lib().fqsl_tune_set(TUNE_KNOBS[knob], int(value)))


def tune_get(knob: str) -> int:
    v = i32()
    check(lib().fqsl_tune_get(TUNE_KNOBS[knob], ctypes.byref(v)))
    return int(v.value)


def last_kernel_ms() -> float:
    v = ctypes.c_float()
    check(lib().fqsl_last_kernel_ms(ctypes.byref(v)))
    return float(v.value)


def last_step_ms() -> float:
    """Stepping-kernel time of the last persistent Scott trace (HIP events)."""
    v = ctypes.c_float()
    check(lib().fqsl_last_step_ms(ctypes.byref(v)))
    return float(v.value)
