"""Drop-in for ``fastqslpy.kernels`` (fastqslpy/kernels.py:10-28).

``compileTraceBlineAdaptive()`` / ``compileTraceBlineRK4()`` /
``compileTraceBlineScott()`` return callables used exactly like the CuPy
``RawKernel`` the reference returns::

    TraceAllBline = kernels.compileTraceBlineAdaptive()
    TraceAllBline(grid, block, (Bx, By, Bz, BshapeN, dummy, dummy, dummy, dummy, flag_twist,
                                x_inp, y_inp, z_inp, inp_norm, ...,  s_len, N, line_len, tol_coef))

The argument tuple has the reference kernel's parameter order
(TraceBlineAdaptive.cu:429-437, TraceBline.cu:383-391,
TraceBlineScott.cu:521-528, TraceBlineAdaptiveStretch.cu:422-431 for
``compileTraceBlineAdaptiveStretch()``).  Arrays may be torch CUDA tensors (used in place,
as CuPy arrays are) or NumPy arrays (copied to the device and the outputs
copied back into them); scalars are one-element arrays as in the reference.
Memory is read as laid out, exactly like the reference: field arrays must
hold x fastest (the ``np.transpose(sav, (2, 1, 0))`` layout).  ``grid`` and
``block`` are accepted and ignored -- the library sizes its own launch.  With
``curB_flag`` set, ``curB_x/y/z`` (float32, B's layout) give the current
density and ``twist`` (float64, one per line) receives the twist number.

Nothing is compiled at run time: the kernels are in the prebuilt
``libfqsl.so``; these functions only bind it.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_F32, _F64, _I32, _U64, _BOOL = np.float32, np.float64, np.int32, np.uint64, np.bool_

# (name, dtype, is_output) per reference parameter
_COMMON_HEAD = [("Bx", _F32, False), ("By", _F32, False), ("Bz", _F32, False),
                ("BshapeN", _I32, False), ("curB_x", _F32, False), ("curB_y", _F32, False),
                ("curB_z", _F32, False), ("twist", _F64, True), ("curB_flag", _BOOL, False),
                ("inp_x", _F32, False), ("inp_y", _F32, False), ("inp_z", _F32, False),
                ("inp_cross_dir", _F32, False),
                ("start_x", _F32, True), ("start_y", _F32, True), ("start_z", _F32, True),
                ("flag_start", _I32, True),
                ("end_x", _F32, True), ("end_y", _F32, True), ("end_z", _F32, True),
                ("flag_end", _I32, True),
                ("B_this_x", _F32, True), ("B_this_y", _F32, True), ("B_this_z", _F32, True)]

PLAIN_PARAMS = _COMMON_HEAD + [
    ("B_flag", _I32, True),
    ("B_start_x", _F32, True), ("B_start_y", _F32, True), ("B_start_z", _F32, True),
    ("B_end_x", _F32, True), ("B_end_y", _F32, True), ("B_end_z", _F32, True),
    ("s_len", _F32, False), ("N", _U64, False), ("LineLen", _F64, True), ("tol_coef", _F32, False)]

# TraceAllBline of TraceBlineAdaptiveStretch.cu:422-431: the axes and flag_uni
# after BshapeN
STRETCH_PARAMS = (_COMMON_HEAD[:4] +
                  [("x_arr", _F32, False), ("y_arr", _F32, False), ("z_arr", _F32, False),
                   ("flag_uni", _BOOL, False)] +
                  _COMMON_HEAD[4:] + PLAIN_PARAMS[len(_COMMON_HEAD):])

SCOTT_PARAMS = _COMMON_HEAD + [
    ("q_0", _F32, True), ("q_perp", _F32, True),
    ("s_len", _F32, False), ("N", _U64, False), ("LineLen", _F64, True), ("tol_coef", _F32, False)]


class _Binding:
    """Holds device views of the call's arrays and copies host outputs back."""

    def __init__(self, params, args):
        import torch
        self.torch = torch
        self.ptrs = []
        self.copy_back = []
        self.keep = []
        for (name, dt, is_out), a in zip(params, args):
            if isinstance(a, torch.Tensor):
                if not a.is_cuda:
                    raise ValueError(f"{name}: CPU tensors are not accepted; pass a CUDA tensor or NumPy array")
                want = getattr(torch, np.dtype(dt).name if dt is not _BOOL else "bool")
                ok = a.dtype == want or (name == "N" and a.dtype in (torch.int64, torch.uint64))
                if not ok:
                    raise TypeError(f"{name}: expected {np.dtype(dt).name}, got {a.dtype}")
                self.ptrs.append(a.data_ptr())
                continue
            arr = np.asarray(a)
            if arr.dtype != dt:
                raise TypeError(f"{name}: expected {np.dtype(dt).name}, got {arr.dtype}")
            if not (arr.flags.c_contiguous or arr.flags.f_contiguous):
                raise ValueError(f"{name}: array must be contiguous (its memory is read as laid out)")
            raw = arr.ravel(order="K")
            t = torch.from_numpy(raw.view(np.uint8) if dt is _BOOL else raw).cuda()
            self.keep.append(t)
            self.ptrs.append(t.data_ptr())
            if is_out:
                self.copy_back.append((raw, t))

    def finish(self, stream):
        stream.synchronize()
        for raw, t in self.copy_back:
            host = t.cpu().numpy()
            raw[...] = host.view(raw.dtype) if host.dtype != raw.dtype else host


class _Kernel:
    def __init__(self, name, params, method=None, accept_extra_tol=False, stretch=False):
        self.name = name
        self.params = params
        self.method = method
        self.accept_extra_tol = accept_extra_tol
        self.stretch = stretch
        names = [p[0] for p in params]
        self._flag_i = names.index("curB_flag")
        self._twist_i = names.index("twist")
        # run-time selection of the tracer (not part of the reference call)
        self.precision = "fp32c"
        self.schedule = "persistent"

    def __call__(self, grid, block, args, stream=None):
        import torch
        args = tuple(args)
        n_expected = len(self.params)
        if self.accept_extra_tol and len(args) == n_expected - 1:
            args = args + (np.ones(1, np.float32),)  # TraceBline.cu declares no tol_coef
        if len(args) != n_expected:
            raise TypeError(f"{self.name} takes {n_expected} arguments ({len(args)} given)")
        flag = args[self._flag_i]
        flag_val = bool(flag.reshape(-1)[0].item()) if hasattr(flag, "reshape") else bool(flag)
        if not flag_val:
            # twist off: the notebooks pass float32 dummies for curB and twist,
            # which the library never reads
            twist_i = self._twist_i
            args = args[:twist_i] + (np.zeros(1, np.float64),) + args[twist_i + 1:]
        b = _Binding(self.params, args)
        st = stream if stream is not None else torch.cuda.current_stream()
        L = _lib.lib()
        prec = _lib.PRECISIONS[self.precision]
        sched = _lib.SCHEDULES[self.schedule]
        if self.stretch:
            rc = L.fqsl_TraceAllBlineStretch(*b.ptrs, prec, sched, ctypes.c_void_p(st.cuda_stream))
        elif self.method is None:
            rc = L.fqsl_TraceAllBlineScott(*b.ptrs, prec, sched, ctypes.c_void_p(st.cuda_stream))
        else:
            rc = L.fqsl_TraceAllBline(self.method, *b.ptrs, prec, sched, ctypes.c_void_p(st.cuda_stream))
        _lib.check(rc)
        if b.copy_back:
            b.finish(st)
        self._last = b  # device copies stay alive until the next call

    # the CuPy RawKernel attribute the notebooks never use, kept for printing
    def __repr__(self):
        return f"<fastqsl_amd kernel {self.name} ({self.precision}, {self.schedule})>"


def compileTraceBlineAdaptive():
    """TraceAllBline of TraceBlineAdaptive.cu (RKF45), kernels.py:10-18."""
    return _Kernel("TraceAllBline", PLAIN_PARAMS, method=_lib.METHODS["rkf45"])


def compileTraceBlineRK4():
    """TraceAllBline of TraceBline.cu (fixed-step RK4), kernels.py:21-28.  The
    reference kernel has no tol_coef parameter; a trailing one is accepted
    and ignored (demo/Q_RK4.ipynb passes it)."""
    return _Kernel("TraceAllBline", PLAIN_PARAMS, method=_lib.METHODS["rk4"], accept_extra_tol=True)


def compileTraceBlineAdaptiveStretch():
    """TraceAllBline of TraceBlineAdaptiveStretch.cu (RKF45 on a stretched
    grid; fastqslpy/kernels.py lists the module but has no loader for it, and
    demo/Q_RKF45CuPyStretch.ipynb cell [9] compiles the file directly).  Its
    39 arguments put the axes x_arr, y_arr, z_arr (float32, one coordinate per
    node) and flag_uni (bool: GetIndexUniStretch instead of GetIndexBiSearch)
    after BshapeN; seeds and results are in the axes' coordinates.  The call
    reads BshapeN, flag_uni, curB_flag and the axes to the host (one stream
    synchronisation)."""
    return _Kernel("TraceAllBline", STRETCH_PARAMS, method=_lib.METHODS["rkf45"], stretch=True)


def compileTraceBlineScott():
    """TraceAllBlineScott of TraceBlineScott.cu (RKF45 with the Scott
    Jacobian; the reference has no loader for it and the notebooks compile
    the file directly, demo/Q_RKF45CuPyScott.ipynb cell [4]).  The per-step
    tolerance is 1e-3 * tol_coef."""
    return _Kernel("TraceAllBlineScott", SCOTT_PARAMS)
