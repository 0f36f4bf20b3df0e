"""The C ABI without a GPU: libfqsl.so loads, exports every function
include/fqsl.h declares, and the Python bindings agree with the header on
argument counts, struct layouts and constants.  Nothing here launches a
kernel or touches a device."""
import ctypes
import os
import re
import subprocess

import pytest

from fastqsl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fqsl.h")


def _header_text():
    s = open(HEADER).read()
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def declared_functions():
    """name -> list of parameter declarations, from the prototypes in the header."""
    s = _header_text()
    out = {}
    for m in re.finditer(r"(?:^|;|\n)\s*(?:const\s+)?[A-Za-z_]\w*\s*\**\s*(fqsl_\w+)\s*\(([^)]*)\)\s*;", s):
        params = [p.strip() for p in m.group(2).split(",")]
        out[m.group(1)] = [] if params == ["void"] else params
    return out


def header_macros():
    return {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define\s+(FQSL_\w+)\s+(-?\d+)", _header_text())}


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(_lib.LIBPATH):
        from fastqsl_amd import build
        build.build()
    return _lib.lib()


def test_every_declared_symbol_is_exported_and_bound(L):
    decl = declared_functions()
    assert len(decl) >= 16, sorted(decl)
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert set(decl) == set(_lib.PROTOTYPES), set(decl) ^ set(_lib.PROTOTYPES)
    for name, params in decl.items():
        assert len(_lib.PROTOTYPES[name][1]) == len(params), (name, params)


def test_reference_shaped_entry_points_take_the_reference_parameter_lists():
    """fqsl_TraceAllBlineScott: TraceBlineScott.cu:521-528 (30 parameters);
    fqsl_TraceAllBline: TraceBlineAdaptive.cu:429-437 (35), after a method
    selector; both then take precision, schedule and a stream."""
    from fastqsl_amd import kernels
    decl = declared_functions()
    scott = decl["fqsl_TraceAllBlineScott"]
    plain = decl["fqsl_TraceAllBline"]
    assert len(scott) == len(kernels.SCOTT_PARAMS) + 3 == 33
    assert len(plain) == 1 + len(kernels.PLAIN_PARAMS) + 3 == 39
    ident = lambda p: re.findall(r"\w+", p)[-1]
    assert [ident(p) for p in scott[:30]] == [n for n, _, _ in kernels.SCOTT_PARAMS]
    assert [ident(p) for p in plain[1:36]] == [n for n, _, _ in kernels.PLAIN_PARAMS]


def test_constants_match_the_header():
    mac = header_macros()
    assert mac["FQSL_OK"] == _lib.FQSL_OK
    assert (mac["FQSL_DTYPE_F32"], mac["FQSL_DTYPE_F64"]) == (_lib.DTYPE_F32, _lib.DTYPE_F64)
    assert {"rkf45": mac["FQSL_METHOD_RKF45"], "rk4": mac["FQSL_METHOD_RK4"]} == _lib.METHODS
    assert {"fp32c": mac["FQSL_PREC_FP32C"], "fp64": mac["FQSL_PREC_FP64"],
            "mixed": mac["FQSL_PREC_MIXED"]} == _lib.PRECISIONS
    assert {"persistent": mac["FQSL_SCHED_PERSISTENT"],
            "per_line": mac["FQSL_SCHED_PER_LINE"]} == _lib.SCHEDULES
    assert (mac["FQSL_LAYOUT_SOA32"], mac["FQSL_LAYOUT_SOA64"], mac["FQSL_LAYOUT_CELL32"],
            mac["FQSL_LAYOUT_ALL"]) == (_lib.LAYOUT_SOA32, _lib.LAYOUT_SOA64, _lib.LAYOUT_CELL32,
                                        _lib.LAYOUT_ALL)
    assert (mac["FQSL_PLANE_Z"], mac["FQSL_PLANE_ABC"]) == (_lib.PLANE_Z, _lib.PLANE_ABC)
    knobs = {k: mac["FQSL_TUNE_" + k.upper()] for k in _lib.TUNE_KNOBS}
    assert knobs == _lib.TUNE_KNOBS and mac["FQSL_TUNE_COUNT"] == len(knobs)


STRUCTS = ["fqsl_scott_params", "fqsl_scott_out", "fqsl_plain_params", "fqsl_plain_out",
           "fqsl_scott_request", "fqsl_plane", "fqsl_neighbour_params"]


def test_struct_layouts_match_a_c_compiler(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fqsl.h"', "int main(void) {"]
    for s in STRUCTS:
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f, _ in getattr(_lib, s)._fields_:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if ln:
            s, f, v = ln.split()
            got[(s, f)] = int(v)
    for s in STRUCTS:
        cs = getattr(_lib, s)
        assert ctypes.sizeof(cs) == got[(s, "size")], s
        for f, _ in cs._fields_:
            assert getattr(cs, f).offset == got[(s, f)], (s, f)


def test_host_only_calls(L):
    """Calls that do not reach the device: version string, tuning defaults,
    argument checks of the tuning knobs and the error message."""
    assert L.fqsl_version().decode().strip()
    assert _lib.tune_get("scott_waves") == 3
    assert _lib.tune_get("grid_pct") == 100
    assert _lib.tune_get("queues") == 1
    assert _lib.tune_get("refill_min") == 8
    assert _lib.tune_get("order") == 1 and _lib.tune_get("order_buckets") == 64
    assert _lib.tune_get("stretch_lds") == 1
    assert L.fqsl_tune_set(_lib.TUNE_KNOBS["stretch_lds"], 2) != _lib.FQSL_OK
    assert L.fqsl_tune_set(99, 1) != _lib.FQSL_OK
    # server and plane-map argument checks that return before any device call
    h = ctypes.c_void_p()
    assert L.fqsl_server_create(0, 99, 0, ctypes.byref(h)) != _lib.FQSL_OK and not h.value
    done = ctypes.c_int(-1)
    assert L.fqsl_poll(None, 1, ctypes.byref(done)) != _lib.FQSL_OK
    assert L.fqsl_wait(None, 1) != _lib.FQSL_OK
    rid = ctypes.c_uint64()
    assert L.fqsl_submit(None, None, None, ctypes.byref(rid)) != _lib.FQSL_OK
    assert L.fqsl_server_destroy(None) == _lib.FQSL_OK
    pl = _lib.fqsl_plane(kind=7, n_u=4, n_v=4)
    prm = _lib.fqsl_scott_params()
    out = _lib.fqsl_scott_out()
    assert L.fqsl_qmap_plane(None, ctypes.byref(pl), ctypes.byref(prm), ctypes.byref(out), None) != _lib.FQSL_OK
    npm = _lib.fqsl_neighbour_params()
    assert L.fqsl_qmap_plane_neighbour(None, ctypes.byref(pl), ctypes.byref(npm), None, None, None) != _lib.FQSL_OK
    assert L.fqsl_last_error().decode()
    v = ctypes.c_int(-1)
    assert L.fqsl_tune_get(-1, ctypes.byref(v)) != _lib.FQSL_OK


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIBPATH", str(tmp_path / "libfqsl.so"))
    with pytest.raises(_lib.FqslError, match="no CPU fallback"):
        _lib.lib()


def test_batch_workspace_size_is_host_arithmetic(L):
    """fqsl_scott_batch_workspace_bytes only groups and counts (no device
    access): grows by the per-line scratch, and by one counter per
    (field, method) group; create rejects a request without a field."""
    R = _lib.fqsl_scott_request

    def table(specs):
        t = (R * len(specs))()
        for k, (field, method, n) in enumerate(specs):
            t[k] = R(field=field, method=method, reserved=0, step=0.25, tol=1e-6, n=n)
        return t

    fp32, fp64 = _lib.PRECISIONS["fp32c"], _lib.PRECISIONS["fp64"]
    a = table([(0x1000, 0, 1000), (0x1000, 0, 24)])
    b = table([(0x1000, 0, 1000), (0x1000, 0, 1024)])
    w = lambda t, p: L.fqsl_scott_batch_workspace_bytes(t, len(t), p)
    assert w(a, fp32) > 0 and w(b, fp32) > w(a, fp32)
    assert w(b, fp32) - w(a, fp32) >= 1000 * (8 * 4 + 96) - 512
    assert w(b, fp64) > w(b, fp32)
    assert L.fqsl_scott_batch_workspace_bytes(a, 0, fp32) == 0
    h = ctypes.c_void_p()
    bad = table([(None, 0, 10)])
    assert L.fqsl_scott_batch_create(bad, 1, fp32, None, 0, None, ctypes.byref(h)) != _lib.FQSL_OK
    assert "NULL field" in L.fqsl_last_error().decode()
