"""fqsl_trace_scott_noalloc through the C ABI: the same bits as
fqsl_trace_scott on the same seeds, on two calls in a row (the second with
fewer lines, so it runs over what the first left in the scratch array), in
both precisions; sizes past the array's maximum and the per-line schedule are
refused before anything is launched."""
import ctypes

import numpy as np
import pytest

from fastqsl_amd import synth

pytestmark = pytest.mark.gpu

OUTS = ("start_x", "start_y", "start_z", "end_x", "end_y", "end_z", "flag_start", "flag_end",
        "b_this_x", "b_this_y", "b_this_z", "q_0", "q_perp", "line_len", "evals")


@pytest.fixture(scope="module")
def fq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd
    fastqsl_amd.lib()
    return fastqsl_amd


def _seeds(n, shape, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    hi = torch.tensor([shape[0] - 1, shape[1] - 1, shape[2] - 1], dtype=torch.float32, device="cuda")
    p = torch.rand((n, 3), generator=g, device="cuda") * hi
    p[: n // 4, 2] = 0.0          # a quarter on the bottom face, as plane maps seed them
    return [p[:, k].contiguous() for k in range(3)]


def _call(fq, entry, field, seeds, precision, method, tol):
    import torch
    from fastqsl_amd import _lib
    from fastqsl_amd.trace import _dtype_of
    n = seeds[0].numel()
    res = {k: torch.full((n,), -3, dtype=_dtype_of(k), device="cuda") for k in OUTS}
    o = _lib.fqsl_scott_out(**{k: res[k].data_ptr() for k in OUTS})
    p = _lib.fqsl_scott_params(method=_lib.METHODS[method], precision=_lib.PRECISIONS[precision],
                               schedule=_lib.SCHEDULES["persistent"], reserved=0, step=0.25, tol=tol)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(_lib.lib(), entry)(field.handle, ctypes.byref(p), n, seeds[0].data_ptr(),
                                    seeds[1].data_ptr(), seeds[2].data_ptr(), ctypes.byref(o), s)
    _lib.check(rc)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in res.items()}


def _bits(a):
    return a.view(np.uint32 if a.itemsize == 4 else np.uint64)


@pytest.mark.parametrize("precision,sizes,method,tol", [
    ("fp32c", (1 << 20, 65536 + 17), "rkf45", 1e-6),
    ("fp64", (1 << 19, 40000 + 3), "rkf45", 1e-6),
    ("fp32c", (200000, 70001), "rk4", 0.0),
])
def test_noalloc_matches_trace_scott(fq, precision, sizes, method, tol):
    shape = (96, 80, 64)
    arrs = synth.multipole(shape=shape, seed=3, device="cuda")
    with fq.Field(*arrs) as f:
        for call, n in enumerate(sizes):
            print(f"no-alloc {precision} {method}: call {call + 1}, {n} lines", flush=True)
            seeds = _seeds(n, shape, seed=11 + call)
            a = _call(fq, "fqsl_trace_scott", f, seeds, precision, method, tol)
            b = _call(fq, "fqsl_trace_scott_noalloc", f, seeds, precision, method, tol)
            assert (a["evals"] > 0).mean() > 0.5
            for k in OUTS:
                np.testing.assert_array_equal(_bits(b[k]), _bits(a[k]), err_msg=k)


def test_noalloc_refuses_what_it_cannot_hold(fq):
    import torch
    from fastqsl_amd import _lib
    arrs = synth.multipole(shape=(32, 32, 32), seed=1, device="cuda")
    x = torch.zeros(8, dtype=torch.float32, device="cuda")
    q = torch.zeros(8, dtype=torch.float32, device="cuda")
    o = _lib.fqsl_scott_out(q_0=q.data_ptr())
    with fq.Field(*arrs) as f:
        for prec, n, sched in (("fp32c", (1 << 20) + 1, "persistent"), ("fp64", (1 << 19) + 1, "persistent"),
                               ("fp32c", 8, "per_line")):
            p = _lib.fqsl_scott_params(method=0, precision=_lib.PRECISIONS[prec],
                                       schedule=_lib.SCHEDULES[sched], reserved=0, step=0.25, tol=1e-6)
            rc = _lib.lib().fqsl_trace_scott_noalloc(f.handle, ctypes.byref(p), n, x.data_ptr(), x.data_ptr(),
                                                     x.data_ptr(), ctypes.byref(o), None)
            assert rc == _lib.FQSL_OK + 1, (prec, n, sched)   # FQSL_ERR_INVALID, nothing launched
            msg = _lib.lib().fqsl_last_error().decode()
            assert ("no-alloc maximum" in msg) or ("persistent schedule only" in msg), msg
