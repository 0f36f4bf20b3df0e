"""The serving loop: each request's result is the same alone and in the mix.

Requests over three resident snapshots arrive faster than they are served,
so the server admits them in batches on two streams while earlier batches
are still on the device; every request's outputs must equal those of tracing
it alone, bit for bit.  A tight HBM budget (one batch's worth) must still
drain the stream without deadlock and give the same results.
"""
import numpy as np
import pytest

from fastqsl_amd import synth

pytestmark = pytest.mark.gpu
SHAPE = (96, 80, 64)


@pytest.fixture(scope="module")
def snaps():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd as fq
    fq.lib()
    fs = [fq.Field(*synth.snapshot(i, shape=SHAPE, device="cuda")) for i in range(3)]
    yield fs
    for f in fs:
        f.close()


def _alone(fq, f, r, precision):
    import torch
    x, y, z = r.seeds()
    a = fq.trace_scott(f, x, y, z, method=r.method, step=r.step, tol=r.tol, precision=precision)
    torch.cuda.synchronize()
    return a


@pytest.mark.parametrize("budget_lines", [None, 9000])
def test_same_alone_and_in_the_mix(snaps, budget_lines):
    import torch
    import fastqsl_amd as fq
    from fastqsl_amd import serve
    reqs = serve.poisson_requests(rate=3000, duration=0.01, nsnap=3, shape=SHAPE, seed=4,
                                  side=(16, 96), min_count=24)
    budget = (48 << 30) if budget_lines is None else budget_lines * serve.LINE_BYTES["fp32c"]
    srv = serve.QServer(snaps, precision="fp32c", streams=2, max_batch_lines=12000,
                        hbm_budget=budget, keep_results=True)
    st = srv.run(reqs)
    print({k: v for k, v in st.items() if k != "latency_ms"}, flush=True)
    assert st["requests"] == len(reqs) and 1 < st["batches"] < len(reqs)
    assert np.all(st["latency_ms"] > 0)
    for r in reqs:
        a = _alone(fq, snaps[r.snapshot], r, "fp32c")
        for k in serve.OUTPUTS:
            assert np.array_equal(r.result[k].cpu().numpy(), a[k].cpu().numpy(), equal_nan=True), (r.rid, k)


def test_serial_fp64_baseline_matches_alone(snaps):
    import fastqsl_amd as fq
    from fastqsl_amd import serve
    reqs = serve.poisson_requests(rate=500, duration=0.01, nsnap=3, shape=SHAPE, seed=5,
                                  side=(16, 48), min_count=6)
    srv = serve.QServer(snaps, precision="fp64", mode="serial", schedule="per_line", keep_results=True)
    st = srv.run(reqs)
    assert st["batches"] == len(reqs)
    for r in reqs:
        a = _alone(fq, snaps[r.snapshot], r, "fp64")
        for k in serve.OUTPUTS:
            assert np.array_equal(r.result[k].cpu().numpy(), a[k].cpu().numpy(), equal_nan=True), (r.rid, k)
