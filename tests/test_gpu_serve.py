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


def test_c5_at_full_size():
    """The serving workload at its stated size: 16 resident 320x320x256
    snapshots of the random multipole (1-5 % Fourier-mode perturbations),
    Poisson arrivals of z = 0 maps of 128^2..2048^2 pixels at 20 requests/s
    (below saturation) for 1 s, continuous batching on two streams; every
    request's result equals its trace alone, bit for bit."""
    import torch
    import fastqsl_amd as fq
    from fastqsl_amd import serve
    shape = (320, 320, 256)
    fs = []
    try:
        for i in range(16):
            fs.append(fq.Field(*synth.snapshot(i, shape=shape, device="cuda"), layouts=("cell32",)))
        print(f"16 snapshots resident: {sum(f.nbytes for f in fs) / 1e9:.1f} GB", flush=True)
        reqs = serve.poisson_requests(rate=20, duration=1.0, nsnap=16, shape=shape, seed=20, min_count=12)
        srv = serve.QServer(fs, precision="fp32c", streams=2, keep_results=True)
        st = srv.run(reqs)
        print({k: v for k, v in st.items() if k != "latency_ms"}, flush=True)
        assert st["requests"] == len(reqs) and st["lines"] > 1e7
        assert min(r.n for r in reqs) >= 128 and max(r.n for r in reqs) <= 2048
        for r in reqs:
            x, y, z = r.seeds()
            a = fq.trace_scott(fs[r.snapshot], x, y, z, method=r.method, step=r.step, tol=r.tol,
                               precision="fp32c", outputs=serve.OUTPUTS)
            for k in serve.OUTPUTS:
                u, v = r.result[k], a[k]
                if u.dtype == torch.float32:
                    u, v = u.view(torch.int32), v.view(torch.int32)
                assert bool(torch.equal(u, v)), (r.rid, k)
            r.result = None
    finally:
        for f in fs:
            f.close()
        torch.cuda.empty_cache()
