"""Batches of uneven requests over several resident fields.

Each request's result must be the same alone and in the mix: a batch of
requests with different fields, methods, tolerances and sizes (one of them
empty) gives, per request, exactly the outputs of tracing that request on its
own -- bit for bit, since a line's result does not depend on the lane or
launch that traces it.  The same holds for every replay of a graph-captured
batch, with eager work between the replays.
"""
import numpy as np
import pytest

from fastqsl_amd import synth

pytestmark = pytest.mark.gpu

KEYS = ("q_0", "q_perp", "flag_start", "flag_end", "evals", "end_x", "end_y", "end_z", "line_len")


@pytest.fixture(scope="module")
def fields():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd as fq
    fq.lib()
    arrs = [synth.quadrupole_64(dtype=np.float64, device="cuda"),
            synth.multipole(shape=(96, 80, 64), seed=7, device="cuda")]
    fs = [fq.Field(*a) for a in arrs]
    yield fs
    for f in fs:
        f.close()


def _requests(fields, count, seed):
    import torch
    reqs = synth.batch_requests([f.shape for f in fields], count=count, seed=seed, side=(8, 96))
    out = []
    for r in reqs:
        x, y, z = synth.plane_seeds_abc_torch(r["A"], r["B"], r["C"], r["n"], r["n"])
        out.append(dict(field=fields[r["field"]], x=x, y=y, z=z, method=r["method"],
                        tol=r["tol"] or 1e-6, step=r["step"]))
    # an empty request and a tiny one
    e = torch.empty(0, dtype=torch.float32, device="cuda")
    out.insert(3, dict(field=fields[0], x=e, y=e, z=e, method="rkf45", tol=1e-6, step=0.25))
    t = torch.tensor([20.5], dtype=torch.float32, device="cuda")
    out.insert(7, dict(field=fields[1], x=t, y=t, z=t * 0, method="rk4", tol=1e-6, step=0.5))
    return out


def _alone(fq, rq, precision):
    import torch
    r = fq.trace_scott(rq["field"], rq["x"], rq["y"], rq["z"], method=rq["method"], step=rq["step"],
                       tol=rq["tol"], precision=precision)
    torch.cuda.synchronize()
    return {k: r[k].cpu().numpy() for k in KEYS}


def _check_same(batch, alone, where):
    import torch
    torch.cuda.synchronize()
    for k, (res, ref) in enumerate(zip(batch.results, alone)):
        for key in KEYS:
            assert np.array_equal(res[key].cpu().numpy(), ref[key], equal_nan=True), (where, k, key)


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_batch_equals_requests_alone(fields, precision):
    import fastqsl_amd as fq
    from fastqsl_amd.batch import ScottBatch
    reqs = _requests(fields, 24, seed=1)
    print(f"batch of {len(reqs)} requests, precision {precision}", flush=True)
    alone = [_alone(fq, rq, precision) for rq in reqs]
    with ScottBatch(reqs, precision=precision, outputs=KEYS) as b:
        assert len({(id(r["field"]), r["method"]) for r in reqs}) >= 3
        b.run()
        _check_same(b, alone, "eager")
        t = b.times_ms()
        n = np.array([r["n"] for r in b.requests])
        assert np.all(t[n == 0] == 0) and np.all(t[n > 0] > 0)
        assert t.max() <= fq.last_kernel_ms() * 1.05 + 0.05


def test_graph_replays_equal_eager(fields):
    import torch
    import fastqsl_amd as fq
    from fastqsl_amd.batch import ScottBatch
    reqs = _requests(fields, 16, seed=2)
    alone = [_alone(fq, rq, "fp32c") for rq in reqs]
    with ScottBatch(reqs, precision="fp32c", outputs=KEYS) as b:
        g = b.graph()
        for rep in range(3):
            for res in b.results:                       # outputs are rewritten by each replay
                for v in res.values():
                    v.zero_()
            g.replay()
            torch.cuda.synchronize()
            _check_same(b, alone, f"replay {rep}")
            # eager work between the replays
            _alone(fq, reqs[rep], "fp32c")
            times = b.times_ms()
            assert np.all(times[[r["n"] > 0 for r in b.requests]] > 0)
