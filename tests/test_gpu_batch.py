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


def _same_bits(a, b):
    """Bitwise equality of two device tensors (NaNs included)."""
    import torch
    if a.dtype in (torch.float32, torch.int32):
        return bool(torch.equal(a.view(torch.int32), b.view(torch.int32)))
    if a.dtype == torch.float64:
        return bool(torch.equal(a.view(torch.int64), b.view(torch.int64)))
    return bool(torch.equal(a, b))


def test_c4_at_full_size(oracle):
    """The batch workload at its stated size: 256 requests (arbitrary planes of
    64^2..1024^2 pixels, RK4 or RKF45 at tol 1e-7..1e-5) over the quadrupole
    64^3 (float64), the random multipole 256^3 (seed 0) and a 320x320x256
    multipole (seed 7), ~5e7 lines, traced as one graph-captured batch.  Every
    request equals its trace alone bit for bit; on 8 lines per request the
    RK4 lines match the FP64 oracle and the RKF45 lines are as close to the
    converged answer (tol 1e-10) as the oracle at the request's tolerance."""
    import torch
    import fastqsl_amd as fq
    from fastqsl_amd.batch import ScottBatch
    from tests import parity
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    arrs = [synth.quadrupole_64(dtype=np.float64, device="cuda"), synth.multipole_256(seed=0, device="cuda"),
            synth.multipole(shape=(320, 320, 256), seed=7, device="cuda")]
    fields = [fq.Field(*a) for a in arrs]
    try:
        specs = synth.batch_requests([f.shape for f in fields], count=256, seed=0)
        reqs = []
        for r in specs:
            x, y, z = synth.plane_seeds_abc_torch(r["A"], r["B"], r["C"], r["n"], r["n"])
            reqs.append(dict(field=fields[r["field"]], x=x, y=y, z=z, method=r["method"],
                             tol=r["tol"] or 1e-6, step=r["step"]))
        nlines = sum(r["n"] ** 2 for r in specs)
        assert len(reqs) == 256 and nlines > 4e7
        print(f"C4: 256 requests, {nlines} lines", flush=True)
        with ScottBatch(reqs, precision="fp32c", outputs=KEYS) as b:
            g = b.graph()
            g.replay()
            torch.cuda.synchronize()
            print("graph replayed", flush=True)
            for k, rq in enumerate(reqs):
                a = fq.trace_scott(rq["field"], rq["x"], rq["y"], rq["z"], method=rq["method"], step=rq["step"],
                                   tol=rq["tol"], precision="fp32c", outputs=KEYS)
                for key in KEYS:
                    assert _same_bits(b.results[k][key], a[key]), (k, key)
            print("every request equals its trace alone", flush=True)
            # a sample against the oracle
            f64 = [tuple(np.asfortranarray(np.asarray(c.cpu().numpy() if hasattr(c, "cpu") else c, np.float64))
                         for c in ar) for ar in arrs]
            rng = np.random.default_rng(2)
            rk4_cls, rk4_d, err_gpu, err_orc = [], [], [], []
            for k, (rq, sp) in enumerate(zip(reqs, specs)):
                n = sp["n"] ** 2
                li = np.sort(rng.choice(n, size=min(8, n), replace=False))
                ti = torch.as_tensor(li, device="cuda")
                seeds = torch.stack([rq["x"][ti], rq["y"][ti], rq["z"][ti]], 1).cpu().numpy()
                res = {key: b.results[k][key][ti].cpu().numpy() for key in ("flag_start", "flag_end", "q_0")}
                ref = oracle.trace_scott(f64[sp["field"]], seeds, method=rq["method"], s0=rq["step"], tol=rq["tol"])
                cls = parity.class_match(res["flag_start"], res["flag_end"], ref["flag_start"], ref["flag_end"])
                if rq["method"] == "rk4":
                    rk4_cls.append(cls)
                    rk4_d.append(np.where(cls, parity.log_q_diff(res["q_0"], ref["q0"]), np.inf))
                else:
                    tr = oracle.trace_scott(f64[sp["field"]], seeds, method="rkf45", s0=rq["step"], tol=1e-10)
                    same = parity.class_match(ref["flag_start"], ref["flag_end"], tr["flag_start"], tr["flag_end"])
                    ct = parity.class_match(res["flag_start"], res["flag_end"], tr["flag_start"], tr["flag_end"])
                    m = same & ct
                    err_gpu.append(parity.log_q_diff(res["q_0"], tr["q0"])[m])
                    err_orc.append(parity.log_q_diff(ref["q0"], tr["q0"])[m])
            rk4_cls, rk4_d = np.concatenate(rk4_cls), np.concatenate(rk4_d)
            eg, eo = np.concatenate(err_gpu), np.concatenate(err_orc)
            fin = np.isfinite(eg) & np.isfinite(eo)
            print(f"RK4 sample: {rk4_cls.size} lines, class match {rk4_cls.mean():.4f}, "
                  f"within 1e-3 {np.mean(rk4_d <= 1e-3):.4f}; RKF45 sample vs converged: median gpu "
                  f"{np.median(eg[fin]):.2e} oracle {np.median(eo[fin]):.2e}", flush=True)
            assert rk4_cls.mean() >= 0.99 and np.mean(rk4_d <= 1e-3) >= 0.98
            assert np.median(eg[fin]) <= 1.5 * np.median(eo[fin]) + 1e-6
            assert np.quantile(eg[fin], 0.99) <= 1.5 * np.quantile(eo[fin], 0.99) + 1e-3
    finally:
        for f in fields:
            f.close()
        del arrs
        torch.cuda.empty_cache()
