"""The library's own plane map and request server, called through the C ABI.

* ``fqsl_qmap_plane`` forms the seeds of a constant-z plane and of a plane
  through three points on the device; every output equals ``fqsl_trace_scott``
  on the seeds ``synth.plane_seeds_z`` / ``plane_seeds_abc`` build on the host,
  bit for bit.
* ``fqsl_qmap_plane_neighbour`` (the whole neighbour-difference pipeline on
  one stream) equals the same steps run from Python on host-built seeds.
* ``fqsl_submit`` / ``fqsl_poll`` / ``fqsl_wait``: requests over two fields,
  both methods and uneven sizes, submitted from two host threads, are traced
  by the server's consumer in batches; each request's outputs equal tracing it
  alone, bit for bit.  Unknown and already-waited ids fail, also for a
  second thread waiting on the same id.
"""
import threading

import numpy as np
import pytest

from fastqsl_amd import synth

pytestmark = pytest.mark.gpu

KEYS = ("q_0", "q_perp", "flag_start", "flag_end", "evals", "start_x", "end_y", "b_this_z", "line_len")


@pytest.fixture(scope="module")
def fq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd
    fastqsl_amd.lib()
    return fastqsl_amd


@pytest.fixture(scope="module")
def fields(fq):
    fs = [fq.Field(*synth.quadrupole_64(device="cuda")),
          fq.Field(*synth.snapshot(1, shape=(72, 64, 56), device="cuda"))]
    yield fs
    for f in fs:
        f.close()


def _same(a, b, keys=KEYS):
    for k in keys:
        x, y = a[k].cpu().numpy().reshape(-1), b[k].cpu().numpy().reshape(-1)
        assert np.array_equal(x, y, equal_nan=True), k


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_qmap_plane_matches_trace_scott(fq, fields, precision):
    import torch
    from fastqsl_amd import qmap
    f = fields[0]
    cases = [(("z", 2.0, 60.5, 3.25, 61.0, 0.0), 70, 50),
             (("z", 10.0, 40.0, 5.0, 50.0, 7.3), 33, 1),
             (("abc", (3.0, 4.0, 0.5), (58.0, 9.0, 2.0), (6.0, 55.0, 30.0)), 48, 40)]
    for plane, nu, nv in cases:
        got = qmap.scott_qmap_plane(f, plane, nu, nv, method="rkf45", step=0.25, tol=1e-6, precision=precision,
                                    schedule="per_line" if precision == "fp64" else "persistent")
        if plane[0] == "z":
            seeds, shape = synth.plane_seeds_z(nu, nv, plane[1:3], plane[3:5], z=plane[5])
        else:
            seeds, shape, _ = synth.plane_seeds_abc(*plane[1:], nu, nv)
        assert shape == (nv, nu)
        want = fq.trace_scott(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", step=0.25, tol=1e-6,
                              precision=precision, schedule="per_line" if precision == "fp64" else "persistent")
        torch.cuda.synchronize()
        assert got["q_0"].shape == (nv, nu)
        _same(got, want)


def test_qmap_plane_rejects_bad_planes(fq, fields):
    from fastqsl_amd import qmap
    with pytest.raises(ValueError):
        qmap.scott_qmap_plane(fields[0], ("xz", 0, 1, 0, 1, 0), 4, 4)
    with pytest.raises(fq.FqslError):
        qmap.scott_qmap_plane(fields[0], ("z", 0.0, float("nan"), 0.0, 1.0, 0.0), 4, 4)


def test_native_server_matches_alone(fq, fields):
    import torch
    from fastqsl_amd.serve import NativeServer
    rng = np.random.default_rng(5)
    jobs = []
    for k in range(10):
        f = fields[k % 2]
        nx, ny, nz = f.shape
        n = int(rng.integers(1, 3000))
        seeds = np.column_stack([rng.uniform(1, nx - 2, n), rng.uniform(1, ny - 2, n),
                                 rng.uniform(0, nz / 3, n)]).astype(np.float32)
        method = "rk4" if k % 3 == 0 else "rkf45"
        jobs.append((f, seeds, method, float(10.0 ** rng.uniform(-7, -5))))
    srv = NativeServer("fp32c", max_lines=5000)
    ids = [None] * len(jobs)

    def submit(part):
        for k in part:
            f, s, m, tol = jobs[k]
            ids[k] = srv.submit(f, s[:, 0], s[:, 1], s[:, 2], method=m, step=0.25, tol=tol,
                                outputs=("q_0", "q_perp", "flag_start", "flag_end", "evals"))

    th = [threading.Thread(target=submit, args=(range(i, len(jobs), 2),)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert sorted(ids) == list(range(1, len(jobs) + 1))
    while not all(srv.poll(i) for i in ids):
        pass
    res = [srv.wait(i) for i in ids]
    with pytest.raises(fq.FqslError):
        srv.wait(ids[0])          # forgotten once waited for
    with pytest.raises(fq.FqslError):
        srv.poll(10 ** 6)
    srv.close()
    for (f, s, m, tol), r in zip(jobs, res):
        a = fq.trace_scott(f, s[:, 0], s[:, 1], s[:, 2], method=m, step=0.25, tol=tol, precision="fp32c")
        torch.cuda.synchronize()
        _same(r, a, ("q_0", "q_perp", "flag_start", "flag_end", "evals"))


def test_native_server_one_wait_per_id(fq, fields):
    """Two threads waiting on one request id: one gets the outputs, the
    other an error (the id is forgotten once waited for), and neither
    blocks."""
    import torch
    from fastqsl_amd.serve import NativeServer
    rng = np.random.default_rng(11)
    seeds = np.column_stack([rng.uniform(2, 60, 4000), rng.uniform(2, 60, 4000),
                             rng.uniform(0, 20, 4000)]).astype(np.float32)
    srv = NativeServer("fp32c")
    rid = srv.submit(fields[0], seeds[:, 0], seeds[:, 1], seeds[:, 2], outputs=("q_perp", "evals"))
    got = []

    def waiter():
        try:
            got.append(("ok", srv.wait(rid)))
        except fq.FqslError:
            got.append(("error", None))

    th = [threading.Thread(target=waiter) for _ in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in th)
    assert sorted(g[0] for g in got) == ["error", "ok"]
    res = next(g[1] for g in got if g[0] == "ok")
    srv.close()
    want = fq.trace_scott(fields[0], seeds[:, 0], seeds[:, 1], seeds[:, 2], precision="fp32c",
                          outputs=("q_perp", "evals"))
    torch.cuda.synchronize()
    _same(res, want, ("q_perp", "evals"))


@pytest.mark.parametrize("kind", ["z", "abc"])
def test_neighbour_qmap_plane_matches_the_python_pipeline(fq, fields, kind):
    """fqsl_qmap_plane_neighbour (seeds, plain traces, QCalcPlane,
    CookPseudoLine, pseudo traces, Qcut, all on one stream) against the same
    steps run from Python (qmap.neighbour_qmap) on the host-built seeds."""
    import torch
    from fastqsl_amd import qmap
    f = fields[1]
    nu, nv = 90, 70
    stride = (69.0 - 2.0) / (nu - 1)
    if kind == "z":
        plane = ("z", 2.0, 69.0, 3.0, 60.0, 1e-4)
        seeds, shape = synth.plane_seeds_z(nu, nv, (2.0, 69.0), (3.0, 60.0), z=1e-4)
    else:
        plane = ("abc", (2.0, 3.0, 1e-4), (69.0, 3.0, 1e-4), (2.0, 60.0, 1e-4))
        seeds, shape, _ = synth.plane_seeds_abc(*plane[1:], nu, nv)
    got = qmap.neighbour_qmap_plane(f, plane, nu, nv, stride, step=0.125, tol_coef=0.5)
    want = qmap.neighbour_qmap(f, seeds, shape, stride, step=0.125, tol_coef=0.5)
    torch.cuda.synchronize()
    assert int(got["n_pseudo"]) == int(want["n_pseudo"]) > 0
    a, b = got["q_all"].cpu().numpy(), want["q_all"].cpu().numpy()
    assert a.shape == (nv, nu)
    assert np.array_equal(a, b, equal_nan=True)
    assert np.isfinite(a[1:-1, 1:-1]).mean() > 0.5
