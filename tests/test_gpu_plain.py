"""GPU parity of the neighbour-difference path and of the reference-shaped
drop-in entry points.

* Position-only tracer (TraceAllBline) against the FP64 oracle: RK4 lines are
  deterministic, so classes and footpoints must agree on well-conditioned
  lines; RKF45 is held to the oracle's own round-off sensitivity.
* QCalcPlane, CookPseudoLine and Qcut on the GPU against the NumPy float32
  restatement (oracle/qplane.py) on identical inputs: bit for bit.
* The full pipeline on the linear X-point field against the closed form.
* ``fastqsl_amd.kernels`` called with the reference's argument tuples
  (torch and NumPy arrays) against the direct API.
"""
import numpy as np
import pytest

from fastqsl_amd import synth
from oracle import qplane
from tests import parity

pytestmark = pytest.mark.gpu

TRACE_KEYS = ("start_x", "start_y", "start_z", "end_x", "end_y", "end_z", "flag_start",
              "flag_end", "b_this_x", "b_this_y", "b_this_z", "b_flag", "b_start_x", "b_start_y",
              "b_start_z", "b_end_x", "b_end_y", "b_end_z", "line_len", "evals")


@pytest.fixture(scope="module")
def fq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd
    fastqsl_amd.lib()
    return fastqsl_amd


def _np(r):
    import torch
    torch.cuda.synchronize()
    return {k: r[k].cpu().numpy() for k in TRACE_KEYS if k in r}


def _xyz(r, pre):
    return np.column_stack([r[pre + "_x"], r[pre + "_y"], r[pre + "_z"]])


@pytest.fixture(scope="module")
def quad():
    field = synth.quadrupole_64(dtype=np.float64, device="cpu")
    seeds, shape = synth.plane_seeds_z(96, 96, (2, 61), (2, 61), z=1e-4)
    return field, seeds, shape


def test_plain_rk4_matches_oracle(fq, oracle, quad):
    field, seeds, shape = quad
    ref = oracle.trace_plain(field, seeds, method="rk4", s0=0.25)
    refp = [oracle.trace_plain(field, seeds.astype(np.float64) + sgn * np.array([1e-6, -1e-6, 0]),
                               method="rk4", s0=0.25) for sgn in (1.0, -1.0)]
    cond = parity.conditioned(ref, *refp, key="end")
    print("well-conditioned:", cond.mean())
    assert cond.mean() > 0.99
    with fq.Field(*field) as f:
        for prec, tol in (("fp64", 1e-4), ("fp32c", 2e-3)):
            for sched in ("persistent", "per_line"):
                r = _np(fq.trace_plain(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rk4", step=0.25,
                                       precision=prec, schedule=sched))
                cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
                print(prec, sched, "class match", cls.mean(), "conditioned", cls[cond].mean())
                assert cls[cond].mean() >= (1.0 if prec == "fp64" else 0.9999)
                m = cond & cls
                de = np.abs(_xyz(r, "end") - ref["end"]).max(axis=1)[m]
                ds = np.abs(_xyz(r, "start") - ref["start"]).max(axis=1)[m]
                print("  endpoint error max %.2e / %.2e" % (de.max(), ds.max()))
                assert de.max() < tol and ds.max() < tol
                np.testing.assert_allclose(_xyz(r, "b_this"), ref["b_this"], rtol=2e-5, atol=1e-6)
                np.testing.assert_allclose(_xyz(r, "b_end")[m], ref["b_end"][m], rtol=1e-2, atol=1e-3)
                assert np.mean(r["b_flag"] == ref["b_flag"]) > 0.9995
                if prec == "fp64":
                    assert np.array_equal(r["evals"][m].astype(np.int64), ref["evals"][m])


def test_plain_rkf45_as_accurate_as_the_oracle(fq, oracle):
    """RKF45 with the reference's tolerance rule (1e-3 |b.n|^1.5 cells, up to
    1e-3): footpoints and classes are only as good as that tolerance, so the
    GPU is held to the oracle's own agreement with a converged trace
    (tol_coef 1e-4)."""
    field = synth.multipole_256(seed=0, device="cuda")
    rng = np.random.default_rng(3)
    seeds = np.column_stack([rng.uniform(0, 255, 4096), rng.uniform(0, 255, 4096),
                             np.full(4096, 1e-4)]).astype(np.float32)
    ref = oracle.trace_plain(field, seeds, method="rkf45", s0=0.125, tol_coef=1.0)
    truth = oracle.trace_plain(field, seeds, method="rkf45", s0=0.125, tol_coef=1e-4)
    c_orc = parity.class_match(ref["flag_start"], ref["flag_end"], truth["flag_start"], truth["flag_end"])
    e_orc = np.abs(ref["end"] - truth["end"]).max(axis=1)
    print("oracle vs converged: class %.5f end err median %.2e p99 %.2e"
          % (c_orc.mean(), np.median(e_orc[c_orc]), np.quantile(e_orc[c_orc], 0.99)))
    with fq.Field(*field) as f:
        for prec in ("fp64", "fp32c"):
            r = _np(fq.trace_plain(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", step=0.125,
                                   tol_coef=1.0, precision=prec))
            c = parity.class_match(r["flag_start"], r["flag_end"], truth["flag_start"], truth["flag_end"])
            e = np.abs(_xyz(r, "end") - truth["end"]).max(axis=1)
            m = c & c_orc
            print(prec, "vs converged: class %.5f end err median %.2e p99 %.2e"
                  % (c.mean(), np.median(e[m]), np.quantile(e[m], 0.99)))
            assert c.mean() >= c_orc.mean() - 2e-3
            assert np.median(e[m]) <= 1.5 * np.median(e_orc[m]) + 1e-5
            assert np.quantile(e[m], 0.99) <= 1.5 * np.quantile(e_orc[m], 0.99) + 1e-5


@pytest.fixture(scope="module")
def plane_traces(fq):
    """FP32 traces on a 160 x 128 grid over the random multipole's z = 0 plane."""
    field = synth.multipole((96, 96, 96), seed=5, device="cuda")
    seeds, shape = synth.plane_seeds_z(160, 128, (10, 85), (10, 73), z=1e-4)
    with fq.Field(*field) as f:
        r = _np(fq.trace_plain(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", step=0.125))
    return field, seeds, shape, {k: v.reshape(shape) for k, v in r.items()}


def test_qcalc_plane_bit_exact(fq, plane_traces):
    import torch
    field, seeds, shape, g = plane_traces
    stride = 75.0 / 159
    for cross in ((0.0, 0.0, 1.0), (0.3, -0.2, 0.93)):
        args = (g["end_x"], g["end_y"], g["end_z"], g["flag_end"], g["start_x"], g["start_y"],
                g["start_z"], g["flag_start"], g["b_start_x"], g["b_start_y"], g["b_start_z"],
                g["b_end_x"], g["b_end_y"], g["b_end_z"], g["b_this_x"], g["b_this_y"], g["b_this_z"],
                g["b_flag"], stride)
        ref = qplane.qcalc_plane(*args, cross_dir=cross)
        got = fq.FastQSL.QCalcPlane(*args, CrossBC_dir=np.asarray(cross, np.float32))
        assert got.shape == ref.shape == (shape[0] - 2, shape[1] - 2)
        np.testing.assert_array_equal(got, ref)
        assert np.isnan(ref).any() and np.isfinite(ref).mean() > 0.5
        # CUDA-tensor inputs give a CUDA tensor back
        gt = fq.FastQSL.QCalcPlane(*[torch.as_tensor(a).cuda() if isinstance(a, np.ndarray) else a
                                     for a in args], CrossBC_dir=cross)
        assert gt.is_cuda
        np.testing.assert_array_equal(gt.cpu().numpy(), ref)


def test_pseudo_lines_and_qcut_bit_exact(fq, plane_traces):
    field, seeds, shape, g = plane_traces
    stride = np.float32(0.25)
    flagged = int((g["b_flag"] == 1).sum())
    print("pseudo pixels:", flagged)
    assert flagged > 0
    ref_seeds, ref_bzs, ref_bze = qplane.pseudo_seeds(g["start_x"], g["start_y"], g["start_z"],
                                                      g["flag_start"], g["b_start_z"], g["b_end_z"],
                                                      g["b_flag"], stride)
    tup = fq.FastQSL.CookPseudoLine(g["end_x"], g["end_y"], g["end_z"], g["flag_end"],
                                    g["start_x"], g["start_y"], g["start_z"], g["flag_start"],
                                    g["b_start_x"], g["b_start_y"], g["b_start_z"],
                                    g["b_end_x"], g["b_end_y"], g["b_end_z"],
                                    g["b_this_z"], g["b_flag"], stride)
    assert len(tup) == 25
    np.testing.assert_array_equal(np.column_stack(tup[0:3]), ref_seeds)
    np.testing.assert_array_equal(tup[23], ref_bzs)
    np.testing.assert_array_equal(tup[24], ref_bze)
    assert int(tup[21][0]) == 4 * flagged
    # trace the pseudo seeds, then Qcut + merge on both sides
    with fq.Field(*field) as f:
        r = _np(fq.trace_plain(f, tup[0], tup[1], tup[2], method="rkf45", step=0.125, tol_coef=0.1))
    st, en = _xyz(r, "start"), _xyz(r, "end")
    ref_q = qplane.qcut(st, r["flag_start"], en, r["flag_end"], ref_bzs, ref_bze, stride)
    got_q = fq.FastQSL.Qcut(r["start_x"], r["start_y"], r["start_z"], r["flag_start"], r["end_x"],
                            r["end_y"], r["end_z"], r["flag_end"], tup[23], tup[24], stride)
    np.testing.assert_array_equal(got_q, ref_q)
    import torch
    q_inner = qplane.qcalc_plane(g["end_x"], g["end_y"], g["end_z"], g["flag_end"], g["start_x"],
                                 g["start_y"], g["start_z"], g["flag_start"], g["b_start_x"],
                                 g["b_start_y"], g["b_start_z"], g["b_end_x"], g["b_end_y"],
                                 g["b_end_z"], g["b_this_x"], g["b_this_y"], g["b_this_z"],
                                 g["b_flag"], 0.25)
    ref_final = qplane.merge(q_inner, ref_q, g["b_flag"])
    q_all = torch.zeros(shape, dtype=torch.float32, device="cuda")
    q_all[1:-1, 1:-1] = torch.as_tensor(q_inner).cuda()
    fq.FastQSL.Qcut(r["start_x"], r["start_y"], r["start_z"], r["flag_start"], r["end_x"], r["end_y"],
                    r["end_z"], r["flag_end"], tup[23], tup[24], stride, Q_all=q_all, B_flag_arr=g["b_flag"])
    np.testing.assert_array_equal(q_all[1:-1, 1:-1].cpu().numpy(), ref_final)


def test_neighbour_pipeline_hyperbolic(fq):
    shape = (41, 41, 33)
    k = 0.03
    field = synth.hyperbolic(shape, k=k, b0=1.0)
    seeds, sh = synth.plane_seeds_z(24, 24, (17, 23), (17, 23), z=1e-4)
    with fq.Field(*field) as f:
        for method in ("rk4", "rkf45"):
            out = fq.qmap.neighbour_qmap(f, seeds, sh, stride=6.0 / 23, method=method, step=0.125,
                                         tol_coef=1e-4)
            q = out["q"].cpu().numpy()
            assert int(out["n_pseudo"]) == 0
            np.testing.assert_allclose(q, 2 * np.cosh(2 * k * shape[2]), rtol=3e-3)


def _scott_args(field32, seeds, torch_arrays):
    import torch
    bx, by, bz = (np.asfortranarray(b.astype(np.float32)) for b in field32)
    n = seeds.shape[0]
    f32 = lambda: np.zeros(n, np.float32)
    a = dict(Bx=bx, By=by, Bz=bz, BshapeN=np.array(bx.shape, np.int32), d1=np.zeros((1, 1), np.float32),
             d2=np.zeros((1, 1), np.float32), d3=np.zeros((1, 1), np.float32),
             d4=np.zeros((1, 1), np.float32), flag_twist=np.array([False]),
             x=np.ascontiguousarray(seeds[:, 0]), y=np.ascontiguousarray(seeds[:, 1]),
             z=np.ascontiguousarray(seeds[:, 2]), norm=np.array([0, 0, 1], np.float32),
             sx=f32(), sy=f32(), sz=f32(), fs=np.zeros(n, np.int32), ex=f32(), ey=f32(), ez=f32(),
             fe=np.zeros(n, np.int32), btx=f32(), bty=f32(), btz=f32(), q0=f32(), qp=f32(),
             s_len=np.array([0.25], np.float32), N=np.array([n], np.uint64),
             line_len=np.zeros(n, np.float64), tol_coef=np.array([1e-3], np.float32))
    if torch_arrays:
        a = {k: (torch.as_tensor(v.ravel(order="K").view(np.uint8) if v.dtype == np.bool_ else
                                 v.ravel(order="K").astype(np.int64) if v.dtype == np.uint64 else
                                 v.ravel(order="K")).cuda()) for k, v in a.items()}
        a["flag_twist"] = a["flag_twist"].to(torch.bool)
    return a


def test_dropin_scott_kernel_matches_direct_api(fq):
    import torch
    field = synth.multipole((64, 48, 40), seed=2, device="cpu")
    rng = np.random.default_rng(1)
    seeds = np.column_stack([rng.uniform(0, 63, 3000), rng.uniform(0, 47, 3000),
                             np.r_[np.zeros(2000), rng.uniform(0, 39, 1000)]]).astype(np.float32)
    # the drop-in reads the caller's float arrays in place (corner form of the
    # interpolant); the handle API on the same in-place SoA layout must give
    # the same bits.  (The default cell layout holds trilinear coefficients,
    # whose float rounding differs in the last bits.)
    with fq.Field(*field, layouts=("soa32", "soa64")) as f:
        # the drop-in's tolerance is 1e-3 * tol_coef with tol_coef a float32
        tol = 1e-3 * float(np.float32(1e-3))
        ref = fq.trace_scott(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", step=0.25, tol=tol)
        torch.cuda.synchronize()
        ref = {k: v.cpu().numpy() for k, v in ref.items() if k != "_seeds"}
    kern = fq.kernels.compileTraceBlineScott()
    for use_torch in (False, True):
        for sched in ("persistent", "per_line"):
            kern.schedule = sched
            a = _scott_args(field, seeds, use_torch)
            kern((16, 1), (256, 1, 1), tuple(a.values()))
            torch.cuda.synchronize()
            got = {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in a.items()}
            for mine, theirs in (("sx", "start_x"), ("ey", "end_y"), ("fs", "flag_start"), ("fe", "flag_end"),
                                 ("q0", "q_0"), ("qp", "q_perp"), ("btz", "b_this_z"), ("line_len", "line_len")):
                assert np.array_equal(got[mine], ref[theirs], equal_nan=True), (use_torch, sched, mine)
    # with curB_flag set the twist output must be a float64 array (here a float32 dummy)
    a = _scott_args(field, seeds[:10], False)
    a["flag_twist"][0] = True
    with pytest.raises(TypeError, match="twist"):
        kern((1, 1), (64, 1, 1), tuple(a.values()))


def test_dropin_plain_kernels_match_direct_api(fq):
    import torch
    field = synth.multipole((64, 48, 40), seed=3, device="cpu")
    seeds, shape = synth.plane_seeds_z(50, 40, (2, 60), (2, 45), z=1e-4)
    n = seeds.shape[0]
    bx, by, bz = (np.asfortranarray(b.astype(np.float32)) for b in field)
    for which, ctor in (("rkf45", fq.kernels.compileTraceBlineAdaptive), ("rk4", fq.kernels.compileTraceBlineRK4)):
        with fq.Field(*field, layouts=("soa32", "soa64")) as f:   # the drop-in's in-place layout
            ref = fq.trace_plain(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], method=which, step=0.125,
                                 tol_coef=0.5, precision="fp32c")
            torch.cuda.synchronize()
            ref = {k: v.cpu().numpy() for k, v in ref.items() if k != "_seeds"}
        f32 = lambda: np.zeros(n, np.float32)
        outs = dict(sx=f32(), sy=f32(), sz=f32(), fs=np.zeros(n, np.int32), ex=f32(), ey=f32(), ez=f32(),
                    fe=np.zeros(n, np.int32), btx=f32(), bty=f32(), btz=f32(), bf=np.zeros(n, np.int32),
                    bsx=f32(), bsy=f32(), bsz=f32(), bex=f32(), bey=f32(), bez=f32())
        dummy = np.zeros((1, 1), np.float32)
        args = [bx, by, bz, np.array(bx.shape, np.int32), dummy, dummy, dummy, dummy, np.array([False]),
                np.ascontiguousarray(seeds[:, 0]), np.ascontiguousarray(seeds[:, 1]),
                np.ascontiguousarray(seeds[:, 2]), np.array([0, 0, 1], np.float32)]
        args += list(outs.values())
        line_len = np.zeros(n, np.float64)
        args += [np.array([0.125], np.float32), np.array([n], np.uint64), line_len]
        if which == "rkf45":
            args.append(np.array([0.5], np.float32))
        ctor()((28, 1), (128, 1, 1), tuple(args))
        for mine, theirs in (("sx", "start_x"), ("ez", "end_z"), ("fs", "flag_start"), ("fe", "flag_end"),
                             ("bf", "b_flag"), ("bsz", "b_start_z"), ("bez", "b_end_z"), ("btx", "b_this_x")):
            assert np.array_equal(outs[mine], ref[theirs], equal_nan=True), (which, mine)
        assert np.array_equal(line_len, ref["line_len"]), which


@pytest.mark.parametrize("b0", [1.0, 0.0])
@pytest.mark.parametrize("method", ["rk4", "rkf45"])
def test_plain_edge_seeds(fq, oracle, b0, method):
    """Seeds at the box's edges (below the surface, on faces and edges, on
    the strict y = 0 plane, outside, NaN, at a null): the exit classes of
    the oracle in both precisions and schedules (checkFlag,
    TraceBlineAdaptive.cu:276-298; drivers :333-426, TraceBline.cu:300-378);
    FP64 RK4 footpoints to 1e-6 cells with the oracle's evaluation counts."""
    seeds = parity.EDGE_SEEDS
    field = synth.hyperbolic((32, 32, 32), k=0.03, b0=b0)
    ref = oracle.trace_plain(field, seeds, method=method, s0=0.25, tol_coef=1.0)
    print(f"edge seeds b0 {b0} {method}", flush=True)
    with fq.Field(*field) as f:
        for prec in ("fp64", "fp32c"):
            for sched in ("persistent", "per_line"):
                r = _np(fq.trace_plain(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], method=method, step=0.25,
                                       tol_coef=1.0, precision=prec, schedule=sched))
                np.testing.assert_array_equal(r["flag_start"], ref["flag_start"])
                np.testing.assert_array_equal(r["flag_end"], ref["flag_end"])
                np.testing.assert_array_equal(np.isnan(_xyz(r, "end")), np.isnan(ref["end"]))
                if prec == "fp64" and method == "rk4":
                    np.testing.assert_allclose(_xyz(r, "end"), ref["end"], rtol=0, atol=1e-6)
                    np.testing.assert_allclose(_xyz(r, "start"), ref["start"], rtol=0, atol=1e-6)
                    np.testing.assert_array_equal(r["evals"], ref["evals"])


def test_empty_inputs(fq):
    """No seeds: every tracer returns empty outputs without a launch."""
    import torch
    e = torch.empty(0, dtype=torch.float32, device="cuda")
    field = synth.hyperbolic((16, 16, 16), k=0.03, b0=1.0)
    with fq.Field(*field) as f:
        for prec in ("fp64", "fp32c"):
            for sched in ("persistent", "per_line"):
                r = fq.trace_plain(f, e, e, e, precision=prec, schedule=sched)
                s = fq.trace_scott(f, e, e, e, precision=prec, schedule=sched)
                assert all(r[k].numel() == 0 for k in TRACE_KEYS)
                assert s["q_0"].numel() == 0 and s["flag_end"].numel() == 0
        ax = np.linspace(0.0, 15.0, 16)
        f.set_axes(ax, ax, ax)
        r = fq.trace_plain(f, e, e, e, step=0.01)
        assert r["end_x"].numel() == 0
    torch.cuda.synchronize()


@pytest.mark.parametrize("streams", [2, 4])
def test_neighbour_qmaps_over_streams_match_serial(fq, streams):
    """Maps run as one job over two or four streams (qmap.neighbour_qmaps)
    equal the same maps run one at a time, bit for bit (a line's result does
    not depend on the launch or stream that traces it)."""
    import torch
    field = synth.quadrupole_64(dtype=np.float64, device="cpu", n=48)
    jobs = []
    for (x0, x1), z in (((2, 45), 1e-4), ((5, 40), 0.5), ((1, 46), 1e-4)):
        seeds, shape = synth.plane_seeds_z(72, 60, (x0, x1), (x0, x1), z=z)
        jobs.append((seeds, shape, (x1 - x0) / 71.0))
    kw = dict(method="rkf45", step=0.125, tol_coef=0.5, pseudo_tol_scale=0.2)
    with fq.Field(*field) as f:
        serial = [fq.qmap.neighbour_qmap(f, s, sh, st, **kw) for s, sh, st in jobs]
        torch.cuda.synchronize()
        piped = fq.qmap.neighbour_qmaps(f, jobs, streams=streams, **kw)
        torch.cuda.synchronize()
        for a, b in zip(serial, piped):
            assert int(a["n_pseudo"]) == int(b["n_pseudo"])
            np.testing.assert_array_equal(a["q_all"].cpu().numpy(), b["q_all"].cpu().numpy())
            np.testing.assert_array_equal(a["traces"]["end_x"].cpu().numpy(), b["traces"]["end_x"].cpu().numpy())
    assert any(int(a["n_pseudo"]) > 0 for a in serial)


def test_pseudo_pipeline_counts_on_the_device(fq):
    """The no-round-trip pseudo-line path: the compaction's device counts
    equal the host-counted CookPseudoLine core, its leading outputs are the
    same bits, and a trace / Qcut bounded by a device count equals the same
    calls on exactly that many seeds / pixels."""
    import torch
    from fastqsl_amd import FastQSL
    field = synth.quadrupole_64(dtype=np.float64, device="cpu", n=48)
    # 73 x 73 pixels over (2, 45): the middle row and column lie on the
    # polarity inversion lines x, y = 23.5 (Bz = 0 by symmetry), so B_flag is set there
    seeds, shape = synth.plane_seeds_z(73, 73, (2, 45), (2, 45), z=1e-4)
    with fq.Field(*field) as f:
        r = fq.trace_plain(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", step=0.125,
                           tol_coef=0.5)
        args = (r["start_x"], r["start_y"], r["start_z"], r["flag_start"], r["b_start_z"],
                r["b_end_z"], r["b_flag"], 43.0 / 72)
        host = FastQSL.cook_pseudo_lines(*args)
        dev = FastQSL.cook_pseudo_lines_async(*args)
        m = host["pixel"].numel()
        assert m > 0
        assert dev["counts"].cpu().tolist() == [m, 4 * m]
        for k in ("seed_x", "seed_y", "seed_z", "bz0_start", "bz0_end", "pixel"):
            np.testing.assert_array_equal(dev[k][:host[k].numel()].cpu().numpy(), host[k].cpu().numpy())
        kw = dict(method="rkf45", step=0.125, tol_coef=0.05,
                  outputs=("start_x", "start_y", "start_z", "end_x", "end_y", "end_z", "flag_start",
                           "flag_end"))
        for schedule in ("persistent", "per_line"):
            a = fq.trace_plain(f, host["seed_x"], host["seed_y"], host["seed_z"], schedule=schedule, **kw)
            out = {k: torch.full((4 * shape[0] * shape[1],), -7, dtype=v.dtype, device="cuda")
                   for k, v in a.items() if k != "_seeds"}
            b = fq.trace_plain(f, dev["seed_x"], dev["seed_y"], dev["seed_z"], schedule=schedule,
                               n_device=dev["counts"][1:], out=out, **kw)
            for k in out:
                np.testing.assert_array_equal(b[k][:4 * m].cpu().numpy(), a[k].cpu().numpy())
                assert bool((b[k][4 * m:] == -7).all()), k   # nothing past the count is written
        rc = fq.trace_plain(f, host["seed_x"], host["seed_y"], host["seed_z"],
                            outputs=("start_x", "start_y", "start_z", "end_x", "end_y", "end_z",
                                     "flag_start", "flag_end"), method="rkf45", step=0.125, tol_coef=0.05)
        qa = torch.zeros(shape, dtype=torch.float32, device="cuda")
        qb = torch.zeros(shape, dtype=torch.float32, device="cuda")
        cut = [rc[k] for k in ("start_x", "start_y", "start_z", "flag_start", "end_x", "end_y", "end_z",
                               "flag_end")]
        FastQSL.Qcut(*cut, host["bz0_start"], host["bz0_end"], 43.0 / 72, Q_all=qa, pixel=host["pixel"])
        pad = lambda t, n: torch.cat([t, torch.zeros(n - t.numel(), dtype=t.dtype, device="cuda")])
        ntot = shape[0] * shape[1]
        FastQSL.Qcut(*[pad(t, 4 * ntot) for t in cut], pad(host["bz0_start"], ntot),
                     pad(host["bz0_end"], ntot), 43.0 / 72, Q_all=qb,
                     pixel=pad(host["pixel"], ntot), m_device=dev["counts"][:1])
        np.testing.assert_array_equal(qa.cpu().numpy(), qb.cpu().numpy())


def test_neighbour_qmap_tile_order_does_not_change_results(fq):
    """Tracing the main lines in 16 x 16 pixel blocks (the default) or row by
    row gives the same map and traces, bit for bit."""
    import torch
    field = synth.quadrupole_64(dtype=np.float64, device="cpu", n=48)
    seeds, shape = synth.plane_seeds_z(70, 52, (2, 45), (3, 44), z=1e-4)   # not a multiple of the tile
    kw = dict(method="rkf45", step=0.125, tol_coef=0.5, pseudo_tol_scale=0.2)
    with fq.Field(*field) as f:
        a = fq.qmap.neighbour_qmap(f, seeds, shape, 43.0 / 69, tile=0, **kw)
        b = fq.qmap.neighbour_qmap(f, seeds, shape, 43.0 / 69, **kw)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a["q_all"].cpu().numpy(), b["q_all"].cpu().numpy())
        for k in ("end_x", "start_y", "flag_end", "b_flag", "b_end_z", "line_len"):
            np.testing.assert_array_equal(a["traces"][k].cpu().numpy(), b["traces"][k].cpu().numpy())


def test_scott_qmap_tile_order_does_not_change_results(fq):
    """The Scott Q map traced in 16 x 16 pixel blocks (the default) or row by
    row: the same map, bit for bit."""
    import torch
    field = synth.quadrupole_64(dtype=np.float64, device="cpu", n=48)
    seeds, shape = synth.plane_seeds_z(70, 52, (2, 45), (3, 44), z=0.0)
    with fq.Field(*field) as f:
        a = fq.qmap.scott_qmap(f, seeds, shape, tile=0)
        b = fq.qmap.scott_qmap(f, seeds, shape)
        torch.cuda.synchronize()
        for k in a:
            np.testing.assert_array_equal(a[k].cpu().numpy(), b[k].cpu().numpy())
