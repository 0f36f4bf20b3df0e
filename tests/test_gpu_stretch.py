"""Stretched-grid tracer (TraceBlineAdaptiveStretch.cu) on the GPU.

* A linear field on stretched axes: footpoints in closed form
  for every cell rule, both precisions.
* The landing plane of each rule (0 / n_dim for the reference rules, the
  faces for the exact rule).
* Against the FP64 oracle on a random multipole over a stretched grid, with
  the bars of the host build of the same code (tests/test_stretch_oracle.py):
  FP64 to the tolerance level, FP32c to 1e-4 with the exact rule and within
  the oracle's own 10 %-tolerance spread with the reference rules.
* B_this / B_start / B_end and B_flag against the oracle's stretched
  Interp3dxyzn.
* Persistent and per-line schedules bit for bit; the LDS copy of the axis
  lookup against the global tables bit for bit; the 39-argument drop-in
  (flag_uni false / true, curB_flag) bit for bit against the handle API.
* Twist: a unit-magnitude force-free field gives a L / 4 pi.
"""
import functools
import math

import numpy as np
import pytest

from fastqsl_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd
    fastqsl_amd.lib()
    return fastqsl_amd


def _run(fq, field, seeds, precision, schedule="persistent", outputs=None, **kw):
    import torch
    from fastqsl_amd.trace import PLAIN_OUTPUTS
    outs = outputs or PLAIN_OUTPUTS
    r = fq.trace_plain(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rkf45", precision=precision,
                       schedule=schedule, outputs=outs, **kw)
    torch.cuda.synchronize()
    return {k: r[k].cpu().numpy() for k in outs}


def _ends(r):
    s = np.column_stack([r["start_x"], r["start_y"], r["start_z"]]).astype(np.float64)
    e = np.column_stack([r["end_x"], r["end_y"], r["end_z"]]).astype(np.float64)
    return s, e


def _axes_to_n(shape):
    return [synth.stretched_axis(n, 0.0, float(n), kind, 1.5, dtype=np.float64)
            for n, kind in zip(shape, ("sinh", "sinh", "geometric"))]


@pytest.mark.parametrize("precision,tol", [("fp64", 2e-6), ("fp32c", 1e-5)])
@pytest.mark.parametrize("rule", ["bisect", "uniform", "exact"])
def test_linear_field_closed_form(fq, rule, precision, tol):
    shape = (24, 22, 20)
    axes = _axes_to_n(shape)
    k, b0, centre = 0.05, 1.0, (12.0, 11.0)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    f = (k * (X - centre[0]), -k * (Y - centre[1]), np.full(X.shape, b0))
    rng = np.random.default_rng(5)
    z0 = 4.0
    seeds = np.column_stack([centre[0] + rng.uniform(-2, 2, 500), centre[1] + rng.uniform(-2, 2, 500),
                             np.full(500, z0)]).astype(np.float32)
    with fq.Field(*f, axes=axes, lookup=rule) as field:
        assert field.stretched
        r = _run(fq, field, seeds, precision, step=0.05, tol_coef=1e-4)
    s, e = _ends(r)
    assert np.all(r["flag_end"] == 6) and np.all(r["flag_start"] == 5)
    zt = axes[2][-1]
    g = math.exp(k * (zt - z0) / b0)
    sd = seeds.astype(np.float64)
    np.testing.assert_allclose(e[:, 2], zt, atol=1e-5)
    np.testing.assert_allclose(e[:, 0] - centre[0], (sd[:, 0] - centre[0]) * g, atol=tol * 10)
    np.testing.assert_allclose(e[:, 1] - centre[1], (sd[:, 1] - centre[1]) / g, atol=tol * 10)
    g = math.exp(-k * z0 / b0)
    np.testing.assert_allclose(s[:, 2], 0.0, atol=1e-5)
    np.testing.assert_allclose(s[:, 0] - centre[0], (sd[:, 0] - centre[0]) * g, atol=tol * 10)


@pytest.mark.parametrize("rule,plane_lo,plane_hi", [("bisect", 0.0, 40.0), ("uniform", 0.0, 40.0),
                                                    ("exact", -5.0, 5.0)])
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_landing_plane(fq, rule, plane_lo, plane_hi, precision):
    shape = (40, 12, 10)
    axes = [synth.stretched_axis(shape[0], -5.0, 5.0, "sinh", 1.0, dtype=np.float64),
            np.linspace(0.0, 3.0, shape[1]), np.linspace(0.0, 2.0, shape[2])]
    f = tuple(np.full(shape, c) for c in (-1.0, 0.0, 0.0))
    seeds = np.array([[1.3, 1.5, 0.5], [-2.2, 0.7, 1.1], [3.9, 2.2, 0.9]], np.float32)
    with fq.Field(*f, axes=axes, lookup=rule) as field:
        r = _run(fq, field, seeds, precision, ncross=(1.0, 0.0, 0.0), step=0.1)
    s, e = _ends(r)
    np.testing.assert_array_equal(r["flag_end"], 1)
    np.testing.assert_array_equal(r["flag_start"], 2)
    np.testing.assert_allclose(e[:, 0], plane_lo, atol=1e-5)
    np.testing.assert_allclose(s[:, 0], plane_hi, atol=1e-5)
    np.testing.assert_allclose(e[:, 1:], seeds[:, 1:], atol=1e-6)


@functools.lru_cache(maxsize=None)
def _case(rule, n=4000):
    kinds = ("uniform",) * 3 if rule == "uniform" else ("sinh", "sinh", "geometric")
    f, axes = synth.stretched_multipole((60, 56, 48), seed=6, dtype=np.float64, device="cpu", kinds=kinds)
    f = tuple(np.asarray(b, np.float64) for b in f)
    axes = [np.asarray(a, np.float64) for a in axes]
    rng = np.random.default_rng(3)
    seeds = np.column_stack([rng.uniform(-0.8, 0.8, n), rng.uniform(-0.8, 0.8, n),
                             rng.uniform(0.01, 0.5, n)]).astype(np.float32)
    return f, axes, seeds


def _agree(o, start, end, fs, fe):
    """Per line: same classes on both sides, and the larger footpoint distance."""
    same = (o["flag_start"] == fs) & (o["flag_end"] == fe)
    d = np.maximum(np.max(np.abs(o["end"] - end), axis=1), np.max(np.abs(o["start"] - start), axis=1))
    return same, d


@pytest.mark.parametrize("rule", ["bisect", "uniform", "exact"])
def test_fp64_matches_oracle(fq, oracle, rule):
    f, axes, seeds = _case(rule)
    o = oracle.trace_stretch(f, axes, seeds, s0=0.01, tol_coef=1e-2, rule=rule)
    with fq.Field(*f, axes=axes, lookup=rule) as field:
        r = _run(fq, field, seeds, "fp64", step=0.01, tol_coef=1e-2)
    s, e = _ends(r)
    same, d = _agree(o, s, e, r["flag_start"], r["flag_end"])
    face = ((o["flag_start"] >= 1) & (o["flag_start"] <= 6) & (o["flag_end"] >= 1) & (o["flag_end"] <= 6))
    assert face.mean() > 0.5
    assert same.mean() > 0.99 and np.mean(d[face] < 1e-4) > 0.98, (same.mean(), np.mean(d[face] < 1e-4))
    # the seed's B (stretched Interp3dxyzn, raw) and B_flag
    bt = np.column_stack([r["b_this_x"], r["b_this_y"], r["b_this_z"]])
    np.testing.assert_allclose(bt, o["b_this"], rtol=1e-6, atol=1e-6 * np.abs(o["b_this"]).max())
    np.testing.assert_array_equal(r["b_flag"], o["b_flag"])
    # B at the (float32) footpoints, as the reference evaluates it at P_out
    be = np.column_stack([r["b_end_x"], r["b_end_y"], r["b_end_z"]])
    want = oracle.interp_stretch(f, axes, e, rule=rule)
    np.testing.assert_allclose(be, want, rtol=1e-6, atol=1e-6 * np.abs(want).max())


def test_fp32c_exact_rule_matches_oracle(fq, oracle):
    f, axes, seeds = _case("exact", 1500)
    o = oracle.trace_stretch(f, axes, seeds, s0=0.01, tol_coef=1e-4, rule="exact")
    with fq.Field(*f, axes=axes, lookup="exact") as field:
        r = _run(fq, field, seeds, "fp32c", step=0.01, tol_coef=1e-4)
    s, e = _ends(r)
    same, d = _agree(o, s, e, r["flag_start"], r["flag_end"])
    assert same.mean() > 0.99 and np.mean(same & (d < 1e-4)) > 0.97, (same.mean(), np.mean(same & (d < 1e-4)))


@pytest.mark.parametrize("rule", ["bisect", "uniform"])
def test_fp32c_reference_rules_within_spread(fq, oracle, rule):
    """The bar of the host build: the oracle's own spread when its tolerance
    moves by 10 % (the FP32 stages put percent-level noise on the error
    estimate at this tolerance, and the reference rules' interpolant jumps at
    some nodes; on the bisection rule FP32c sits between the 3 % and the 10 %
    spread, on the others at the 1 % spread)."""
    f, axes, seeds = _case(rule, 1500)
    o = oracle.trace_stretch(f, axes, seeds, s0=0.01, tol_coef=1e-3, rule=rule)
    with fq.Field(*f, axes=axes, lookup=rule) as field:
        r = _run(fq, field, seeds, "fp32c", step=0.01, tol_coef=1e-3)
    s, e = _ends(r)
    same, d = _agree(o, s, e, r["flag_start"], r["flag_end"])
    spread = []
    for eps in (1e-1, -1e-1):
        p = oracle.trace_stretch(f, axes, seeds, s0=0.01, tol_coef=1e-3 * (1 + eps), rule=rule)
        ps, pd = _agree(o, p["start"], p["end"], p["flag_start"], p["flag_end"])
        spread.append((ps.mean(), np.mean(ps & (pd < 1e-2))))
    assert same.mean() >= min(c for c, _ in spread) - 0.03, (same.mean(), spread)
    close = np.mean(same & (d < 1e-2))
    assert close >= min(c for _, c in spread) - 0.05, (close, spread)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_schedules_bitwise_identical(fq, precision):
    f, axes, seeds = _case("bisect", 3000)
    with fq.Field(*f, axes=axes, lookup="bisect") as field:
        a = _run(fq, field, seeds, precision, "persistent", step=0.01, tol_coef=1e-2)
        b = _run(fq, field, seeds, precision, "per_line", step=0.01, tol_coef=1e-2)
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


@pytest.mark.parametrize("rule", ["bisect", "uniform", "exact"])
def test_lds_tables_bitwise_identical(fq, rule):
    """The FP32 kernels read the axis lookup from a per-workgroup LDS copy of
    a compact table (knob stretch_lds = 1, the default); every output,
    twist included, equals the global-table lookup bit for bit on both
    schedules and both float layouts."""
    from fastqsl_amd import _lib
    from fastqsl_amd.trace import PLAIN_OUTPUTS
    f, axes, seeds = _case(rule, 3000)
    cur = tuple(0.3 * b for b in f[::-1])
    outs = tuple(PLAIN_OUTPUTS) + ("twist",)
    before = _lib.tune_get("stretch_lds")
    try:
        for layouts in (("cell32", "soa64"), ("soa32", "soa64")):
            with fq.Field(*f, axes=axes, lookup=rule, layouts=layouts, current=cur) as field:
                for schedule in ("persistent", "per_line"):
                    res = []
                    for lds in (1, 0):
                        _lib.tune_set("stretch_lds", lds)
                        res.append(_run(fq, field, seeds, "fp32c", schedule, outputs=outs, step=0.01,
                                        tol_coef=1e-2))
                    for k in outs:
                        assert np.array_equal(res[0][k], res[1][k], equal_nan=True), (layouts, schedule, k)
    finally:
        _lib.tune_set("stretch_lds", before)


@pytest.mark.parametrize("uni", [False, True])
def test_dropin_matches_handle_api(fq, uni):
    import torch
    from fastqsl_amd import FastQSL
    kinds = ("uniform",) * 3 if uni else ("sinh", "sinh", "geometric")
    field, axes = synth.stretched_multipole((48, 44, 40), seed=2, device="cpu", kinds=kinds)
    axes = [np.asarray(a, np.float32) for a in axes]
    rng = np.random.default_rng(8)
    n = 3000
    seeds = np.column_stack([rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n),
                             np.full(n, 1e-3)]).astype(np.float32)
    bx, by, bz = (np.asfortranarray(b.astype(np.float32)) for b in field)
    J = FastQSL.curl_of_B_grid(bx, by, bz, *(a.astype(np.float64) for a in axes))
    jx, jy, jz = (np.asfortranarray(c.astype(np.float32)) for c in J)
    lookup = "uniform" if uni else "bisect"
    outs = ("start_x", "start_y", "start_z", "end_x", "end_y", "end_z", "flag_start", "flag_end",
            "b_this_x", "b_this_y", "b_this_z", "b_flag", "b_start_x", "b_start_y", "b_start_z",
            "b_end_x", "b_end_y", "b_end_z", "line_len", "twist")
    with fq.Field(bx, by, bz, layouts=("soa32", "soa64"), axes=axes, lookup=lookup,
                  current=(jx, jy, jz)) as f:   # the drop-in's in-place layout
        # the drop-in reads s_len as the float32 the reference passes
        ref = _run(fq, f, seeds, "fp32c", outputs=outs, step=float(np.float32(0.01)), tol_coef=0.5)
    f32 = lambda: np.zeros(n, np.float32)
    o = dict(sx=f32(), sy=f32(), sz=f32(), fs=np.zeros(n, np.int32), ex=f32(), ey=f32(), ez=f32(),
             fe=np.zeros(n, np.int32), btx=f32(), bty=f32(), btz=f32(), bf=np.zeros(n, np.int32),
             bsx=f32(), bsy=f32(), bsz=f32(), bex=f32(), bey=f32(), bez=f32())
    twist = np.zeros(n, np.float64)
    line_len = np.zeros(n, np.float64)
    args = [bx, by, bz, np.array(bx.shape, np.int32), *axes, np.array([uni]), jx, jy, jz, twist,
            np.array([True]), *(np.ascontiguousarray(seeds[:, q]) for q in range(3)),
            np.array([0, 0, 1], np.float32)]
    args += list(o.values())
    args += [np.array([0.01], np.float32), np.array([n], np.uint64), line_len, np.array([0.5], np.float32)]
    kern = fq.kernels.compileTraceBlineAdaptiveStretch()
    kern((28, 1), (128, 1, 1), tuple(args))
    kern((28, 1), (128, 1, 1), tuple(args))  # a second call reuses the cached axis tables
    torch.cuda.synchronize()
    for mine, theirs in (("sx", "start_x"), ("ey", "end_y"), ("ez", "end_z"), ("fs", "flag_start"),
                         ("fe", "flag_end"), ("bf", "b_flag"), ("bsz", "b_start_z"), ("bex", "b_end_x"),
                         ("btx", "b_this_x")):
        assert np.array_equal(o[mine], ref[theirs], equal_nan=True), mine
    assert np.array_equal(line_len, ref["line_len"])
    assert np.array_equal(twist, ref["twist"])
    assert np.any(twist != 0)


def test_dropin_tables_shared_by_two_threads(fq):
    """The stretched drop-in keeps one set of lookup tables per device and
    rebuilds it when the axes change.  Two host threads, each on its own
    stream with its own axes, alternate on it: every call replaces the other
    thread's tables, whose launch must have been enqueued (and is waited for)
    before they are freed.  Every result equals the same call made alone."""
    import threading
    import torch
    cases = []
    for strength, seed in ((1.5, 2), (0.7, 4)):
        field, axes = synth.stretched_multipole((48, 44, 40), seed=seed, device="cpu",
                                                strength=(strength, strength, 2.0))
        axes = [np.asarray(a, np.float32) for a in axes]
        bx, by, bz = (np.asfortranarray(b.astype(np.float32)) for b in field)
        rng = np.random.default_rng(seed)
        n = 1500
        seeds = np.column_stack([rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n),
                                 np.full(n, 1e-3)]).astype(np.float32)
        cases.append((bx, by, bz, axes, seeds))

    def call(kern, case):
        bx, by, bz, axes, seeds = case
        n = seeds.shape[0]
        f32 = lambda: np.zeros(n, np.float32)
        o = [f32(), f32(), f32(), np.zeros(n, np.int32), f32(), f32(), f32(), np.zeros(n, np.int32),
             f32(), f32(), f32(), np.zeros(n, np.int32), f32(), f32(), f32(), f32(), f32(), f32()]
        line_len = np.zeros(n, np.float64)
        dummy = np.zeros(1, np.float32)
        args = [bx, by, bz, np.array(bx.shape, np.int32), *axes, np.array([False]), dummy, dummy, dummy,
                np.zeros(1, np.float64), np.array([False]), *(np.ascontiguousarray(seeds[:, q]) for q in range(3)),
                np.array([0, 0, 1], np.float32), *o, np.array([0.01], np.float32), np.array([n], np.uint64),
                line_len, np.array([0.5], np.float32)]
        kern((28, 1), (128, 1, 1), tuple(args))
        return o + [line_len]

    alone = [call(fq.kernels.compileTraceBlineAdaptiveStretch(), c) for c in cases]
    got = [[], []]
    errors = []

    def worker(t):
        try:
            kern = fq.kernels.compileTraceBlineAdaptiveStretch()
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                for _ in range(4):
                    got[t].append(call(kern, cases[t]))
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(t,)) for t in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    for t in range(2):
        assert len(got[t]) == 4
        for res in got[t]:
            for a, b in zip(res, alone[t]):
                assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_twist_in_a_unit_force_free_field(fq, precision):
    shape = (20, 20, 24)
    axes = [synth.stretched_axis(n, 0.0, 10.0, kind, 1.0, dtype=np.float64)
            for n, kind in zip(shape, ("sinh", "tanh", "geometric"))]
    a = 0.1
    Z = np.broadcast_to(axes[2][None, None, :], shape)
    f = (np.sin(a * Z), np.cos(a * Z), np.zeros(shape))
    rng = np.random.default_rng(4)
    seeds = np.column_stack([rng.uniform(2, 8, 300), rng.uniform(2, 8, 300),
                             rng.uniform(1, 9, 300)]).astype(np.float32)
    with fq.Field(*f, axes=axes, lookup="exact", current=tuple(a * c for c in f)) as field:
        r = _run(fq, field, seeds, precision, outputs=("flag_start", "line_len", "twist"),
                 ncross=(0.0, 1.0, 0.0), step=0.05, tol_coef=1e-2)
    face = (r["flag_start"] >= 1) & (r["flag_start"] <= 4)
    assert face.sum() >= 30
    np.testing.assert_allclose(r["twist"], a * r["line_len"] / (4 * math.pi), rtol=5e-4)


def test_stretched_field_errors(fq):
    shape = (8, 7, 6)
    f = tuple(np.ones(shape) for _ in range(3))
    axes = [np.arange(n, dtype=np.float64) * 0.5 for n in shape]
    seeds = np.array([[1.0, 1.0, 1.0]], np.float32)
    with fq.Field(*f, axes=axes) as field:
        with pytest.raises(fq.FqslError):
            fq.trace_plain(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rk4")
        with pytest.raises(fq.FqslError):
            fq.trace_scott(field, seeds[:, 0], seeds[:, 1], seeds[:, 2])
        bad = [a.copy() for a in axes]
        bad[1][3] = bad[1][2]
        with pytest.raises(fq.FqslError):
            field.set_axes(*bad)
        assert field.stretched  # a rejected call keeps the earlier axes
        field.set_axes(None, None, None)
        assert not field.stretched
        r = fq.trace_scott(field, seeds[:, 0], seeds[:, 1], seeds[:, 2], method="rk4")
        assert r["flag_end"].numel() == 1


@pytest.mark.parametrize("b0", [1.0, 0.0])
@pytest.mark.parametrize("rule", ["bisect", "uniform", "exact"])
def test_edge_seeds(fq, oracle, rule, b0):
    """Seeds at the edges of a stretched box over [0, 31]^3 (below the
    surface, on faces and edges, the strict y = 0 plane, outside, NaN): the
    oracle's exit classes (checkFlag, TraceBlineAdaptiveStretch.cu:298-321)
    in FP64 on every seed and in FP32c away from the null of the b0 = 0
    field (where B rounds to zero in float, class 42)."""
    from tests import parity
    seeds = parity.EDGE_SEEDS
    ax = synth.stretched_axis(32, 0.0, 31.0, "sinh", 1.5, dtype=np.float64)
    axes = [ax, ax, synth.stretched_axis(32, 0.0, 31.0, "geometric", 3.0, dtype=np.float64)]
    f = synth.hyperbolic((32, 32, 32), k=0.03, b0=b0)
    ref = oracle.trace_stretch(f, axes, seeds, s0=0.01, tol_coef=1.0, rule=rule)
    keep = np.ones(len(seeds), bool)
    print(f"stretched edge seeds {rule} b0 {b0}", flush=True)
    with fq.Field(*f, axes=axes, lookup=rule) as field:
        for precision in ("fp64", "fp32c"):
            r = _run(fq, field, seeds, precision, step=0.01, tol_coef=1.0)
            m = keep if (precision == "fp64" or b0 != 0.0) else keep & (np.arange(len(seeds)) != 4)
            np.testing.assert_array_equal(r["flag_start"][m], ref["flag_start"][m])
            np.testing.assert_array_equal(r["flag_end"][m], ref["flag_end"][m])
