"""The product's per-line tracer code (fastqsl_amd/csrc/fqsl_tracer.h), built
for the host, against the FP64 oracle -- no GPU needed.

The HIP kernels run exactly these functions per lane, so this pins their
logic (the one-stencil difference quotient, the cell-stencil layout, the
compensated FP32 position, the line driver and the Q assembly) on the CPU;
the GPU tests then only have to show that the device build computes the same.
"""
import numpy as np
import pytest

from fastqsl_amd import synth
from tests import parity
from tests.hostcheck import hc


@pytest.fixture(scope="module")
def quad(oracle):
    f = synth.quadrupole_64(dtype=np.float64, device="cpu")
    rng = np.random.default_rng(11)
    seeds = np.column_stack([rng.uniform(0, 63, 400), rng.uniform(0, 63, 400),
                             np.zeros(400)]).astype(np.float32)
    ref = oracle.trace_scott(f, seeds, method="rk4", s0=0.25)
    refp = [oracle.trace_scott(f, seeds.astype(np.float64) + sgn * np.array([1e-6, -1e-6, 0.0]),
                               method="rk4", s0=0.25) for sgn in (1.0, -1.0)]
    return f, seeds, ref, parity.conditioned(ref, *refp)


def test_one_stencil_quotient_equals_the_seven_stencil_quotient(oracle):
    """grad b from one 8-corner stencil (closed-form +-h coverage, neighbour
    cell across a plane, face clamps) vs the oracle's literal central
    difference of seven clamped interpolations (TraceBlineScott.cu:204-226)."""
    rng = np.random.default_rng(12)
    shape = (9, 8, 7)
    f = tuple(np.asfortranarray(rng.normal(size=shape) + 2.0) for _ in range(3))
    hi = np.array(shape, float) - 1
    pts = [rng.uniform(0, hi, size=(300, 3))]
    # within h = 1e-3 of an interior plane or a face, and exactly on them
    q = rng.uniform(0, hi, size=(300, 3))
    ax = rng.integers(0, 3, 300)
    q[np.arange(300), ax] = np.round(q[np.arange(300), ax]) + rng.choice([-6e-4, 0.0, 4e-4, 1e-3], 300)
    pts.append(np.clip(q, 0, hi))
    pts = np.concatenate(pts)
    got = hc.grad_b(f, pts)
    want = oracle.grad_b(f, pts, mode=0)
    scale = np.abs(want).max(axis=(1, 2), keepdims=True)
    assert np.max(np.abs(got - want) / scale) < 1e-10


def test_fp64_tracer_matches_the_oracle(quad):
    f, seeds, ref, cond = quad
    r = hc.trace_scott(f, seeds, precision=1, method=1, s0=0.25)
    cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
    assert cls[cond].all()
    d = parity.log_q_diff(r["q0"], ref["q0"])
    assert np.max(d[cond]) < 1e-9
    np.testing.assert_array_equal(r["evals"], ref["evals"])


def test_fp32c_tracer_matches_the_oracle(quad):
    f, seeds, ref, cond = quad
    r = hc.trace_scott(f, seeds, precision=0, method=1, s0=0.25)
    cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
    assert cls[cond].mean() >= 0.9975
    d = parity.log_q_diff(r["q0"], ref["q0"])
    assert np.mean(d[cond] <= 1e-3) >= 0.995
    assert np.median(d[cond]) < 1e-5


@pytest.mark.parametrize("precision", [0, 1])
def test_hyperbolic_closed_form(precision):
    shape = (41, 41, 33)
    k = 0.03
    f = synth.hyperbolic(shape, k=k, b0=1.0)
    rng = np.random.default_rng(13)
    seeds = np.column_stack([rng.uniform(17, 23, 30), rng.uniform(17, 23, 30),
                             np.zeros(30)]).astype(np.float32)
    for method, s0, tol in ((1, 0.125, 0.0), (0, 0.5, 1e-9)):
        r = hc.trace_scott(f, seeds, precision=precision, method=method, s0=s0, tol=tol)
        assert np.all(r["flag_start"] == 5) and np.all(r["flag_end"] == 6)
        np.testing.assert_allclose(r["q0"], 2 * np.cosh(2 * k * shape[2]), rtol=2e-3)


@pytest.mark.parametrize("precision", [0, 1])
def test_plain_tracer_matches_the_oracle(oracle, quad, precision):
    f, seeds, _, _ = quad
    s = seeds[:120].astype(np.float32)
    s[:, 2] = 0.5                                   # above the surface: traced both ways
    ref = oracle.trace_plain(f, s, method="rk4", s0=0.25)
    refp = [oracle.trace_plain(f, s.astype(np.float64) + sgn * np.array([1e-6, -1e-6, 0.0]),
                               method="rk4", s0=0.25) for sgn in (1.0, -1.0)]
    cond = parity.conditioned(ref, *refp, key="end")
    r = hc.trace_plain(f, s, precision=precision, method=1, s0=0.25)
    cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
    assert cls[cond].all()
    tol = 1e-8 if precision == 1 else 2e-3
    assert np.abs(r["end"] - ref["end"])[cond].max() < tol
    assert np.abs(r["start"] - ref["start"])[cond].max() < tol


def test_controller_power_is_double_accurate():
    """The FP32 path evaluates the step controller's (tol/err/11.09)^0.2 as
    ctl * e2^(-1/10) with a float estimate refined by two Newton steps; it
    must agree with the double power to double rounding over the range, and
    keep pow's special values (0 -> inf, NaN -> NaN).  A float-only estimate
    would be off by up to ~1e-6, one Newton step by ~3e-12."""
    rng = np.random.default_rng(14)
    e2 = 10.0 ** rng.uniform(-29.9, 29.9, 20000)
    got = hc.pow_minus_tenth(e2)
    np.testing.assert_allclose(got, e2 ** -0.1, rtol=1e-15, atol=0)
    sp = hc.pow_minus_tenth([0.0, np.nan, np.inf, 1e-40, 1e40])
    assert sp[0] == np.inf and np.isnan(sp[1]) and sp[2] == 0.0
    np.testing.assert_allclose(sp[3:], np.array([1e-40, 1e40]) ** -0.1, rtol=1e-15)


@pytest.mark.parametrize("precision,rtol", [(1, 1e-9), (0, 2e-4)])
def test_twist_matches_the_oracle(oracle, precision, rtol):
    """The tracer header's twist (J, B raw at each step's end point, landing
    term with the raw B) against the oracle, with J from the grid curl."""
    from oracle import curl
    f = synth.multipole((48, 40, 36), seed=6, dtype=np.float64, device="cpu")
    cur = curl.curl_of_B_grid(*f, *(np.arange(n, dtype=np.float64) for n in f[0].shape))
    rng = np.random.default_rng(15)
    seeds = np.column_stack([rng.uniform(2, 45, 200), rng.uniform(2, 37, 200),
                             np.r_[np.zeros(150), rng.uniform(0, 30, 50)]]).astype(np.float32)
    ref = oracle.trace_scott(f, seeds, method="rk4", s0=0.25, current=cur)
    refp = [oracle.trace_scott(f, seeds.astype(np.float64) + sg * np.array([1e-6, -1e-6, 0.0]), method="rk4",
                               s0=0.25, current=cur) for sg in (1.0, -1.0)]
    cond = parity.conditioned(ref, *refp)
    for p in refp:
        cond &= np.abs(p["twist"] - ref["twist"]) <= 1e-6 * (1 + np.abs(ref["twist"]))
    r = hc.trace_scott(f, seeds, precision=precision, method=1, s0=0.25, current=cur)
    ok = cond & parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
    assert ok.mean() > 0.95
    scale = np.maximum(np.abs(ref["twist"]), 1e-3)
    assert np.quantile(np.abs(r["twist"] - ref["twist"])[ok] / scale[ok], 0.99) < rtol
    assert np.mean(np.abs(r["twist"]) > 1e-6) > 0.5



@pytest.mark.parametrize("b0", [1.0, 0.0])
@pytest.mark.parametrize("method", ["rk4", "rkf45"])
def test_plain_tracer_edge_seeds(oracle, b0, method):
    """Exit classes of the plain driver at the edges (TraceBlineAdaptive.cu
    :333-426, TraceBline.cu:300-378, checkFlag :276-298): identical to the
    oracle in both precisions; in FP64 the same evaluation counts and
    footpoints to round-off."""
    f = synth.hyperbolic((32, 32, 32), k=0.03, b0=b0)
    ref = oracle.trace_plain(f, parity.EDGE_SEEDS, method=method, s0=0.25, tol_coef=1.0)
    for precision in (1, 0):
        r = hc.trace_plain(f, parity.EDGE_SEEDS, precision=precision, method={"rk4": 1, "rkf45": 0}[method],
                           s0=0.25, tol_coef=1.0)
        np.testing.assert_array_equal(r["flag_start"], ref["flag_start"])
        np.testing.assert_array_equal(r["flag_end"], ref["flag_end"])
        np.testing.assert_array_equal(np.isnan(r["end"]), np.isnan(ref["end"]))
        if precision == 1:
            np.testing.assert_allclose(r["end"], ref["end"], rtol=0, atol=1e-7)
            np.testing.assert_allclose(r["start"], ref["start"], rtol=0, atol=1e-7)
            np.testing.assert_array_equal(r["evals"], ref["evals"])


ROT_TOLS = (1e-6, 1e-8)


@pytest.fixture(scope="module")
def rotation_case():
    """RKF45 where the controller cannot fork: a linear rotating field, which
    the trilinear interpolant reproduces exactly, with no kinks at cell
    faces, so the error estimate varies smoothly along every line and stays
    far above round-off; seeds on z = 0 and inside the box."""
    f = synth.rotation((40, 40, 48), omega=0.05, b0=1.0)
    rng = np.random.default_rng(3)
    n = 2000
    seeds = np.column_stack([rng.uniform(8, 31, n), rng.uniform(8, 31, n),
                             np.r_[np.zeros(n // 2), rng.uniform(0, 40, n - n // 2)]]).astype(np.float32)
    return f, seeds


@pytest.mark.parametrize("tol", ROT_TOLS)
def test_fp64_rkf45_line_by_line(oracle, rotation_case, tol):
    """The FP64 build takes the oracle's RKF45 step sequence on every line
    (accepted and rejected attempts: the same evaluation count) and lands on
    the same Q."""
    f, seeds = rotation_case
    ref = oracle.trace_scott(f, seeds, method="rkf45", s0=0.25, tol=tol)
    r = hc.trace_scott(f, seeds, precision=1, method=0, s0=0.25, tol=tol)
    np.testing.assert_array_equal(r["evals"], ref["evals"])
    assert np.array_equal(r["flag_end"], ref["flag_end"]) and np.array_equal(r["flag_start"], ref["flag_start"])
    d = parity.log_q_diff(r["q0"], ref["q0"])
    assert np.median(d) <= 1e-9 and d.max() <= 1e-8, (np.median(d), d.max())


def test_mixed_tracer_follows_the_oracles_controller(oracle):
    """The mixed variant (FP64 position path, FP32 Jacobian) runs the FP64
    controller on an FP64 error norm: on a random multipole its evaluation
    counts follow the oracle's (the FP32 path's do not: its error estimate's
    round-off floor restrains step growth) and its Q is as close to the
    oracle's as the FP64 path."""
    f = synth.multipole((64, 64, 64), seed=4, device="cpu")
    rng = np.random.default_rng(9)
    seeds = np.column_stack([rng.uniform(0, 63, 3000), rng.uniform(0, 63, 3000), np.zeros(3000)]).astype(np.float32)
    ref = oracle.trace_scott(f, seeds, method="rkf45", s0=0.25, tol=1e-6)
    mixed = hc.trace_scott(f, seeds, precision=2, method=0, s0=0.25, tol=1e-6)
    fp32 = hc.trace_scott(f, seeds, precision=0, method=0, s0=0.25, tol=1e-6)
    ev_m = mixed["evals"].sum() / ref["evals"].sum()
    ev_f = fp32["evals"].sum() / ref["evals"].sum()
    print("evals vs oracle: mixed %.4f fp32c %.4f" % (ev_m, ev_f))
    assert 0.99 < ev_m < 1.01
    assert np.mean(mixed["evals"] == ref["evals"]) > 0.5
    cls = parity.class_match(mixed["flag_start"], mixed["flag_end"], ref["flag_start"], ref["flag_end"])
    d = parity.log_q_diff(mixed["q0"], ref["q0"])
    assert cls.mean() >= 0.995 and np.median(d[cls]) < 1e-4


def test_mixed_tracer_closed_form():
    """Q = 2 cosh(2 k N / B0) on the linear X-point field through the mixed
    variant (RK4 and RKF45)."""
    shape, k = (41, 41, 33), 0.03
    f = synth.hyperbolic(shape, k=k, b0=1.0)
    g = np.linspace(14, 26, 9, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)]).astype(np.float32)
    for method, tol in ((1, 0.0), (0, 1e-10)):   # RKF45 at a tolerance FP32 could not resolve
        r = hc.trace_scott(f, seeds, precision=2, method=method, s0=0.25, tol=tol)
        top = (r["flag_end"] == 6) & np.isfinite(r["q0"])   # the seed on the axis has no frame (NaN)
        assert top.mean() > 0.9
        np.testing.assert_allclose(r["q0"][top], 2 * np.cosh(2 * k * shape[2]), rtol=2e-3)
