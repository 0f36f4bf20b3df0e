"""GPU parity of the Scott tracer against the FP64 oracle.

* RK4 fixed step (the quadrupole workload, 64^3 float64 field, 256^2 seeds),
  on every pixel off the field's two mirror planes x = y and x + y = 63
  (65,024 of 65,536): a line seeded in a mirror plane stays in it in exact
  arithmetic and runs into the central null, so round-off of any size
  decides its fate.  FP64 path -- classes identical, |dlog10 Q| <= 1e-5
  (float32 output), the oracle's evaluation counts; FP32-compensated and
  mixed paths -- classes match on >= 99.99 %, |dlog10 Q| <= 1e-3 on
  >= 99.9 %.
* RKF45 (random multipole, 256^3 float32, tol 1e-6, 65,536 seeds of the
  1024^2 plane): the reference's controller forks on round-off, so every
  bar is relative to the oracle's own spread, measured on the same lines.
  Mixed and FP64 paths (the FP64 controller on an FP64 error norm): class
  mismatches against the oracle at most the oracle's own under one-float32-
  ulp seed moves plus 1e-4 of the lines; as close to the converged Q (tol
  1e-10) as the oracle, p99.9 included.  FP32-compensated path (error norm
  from float stages, whose round-off floor restrains the step growth): class
  mismatches against the oracle at most those of the reference's own float
  arithmetic (the oracle's float model) plus 1e-4 of the lines, and against
  the converged answer no less accurate than the oracle.
"""
import numpy as np
import pytest

from fastqsl_amd import synth
from tests import parity

pytestmark = pytest.mark.gpu


def _np(r, keys=("start_x", "start_y", "start_z", "end_x", "end_y", "end_z", "flag_start",
                 "flag_end", "q_0", "q_perp", "line_len", "evals", "b_this_x", "b_this_y",
                 "b_this_z")):
    import torch
    torch.cuda.synchronize()
    return {k: r[k].cpu().numpy() for k in keys if k in r}


@pytest.fixture(scope="module")
def fq():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd
    fastqsl_amd.lib()
    return fastqsl_amd


@pytest.fixture(scope="module")
def c1(oracle):
    field = synth.quadrupole_64(dtype=np.float64, device="cpu")
    g = np.linspace(0, 63, 256, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)]).astype(np.float32)
    ref = oracle.trace_scott(field, seeds, method="rk4", s0=0.25)
    # off the mirror planes x = y and x + y = 63 of the alternating quadrupole
    row, col = np.divmod(np.arange(seeds.shape[0]), 256)
    cond = (row != col) & (row + col != 255)
    assert cond.sum() == 65536 - 512
    return field, seeds, ref, cond


def _trace(fq, f, seeds, **kw):
    return _np(fq.trace_scott(f, seeds[:, 0], seeds[:, 1], seeds[:, 2], **kw))


def test_c1_fp64_matches_oracle(fq, c1):
    field, seeds, ref, cond = c1
    with fq.Field(*field) as f:
        for sched in ("per_line", "persistent"):
            r = _trace(fq, f, seeds, method="rk4", step=0.25, precision="fp64", schedule=sched)
            cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
            d = parity.log_q_diff(r["q_0"], ref["q0"])
            dp = parity.log_q_diff(r["q_perp"], ref["qperp"])
            print(sched, parity.summary(d, cls, cond), "all:", parity.summary(d, cls))
            assert cls[cond].all()
            assert d[cond].max() <= 1e-5 and dp[cond].max() <= 1e-5
            end = np.column_stack([r["end_x"], r["end_y"], r["end_z"]])
            start = np.column_stack([r["start_x"], r["start_y"], r["start_z"]])
            assert np.abs(end - ref["end"])[cond].max() < 1e-4
            assert np.abs(start - ref["start"])[cond].max() < 1e-4
            np.testing.assert_allclose(r["line_len"][cond], ref["length"][cond], rtol=1e-8)
            assert np.array_equal(r["evals"][cond].astype(np.int64), ref["evals"][cond])
            bt = np.column_stack([r["b_this_x"], r["b_this_y"], r["b_this_z"]])
            np.testing.assert_allclose(bt, ref["b_this"], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("precision", ["fp32c", "mixed"])
def test_c1_fp32_parity(fq, c1, precision):
    field, seeds, ref, cond = c1
    with fq.Field(*field) as f:
        r = _trace(fq, f, seeds, method="rk4", step=0.25, precision=precision, schedule="persistent")
    cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
    d = parity.log_q_diff(r["q_0"], ref["q0"])
    s = parity.summary(d, cls, cond)
    print(precision, "off the mirror planes:", s, "all:", parity.summary(d, cls))
    assert s["class_match"] >= 0.9999
    assert s["frac_le_1e3"] >= 0.999
    end = np.column_stack([r["end_x"], r["end_y"], r["end_z"]])
    err = np.abs(end - ref["end"]).max(axis=1)[cond & cls]
    print("endpoint error (cells): p99.9 %.2e max %.2e" % (np.quantile(err, 0.999), err.max()))


def test_schedules_bitwise_identical(fq, c1):
    """A line's result does not depend on which lane or launch traced it."""
    field, seeds, ref, cond = c1
    with fq.Field(*field) as f:
        for prec in ("fp32c", "fp64"):
            a = _trace(fq, f, seeds, method="rk4", step=0.25, precision=prec, schedule="persistent")
            b = _trace(fq, f, seeds, method="rk4", step=0.25, precision=prec, schedule="per_line")
            perm = np.random.default_rng(0).permutation(len(seeds))
            c = _trace(fq, f, seeds[perm], method="rk4", step=0.25, precision=prec, schedule="persistent")
            inv = np.argsort(perm)
            for k in a:
                assert np.array_equal(a[k], b[k], equal_nan=True), (prec, k)
                assert np.array_equal(a[k], c[k][inv], equal_nan=True), (prec, k)


def test_rolled_and_unrolled_stages_bitwise_identical(fq):
    """The three-wave FP32c kernels run the RKF45 and RK4 stages as one loop
    (stage derivatives in LDS), the two-wave ones unrolled; launches above
    one two-wave grid's lanes take the three-wave kernel under the default
    knob.  Both forms give the same bits, in the persistent and the
    one-thread-per-line schedule."""
    from fastqsl_amd import _lib
    bx, by, bz = synth.multipole_256(seed=0, device="cuda")
    g = np.linspace(0.0, 255.0, 512, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size, np.float32)])   # 262,144 lines
    assert len(seeds) > 256 * 4 * 2 * 64
    keep = _lib.tune_get("scott_waves")
    try:
        with fq.Field(bx, by, bz, layouts=("soa64", "cell32")) as f:
            for kw in (dict(method="rkf45", step=0.25, tol=1e-6, schedule="persistent"),
                       dict(method="rk4", step=0.25, schedule="persistent"),
                       dict(method="rkf45", step=0.25, tol=1e-6, schedule="per_line")):
                out = {}
                for w in (2, 3):
                    _lib.tune_set("scott_waves", w)
                    out[w] = _trace(fq, f, seeds, precision="fp32c", **kw)
                for k in out[2]:
                    assert np.array_equal(out[2][k], out[3][k], equal_nan=True), (kw, k)
    finally:
        _lib.tune_set("scott_waves", keep)


C2_LINES = 1 << 16


@pytest.fixture(scope="module")
def c2(oracle):
    """65,536 seeds drawn from the 1024^2 z = 0 plane of the random multipole;
    the oracle at tol 1e-6, again from seeds moved by one float32 ulp, in its
    float model, and converged (tol 1e-10)."""
    field = synth.multipole_256(seed=0, device="cuda")
    n = 1024
    rng = np.random.default_rng(0)
    ij = rng.choice(n * n, size=C2_LINES, replace=False)
    g = np.linspace(0, 255, n, dtype=np.float32)
    seeds = np.column_stack([g[ij % n], g[ij // n], np.zeros(ij.size, np.float32)]).astype(np.float32)
    kw = dict(method="rkf45", s0=0.25)
    ref = oracle.trace_scott(field, seeds, tol=1e-6, **kw)
    ulp = oracle.trace_scott(field, parity.ulp_moved(seeds, (255, 255), np.random.default_rng(5)), tol=1e-6, **kw)
    f32 = oracle.trace_scott_f32(field, seeds, s0=0.25, tol=1e-6)
    truth = oracle.trace_scott(field, seeds, tol=1e-10, **kw)
    return field, seeds, ref, ulp, f32, truth


def _mismatches(a, b):
    fs_a = a["flag_start"]
    fe_a = a["flag_end"]
    return int((~parity.class_match(fs_a, fe_a, b["flag_start"], b["flag_end"])).sum())


def _vs_truth(q, flags, truth):
    """|dlog10 Q| against the converged answer on the lines whose classes match it."""
    cls = parity.class_match(flags[0], flags[1], truth["flag_start"], truth["flag_end"])
    d = parity.log_q_diff(q, truth["q0"])[cls]
    d = d[np.isfinite(d)]
    return int((~cls).sum()), float(np.median(d)), float(np.quantile(d, 0.999))


def _gpu_c2(fq, field, seeds, precision):
    with fq.Field(*field) as f:
        r = _trace(fq, f, seeds, method="rkf45", step=0.25, tol=1e-6, precision=precision,
                   schedule="per_line" if precision == "fp64" else "persistent")
    r["q0"] = r["q_0"]
    return r


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_c2_class_parity_within_the_oracles_own_spread(fq, c2, precision):
    """Paths that run the FP64 controller on an FP64 error norm: within the
    oracle's one-ulp spread plus 1e-4 of the lines, and as close to the
    converged answer."""
    field, seeds, ref, ulp, f32, truth = c2
    N = seeds.shape[0]
    floor = _mismatches(ulp, ref)
    r = _gpu_c2(fq, field, seeds, precision)
    mis = _mismatches(r, ref)
    d = parity.log_q_diff(r["q_0"], ref["q0"])
    d_ulp = parity.log_q_diff(ulp["q0"], ref["q0"])
    t_orc = _vs_truth(ref["q0"], (ref["flag_start"], ref["flag_end"]), truth)
    t_gpu = _vs_truth(r["q_0"], (r["flag_start"], r["flag_end"]), truth)
    print(f"{precision}: class mismatches {mis} (one-ulp floor {floor}, budget {floor + 1e-4 * N:.1f}); "
          f"within 1e-3 {np.mean(d <= 1e-3):.4f} (floor {np.mean(d_ulp <= 1e-3):.4f}); vs converged "
          f"mismatches/median/p99.9 {t_gpu} (oracle {t_orc})")
    assert mis <= floor + 1e-4 * N
    assert np.mean(d <= 1e-3) >= np.mean(d_ulp <= 1e-3) - 0.02
    assert t_gpu[0] <= t_orc[0] + 1e-4 * N
    assert t_gpu[1] <= 1.2 * t_orc[1] + 1e-6
    assert t_gpu[2] <= 1.25 * t_orc[2] + 1e-3
    ev = r["evals"].astype(np.float64).sum() / ref["evals"].sum()
    print(precision, "evals ratio vs oracle %.4f" % ev)
    assert 0.98 < ev < 1.02


def test_c2_fp32c_accuracy(fq, c2):
    """The FP32-compensated path: its error norm is formed from float stages,
    whose round-off floor (~1e-8 x the step) restrains the step growth the
    FP64 controller takes when the true error is tiny, so its step sequences
    are the reference's float arithmetic's, not the FP64 oracle's.  Against
    the oracle: no further than the reference's own float arithmetic (the
    oracle's float model) plus 1e-4 of the lines; against the converged
    answer: at least as accurate as the oracle, p99.9 included."""
    field, seeds, ref, ulp, f32, truth = c2
    N = seeds.shape[0]
    r = _gpu_c2(fq, field, seeds, "fp32c")
    mis = _mismatches(r, ref)
    mis_f32 = _mismatches(f32, ref)
    t_orc = _vs_truth(ref["q0"], (ref["flag_start"], ref["flag_end"]), truth)
    t_gpu = _vs_truth(r["q_0"], (r["flag_start"], r["flag_end"]), truth)
    d = parity.log_q_diff(r["q_0"], ref["q0"])
    d_f32 = parity.log_q_diff(f32["q0"], ref["q0"])
    print(f"fp32c: class mismatches vs oracle {mis} (float model {mis_f32}, one-ulp floor {_mismatches(ulp, ref)}); "
          f"within 1e-3 {np.mean(d <= 1e-3):.4f} (float model {np.mean(d_f32 <= 1e-3):.4f}); vs converged "
          f"mismatches/median/p99.9 {t_gpu} (oracle {t_orc})")
    assert mis <= mis_f32 + 1e-4 * N
    assert np.mean(d <= 1e-3) >= np.mean(d_f32 <= 1e-3)
    assert t_gpu[0] <= t_orc[0] + 1e-4 * N
    assert t_gpu[1] <= 1.2 * t_orc[1] + 1e-6
    assert t_gpu[2] <= 1.25 * t_orc[2] + 1e-3
    ev = r["evals"].astype(np.float64).sum() / ref["evals"].sum()
    print("fp32c evals ratio vs oracle %.4f (float model %.4f)" % (ev, f32["evals"].sum() / ref["evals"].sum()))
    assert 0.6 < ev < 1.1


def test_rkf45_line_by_line_where_the_controller_cannot_fork(fq, oracle):
    """A linear rotating field (reproduced exactly by the interpolant, no
    kinks at cell faces): the FP64 path takes the oracle's RKF45 step
    sequence on every line -- the same evaluation count -- and its Q."""
    field = synth.rotation((40, 40, 48), omega=0.05, b0=1.0)
    rng = np.random.default_rng(3)
    n = 4000
    seeds = np.column_stack([rng.uniform(8, 31, n), rng.uniform(8, 31, n),
                             np.r_[np.zeros(n // 2), rng.uniform(0, 40, n - n // 2)]]).astype(np.float32)
    with fq.Field(*field) as f:
        for tol in (1e-6, 1e-8):
            ref = oracle.trace_scott(field, seeds, method="rkf45", s0=0.25, tol=tol)
            for sched in ("per_line", "persistent"):
                r = _trace(fq, f, seeds, method="rkf45", step=0.25, tol=tol, precision="fp64", schedule=sched)
                np.testing.assert_array_equal(r["evals"].astype(np.int64), ref["evals"])
                assert np.array_equal(r["flag_end"], ref["flag_end"])
                d = parity.log_q_diff(r["q_0"], ref["q0"])
                print(f"tol {tol:g} {sched}: |dlog10 Q| median {np.median(d):.2e} max {d.max():.2e}")
                # q_0 is written as float32: its rounding (~3e-8 in log10) is the floor here;
                # the FP64 line length (the sum of the accepted steps) pins the sequence tighter
                assert d.max() <= 1e-7
                np.testing.assert_allclose(r["line_len"], ref["length"], rtol=1e-9)


def test_hyperbolic_closed_form_gpu(fq):
    shape = (41, 41, 33)
    k = 0.03
    field = synth.hyperbolic(shape, k=k, b0=1.0)
    rng = np.random.default_rng(4)
    seeds = np.column_stack([rng.uniform(17, 23, 64), rng.uniform(17, 23, 64), np.zeros(64)]).astype(np.float32)
    seeds[:16, 2] = rng.uniform(1, 20, 16)
    Qc = 2 * np.cosh(2 * k * shape[2])
    with fq.Field(*field) as f:
        for prec in ("fp64", "fp32c"):
            for method, step, tol in (("rk4", 0.125, 0.0), ("rkf45", 0.5, 1e-9)):
                r = _trace(fq, f, seeds, method=method, step=step, tol=max(tol, 1e-12), precision=prec)
                assert np.all(r["flag_start"] == 5) and np.all(r["flag_end"] == 6)
                np.testing.assert_allclose(r["q_0"], Qc, rtol=2e-3)
                np.testing.assert_allclose(r["end_z"], shape[2], atol=1e-5)


def test_edge_cases_gpu(fq, oracle):
    shape = (32, 32, 32)
    hyp = synth.hyperbolic(shape, k=0.03, b0=1.0)
    null = synth.hyperbolic(shape, k=0.03, b0=0.0)
    seeds = np.array([[10.0, 10.0, -0.5],     # below the surface
                      [0.0, 5.0, 0.0],        # on the x = 0 face
                      [31.0, 12.0, 3.0],      # on the x max face
                      [5.0, 0.0, 2.0],        # y = 0 (outside by the strict rule)
                      [15.5, 15.5, 4.0]], np.float32)   # the null (second field)
    for field in (hyp, null):
        ref = oracle.trace_scott(field, seeds, method="rk4", s0=0.25)
        with fq.Field(*field) as f:
            for prec in ("fp64", "fp32c"):
                r = _trace(fq, f, seeds, method="rk4", step=0.25, precision=prec)
                np.testing.assert_array_equal(r["flag_start"], ref["flag_start"])
                np.testing.assert_array_equal(r["flag_end"], ref["flag_end"])
                np.testing.assert_array_equal(np.isnan(r["q_0"]), np.isnan(ref["q0"]))
                np.testing.assert_array_equal(r["q_0"] == 0, ref["q0"] == 0)
