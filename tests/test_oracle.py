"""Pins the FP64 CPU oracle against mathematics that does not depend on it.

Each check is chosen so that a plausible slip in the restatement (a wrong
corner index, a transposed Jacobian, a dropped tableau term, a sign error in
the Q projection, a wrong landing plane) fails it:

* trilinear interpolation reproduces affine fields exactly and follows both
  clamp rules (Interp3dxyzn / Interp3d, TraceBlineScott.cu:47-201);
* the difference-quotient grad b agrees with the closed-form derivative away
  from cell faces (grad_unit_vec_B, TraceBlineScott.cu:204-226);
* checkFlag classes for enumerated points (TraceBlineScott.cu:310-332);
* RK4 / RKF45 local error orders on a helical field with a closed-form
  solution (RKF45_Scott :256-290; RK4, TraceBline.cu:226-234);
* Q = 2 cosh(2 k N / B0) for the linear X-point field, through both the Scott
  path and the neighbour-difference path (FastQSL.py:65-117);
* Q = 2 for a uniform oblique field, with closed-form exit points;
* 90-degree rotation symmetry of the alternating quadrupole;
* the decision register's edge cases (below-surface seeds, limits, nulls).
"""
import numpy as np
import pytest

from fastqsl_amd import synth


def affine_field(shape, A, c):
    nx, ny, nz = shape
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    P = np.stack([i, j, k], axis=-1).astype(np.float64)
    B = P @ np.asarray(A, np.float64).T + np.asarray(c, np.float64)
    return tuple(np.asfortranarray(B[..., a]) for a in range(3))


def test_interp_exact_on_affine_and_clamps(oracle):
    rng = np.random.default_rng(1)
    shape = (9, 7, 6)
    A = rng.normal(size=(3, 3))
    c = rng.normal(size=3)
    f = affine_field(shape, A, c)
    lo = np.zeros(3)
    hi = np.array(shape, float) - 1
    pts = rng.uniform(-3, np.array(shape) + 2, size=(500, 3))
    got = oracle.interp(f, pts, mode=0)
    clamped = np.clip(pts, lo, hi)                       # Interp3dxyzn: constant outside
    np.testing.assert_allclose(got, clamped @ A.T + c, rtol=0, atol=1e-12)
    got_n = oracle.interp(f, pts, mode=1)
    ref = clamped @ A.T + c
    np.testing.assert_allclose(got_n, ref / np.linalg.norm(ref, axis=1, keepdims=True), atol=1e-12)
    got_c = oracle.interp(f, pts, mode=2)                 # Interp3d: pulled into [1e-4, N-1-1e-4]
    cl2 = np.where(pts > 0, pts, 1e-4)
    cl2 = np.where(cl2 < hi, cl2, hi - 1e-4)
    np.testing.assert_allclose(got_c, cl2 @ A.T + c, atol=1e-12)


def test_interp_matches_independent_trilinear(oracle):
    from scipy.interpolate import RegularGridInterpolator
    rng = np.random.default_rng(2)
    shape = (6, 5, 7)
    f = tuple(np.asfortranarray(rng.normal(size=shape)) for _ in range(3))
    pts = rng.uniform(0, np.array(shape) - 1, size=(300, 3))
    axes = [np.arange(n, dtype=float) for n in shape]
    ref = np.stack([RegularGridInterpolator(axes, b)(pts) for b in f], axis=1)
    np.testing.assert_allclose(oracle.interp(f, pts), ref, atol=1e-12)


def test_fd_gradient_matches_closed_form_inside_cells(oracle):
    f = synth.quadrupole_64(device="cpu", n=24)
    rng = np.random.default_rng(3)
    pts = rng.uniform(1, 22, size=(400, 3))
    frac = pts - np.floor(pts)
    ok = np.all((frac > 3e-3) & (frac < 1 - 3e-3), axis=1)
    pts = pts[ok]
    fd = oracle.grad_b(f, pts, mode=0)
    an = oracle.grad_b(f, pts, mode=1)
    scale = np.abs(an).max(axis=(1, 2), keepdims=True)
    assert np.max(np.abs(fd - an) / scale) < 1e-5
    # a transposed Jacobian would fail this (d b / d x_j has no symmetry)
    assert np.max(np.abs(fd - np.swapaxes(an, 1, 2)) / scale) > 1e-2


def test_fd_gradient_face_rules(oracle):
    """At x = 0 the clamped difference quotient is half the one-sided
    derivative; across an interior grid plane it averages the two cells."""
    shape = (6, 6, 6)
    A = np.array([[0.5, 0.0, 0.0], [0.0, 0.2, 0.0], [0.1, 0.0, 0.0]])
    c = np.array([1.0, 2.0, 3.0])
    f = affine_field(shape, A, c)
    p = np.array([[0.0, 2.5, 2.5]])
    fd = oracle.grad_b(f, p, mode=0)[0]
    # (b(h) - b(0)) / 2h: half the one-sided quotient, i.e. half the slope at h/2
    an_mid = oracle.grad_b(f, p + np.array([[5e-4, 0, 0]]), mode=1)[0]
    an = oracle.grad_b(f, p + np.array([[1e-9, 0, 0]]), mode=1)[0]
    np.testing.assert_allclose(fd[0], 0.5 * an_mid[0], rtol=1e-5)
    np.testing.assert_allclose(fd[1:], an[1:], rtol=2e-6)
    # across the interior plane x = 2 the quotient blends the two cells by overlap
    q = np.array([[2.0 + 4e-4, 2.5, 2.5]])
    fdq = oracle.grad_b(f, q, mode=0)[0]
    left = oracle.grad_b(f, q - np.array([[7e-4, 0, 0]]), mode=1)[0]
    right = oracle.grad_b(f, q + np.array([[3e-4, 0, 0]]), mode=1)[0]
    np.testing.assert_allclose(fdq[0], 0.3 * left[0] + 0.7 * right[0], rtol=1e-4)


def test_check_flag_classes(oracle):
    n = (10, 8, 6)
    cases = [
        ([1, 1, 1], 0), ([0, 1, 0], 0), ([9, 7, 5], 0),
        ([1, 0, 1], 42),                  # y == 0 is outside (strict y > 0) and on no face
        ([-0.1, 3, 3], 1), ([9.5, 3, 3], 2), ([3, -0.1, 3], 3), ([3, 7.2, 3], 4),
        ([3, 3, -0.1], 5), ([3, 3, 5.5], 6),
        ([-0.1, -0.1, 3], 42), ([9.5, 3, 5.5], 42), ([np.nan, 3, 3], 42),
        ([9.5, 7.0, 3], 42),              # on the y max line of the x max face
    ]
    pts = np.array([c[0] for c in cases], float)
    np.testing.assert_array_equal(oracle.check_flag(n, pts), [c[1] for c in cases])


def helix_field(shape, w):
    nx, ny, nz = shape
    xc, yc = (nx - 1) / 2, (ny - 1) / 2
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return (np.asfortranarray(-w * (j - yc)), np.asfortranarray(w * (i - xc)),
            np.asfortranarray(np.ones(shape))), (xc, yc)


@pytest.mark.parametrize("method,order", [("rk4", 5), ("rkf45", 6)])
def test_local_error_order(oracle, method, order):
    w = 0.1
    f, (xc, yc) = helix_field((32, 32, 32), w)
    r0, th0, z0 = 5.0, 0.3, 4.0
    x0 = np.array([xc + r0 * np.cos(th0), yc + r0 * np.sin(th0), z0])
    Bm = np.sqrt(1 + (w * r0) ** 2)
    errs = []
    for h in (1.6, 0.8, 0.4):
        y = np.concatenate([x0, [1, 0, 0], [0, 1, 0]])
        out, _ = oracle.scott_step(f, y, h, method=method)
        th = th0 + w * h / Bm
        exact = np.array([xc + r0 * np.cos(th), yc + r0 * np.sin(th), z0 + h / Bm])
        errs.append(np.linalg.norm(out[:3] - exact))
    ratios = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(ratios - order) < 0.35), (errs, ratios)


def test_rkf45_error_estimate_order(oracle):
    f, (xc, yc) = helix_field((32, 32, 32), 0.1)
    x0 = np.array([xc + 5.0, yc, 4.0])
    es = [oracle.scott_step(f, np.concatenate([x0, [1, 0, 0], [0, 1, 0]]), h)[1]
          for h in (1.6, 0.8, 0.4)]
    ratios = np.log2(np.array(es[:-1]) / np.array(es[1:]))
    assert np.all(np.abs(ratios - 5) < 0.35), (es, ratios)


@pytest.mark.parametrize("method,s0,tol", [("rk4", 0.125, 0.0), ("rkf45", 0.5, 1e-10)])
def test_scott_hyperbolic_closed_form(oracle, method, s0, tol):
    shape = (41, 41, 33)
    k = 0.03
    f = synth.hyperbolic(shape, k=k, b0=1.0)
    rng = np.random.default_rng(4)
    seeds = np.column_stack([rng.uniform(17, 23, 40), rng.uniform(17, 23, 40), np.zeros(40)])
    seeds[:10, 2] = rng.uniform(1, 20, 10)          # interior seeds trace both ways
    r = oracle.trace_scott(f, seeds, method=method, s0=s0, tol=tol)
    assert np.all(r["flag_start"] == 5) and np.all(r["flag_end"] == 6)
    Qc = 2 * np.cosh(2 * k * shape[2])               # max-face landing plane is z = N
    np.testing.assert_allclose(r["q0"], Qc, rtol=1.5e-3)
    np.testing.assert_allclose(r["end"][:, 2], shape[2], atol=1e-9)
    xc = (shape[0] - 1) / 2
    # footpoint map x - xc -> (x - xc) exp(k N): checks the landing and the direction
    x_start = r["start"][:, 0] - xc
    np.testing.assert_allclose(r["end"][:, 0] - xc, x_start * np.exp(k * shape[2]), rtol=2e-3)
    # backward trace from interior seeds lands on z = 0
    np.testing.assert_allclose(r["start"][:, 2], 0.0, atol=1e-9)


def test_scott_uniform_field_q_is_two(oracle):
    shape = (20, 24, 18)
    b = np.array([0.3, 0.2, 1.0])
    f = synth.uniform(shape, b)
    seeds = np.array([[5.0, 5.0, 0.0], [3.0, 4.0, 7.0], [10.0, 12.0, 3.0]])
    r = oracle.trace_scott(f, seeds, method="rk4", s0=0.25)
    np.testing.assert_allclose(r["q0"], 2.0, rtol=1e-9)
    np.testing.assert_allclose(r["qperp"], 2.0, rtol=1e-9)
    u = b / np.linalg.norm(b)
    for i, p in enumerate(seeds):
        # exit on the z max face: landed at z = N (not N-1)
        t = (shape[2] - p[2]) / u[2]
        np.testing.assert_allclose(r["end"][i], p + t * u, atol=1e-9)
        assert r["flag_end"][i] == 6
        if p[2] > 0:
            t0 = -p[2] / u[2]
            np.testing.assert_allclose(r["start"][i], p + t0 * u, atol=1e-9)
            assert r["flag_start"][i] == 5
    # line length: in-box path plus the landing term |p_mid - p1| / |b_dim|
    assert np.all(r["length"] > 0)


@pytest.mark.parametrize("method,s0,tol", [("rk4", 0.125, 0.0), ("rkf45", 0.5, 1e-9)])
def test_scott_monopole_conformal_map(oracle, method, s0, tol):
    """A single sub-surface monopole maps z = 0 to the top face by a radial
    projection, which is conformal, so Q = 2.  The reference lands max-face
    exits on the plane z = N, one cell beyond the last grid plane, where the
    clamped field is B(x, y, N-1): there Bx/Bz = x/D (D = distance from the
    source to plane N-1, in cells), so the slab scales the map by 1 + 1/D
    while Bn at the landing point stays B(N-1)'s.  The Q formula then gives
    2 (1 + 1/D)^2 -- a 4 % offset that a landing on N-1, or an inverted Bn
    ratio, would miss (residual: trilinear error, (h/r)^2 ~ 4e-4)."""
    n, depth = 33, 4.0                               # box [-1,1]^2 x [0,2], h = 1/16
    f = synth.point_sources((n, n, n), [1.0], [(0.0, 0.0, -depth)], dtype=np.float64,
                            device="cpu")
    c, w = (n - 1) / 2, (n - 1) / 2 * 0.35
    rng = np.random.default_rng(5)
    seeds = np.column_stack([rng.uniform(c - w, c + w, 40), rng.uniform(c - w, c + w, 40),
                             np.zeros(40)])
    r = oracle.trace_scott(f, seeds, method=method, s0=s0, tol=tol)
    assert np.all(r["flag_start"] == 5) and np.all(r["flag_end"] == 6)
    D = (depth + 2.0) * (n - 1) / 2
    Qc = 2 * (1 + 1 / D) ** 2
    np.testing.assert_allclose(r["q0"], Qc, rtol=1.5e-3)
    np.testing.assert_allclose(r["qperp"], Qc, rtol=1.5e-3)
    # radial projection: the exit point lies on the ray from the source
    src = np.array([c, c, -depth * (n - 1) / 2])
    ray = (r["end"] - src) / np.linalg.norm(r["end"] - src, axis=1, keepdims=True)
    ray0 = (r["start"] - src) / np.linalg.norm(r["start"] - src, axis=1, keepdims=True)
    assert np.max(np.linalg.norm(ray - ray0, axis=1)) < 2e-3


def test_scott_uniform_rkf45_grows_step_to_edge_exit(oracle):
    """With a zero error estimate the controller grows s to its 100-cell cap and
    the next step leaves through an edge: class 42, Q = 0 (reference rule)."""
    f = synth.uniform((20, 20, 20), (0.3, 0.2, 1.0))
    r = oracle.trace_scott(f, np.array([[5.0, 5.0, 0.0], [3.0, 4.0, 7.0]]), method="rkf45",
                           s0=0.5, tol=1e-6)
    assert list(r["flag_end"]) == [42, 42]
    assert r["flag_start"][0] == 5 and r["flag_start"][1] == 42
    np.testing.assert_array_equal(r["q0"], 0.0)
    np.testing.assert_allclose(r["end"][1], [3.0, 4.0, 7.0])   # never traced forward


def test_plain_path_hyperbolic_closed_form(oracle):
    from oracle import qplane
    shape = (41, 41, 33)
    k = 0.03
    f = synth.hyperbolic(shape, k=k, b0=1.0)
    seeds, sh = synth.plane_seeds_z(12, 12, (17, 23), (17, 23), z=1e-4, dtype=np.float64)
    stride = 6.0 / 11
    for method in ("rk4", "rkf45"):
        r = oracle.trace_plain(f, seeds, method=method, s0=0.125, tol_coef=1e-4)
        g = lambda a, c=None: (a if c is None else a[:, c]).reshape(sh)
        Q = qplane.qcalc_plane(g(r["end"], 0), g(r["end"], 1), g(r["end"], 2), g(r["flag_end"]),
                               g(r["start"], 0), g(r["start"], 1), g(r["start"], 2),
                               g(r["flag_start"]), g(r["b_start"], 0), g(r["b_start"], 1),
                               g(r["b_start"], 2), g(r["b_end"], 0), g(r["b_end"], 1),
                               g(r["b_end"], 2), g(r["b_this"], 0), g(r["b_this"], 1),
                               g(r["b_this"], 2), g(r["b_flag"]), stride)
        np.testing.assert_allclose(Q, 2 * np.cosh(2 * k * shape[2]), rtol=3e-3)


def test_quadrupole_rotation_symmetry(oracle):
    """B(R r) = -R B(r) for the alternating quadrupole, so Q(x, y) = Q(63-y, x)."""
    f = synth.quadrupole_64(device="cpu")
    n = 64
    g = np.linspace(1.0, 62.0, 24)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)])
    rot = np.column_stack([(n - 1) - seeds[:, 1], seeds[:, 0], seeds[:, 2]])
    a = oracle.trace_scott(f, seeds, method="rk4", s0=0.25)
    b = oracle.trace_scott(f, rot, method="rk4", s0=0.25)
    # the max-face landing plane sits at index N, not N-1, which breaks the
    # symmetry for lines that leave through a side face: compare z-to-z lines
    zz = lambda r: np.isin(r["flag_start"], (5, 6)) & np.isin(r["flag_end"], (5, 6))
    # lines grazing the central null (Q > 1e4) amplify round-off by Q itself
    good = zz(a) & zz(b) & np.isfinite(a["q0"]) & (a["q0"] < 1e4)
    assert good.sum() > 200
    rel = np.abs(a["q0"][good] - b["q0"][good]) / a["q0"][good]
    assert np.all(rel < 1e-8)
    # Q >= 2 holds for the exact map; RK4 at s = 0.25 across cell faces of a
    # 64^3 trilinear field undershoots by up to ~1% (it converges to >= 2 as
    # s -> 0 and N grows)
    assert np.all(a["q0"][good] > 1.97)
    side = ~zz(a) & np.isfinite(a["q0"])
    assert np.any(side)


def test_edge_cases(oracle):
    shape = (32, 32, 32)
    f = synth.hyperbolic(shape, k=0.03, b0=1.0)
    # below the surface: class -1, Q NaN
    r = oracle.trace_scott(f, np.array([[10.0, 10.0, -0.5]]))
    assert r["flag_start"][0] == -1 and r["flag_end"][0] == -1 and np.isnan(r["q0"][0])
    # a null (B = 0 at the X point when b0 = 0): class 42, Q = 0
    f0 = synth.hyperbolic(shape, k=0.03, b0=0.0)
    r = oracle.trace_scott(f0, np.array([[15.5, 15.5, 4.0]]), method="rk4")
    assert r["flag_start"][0] == 42 and r["q0"][0] == 0.0
    # length limit: a tightly wound helix never leaves the box
    hf, (xc, yc) = helix_field(shape, 1.0)
    hf = (hf[0], hf[1], np.asfortranarray(np.full(shape, 0.01)))
    r = oracle.trace_scott(hf, np.array([[xc + 4, yc, 5.0]]), method="rk4", s0=0.5)
    assert r["flag_start"][0] == 0 and r["flag_end"][0] == 0 and np.isnan(r["q0"][0])
    assert r["length"][0] >= 2 * sum(shape)
    np.testing.assert_allclose(r["end"][0], [xc + 4, yc, 5.0])    # never traced forward


# ------------------------------------------ the reference-precision (float) Scott tracer
def test_f32_model_closed_form_and_uniform(oracle):
    """The float restatement of TraceBlineScott (trace_scott_f32): Q = 2 cosh(2kN/B0)
    on the linear X-point field, and Q = 2 with the closed-form exit points
    of a uniform oblique field, to float accuracy."""
    from fastqsl_amd import synth
    shape, k = (41, 41, 33), 0.03
    f = synth.hyperbolic(shape, k=k, b0=1.0)
    g = np.linspace(14, 26, 9, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    seeds = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)]).astype(np.float32)
    r = oracle.trace_scott_f32(f, seeds, s0=0.25, tol=1e-6)
    # the seed on the axis (b0 || z) has no frame: Q = NaN there (decision register)
    top = (r["flag_end"] == 6) & np.isfinite(r["q0"])
    assert top.mean() > 0.9 and (r["flag_start"] == 5).all()
    # the float seven-stencil quotient carries ~1e-4 relative noise per step
    rel = np.abs(r["q0"][top] / (2 * np.cosh(2 * k * shape[2])) - 1)
    assert np.median(rel) < 2e-3 and rel.max() < 1e-2
    # closed-form footpoints: x - xc = (x0 - xc) exp(k z / b0) at z = N (the landing plane)
    xc = (shape[0] - 1) / 2.0
    np.testing.assert_allclose(r["end"][top, 0] - xc, (seeds[top, 0] - xc) * np.exp(k * shape[2]), rtol=1e-2)
    np.testing.assert_allclose(r["end"][top, 2], shape[2], atol=1e-5)
    u = synth.uniform((20, 20, 20), (0.3, -0.2, 1.0))
    su = np.array([[5.0, 7.0, 0.0], [10.0, 12.5, 0.0], [3.0, 15.0, 0.0]], np.float32)
    ru = oracle.trace_scott_f32(u, su, s0=0.25, tol=1e-6)
    assert (ru["flag_end"] == 6).all()
    np.testing.assert_allclose(ru["q0"], 2.0, rtol=1e-3)   # float weights: B flat to ~1e-7, J to ~5e-5
    np.testing.assert_allclose(ru["end"][:, 0], su[:, 0] + 0.3 * 20, atol=1e-4)
    np.testing.assert_allclose(ru["end"][:, 1], su[:, 1] - 0.2 * 20, atol=1e-4)


def test_f32_model_is_the_fp64_tracer_in_float(oracle):
    """On a random multipole the float model stays within float-arithmetic
    distance of the FP64 oracle (a dropped term or a wrong index would show
    as a gross difference): the same classes, footpoints within 1e-3 cells,
    Q within the noise of a float seven-stencil quotient (~1e-3); and its
    RKF45 evaluation counts sit below the oracle's, as an error norm with a
    float round-off floor predicts."""
    from fastqsl_amd import synth
    from tests import parity
    f = synth.multipole((48, 48, 48), seed=2, device="cpu")
    rng = np.random.default_rng(4)
    seeds = np.column_stack([rng.uniform(2, 45, 2000), rng.uniform(2, 45, 2000), np.zeros(2000)]).astype(np.float32)
    ref = oracle.trace_scott(f, seeds, method="rkf45", s0=0.25, tol=1e-6)
    r = oracle.trace_scott_f32(f, seeds, s0=0.25, tol=1e-6)
    cls = parity.class_match(r["flag_start"], r["flag_end"], ref["flag_start"], ref["flag_end"])
    assert cls.mean() > 0.99
    e = np.abs(r["end"] - ref["end"]).max(axis=1)[cls]
    assert np.median(e) < 1e-3
    d = parity.log_q_diff(r["q0"], ref["q0"])[cls]
    assert np.median(d) < 1e-2
    assert 0.5 < r["evals"].sum() / ref["evals"].sum() < 1.0
