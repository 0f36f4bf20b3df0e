/*
 * fqsl_oracle.c -- FP64 CPU restatement of FastQSL's field-line tracers.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * kernels in fastqsl_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the
 * timed CPU baseline).  The product path never calls it.
 *
 * Everything is double precision, except the *_f functions, which restate
 * the Scott tracer in the reference's own float arithmetic.  The field is three SoA arrays with x
 * fastest: B[k*ny*nx + j*nx + i] (TraceBlineScott.cu:36-39).
 *
 * What follows the reference, function by function:
 *   interp_vec        Interp3dxyzn            TraceBlineScott.cu:95-201
 *   interp_clamped    Interp3d                TraceBlineScott.cu:47-93
 *   grad_b_fd         grad_unit_vec_B         TraceBlineScott.cu:204-226
 *   scott_rhs         f_scott                 TraceBlineScott.cu:230-249
 *   scott_rkf45       RKF45_Scott             TraceBlineScott.cu:256-290
 *   scott_rk4         RK4 tableau of TraceBline.cu:226-234 applied to f_scott
 *   check_flag        checkFlag               TraceBlineScott.cu:310-332
 *   trace_scott_line  TraceBlineScott         TraceBlineScott.cu:334-519
 *   trace_scott_line_f32, interp_vec_f, grad_b_fd_f, scott_rhs_f, scott_rkf45_f
 *                     the same in the reference's float arithmetic
 *                                             TraceBlineScott.cu:95-508
 *   plain_rkf45       RKF45                   TraceBlineAdaptive.cu:215-254
 *   plain_rk4         RK4                     TraceBline.cu:226-234
 *   rk4_boundary      RK4_boundary            TraceBline.cu:253-274
 *   trace_plain_dir   TraceBlineAdap / TraceBline
 *                                             TraceBlineAdaptive.cu:333-426,
 *                                             TraceBline.cu:300-378
 *   orc_trace_plain   TraceAllBline           TraceBlineAdaptive.cu:429-510
 *   lookup_bisect     GetIndexBiSearch        TraceBlineAdaptiveStretch.cu:42-65
 *   lookup_uniform    GetIndexUniStretch      TraceBlineAdaptiveStretch.cu:68-75
 *   stretch_cell      Interp3dxyzn clamp and weights
 *                                             TraceBlineAdaptiveStretch.cu:82-124
 *   trace_stretch_dir TraceBlineAdap          TraceBlineAdaptiveStretch.cu:323-419
 *   orc_trace_stretch TraceAllBline           TraceBlineAdaptiveStretch.cu:422-508
 *
 * Where the reference reads uninitialised memory or leaves outputs undefined,
 * this file takes one documented resolution, shared with the GPU kernels:
 *   - Scott tolerance is an explicit absolute tolerance (TraceBlineScott.cu:378
 *     multiplies an uninitialised tol_this).
 *   - The Scott landing arc-length term uses the unit vector b at the last
 *     in-box state (TraceBlineScott.cu:448 reads an uninitialised B_P1) when
 *     twist is off; with twist on B_P1 is the raw B at the step's end point
 *     (:425), which the landing length and twist terms then use.
 *   - Twist (J.B/|B|^2/4pi along the line, TraceBline.cu:325-365,
 *     TraceBlineAdaptive.cu:373-384, TraceBlineScott.cu:423-457): a face exit
 *     whose step has no extent along the exit axis adds no landing twist (the
 *     reference's p_mid is unset there); the RKF45 plain tracer adds none at
 *     all (its landing term is commented out); a seed below z=-1e-8 has
 *     twist 0.
 *   - A direction that stops on the length or step limit keeps class 0 and its
 *     last state as endpoint; the line's Q is NaN.
 *   - A seed below z=-1e-8 gets class -1 on both sides and Q = NaN.
 *   - A class-42 exit sets both classes of an unfinished line to 42, keeps the
 *     last in-box state as that side's endpoint (the seed for a side never
 *     traced) and sets Q = 0 (TraceBlineScott.cu:460-465).
 *   - Stretched grids: the reference's cell rules and its landing plane
 *     (coordinate 0 / n_dim, not the box faces) are reproduced; the exact
 *     rule (the cell that holds the point, landing on the faces) is an
 *     alternative the caller selects.
 *   - Every line stops after FQSL_ATTEMPT_CAP integrator attempts (accepted or
 *     rejected) with class 0; the reference has no such cap, and no line of a
 *     field with finite values reaches it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#ifdef _OPENMP
#include <omp.h>
#endif

#define FD_H 1e-3              /* TraceBlineScott.cu:210 */
#define SCOTT_STEP_LIM 10000   /* N_STEP_LIM, TraceBlineScott.cu:13 */
#define SCOTT_LEN_RATIO 2.0    /* MAX_STEP_RATIO, TraceBlineScott.cu:11 */
#define RKF45_PLAIN_STEP_LIM 500000 /* MAX_NUM_STEP, TraceBlineAdaptive.cu:13 */
#define RKF45_PLAIN_LEN_RATIO 3.0   /* MAX_STEP_RATIO, TraceBlineAdaptive.cu:11 */
#define RK4_PLAIN_STEP_LIM 100000   /* MAX_NUM_STEP, TraceBline.cu:12 */
#define RK4_PLAIN_LEN_RATIO 2.0     /* MAX_STEP_RATIO, TraceBline.cu:10 */
#define TOL_BASE 1e-3               /* TOL, TraceBlineAdaptive.cu:12 */
#define FQSL_ATTEMPT_CAP 1000000L

enum { M_RKF45 = 0, M_RK4 = 1 };

/* A stretched grid: node coordinates per axis (strictly increasing) and the
 * rule that picks the interpolation cell (TraceBlineAdaptiveStretch.cu). */
enum { LOOKUP_BISECT = 0, LOOKUP_UNIFORM = 1, LOOKUP_EXACT = 2 };
typedef struct {
    const double *a[3];
    int rule;
} axes_t;

typedef struct {
    const double *bx, *by, *bz;
    int n[3];
    const axes_t *ax;   /* NULL: grid-index coordinates */
} field_t;

typedef struct { double x[3], u[3], v[3]; } state9;

/* ---------------------------------------------------------------- helpers */
static double dot3(const double *a, const double *b) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static double len3(const double *a) { return sqrt(dot3(a, a)); }

/* Interp3dxyzn (TraceBlineScott.cu:95-201): trilinear interpolation of the
 * three components with shared weights.  A coordinate <= 0 uses cell 0 at
 * weight 0; a coordinate >= N-1 uses cell N-2 at weight 1 (constant
 * extrapolation outside the box). */
static void axis_cell(double p, int n, int *idx, double *r) {
    double fl = floor(p);
    double rr = p - fl;
    int ii;
    if (p != p) { *idx = 0; *r = rr; return; }  /* NaN: keep index in range, value stays NaN */
    if (p <= 0.0) { *idx = 0; *r = 0.0; return; }
    if (p >= n - 1) { *idx = n - 2; *r = 1.0; return; }
    ii = (int)fl;
    if (ii > n - 2) ii = n - 2;
    *idx = ii; *r = rr;
}

/* GetIndexBiSearch (TraceBlineAdaptiveStretch.cu:42-65): bisection that
 * stops once the bracket spans at most three cells and then decides between
 * the bracket's middle node and the one below it only, so a point in the
 * bracket's top cell is given the cell below (and extrapolated from it). */
static int lookup_bisect(double p, int n, const double *a) {
    int lo = 0, hi = n - 1, mid = n / 2, r;
    while (hi - lo > 3) {
        if (p > a[mid]) lo = mid; else hi = mid;
        mid = (lo + hi) / 2;
    }
    r = (p > a[mid]) ? mid : mid - 1;
    return r > n - 2 ? n - 2 : r;
}

/* GetIndexUniStretch (TraceBlineAdaptiveStretch.cu:68-75): the cell index of
 * an evenly spaced axis, with the span divided by n (not n - 1). */
static int lookup_uniform(double p, int n, const double *a) {
    double r = floor((p - a[0]) / ((a[n - 1] - a[0]) / (double)n));
    if (r < 0.0) return 0;
    if (r > n - 2) return n - 2;
    return (int)r;
}

/* The cell that contains p: a[c] < p <= a[c+1], c in [0, n-2] (linear scan;
 * test infrastructure). */
static int lookup_exact(double p, int n, const double *a) {
    int c = 0;
    while (c < n - 2 && p > a[c + 1]) ++c;
    return c;
}

/* Interp3dxyzn of the stretched tracer (TraceBlineAdaptiveStretch.cu:82-194):
 * the point is first pulled into [a0 + 1e-6, a_{n-1} - 1e-6] (NaN goes to the
 * low bound), then the rule picks the cell c and the weight is
 * (p - a[c]) / (a[c+1] - a[c]), outside [0, 1] when the cell does not hold p. */
static void stretch_cell(const axes_t *ax, int axis, double p, int n, int *idx, double *r) {
    const double *a = ax->a[axis];
    int c;
    p = (p > a[0]) ? p : a[0] + 1e-6;
    p = (p < a[n - 1]) ? p : a[n - 1] - 1e-6;
    if (ax->rule == LOOKUP_BISECT) c = lookup_bisect(p, n, a);
    else if (ax->rule == LOOKUP_UNIFORM) c = lookup_uniform(p, n, a);
    else c = lookup_exact(p, n, a);
    *idx = c;
    *r = (p - a[c]) / (a[c + 1] - a[c]);
}

static void interp_vec(const field_t *f, const double *p, double *out, int normalize) {
    int ix, iy, iz;
    double rx, ry, rz, w[8];
    long nx = f->n[0], nxy = (long)f->n[0] * f->n[1];
    long base;
    long off[8];
    int c;
    if (f->ax) {
        stretch_cell(f->ax, 0, p[0], f->n[0], &ix, &rx);
        stretch_cell(f->ax, 1, p[1], f->n[1], &iy, &ry);
        stretch_cell(f->ax, 2, p[2], f->n[2], &iz, &rz);
    } else {
        axis_cell(p[0], f->n[0], &ix, &rx);
        axis_cell(p[1], f->n[1], &iy, &ry);
        axis_cell(p[2], f->n[2], &iz, &rz);
    }
    w[0] = (1. - rx) * (1. - ry) * (1. - rz);
    w[1] = (1. - rx) * (1. - ry) * (rz);
    w[2] = (1. - rx) * (ry) * (1. - rz);
    w[3] = (1. - rx) * (ry) * (rz);
    w[4] = (rx) * (1. - ry) * (1. - rz);
    w[5] = (rx) * (1. - ry) * (rz);
    w[6] = (rx) * (ry) * (1. - rz);
    w[7] = (rx) * (ry) * (rz);
    base = iz * nxy + iy * nx + ix;
    off[0] = base;                 off[1] = base + nxy;
    off[2] = base + nx;            off[3] = base + nx + nxy;
    off[4] = base + 1;             off[5] = base + 1 + nxy;
    off[6] = base + 1 + nx;        off[7] = base + 1 + nx + nxy;
    out[0] = out[1] = out[2] = 0.0;
    for (c = 0; c < 8; ++c) {
        out[0] += f->bx[off[c]] * w[c];
        out[1] += f->by[off[c]] * w[c];
        out[2] += f->bz[off[c]] * w[c];
    }
    if (normalize) {
        double nb = len3(out);
        out[0] /= nb; out[1] /= nb; out[2] /= nb;
    }
}

/* Interp3d (TraceBlineScott.cu:47-93): scalar trilinear with the point pulled
 * into [1e-4, N-1-1e-4] per axis first (a NaN coordinate becomes 1e-4). */
static void interp_clamped(const field_t *f, const double *p, double *out) {
    double q[3];
    int a;
    for (a = 0; a < 3; ++a) {
        double v = p[a];
        v = (v > 0) ? v : 0.0001;
        v = (v < f->n[a] - 1) ? v : ((double)f->n[a] - 1 - 0.0001);
        q[a] = v;
    }
    interp_vec(f, q, out, 0);
}

/* grad_unit_vec_B (TraceBlineScott.cu:204-226): J[j][i] = d b_i / d x_j by a
 * central difference of the normalised, clamped interpolant with h = 1e-3. */
static void grad_b_fd(const field_t *f, const double *p, double J[3][3]) {
    int j, i;
    for (j = 0; j < 3; ++j) {
        double p1[3] = {p[0], p[1], p[2]}, p2[3] = {p[0], p[1], p[2]};
        double b1[3], b2[3];
        p1[j] -= FD_H; p2[j] += FD_H;
        interp_vec(f, p1, b1, 1);
        interp_vec(f, p2, b2, 1);
        for (i = 0; i < 3; ++i) J[j][i] = (b2[i] - b1[i]) / (2 * FD_H);
    }
}

/* Closed-form derivative of b = B/|B| inside one cell (no difference
 * quotient).  Used only to cross-check grad_b_fd away from cell faces. */
static void grad_b_analytic(const field_t *f, const double *p, double J[3][3]) {
    int ix, iy, iz, j, i, c;
    double r[3], B[3] = {0, 0, 0}, dB[3][3] = {{0}};
    long nx = f->n[0], nxy = (long)f->n[0] * f->n[1];
    const double *comp[3] = {f->bx, f->by, f->bz};
    axis_cell(p[0], f->n[0], &ix, &r[0]);
    axis_cell(p[1], f->n[1], &iy, &r[1]);
    axis_cell(p[2], f->n[2], &iz, &r[2]);
    for (c = 0; c < 8; ++c) {
        int a = (c >> 2) & 1, b = (c >> 1) & 1, cc = c & 1;
        double wa = a ? r[0] : 1 - r[0], wb = b ? r[1] : 1 - r[1], wc = cc ? r[2] : 1 - r[2];
        double sa = a ? 1 : -1, sb = b ? 1 : -1, sc = cc ? 1 : -1;
        long o = (iz + cc) * nxy + (iy + b) * nx + (ix + a);
        for (i = 0; i < 3; ++i) {
            double v = comp[i][o];
            B[i] += v * wa * wb * wc;
            dB[0][i] += v * sa * wb * wc;
            dB[1][i] += v * wa * sb * wc;
            dB[2][i] += v * wa * wb * sc;
        }
    }
    {
        double nb = len3(B), b[3] = {B[0] / nb, B[1] / nb, B[2] / nb};
        for (j = 0; j < 3; ++j) {
            double bd = dot3(b, dB[j]);
            for (i = 0; i < 3; ++i) J[j][i] = (dB[j][i] - b[i] * bd) / nb;
        }
    }
}

/* checkFlag (TraceBlineScott.cu:310-332): 0 inside (note the strict y > 0),
 * 1..6 = x min/max, y min/max, z min/max, 42 otherwise (edges, NaN). */
static int check_flag_box(const double *lo, const double *hi, const double *p) {
    double X = p[0], Y = p[1], Z = p[2];
    double x0 = lo[0], y0 = lo[1], z0 = lo[2], mx = hi[0], my = hi[1], mz = hi[2];
    int f = 42;
    if (X >= x0 && Y > y0 && Z >= z0 && X <= mx && Y <= my && Z <= mz) return 0;
    if (X < x0 && Y >= y0 && Z >= z0 && X < mx && Y < my && Z < mz) f = 1;
    if (X >= x0 && Y >= y0 && Z >= z0 && X >= mx && Y < my && Z < mz) f = 2;
    if (X >= x0 && Y < y0 && Z >= z0 && X < mx && Y < my && Z < mz) f = 3;
    if (X >= x0 && Y >= y0 && Z >= z0 && X < mx && Y >= my && Z < mz) f = 4;
    if (X >= x0 && Y >= y0 && Z < z0 && X < mx && Y < my && Z < mz) f = 5;
    if (X >= x0 && Y >= y0 && Z >= z0 && X < mx && Y < my && Z >= mz) f = 6;
    return f;
}

static int check_flag(const int *n, const double *p) {
    const double lo[3] = {0., 0., 0.}, hi[3] = {n[0] - 1., n[1] - 1., n[2] - 1.};
    return check_flag_box(lo, hi, p);
}

/* ------------------------------------------------------------ Scott RHS */
/* f_scott (TraceBlineScott.cu:230-249): x' = b, u' = (u.grad) b, v' = (v.grad) b. */
static void scott_rhs(const field_t *f, const state9 *y, state9 *dy, int grad_mode) {
    double J[3][3];
    int i;
    interp_vec(f, y->x, dy->x, 1);
    if (grad_mode == 0) grad_b_fd(f, y->x, J); else grad_b_analytic(f, y->x, J);
    for (i = 0; i < 3; ++i) {
        dy->u[i] = y->u[0] * J[0][i] + y->u[1] * J[1][i] + y->u[2] * J[2][i];
        dy->v[i] = y->v[0] * J[0][i] + y->v[1] * J[1][i] + y->v[2] * J[2][i];
    }
}

static void s9_axpy(state9 *out, const state9 *y, int nk, const double *coef, state9 *const *k) {
    /* out = y + sum_j coef[j] * k[j] */
    const double *py = (const double *)y;
    double *po = (double *)out;
    int c, j;
    for (c = 0; c < 9; ++c) {
        double acc = 0.0;
        for (j = 0; j < nk; ++j) acc += coef[j] * ((const double *)k[j])[c];
        po[c] = py[c] + acc;
    }
}

static void s9_scale(state9 *k, double s) {
    double *p = (double *)k;
    int c;
    for (c = 0; c < 9; ++c) p[c] *= s;
}

/* Fehlberg 4(5) tableau (TraceBlineScott.cu:267-275). */
static const double A2[1] = {1. / 4.};
static const double A3[2] = {3. / 32., 9. / 32.};
static const double A4[3] = {1932. / 2197., -7200. / 2197., 7296. / 2197.};
static const double A5[4] = {439. / 216., -8., 3680. / 513., -845. / 4104.};
static const double A6[5] = {-8. / 27., 2., -3544. / 2565., 1859. / 4104., -11. / 40.};
static const double B5[6] = {16. / 135., 0., 6656. / 12825., 28561. / 56430., -9. / 50., 2. / 55.};
static const double CE[6] = {1. / 360., 0., -128. / 4275., -2197. / 75240., 1. / 50., 2. / 55.};

/* RKF45_Scott (TraceBlineScott.cu:256-290): k_i = s f(y + sum a_ij k_j),
 * advance with the 5th-order weights, error = |sum ce_i k_i.x| (position). */
static double scott_rkf45(const field_t *f, const state9 *y, state9 *out, double s, int gm) {
    state9 k1, k2, k3, k4, k5, k6, t;
    state9 *kk[6] = {&k1, &k2, &k3, &k4, &k5, &k6};
    double e[3];
    int i;
    scott_rhs(f, y, &k1, gm); s9_scale(&k1, s);
    s9_axpy(&t, y, 1, A2, kk); scott_rhs(f, &t, &k2, gm); s9_scale(&k2, s);
    s9_axpy(&t, y, 2, A3, kk); scott_rhs(f, &t, &k3, gm); s9_scale(&k3, s);
    s9_axpy(&t, y, 3, A4, kk); scott_rhs(f, &t, &k4, gm); s9_scale(&k4, s);
    s9_axpy(&t, y, 4, A5, kk); scott_rhs(f, &t, &k5, gm); s9_scale(&k5, s);
    s9_axpy(&t, y, 5, A6, kk); scott_rhs(f, &t, &k6, gm); s9_scale(&k6, s);
    s9_axpy(out, y, 6, B5, kk);
    for (i = 0; i < 3; ++i)
        e[i] = CE[0] * k1.x[i] + CE[2] * k3.x[i] + CE[3] * k4.x[i] + CE[4] * k5.x[i] + CE[5] * k6.x[i];
    return len3(e);
}

/* RK4 (TraceBline.cu:226-234) applied to the Scott 9-vector. */
static void scott_rk4(const field_t *f, const state9 *y, state9 *out, double s, int gm) {
    state9 k1, k2, k3, k4, t;
    const double *py = (const double *)y;
    double *pt = (double *)&t, *po = (double *)out;
    int c;
    scott_rhs(f, y, &k1, gm);
    for (c = 0; c < 9; ++c) pt[c] = py[c] + s * ((double *)&k1)[c] / 2.;
    scott_rhs(f, &t, &k2, gm);
    for (c = 0; c < 9; ++c) pt[c] = py[c] + s * ((double *)&k2)[c] / 2.;
    scott_rhs(f, &t, &k3, gm);
    for (c = 0; c < 9; ++c) pt[c] = py[c] + s * ((double *)&k3)[c];
    scott_rhs(f, &t, &k4, gm);
    for (c = 0; c < 9; ++c)
        po[c] = py[c] + 1. / 6. * s * (((double *)&k1)[c] + 2. * ((double *)&k2)[c] +
                                       2. * ((double *)&k3)[c] + ((double *)&k4)[c]);
}

/* ------------------------------------------------------- Scott driver */
typedef struct {
    double start[3], end[3];
    int flag_start, flag_end;
    double q0, qperp, len, twist;
    double bthis[3];
    long evals, steps;
} scott_result;

/* J.B / |B|^2 / 4 pi at p from the current-density field cur and B (both the
 * raw Interp3dxyzn, TraceBlineScott.cu:424-426), and that B. */
static double twist_density(const field_t *f, const field_t *cur, const double *p, double *Bout) {
    double J[3];
    interp_vec(cur, p, J, 0);
    interp_vec(f, p, Bout, 0);
    return dot3(J, Bout) / dot3(Bout, Bout) / 4.0 / M_PI;
}

static void normalize3(double *a) {
    double n = len3(a);
    a[0] /= n; a[1] /= n; a[2] /= n;
}

/* TraceBlineScott (TraceBlineScott.cu:334-519) with the Q assembly at :475-508. */
static void trace_scott_line(const field_t *f, const field_t *cur, const double *p0, int method,
                             double s0, double tol, int gm, scott_result *res) {
    const int *n = f->n;
    double len_lim = SCOTT_LEN_RATIO * (n[0] + n[1] + n[2]);
    double B0[3], b0[3], Nb0, u0[3], v0[3];
    double s = s0, len = 0.0;
    long steps = 0, attempts = 0, evals = 0;
    int z0, dir, flags[2] = {0, 0};
    state9 ends[2];
    int aborted = 0, abort_side = -1;
    double twist = 0.0;

    res->twist = 0.0;
    interp_clamped(f, p0, res->bthis);           /* TraceBlineScott.cu:577-579 */
    memset(ends, 0, sizeof(ends));
    if (p0[2] < -1e-8) {                           /* TraceBlineScott.cu:366 */
        memcpy(res->start, p0, sizeof(double) * 3);
        memcpy(res->end, p0, sizeof(double) * 3);
        res->flag_start = res->flag_end = -1;
        res->q0 = res->qperp = NAN;
        res->len = 0; res->evals = 0; res->steps = 0;
        return;
    }
    interp_vec(f, p0, B0, 0);
    Nb0 = len3(B0);
    interp_vec(f, p0, b0, 1);
    /* v0 = normalize((b.y,-b.x,0) - ((b.y,-b.x,0).b) b), u0 = normalize(b x v0)
     * (TraceBlineScott.cu:380-386) */
    v0[0] = b0[1]; v0[1] = -b0[0]; v0[2] = 0.0;
    {
        double d = dot3(v0, b0);
        v0[0] -= d * b0[0]; v0[1] -= d * b0[1]; v0[2] -= d * b0[2];
        normalize3(v0);
    }
    u0[0] = b0[1] * v0[2] - b0[2] * v0[1];
    u0[1] = b0[2] * v0[0] - b0[0] * v0[2];
    u0[2] = b0[0] * v0[1] - b0[1] * v0[0];
    normalize3(u0);
    z0 = p0[2] < 1e-6;                             /* TraceBlineScott.cu:388 */

    /* backward (-1) first, then forward (+1); s, len and steps carry over
     * (TraceBlineScott.cu:389-405) */
    for (dir = -1; dir < 2 && !aborted; dir += 2) {
        int side = dir < 0 ? 0 : 1;
        state9 a, b;
        int flag = 0;
        memcpy(a.x, p0, sizeof(double) * 3);
        memcpy(a.u, u0, sizeof(double) * 3);
        memcpy(a.v, v0, sizeof(double) * 3);
        if (z0 && b0[2] * dir < 0) {               /* TraceBlineScott.cu:396-401 */
            ends[side] = a; flags[side] = 5;
            continue;
        }
        s = fabs(s);
        while (flag == 0 && len < len_lim && steps < SCOTT_STEP_LIM) {
            double dL, d[3];
            if (attempts >= FQSL_ATTEMPT_CAP) break;
            ++attempts;
            if (method == M_RKF45) {
                double err = scott_rkf45(f, &a, &b, s * dir, gm);
                double scale = pow(tol / err / 11.09, 0.2);
                evals += 6;
                if (scale < 0.618) { s = s * 0.618; continue; }   /* :415-416 */
                s = s * scale;
                if (s > 100.) s = 100.;
                if (s < 1. / 10.) s = 1. / 10.;
            } else {
                scott_rk4(f, &a, &b, s * dir, gm);
                evals += 4;
            }
            d[0] = b.x[0] - a.x[0]; d[1] = b.x[1] - a.x[1]; d[2] = b.x[2] - a.x[2];
            dL = len3(d);
            len += dL;
            /* twist: J.B/|B|^2/4pi at the step's end point times its length
             * (TraceBlineScott.cu:423-427) */
            double tw_rate = 0.0, Bend[3] = {0.0, 0.0, 0.0};
            if (cur) {
                tw_rate = twist_density(f, cur, b.x, Bend);
                twist += tw_rate * dL;
            }
            flag = check_flag(n, b.x);
            if (flag > 0) {
                len -= dL;
                if (cur) twist -= tw_rate * dL;          /* :432-434 */
                if (flag <= 6) {
                    int dim = (flag - 1) / 2;
                    double p1 = a.x[dim], p2 = b.x[dim];
                    if (fabs(p1 - p2) > 0) {
                        double pm = (flag % 2 == 1) ? 0.0 : (double)n[dim];
                        double ba[3], arc;
                        int c;
                        const double *pa = (const double *)&a, *pb = (const double *)&b;
                        double *pe = (double *)&ends[side];
                        for (c = 0; c < 9; ++c) pe[c] = (pa[c] * (p2 - pm) + pb[c] * (pm - p1)) / (p2 - p1);
                        /* the landing term's B: with twist on, the reference's
                         * B_P1 is the raw B at the step's end point (:425); with
                         * it off B_P1 is never set and b at the last in-box state
                         * is used (decision register) */
                        if (cur) memcpy(ba, Bend, sizeof(ba));
                        else interp_vec(f, a.x, ba, 1);
                        arc = fabs(pm - p1) / (1e-4 + fabs(ba[dim]));
                        len += arc;
                        if (cur) twist += tw_rate * arc;  /* :455-457 */
                    } else {
                        /* no landing segment: the reference's p_mid is unset
                         * here; no length and no twist are added */
                        ends[side] = a;
                    }
                    flags[side] = flag;
                } else {
                    ends[side] = a;
                    flags[side] = 42;
                    aborted = 1;
                    abort_side = side;
                }
            }
            a = b;
            ++steps;
            if (aborted) break;
        }
        if (!aborted && flag == 0) {                /* length / step limit */
            ends[side] = a;
            flags[side] = 0;
        }
    }
    if (aborted) {
        /* class 42: q = 0 (TraceBlineScott.cu:460-465); when the backward side
         * aborts, the forward side is never traced and keeps the seed */
        if (abort_side == 0) {
            memcpy(ends[1].x, p0, sizeof(double) * 3);
            flags[1] = 42;
        }
        res->q0 = 0.0; res->qperp = 0.0;
    } else if (flags[0] == 0 || flags[1] == 0) {
        res->q0 = NAN; res->qperp = NAN;
    } else {
        /* Q assembly (TraceBlineScott.cu:475-508) */
        double Bs[3], Be[3], us1[3], vs1[3], ue1[3], ve1[3];
        int ds = (flags[0] - 1) / 2, de = (flags[1] - 1) / 2, i;
        double *us = ends[0].u, *vs = ends[0].v, *ue = ends[1].u, *ve = ends[1].v;
        double BNs, BNe, nbs, nbe, num;
        interp_vec(f, ends[0].x, Bs, 0);
        interp_vec(f, ends[1].x, Be, 0);
        BNs = Bs[ds]; BNe = Be[de];
        for (i = 0; i < 3; ++i) {
            us1[i] = us[i] - us[ds] / Bs[ds] * Bs[i];
            vs1[i] = vs[i] - vs[ds] / Bs[ds] * Bs[i];
            ue1[i] = ue[i] - ue[de] / Be[de] * Be[i];
            ve1[i] = ve[i] - ve[de] / Be[de] * Be[i];
        }
        num = dot3(ue1, ue1) * dot3(vs1, vs1) + dot3(us1, us1) * dot3(ve1, ve1)
              - 2.0 * dot3(ue1, ve1) * dot3(us1, vs1);
        res->q0 = fabs(num) / fabs((Nb0 * Nb0) / (BNs * BNe));
        nbs = len3(Bs); nbe = len3(Be);
        {
            double due = dot3(ue, Be) / nbe, dve = dot3(ve, Be) / nbe;
            double dus = dot3(us, Bs) / nbs, dvs = dot3(vs, Bs) / nbs;
            for (i = 0; i < 3; ++i) {
                ue1[i] = ue[i] - due * (Be[i] / nbe);
                ve1[i] = ve[i] - dve * (Be[i] / nbe);
                us1[i] = us[i] - dus * (Bs[i] / nbs);
                vs1[i] = vs[i] - dvs * (Bs[i] / nbs);
            }
        }
        num = dot3(ue1, ue1) * dot3(vs1, vs1) + dot3(us1, us1) * dot3(ve1, ve1)
              - 2.0 * dot3(ue1, ve1) * dot3(us1, vs1);
        res->qperp = fabs(num) / ((Nb0 * Nb0) / (nbs * nbe));
    }
    memcpy(res->start, ends[0].x, sizeof(double) * 3);
    memcpy(res->end, ends[1].x, sizeof(double) * 3);
    res->flag_start = flags[0];
    res->flag_end = flags[1];
    res->len = len;
    res->twist = twist;
    res->evals = evals;
    res->steps = steps;
}

/* ------------------------------------ Scott driver in the reference's float */
/* The Scott line driver again, in the reference kernel's own arithmetic
 * (TraceBlineScott.cu:95-508 as written): float field values, positions,
 * stages, deviation vectors, error norm, step size and controller (powf);
 * the weight products (1.-r)(1.-r)(1.-r) in double rounded to float, as the
 * double literals make them; the seven-stencil difference quotient in float
 * (the +-0.001 offsets formed in double and rounded); arc length in double.
 * It is compiled without FMA contraction, which nvcc applies by default, so
 * it reproduces the reference's algorithm and precision, not its bits.  Its
 * step sequences, evaluation counts and classes are what an FP32 evaluation
 * of the reference produces.  RKF45 only (the reference's Scott kernel);
 * the decision register above applies unchanged. */
typedef struct { float x[3], u[3], v[3]; } state9f;

static float dot3f(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Interp3dxyzn (TraceBlineScott.cu:95-201) in float. */
static void interp_vec_f(const field_t *f, const float *p, float *out, int normalize) {
    float r[3], w[8];
    int idx[3], a, c;
    long nx = f->n[0], nxy = (long)f->n[0] * f->n[1], base, off[8];
    for (a = 0; a < 3; ++a) {
        r[a] = p[a] - floorf(p[a]);
        idx[a] = (int)floorf(p[a]);
        if (p[a] <= 0.) { r[a] = 0; idx[a] = 0; }
        if (p[a] >= f->n[a] - 1) { r[a] = 1; idx[a] = f->n[a] - 2; }
    }
    w[0] = (float)((1. - r[0]) * (1. - r[1]) * (1. - r[2]));
    w[1] = (float)((1. - r[0]) * (1. - r[1]) * (r[2]));
    w[2] = (float)((1. - r[0]) * (r[1]) * (1. - r[2]));
    w[3] = (float)((1. - r[0]) * (r[1]) * (r[2]));
    w[4] = (float)((r[0]) * (1. - r[1]) * (1. - r[2]));
    w[5] = (float)((r[0]) * (1. - r[1]) * (r[2]));
    w[6] = (float)((r[0]) * (r[1]) * (1. - r[2]));
    w[7] = (float)((r[0]) * (r[1]) * (r[2]));
    base = idx[2] * nxy + idx[1] * nx + idx[0];
    off[0] = base;                 off[1] = base + nxy;
    off[2] = base + nx;            off[3] = base + nx + nxy;
    off[4] = base + 1;             off[5] = base + 1 + nxy;
    off[6] = base + 1 + nx;        off[7] = base + 1 + nx + nxy;
    out[0] = (float)f->bx[off[0]] * w[0];
    out[1] = (float)f->by[off[0]] * w[0];
    out[2] = (float)f->bz[off[0]] * w[0];
    for (c = 1; c < 8; ++c) {
        out[0] = out[0] + (float)f->bx[off[c]] * w[c];
        out[1] = out[1] + (float)f->by[off[c]] * w[c];
        out[2] = out[2] + (float)f->bz[off[c]] * w[c];
    }
    if (normalize) {
        float nb = sqrtf(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
        out[0] = out[0] / nb; out[1] = out[1] / nb; out[2] = out[2] / nb;
    }
}

/* grad_unit_vec_B (TraceBlineScott.cu:204-226) in float. */
static void grad_b_fd_f(const field_t *f, const float *p, float J[3][3]) {
    int j, i;
    for (j = 0; j < 3; ++j) {
        float p1[3] = {p[0], p[1], p[2]}, p2[3] = {p[0], p[1], p[2]};
        float b1[3], b2[3];
        p1[j] = (float)((double)p1[j] - 0.001);
        p2[j] = (float)((double)p2[j] + 0.001);
        interp_vec_f(f, p1, b1, 1);
        interp_vec_f(f, p2, b2, 1);
        for (i = 0; i < 3; ++i) J[j][i] = (b2[i] - b1[i]) / 0.002f;
    }
}

/* f_scott (TraceBlineScott.cu:230-249) in float. */
static void scott_rhs_f(const field_t *f, const state9f *y, state9f *dy) {
    float J[3][3];
    int i;
    interp_vec_f(f, y->x, dy->x, 1);
    grad_b_fd_f(f, y->x, J);
    for (i = 0; i < 3; ++i) {
        dy->u[i] = y->u[0] * J[0][i] + y->u[1] * J[1][i] + y->u[2] * J[2][i];
        dy->v[i] = y->v[0] * J[0][i] + y->v[1] * J[1][i] + y->v[2] * J[2][i];
    }
}

/* out = y + (sum_j coef[j] * k[j]), the float9 expression of RKF45_Scott. */
static void s9f_axpy(state9f *out, const state9f *y, int nk, const float *coef, state9f *const *k) {
    const float *py = (const float *)y;
    float *po = (float *)out;
    int c, j;
    for (c = 0; c < 9; ++c) {
        float acc = coef[0] * ((const float *)k[0])[c];
        for (j = 1; j < nk; ++j) acc = acc + coef[j] * ((const float *)k[j])[c];
        po[c] = py[c] + acc;
    }
}

static void s9f_scale(state9f *k, float s) {
    float *p = (float *)k;
    int c;
    for (c = 0; c < 9; ++c) p[c] = s * p[c];
}

/* RKF45_Scott (TraceBlineScott.cu:256-290) in float: the tableau rounded to
 * float, the 5th-order advance, err = |sum ce_i k_i.x| in float. */
static float scott_rkf45_f(const field_t *f, const state9f *y, state9f *out, float s) {
    static const float a2[1] = {(float)(1. / 4.)};
    static const float a3[2] = {(float)(3. / 32.), (float)(9. / 32.)};
    static const float a4[3] = {(float)(1932. / 2197.), (float)(-7200. / 2197.), (float)(7296. / 2197.)};
    static const float a5[4] = {(float)(439. / 216.), (float)(-8.), (float)(3680. / 513.), (float)(-845. / 4104.)};
    static const float a6[5] = {(float)(-8. / 27.), (float)2., (float)(-3544. / 2565.), (float)(1859. / 4104.),
                                (float)(-11. / 40.)};
    static const float b5[5] = {(float)(16. / 135.), (float)(6656. / 12825.), (float)(28561. / 56430.),
                                (float)(-9. / 50.), (float)(2. / 55.)};
    static const float ce[5] = {(float)(1. / 360.), (float)(-128. / 4275.), (float)(-2197. / 75240.),
                                (float)(1. / 50.), (float)(2. / 55.)};
    state9f k1, k2, k3, k4, k5, k6, t;
    state9f *kk[6] = {&k1, &k2, &k3, &k4, &k5, &k6};
    state9f *kb[5] = {&k1, &k3, &k4, &k5, &k6};
    float e[3];
    int i;
    scott_rhs_f(f, y, &k1); s9f_scale(&k1, s);
    s9f_axpy(&t, y, 1, a2, kk); scott_rhs_f(f, &t, &k2); s9f_scale(&k2, s);
    s9f_axpy(&t, y, 2, a3, kk); scott_rhs_f(f, &t, &k3); s9f_scale(&k3, s);
    s9f_axpy(&t, y, 3, a4, kk); scott_rhs_f(f, &t, &k4); s9f_scale(&k4, s);
    s9f_axpy(&t, y, 4, a5, kk); scott_rhs_f(f, &t, &k5); s9f_scale(&k5, s);
    s9f_axpy(&t, y, 5, a6, kk); scott_rhs_f(f, &t, &k6); s9f_scale(&k6, s);
    s9f_axpy(out, y, 5, b5, kb);
    for (i = 0; i < 3; ++i)
        e[i] = ce[0] * k1.x[i] + ce[1] * k3.x[i] + ce[2] * k4.x[i] + ce[3] * k5.x[i] + ce[4] * k6.x[i];
    return sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
}

static void normalize3f(float *a) {
    float n = 1.0f / sqrtf(dot3f(a, a));
    a[0] *= n; a[1] *= n; a[2] *= n;
}

/* TraceBlineScott (TraceBlineScott.cu:334-508) in float, RKF45. */
static void trace_scott_line_f32(const field_t *f, const double *p0d, double s0, double tol, scott_result *res) {
    const int *n = f->n;
    const double len_lim = SCOTT_LEN_RATIO * (n[0] + n[1] + n[2]);
    const float p0[3] = {(float)p0d[0], (float)p0d[1], (float)p0d[2]};
    const float tolf = (float)tol;
    float b0[3], B0raw[3], Nb0, u0[3], v0[3];
    float s = (float)s0;
    double len = 0.0;
    long steps = 0, attempts = 0, evals = 0;
    int dir, flags[2] = {0, 0}, side;
    state9f ends[2];
    int aborted = 0, abort_side = -1;

    memset(ends, 0, sizeof(ends));
    interp_clamped(f, p0d, res->bthis);
    res->twist = 0.0;
    if (p0[2] < -1e-8) {                           /* TraceBlineScott.cu:366 */
        memcpy(res->start, p0d, sizeof(double) * 3);
        memcpy(res->end, p0d, sizeof(double) * 3);
        res->flag_start = res->flag_end = -1;
        res->q0 = res->qperp = NAN;
        res->len = 0; res->evals = 0; res->steps = 0;
        return;
    }
    interp_vec_f(f, p0, b0, 1);                    /* :368-369 */
    interp_vec_f(f, p0, B0raw, 0);
    Nb0 = sqrtf(dot3f(B0raw, B0raw));
    v0[0] = b0[1]; v0[1] = -b0[0]; v0[2] = 0.0f;  /* :380-386 */
    {
        float d = dot3f(v0, b0);
        v0[0] = v0[0] - d * b0[0]; v0[1] = v0[1] - d * b0[1]; v0[2] = v0[2] - d * b0[2];
        normalize3f(v0);
    }
    u0[0] = b0[1] * v0[2] - b0[2] * v0[1];
    u0[1] = b0[2] * v0[0] - b0[0] * v0[2];
    u0[2] = b0[0] * v0[1] - b0[1] * v0[0];
    normalize3f(u0);
    for (dir = -1; dir < 2 && !aborted; dir += 2) {
        state9f a, b;
        int flag = 0;
        side = dir < 0 ? 0 : 1;
        memcpy(a.x, p0, sizeof(a.x));
        memcpy(a.u, u0, sizeof(a.u));
        memcpy(a.v, v0, sizeof(a.v));
        if (p0[2] < 1e-6 && b0[2] * dir < 0) {     /* :396-401 */
            ends[side] = a; flags[side] = 5;
            continue;
        }
        s = fabsf(s);
        while (flag == 0 && len < len_lim && steps < SCOTT_STEP_LIM) {
            float err, scale, dL, d[3];
            if (attempts >= FQSL_ATTEMPT_CAP) break;
            ++attempts;
            err = scott_rkf45_f(f, &a, &b, s * (float)dir);
            evals += 6;
            scale = powf((float)(tolf / err / 11.09), 0.2f);      /* :413 */
            if (scale < 0.618) { s = (float)(s * 0.618); continue; }
            s = s * scale;
            if (s > 100.) s = 100.f;
            if (s < 1. / 10.) s = (float)(1. / 10.);
            d[0] = b.x[0] - a.x[0]; d[1] = b.x[1] - a.x[1]; d[2] = b.x[2] - a.x[2];
            dL = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            len += dL;
            {
                const double bx[3] = {b.x[0], b.x[1], b.x[2]};
                flag = check_flag(n, bx);
            }
            if (flag > 0) {
                len -= dL;
                if (flag <= 6) {
                    int dim = (flag - 1) / 2;
                    float p1 = a.x[dim], p2 = b.x[dim];
                    if (fabsf(p1 - p2) > 0) {
                        float pm = (flag % 2 == 1) ? 0.0f : (float)n[dim];
                        float ba[3];
                        int c;
                        const float *pa = (const float *)&a, *pb = (const float *)&b;
                        float *pe = (float *)&ends[side];
                        for (c = 0; c < 9; ++c) pe[c] = (pa[c] * (p2 - pm) + pb[c] * (pm - p1)) / (p2 - p1);
                        interp_vec_f(f, a.x, ba, 1);   /* decision register: b at the last in-box state */
                        len += fabsf(pm - p1) / (1e-4 + fabsf(ba[dim]));
                    } else {
                        ends[side] = a;
                    }
                    flags[side] = flag;
                } else {
                    ends[side] = a;
                    flags[side] = 42;
                    aborted = 1;
                    abort_side = side;
                }
            }
            a = b;
            ++steps;
            if (aborted) break;
        }
        if (!aborted && flag == 0) {
            ends[side] = a;
            flags[side] = 0;
        }
    }
    if (aborted) {
        if (abort_side == 0) {
            ends[1].x[0] = p0[0]; ends[1].x[1] = p0[1]; ends[1].x[2] = p0[2];
            flags[1] = 42;
        }
        res->q0 = 0.0; res->qperp = 0.0;
    } else if (flags[0] == 0 || flags[1] == 0) {
        res->q0 = NAN; res->qperp = NAN;
    } else {
        /* Q assembly (TraceBlineScott.cu:475-508) in float; the 2.0 literal
         * makes the last term of each numerator double */
        float Bs[3], Be[3], us1[3], vs1[3], ue1[3], ve1[3], BNs, BNe, nbs, nbe, den;
        int ds = (flags[0] - 1) / 2, de = (flags[1] - 1) / 2, i;
        const float *us = ends[0].u, *vs = ends[0].v, *ue = ends[1].u, *ve = ends[1].v;
        double num;
        interp_vec_f(f, ends[0].x, Bs, 0);
        interp_vec_f(f, ends[1].x, Be, 0);
        BNs = Bs[ds]; BNe = Be[de];
        for (i = 0; i < 3; ++i) {
            us1[i] = us[i] - us[ds] / Bs[ds] * Bs[i];
            vs1[i] = vs[i] - vs[ds] / Bs[ds] * Bs[i];
            ue1[i] = ue[i] - ue[de] / Be[de] * Be[i];
            ve1[i] = ve[i] - ve[de] / Be[de] * Be[i];
        }
        num = (double)(dot3f(ue1, ue1) * dot3f(vs1, vs1) + dot3f(us1, us1) * dot3f(ve1, ve1)) -
              2.0 * dot3f(ue1, ve1) * dot3f(us1, vs1);
        den = fabsf((Nb0 * Nb0) / (BNs * BNe));
        res->q0 = (float)(fabs(num) / den);
        nbs = sqrtf(dot3f(Bs, Bs)); nbe = sqrtf(dot3f(Be, Be));
        {
            float due = dot3f(ue, Be) / nbe, dve = dot3f(ve, Be) / nbe;
            float dus = dot3f(us, Bs) / nbs, dvs = dot3f(vs, Bs) / nbs;
            for (i = 0; i < 3; ++i) {
                ue1[i] = ue[i] - due * (Be[i] / nbe);
                ve1[i] = ve[i] - dve * (Be[i] / nbe);
                us1[i] = us[i] - dus * (Bs[i] / nbs);
                vs1[i] = vs[i] - dvs * (Bs[i] / nbs);
            }
        }
        num = (double)(dot3f(ue1, ue1) * dot3f(vs1, vs1) + dot3f(us1, us1) * dot3f(ve1, ve1)) -
              2.0 * dot3f(ue1, ve1) * dot3f(us1, vs1);
        den = (Nb0 * Nb0) / (nbs * nbe);
        res->qperp = (float)(fabs(num) / den);
    }
    for (side = 0; side < 3; ++side) {
        res->start[side] = ends[0].x[side];
        res->end[side] = ends[1].x[side];
    }
    res->flag_start = flags[0];
    res->flag_end = flags[1];
    res->len = len;
    res->evals = evals;
    res->steps = steps;
}

/* ------------------------------------------------------- plain tracer */
/* RKF45 (TraceBlineAdaptive.cu:215-254) on the position only. */
static double plain_rkf45(const field_t *f, const double *p, double *out, double s) {
    double k[6][3], t[3];
    const double *A[6] = {NULL, A2, A3, A4, A5, A6};
    double e[3];
    int st, j, i;
    for (st = 0; st < 6; ++st) {
        for (i = 0; i < 3; ++i) {
            double acc = 0.0;
            for (j = 0; j < st; ++j) acc += A[st][j] * k[j][i];
            t[i] = p[i] + acc;
        }
        interp_vec(f, t, k[st], 1);
        for (i = 0; i < 3; ++i) k[st][i] *= s;
    }
    for (i = 0; i < 3; ++i) {
        double acc = 0.0;
        for (j = 0; j < 6; ++j) acc += B5[j] * k[j][i];
        out[i] = p[i] + acc;
        e[i] = CE[0] * k[0][i] + CE[2] * k[2][i] + CE[3] * k[3][i] + CE[4] * k[4][i] + CE[5] * k[5][i];
    }
    return len3(e);
}

/* RK4 (TraceBline.cu:226-234). */
static void plain_rk4(const field_t *f, const double *p, double *out, double s) {
    double k1[3], k2[3], k3[3], k4[3], t[3];
    int i;
    interp_vec(f, p, k1, 1);
    for (i = 0; i < 3; ++i) t[i] = p[i] + s * k1[i] / 2.;
    interp_vec(f, t, k2, 1);
    for (i = 0; i < 3; ++i) t[i] = p[i] + s * k2[i] / 2.;
    interp_vec(f, t, k3, 1);
    for (i = 0; i < 3; ++i) t[i] = p[i] + s * k3[i];
    interp_vec(f, t, k4, 1);
    for (i = 0; i < 3; ++i) out[i] = p[i] + 1. / 6. * s * (k1[i] + 2. * k2[i] + 2. * k3[i] + k4[i]);
}

/* RK4_boundary (TraceBline.cu:253-274): RK4 in the exit coordinate with
 * dx/dx_dim = b / b_dim, over the signed distance s to the face. */
static void rk4_boundary(const field_t *f, const double *p, double *out, double s, int dim) {
    double k1[3], k2[3], k3[3], k4[3], t[3], bd;
    int i;
    interp_vec(f, p, k1, 1); bd = k1[dim]; for (i = 0; i < 3; ++i) k1[i] /= bd;
    for (i = 0; i < 3; ++i) t[i] = p[i] + s * k1[i] / 2.;
    interp_vec(f, t, k2, 1); bd = k2[dim]; for (i = 0; i < 3; ++i) k2[i] /= bd;
    for (i = 0; i < 3; ++i) t[i] = p[i] + s * k2[i] / 2.;
    interp_vec(f, t, k3, 1); bd = k3[dim]; for (i = 0; i < 3; ++i) k3[i] /= bd;
    for (i = 0; i < 3; ++i) t[i] = p[i] + s * k3[i];
    interp_vec(f, t, k4, 1); bd = k4[dim]; for (i = 0; i < 3; ++i) k4[i] /= bd;
    for (i = 0; i < 3; ++i) out[i] = p[i] + (1. / 6.) * s * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

/* TraceBlineAdap (TraceBlineAdaptive.cu:333-426) / TraceBline
 * (TraceBline.cu:300-378), one direction; with cur, the twist
 * J.B/|B|^2/4pi integrated along the accepted steps (at each step's start
 * point, TraceBline.cu:325-329, TraceBlineAdaptive.cu:373-377). */
static void trace_plain_dir(const field_t *f, const field_t *cur, const double *p0, int method, double s,
                            double tol_coef, const double *ncross, double dir,
                            double *pout, int *flag_out, double *len_out, long *evals_out,
                            double *twist_out) {
    const int *n = f->n;
    double ratio = method == M_RKF45 ? RKF45_PLAIN_LEN_RATIO : RK4_PLAIN_LEN_RATIO;
    long step_lim = method == M_RKF45 ? RKF45_PLAIN_STEP_LIM : RK4_PLAIN_STEP_LIM;
    double len_lim = ratio * (n[0] + n[1] + n[2]);
    double p1[3], p2[3], b0[3], tol = 0.0, len = 0.0, twist = 0.0;
    long steps = 0, attempts = 0, evals = 0;
    int flag = 0;
    memcpy(p1, p0, sizeof(p1));
    memcpy(pout, p0, sizeof(p1));
    interp_vec(f, p0, b0, 1);
    if (method == M_RKF45) {                        /* TraceBlineAdaptive.cu:354-359 */
        double bn = fabs(dot3(b0, ncross));
        if (bn <= 0.05) tol = TOL_BASE / 8e3;
        else tol = TOL_BASE * pow(bn, 1.5);
        tol = tol * tol_coef;
    }
    while (flag == 0 && len < len_lim && steps < step_lim) {
        double d[3], dl;
        if (attempts >= FQSL_ATTEMPT_CAP) break;
        ++attempts;
        if (method == M_RKF45) {
            double err = plain_rkf45(f, p1, p2, s * dir);
            double scale = pow(tol / err / 11.09, 0.2);
            evals += 6;
            if (scale < 0.618) { s = s * 0.618; continue; }
            s = s * scale;
            if (s > 100.) s = 100.;
            if (s < 1. / 10.) s = 1. / 10.;
        } else {
            plain_rk4(f, p1, p2, s * dir);
            evals += 4;
        }
        d[0] = p1[0] - p2[0]; d[1] = p1[1] - p2[1]; d[2] = p1[2] - p2[2];
        dl = len3(d);
        len += dl;
        double Jp1[3] = {0.0, 0.0, 0.0}, tw_rate = 0.0;
        if (cur) {
            double Bp1[3];
            interp_vec(cur, p1, Jp1, 0);
            interp_vec(f, p1, Bp1, 0);
            tw_rate = dot3(Jp1, Bp1) / dot3(Bp1, Bp1) / 4.0 / M_PI;
            twist += tw_rate * dl;
        }
        flag = check_flag(n, p2);
        if (flag > 0) {
            len -= dl;
            if (cur) twist -= tw_rate * dl;             /* TraceBline.cu:334-336 */
            if (flag <= 6) {
                int dim = (flag - 1) / 2, i;
                double pa = p1[dim], pb = p2[dim];
                if (fabs(pa - pb) > 1e-3) {
                    double pm = (flag % 2 == 1) ? 0.0 : (double)n[dim];
                    double bp1[3], bp2[3];
                    interp_vec(f, p1, bp1, 1);
                    interp_vec(f, p2, bp2, 1);
                    if (fabs(bp1[dim]) < 0.2 || fabs(bp2[dim]) < 0.2) {
                        for (i = 0; i < 3; ++i) pout[i] = (p1[i] * (pb - pm) + p2[i] * (pm - pa)) / (pb - pa);
                    } else {
                        double q[3];
                        rk4_boundary(f, p1, q, pm - pa, dim);
                        evals += 4;
                        memcpy(pout, q, sizeof(q));
                    }
                    len += fabs(pm - pa) / (1e-4 + fabs(bp1[dim]));
                    /* RK4 only (the RKF45 tracer has this term commented out):
                     * the landing arc times J.b/|b|^2 with b the normalised B at
                     * p1, as the reference re-assigns B_P1 (TraceBline.cu:362-365) */
                    if (cur && method == M_RK4)
                        twist += dot3(Jp1, bp1) / dot3(bp1, bp1) / 4.0 / M_PI * fabs(pm - pa) / (1e-4 + fabs(bp1[dim]));
                } else {
                    /* no landing segment; the reference's p_mid is unset here:
                     * no twist is added (decision register) */
                    memcpy(pout, p1, sizeof(p1));
                }
            } else {
                memcpy(pout, p2, sizeof(p2));
            }
        }
        memcpy(p1, p2, sizeof(p1));
        ++steps;
    }
    if (flag == 0) memcpy(pout, p1, sizeof(p1));    /* limit exit: last position */
    *flag_out = flag;
    *len_out = len;
    *evals_out = evals;
    if (twist_out) *twist_out = twist;
}

/* ------------------------------------------------ stretched-grid tracer */
#define STRETCH_LEN_RATIO 2.0   /* MAX_STEP_RATIO, TraceBlineAdaptiveStretch.cu:11 */

/* TraceBlineAdap of the stretched tracer (TraceBlineAdaptiveStretch.cu:
 * 323-419), one direction, positions in the axes' coordinates: the RKF45
 * step and controller of the uniform tracer, the box [a0, a_{n-1}] per axis,
 * len_lim = 2 (sum of the three spans), no step-count limit, and a face
 * landing on the plane p_mid = 0 (min face) or n_dim (max face) as the
 * reference writes it (:385-386) -- or on the face itself (a0 / a_{n-1}) when
 * the axes use the exact lookup.  With cur, the twist of each step at its
 * start point and, unlike the uniform RKF45 tracer, the landing term
 * J.b/|b|^2 with b the normalised B at p1 (:403-406). */
static void trace_stretch_dir(const field_t *f, const field_t *cur, const double *p0, double s,
                              double tol_coef, const double *ncross, double dir,
                              double *pout, int *flag_out, double *len_out, long *evals_out,
                              double *twist_out) {
    const int *n = f->n;
    const axes_t *ax = f->ax;
    double lo[3], hi[3];
    double p1[3], p2[3], b0[3], tol, len = 0.0, twist = 0.0, len_lim;
    long attempts = 0, evals = 0;
    int flag = 0, a;
    for (a = 0; a < 3; ++a) { lo[a] = ax->a[a][0]; hi[a] = ax->a[a][n[a] - 1]; }
    len_lim = STRETCH_LEN_RATIO * ((hi[0] + hi[1] + hi[2]) - (lo[0] + lo[1] + lo[2]));
    memcpy(p1, p0, sizeof(p1));
    memcpy(pout, p0, sizeof(p1));
    interp_vec(f, p0, b0, 1);
    {                                                  /* :345-352 */
        double bn = fabs(dot3(b0, ncross));
        tol = (bn <= 0.05) ? TOL_BASE / 8e3 : TOL_BASE * pow(bn, 1.5);
        tol = tol * tol_coef;
    }
    while (flag == 0 && len < len_lim) {
        double d[3], dl, err, scale;
        if (attempts >= FQSL_ATTEMPT_CAP) break;
        ++attempts;
        err = plain_rkf45(f, p1, p2, s * dir);
        scale = pow(tol / err / 11.09, 0.2);
        evals += 6;
        if (scale < 0.618) { s = s * 0.618; continue; }
        s = s * scale;
        if (s > 100.) s = 100.;
        if (s < 1. / 10.) s = 1. / 10.;
        d[0] = p1[0] - p2[0]; d[1] = p1[1] - p2[1]; d[2] = p1[2] - p2[2];
        dl = len3(d);
        len += dl;
        double Jp1[3] = {0.0, 0.0, 0.0}, tw_rate = 0.0;
        if (cur) {
            double Bp1[3];
            interp_vec(cur, p1, Jp1, 0);
            interp_vec(f, p1, Bp1, 0);
            tw_rate = dot3(Jp1, Bp1) / dot3(Bp1, Bp1) / 4.0 / M_PI;
            twist += tw_rate * dl;
        }
        flag = check_flag_box(lo, hi, p2);
        if (flag > 0) {
            len -= dl;
            if (cur) twist -= tw_rate * dl;
            if (flag <= 6) {
                int dim = (flag - 1) / 2, i;
                double pa = p1[dim], pb = p2[dim];
                if (fabs(pa - pb) > 1e-3) {
                    double pm, bp1[3], bp2[3], arc;
                    if (ax->rule == LOOKUP_EXACT) pm = (flag % 2 == 1) ? lo[dim] : hi[dim];
                    else pm = (flag % 2 == 1) ? 0.0 : (double)n[dim];
                    interp_vec(f, p1, bp1, 1);
                    interp_vec(f, p2, bp2, 1);
                    if (fabs(bp1[dim]) < 0.2 || fabs(bp2[dim]) < 0.2) {
                        for (i = 0; i < 3; ++i) pout[i] = (p1[i] * (pb - pm) + p2[i] * (pm - pa)) / (pb - pa);
                    } else {
                        double q[3];
                        rk4_boundary(f, p1, q, pm - pa, dim);
                        evals += 4;
                        memcpy(pout, q, sizeof(q));
                    }
                    arc = fabs(pm - pa) / (1e-4 + fabs(bp1[dim]));
                    len += arc;
                    if (cur) twist += dot3(Jp1, bp1) / dot3(bp1, bp1) / 4.0 / M_PI * arc;
                } else {
                    /* no landing segment (p_mid unset in the reference): no
                     * landing twist */
                    memcpy(pout, p1, sizeof(p1));
                }
            } else {
                memcpy(pout, p2, sizeof(p2));
            }
        }
        memcpy(p1, p2, sizeof(p1));
    }
    if (flag == 0) memcpy(pout, p1, sizeof(p1));    /* limit exit: last position */
    *flag_out = flag;
    *len_out = len;
    *evals_out = evals;
    if (twist_out) *twist_out = twist;
}

/* ================================================================ C ABI */
static int nthreads_or_default(int nt) {
#ifdef _OPENMP
    return nt > 0 ? nt : omp_get_max_threads();
#else
    (void)nt;
    return 1;
#endif
}

int orc_num_threads(void) { return nthreads_or_default(0); }

/* Vector interpolation at n points (mode 0: Interp3dxyzn raw, 1: normalised,
 * 2: Interp3d clamped scalar rule). */
void orc_interp(const double *bx, const double *by, const double *bz, int nx, int ny, int nz,
                long npts, const double *pts, double *out, int mode) {
    field_t f = {bx, by, bz, {nx, ny, nz}, NULL};
    long i;
    for (i = 0; i < npts; ++i) {
        if (mode == 2) interp_clamped(&f, pts + 3 * i, out + 3 * i);
        else interp_vec(&f, pts + 3 * i, out + 3 * i, mode);
    }
}

/* grad b at n points; out[9*i + 3*j + c] = d b_c / d x_j.  mode 0: the
 * reference's difference quotient, 1: closed form inside the cell. */
void orc_grad_b(const double *bx, const double *by, const double *bz, int nx, int ny, int nz,
                long npts, const double *pts, double *out, int mode) {
    field_t f = {bx, by, bz, {nx, ny, nz}, NULL};
    long i;
    for (i = 0; i < npts; ++i) {
        double J[3][3];
        if (mode == 0) grad_b_fd(&f, pts + 3 * i, J); else grad_b_analytic(&f, pts + 3 * i, J);
        memcpy(out + 9 * i, J, sizeof(J));
    }
}

void orc_check_flag(int nx, int ny, int nz, long npts, const double *pts, int *out) {
    int n[3] = {nx, ny, nz};
    long i;
    for (i = 0; i < npts; ++i) out[i] = check_flag(n, pts + 3 * i);
}

/* Scott tracer over n seeds (TraceAllBlineScott, TraceBlineScott.cu:521-595).
 * method 0 = RKF45 (tol absolute, per step), 1 = RK4 fixed step s0.
 * grad_mode 0 = difference quotient (the reference rule), 1 = closed form. */
int orc_trace_scott(const double *bx, const double *by, const double *bz, int nx, int ny, int nz,
                    long nseeds, const double *seeds, int method, double s0, double tol,
                    int grad_mode, int nthreads,
                    double *start, double *end, int *flag_start, int *flag_end,
                    double *q0, double *qperp, double *len, double *bthis,
                    long *evals, long *steps,
                    const double *jx, const double *jy, const double *jz, double *twist) {
    field_t f = {bx, by, bz, {nx, ny, nz}, NULL};
    field_t cur = {jx, jy, jz, {nx, ny, nz}, NULL};
    long i;
    if (nx < 2 || ny < 2 || nz < 2) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads_or_default(nthreads))
    for (i = 0; i < nseeds; ++i) {
        scott_result r;
        trace_scott_line(&f, jx ? &cur : NULL, seeds + 3 * i, method, s0, tol, grad_mode, &r);
        if (twist) twist[i] = r.twist;
        memcpy(start + 3 * i, r.start, sizeof(r.start));
        memcpy(end + 3 * i, r.end, sizeof(r.end));
        flag_start[i] = r.flag_start;
        flag_end[i] = r.flag_end;
        q0[i] = r.q0;
        qperp[i] = r.qperp;
        len[i] = r.len;
        memcpy(bthis + 3 * i, r.bthis, sizeof(r.bthis));
        if (evals) evals[i] = r.evals;
        if (steps) steps[i] = r.steps;
    }
    return 0;
}

/* The Scott tracer in the reference kernel's float arithmetic
 * (trace_scott_line_f32), RKF45 only, over n seeds. */
int orc_trace_scott_f32(const double *bx, const double *by, const double *bz, int nx, int ny, int nz,
                        long nseeds, const double *seeds, double s0, double tol, int nthreads,
                        double *start, double *end, int *flag_start, int *flag_end,
                        double *q0, double *qperp, double *len, double *bthis, long *evals, long *steps) {
    field_t f = {bx, by, bz, {nx, ny, nz}, NULL};
    long i;
    if (nx < 2 || ny < 2 || nz < 2) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads_or_default(nthreads))
    for (i = 0; i < nseeds; ++i) {
        scott_result r;
        trace_scott_line_f32(&f, seeds + 3 * i, s0, tol, &r);
        memcpy(start + 3 * i, r.start, sizeof(r.start));
        memcpy(end + 3 * i, r.end, sizeof(r.end));
        flag_start[i] = r.flag_start;
        flag_end[i] = r.flag_end;
        q0[i] = r.q0;
        qperp[i] = r.qperp;
        len[i] = r.len;
        memcpy(bthis + 3 * i, r.bthis, sizeof(r.bthis));
        if (evals) evals[i] = r.evals;
        if (steps) steps[i] = r.steps;
    }
    return 0;
}

/* One integrator step of the Scott 9-vector from state y (x,u,v): method 0
 * = RKF45_Scott (returns the error estimate), 1 = RK4.  Used by the order
 * tests only. */
double orc_scott_step(const double *bx, const double *by, const double *bz, int nx, int ny, int nz,
                      const double *y, double *out, double s, int method, int grad_mode) {
    field_t f = {bx, by, bz, {nx, ny, nz}, NULL};
    state9 a, b;
    memcpy(&a, y, sizeof(a));
    if (method == M_RKF45) {
        double e = scott_rkf45(&f, &a, &b, s, grad_mode);
        memcpy(out, &b, sizeof(b));
        return e;
    }
    scott_rk4(&f, &a, &b, s, grad_mode);
    memcpy(out, &b, sizeof(b));
    return 0.0;
}

/* Plain tracer over n seeds (TraceAllBline, TraceBlineAdaptive.cu:429-510 and
 * TraceBline.cu:383-464): forward first, then backward, each with a fresh
 * step s0.  tol_coef is ignored by RK4. */
int orc_trace_plain(const double *bx, const double *by, const double *bz, int nx, int ny, int nz,
                    long nseeds, const double *seeds, const double *ncross, int method, double s0,
                    double tol_coef, int nthreads,
                    double *start, double *end, int *flag_start, int *flag_end,
                    double *bstart, double *bend, double *bthis, int *bflag,
                    double *len, long *evals,
                    const double *jx, const double *jy, const double *jz, double *twist) {
    field_t f = {bx, by, bz, {nx, ny, nz}, NULL};
    field_t curf = {jx, jy, jz, {nx, ny, nz}, NULL};
    const field_t *cur = jx ? &curf : NULL;
    long i;
    if (nx < 2 || ny < 2 || nz < 2) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads_or_default(nthreads))
    for (i = 0; i < nseeds; ++i) {
        const double *p0 = seeds + 3 * i;
        double lf, lb, bt[3], tf = 0.0, tb = 0.0;
        long ef, eb;
        trace_plain_dir(&f, cur, p0, method, s0, tol_coef, ncross, 1.0, end + 3 * i, &flag_end[i], &lf, &ef, &tf);
        interp_clamped(&f, end + 3 * i, bend + 3 * i);
        trace_plain_dir(&f, cur, p0, method, s0, tol_coef, ncross, -1.0, start + 3 * i, &flag_start[i], &lb, &eb, &tb);
        if (twist) twist[i] = tf + tb;
        interp_clamped(&f, start + 3 * i, bstart + 3 * i);
        len[i] = lf + lb;
        interp_clamped(&f, p0, bt);
        memcpy(bthis + 3 * i, bt, sizeof(bt));
        /* B_flag, TraceBlineAdaptive.cu:498-501 */
        bflag[i] = fabs(bt[0] * ncross[0] + bt[1] * ncross[1] + bt[2] * ncross[2]) * 100. < len3(bt) ? 1 : 0;
        if (evals) evals[i] = ef + eb;
    }
    return 0;
}

/* Cell index and weight along one stretched axis at n points
 * (Interp3dxyzn's lookup, TraceBlineAdaptiveStretch.cu:82-124). */
void orc_stretch_cell(const double *a, int n, int rule, long npts, const double *p, int *cell, double *w) {
    axes_t ax = {{a, a, a}, rule};
    long i;
    for (i = 0; i < npts; ++i) stretch_cell(&ax, 0, p[i], n, cell + i, w + i);
}

/* Raw or normalised B on a stretched grid at n points. */
void orc_interp_stretch(const double *bx, const double *by, const double *bz, int nx, int ny, int nz,
                        const double *xa, const double *ya, const double *za, int rule,
                        long npts, const double *pts, double *out, int normalize) {
    axes_t ax = {{xa, ya, za}, rule};
    field_t f = {bx, by, bz, {nx, ny, nz}, &ax};
    long i;
    for (i = 0; i < npts; ++i) interp_vec(&f, pts + 3 * i, out + 3 * i, normalize);
}

/* Stretched-grid RKF45 tracer over n seeds (TraceAllBline,
 * TraceBlineAdaptiveStretch.cu:422-508): forward first, then backward, each
 * from s0; B_start / B_end / B_this by the stretched Interp3dxyzn (raw). */
int orc_trace_stretch(const double *bx, const double *by, const double *bz, int nx, int ny, int nz,
                      const double *xa, const double *ya, const double *za, int rule,
                      long nseeds, const double *seeds, const double *ncross, double s0,
                      double tol_coef, int nthreads,
                      double *start, double *end, int *flag_start, int *flag_end,
                      double *bstart, double *bend, double *bthis, int *bflag,
                      double *len, long *evals,
                      const double *jx, const double *jy, const double *jz, double *twist) {
    axes_t ax = {{xa, ya, za}, rule};
    field_t f = {bx, by, bz, {nx, ny, nz}, &ax};
    field_t curf = {jx, jy, jz, {nx, ny, nz}, &ax};
    const field_t *cur = jx ? &curf : NULL;
    long i;
    if (nx < 2 || ny < 2 || nz < 2) return -1;
    if (rule < LOOKUP_BISECT || rule > LOOKUP_EXACT) return -2;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads_or_default(nthreads))
    for (i = 0; i < nseeds; ++i) {
        const double *p0 = seeds + 3 * i;
        double lf, lb, bt[3], tf = 0.0, tb = 0.0;
        long ef, eb;
        trace_stretch_dir(&f, cur, p0, s0, tol_coef, ncross, 1.0, end + 3 * i, &flag_end[i], &lf, &ef, &tf);
        interp_vec(&f, end + 3 * i, bend + 3 * i, 0);
        trace_stretch_dir(&f, cur, p0, s0, tol_coef, ncross, -1.0, start + 3 * i, &flag_start[i], &lb, &eb, &tb);
        interp_vec(&f, start + 3 * i, bstart + 3 * i, 0);
        if (twist) twist[i] = tf + tb;
        len[i] = lf + lb;
        interp_vec(&f, p0, bt, 0);
        memcpy(bthis + 3 * i, bt, sizeof(bt));
        bflag[i] = fabs(bt[0] * ncross[0] + bt[1] * ncross[1] + bt[2] * ncross[2]) * 100. < len3(bt) ? 1 : 0;
        if (evals) evals[i] = ef + eb;
    }
    return 0;
}
