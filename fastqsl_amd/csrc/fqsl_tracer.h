// fqsl_tracer.h -- per-line field-line tracing math for gfx950.
//
// One struct per traced line holds its whole state; `attempt()` performs one
// integrator attempt (step, controller, exit test, face landing).  The same
// code runs under two schedulers in fqsl_kernels.hip: a persistent kernel in
// which each lane of a wavefront refills itself from a device work queue, and
// a one-thread-per-line kernel.
//
// Precision variants (template parameter R):
//   float  -- field, stages, Jacobian products in FP32; the position is a
//             compensated (hi, lo) float pair and every stage point is formed
//             relative to its cell so interpolation weights keep ~1e-7 cell
//             accuracy at any distance from the origin; arc length, the RKF45
//             error norm, the step controller, landing and Q assembly in FP64.
//   double -- everything in FP64.
// The Scott tracer has a third, mixed variant (R = float, X = double): the
// position path -- b = B/|B| at the stage points, the stage positions, the
// increments and the RKF45 error norm -- in FP64 from the float cell
// coefficients, the Jacobian path (grad b, the deviation vectors u, v) in
// FP32.  Its error estimate has no FP32 round-off floor, so its step-size
// sequence follows the FP64 controller's.
//
// Reference behaviour reproduced here (file:line in peijin94/FastQSL):
//   field indexing Arr[k*ny*nx + j*nx + i]         TraceBlineScott.cu:36-39
//   Interp3dxyzn clamp rule (<=0 -> cell 0, w 0;
//     >=N-1 -> cell N-2, w 1)                     TraceBlineScott.cu:120-131
//   Interp3d clamp rule (into [1e-4, N-1-1e-4])    TraceBlineScott.cu:57-62
//   grad_unit_vec_B difference quotient, h = 1e-3  TraceBlineScott.cu:204-226
//   f_scott right-hand side                        TraceBlineScott.cu:230-249
//   RKF45_Scott (Fehlberg, 5th-order advance,
//     position-only error)                         TraceBlineScott.cu:256-290
//   step controller                                TraceBlineScott.cu:410-419
//   checkFlag                                      TraceBlineScott.cu:310-332
//   TraceBlineScott driver and face landing        TraceBlineScott.cu:334-471
//   Q assembly (q_0, q_perp)                       TraceBlineScott.cu:475-508
//   plain RKF45 / RK4 / RK4_boundary / driver      TraceBlineAdaptive.cu:215-426,
//                                                  TraceBline.cu:226-378
//
// The difference quotient of the reference, (b(x+h e_j) - b(x-h e_j)) / 2h on
// the clamped piecewise-trilinear interpolant, is evaluated exactly but from
// one 8-corner stencil: along each axis the interpolant is piecewise linear,
// so B(x +- h e_j) = B(x) + a+ and B(x) - a- with a+, a- sums of the cell's
// edge derivative over the covered length (plus the neighbour cell's when the
// +-h sample crosses a grid plane, zero outside the box).  The quotient is
// then formed without cancellation from D = a+ + a- and S = B+ + B-:
//   b+ - b- = D (1/n+ + 1/n-)/2 - S (S.D) / (2 n+ n- (n+ + n-)).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#define FQSL_HD __host__ __device__ __forceinline__
#define FQSL_COLD __host__ __device__ __attribute__((noinline))

// Address spaces for the device compiler: the field and lookup tables live
// in global memory (global, not flat, loads), and the small per-grid
// parameter block is read-only for a kernel's lifetime (constant: uniform
// reads become scalar loads).  Host builds see plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define FQSL_GLOBAL __attribute__((address_space(1)))
#define FQSL_CONST __attribute__((address_space(4)))
#else
#define FQSL_GLOBAL
#define FQSL_CONST
#endif

namespace fqsl {

// ---------------------------------------------------------------- constants
constexpr double kFdH = 1e-3;                // TraceBlineScott.cu:210
constexpr int kScottStepLim = 10000;         // N_STEP_LIM, TraceBlineScott.cu:13
constexpr double kScottLenRatio = 2.0;       // MAX_STEP_RATIO, TraceBlineScott.cu:11
constexpr int kPlainRkf45StepLim = 500000;   // MAX_NUM_STEP, TraceBlineAdaptive.cu:13
constexpr double kPlainRkf45LenRatio = 3.0;  // MAX_STEP_RATIO, TraceBlineAdaptive.cu:11
constexpr int kPlainRk4StepLim = 100000;     // MAX_NUM_STEP, TraceBline.cu:12
constexpr double kPlainRk4LenRatio = 2.0;    // MAX_STEP_RATIO, TraceBline.cu:10
constexpr double kTolBase = 1e-3;            // TOL, TraceBlineAdaptive.cu:12
constexpr int kAttemptCap = 1000000;         // hard bound on integrator attempts per line
// Smallest RKF45 tolerance (cells) the FP32 path accepts.  Its error norm is
// formed from float stages, whose round-off puts a floor of ~1e-8 x step
// under the estimate; down to 1e-9 the controller still steers by the true
// error (on a 256^3 multipole it keeps the oracle's accuracy with fewer
// evaluations), below it the floor pins the step near its lower clamp.
constexpr double kFp32cMinTol = 1e-9;

enum Method : int { kRKF45 = 0, kRK4 = 1 };

// class codes written to flag_start / flag_end
constexpr int kClassLimit = 0;
constexpr int kClassBelow = -1;
constexpr int kClassOther = 42;

// Fehlberg 4(5), TraceBlineScott.cu:267-275
struct Tab {
    static constexpr double a21 = 1. / 4.;
    static constexpr double a31 = 3. / 32., a32 = 9. / 32.;
    static constexpr double a41 = 1932. / 2197., a42 = -7200. / 2197., a43 = 7296. / 2197.;
    static constexpr double a51 = 439. / 216., a52 = -8., a53 = 3680. / 513., a54 = -845. / 4104.;
    static constexpr double a61 = -8. / 27., a62 = 2., a63 = -3544. / 2565., a64 = 1859. / 4104.,
                            a65 = -11. / 40.;
    static constexpr double b1 = 16. / 135., b3 = 6656. / 12825., b4 = 28561. / 56430.,
                            b5 = -9. / 50., b6 = 2. / 55.;
    static constexpr double e1 = 1. / 360., e3 = -128. / 4275., e4 = -2197. / 75240.,
                            e5 = 1. / 50., e6 = 2. / 55.;
};

// ------------------------------------------------------------ scalar helpers
// Hardware reciprocal square root / reciprocal (v_rsq_f32, v_rcp_f32: about
// 1 ulp) on the device.
FQSL_HD float rsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
FQSL_HD double rsq(double x) { return 1.0 / sqrt(x); }
FQSL_HD float rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
FQSL_HD double rcp(double x) { return 1.0 / x; }
FQSL_HD float flr(float x) { return floorf(x); }
FQSL_HD double flr(double x) { return floor(x); }

// float -> int that keeps NaN and huge values in range (value semantics are
// restored by the NaN branch of axis_setup)
FQSL_HD int to_int(float v) { return (int)fminf(fmaxf(v, -1.0e8f), 1.0e8f); }
FQSL_HD int to_int(double v) { return (int)fmin(fmax(v, -1.0e8), 1.0e8); }

template <typename R> FQSL_HD R dot3(const R* a, const R* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ----------------------------------------------------------------- fields
// Plain SoA layout, x fastest (the reference's layout).
template <typename T>
struct FieldSoA {
    static constexpr bool kCoef = false;  // load() yields corner values
    const T* bx;
    const T* by;
    const T* bz;
    int n[3];
    // c[comp][a + 2b + 4c] = B_comp at node (i+a, j+b, k+c)
    template <typename R>
    FQSL_HD void load(int i, int j, int k, R c[3][8]) const {
        const long nx = n[0], nxy = (long)n[0] * n[1];
        const long base = (long)k * nxy + (long)j * nx + i;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long o = base + (q & 1) + ((q >> 1) & 1) * nx + ((q >> 2) & 1) * nxy;
            c[0][q] = (R)((const FQSL_GLOBAL T*)bx)[o];
            c[1][q] = (R)((const FQSL_GLOBAL T*)by)[o];
            c[2][q] = (R)((const FQSL_GLOBAL T*)bz)[o];
        }
    }
};

// Trilinear coefficients of one cell from its corners v[a + 2b + 4c]:
//   B(x, y, z) = c0 + c1 x + c2 y + c3 z + c4 xy + c5 yz + c6 xz + c7 xyz
// for (x, y, z) in [0, 1]^3, formed in double and rounded once.
template <typename T>
FQSL_HD void cell_coefs(const T* v, float* c) {
    const double v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3], v4 = v[4], v5 = v[5], v6 = v[6], v7 = v[7];
    c[0] = (float)v0;
    c[1] = (float)(v1 - v0);
    c[2] = (float)(v2 - v0);
    c[3] = (float)(v4 - v0);
    c[4] = (float)((v3 - v2) - (v1 - v0));
    c[5] = (float)((v6 - v4) - (v2 - v0));
    c[6] = (float)((v5 - v4) - (v1 - v0));
    c[7] = (float)(((v7 - v6) - (v5 - v4)) - ((v3 - v2) - (v1 - v0)));
}

// Cell layout: each cell (i, j, k), i < nx-1 ..., stores the trilinear
// coefficients c0..c7 of its three components contiguously as 8 float4
// (128 B, one cache line): [Bx c0-3][Bx c4-7][By c0-3][By c4-7][Bz c0-3]
// [Bz c4-7][pad][pad].  One interpolation is six 16-byte loads from one line,
// and the coefficient form gives B and its three edge derivatives in 11
// FMAs per component (the corner form needs ~22).
struct FieldCell32 {
    static constexpr bool kCoef = true;  // load() yields coefficients
    const float4* cells;
    int n[3];
    // first float of cell (i, j, k): 3 components x 8 coefficients, then padding
    FQSL_HD const FQSL_GLOBAL float* cell(int i, int j, int k) const {
        const long cx = n[0] - 1, cxy = (long)(n[0] - 1) * (n[1] - 1);
        return (const FQSL_GLOBAL float*)(cells + 8 * ((long)k * cxy + (long)j * cx + i));
    }
    template <typename R>
    FQSL_HD void load(int i, int j, int k, R c[3][8]) const {
        const FQSL_GLOBAL float4* p = (const FQSL_GLOBAL float4*)cell(i, j, k);
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            const float4 lo = p[2 * comp], hi = p[2 * comp + 1];
            c[comp][0] = lo.x; c[comp][1] = lo.y; c[comp][2] = lo.z; c[comp][3] = lo.w;
            c[comp][4] = hi.x; c[comp][5] = hi.y; c[comp][6] = hi.z; c[comp][7] = hi.w;
        }
    }
};

// ------------------------------------------------------ axis / stencil setup
// Per-axis view of the point x = i + r (i = floor(x)): the interpolation cell
// c and weight w after the Interp3dxyzn clamp rule, and the lengths that make
// up a+ = B(x+h) - B(x) = ap d(c) + apn d(c+1) and
//     a- = B(x) - B(x-h) = am d(c) + amp d(c-1)
// on the clamped piecewise-linear interpolant.
template <typename R>
struct Axis {
    int c;
    R w, ap, apn, am, amp;
};

template <typename R>
FQSL_HD Axis<R> axis_setup(int i, R r, int n) {
    const R h = (R)kFdH;
    Axis<R> a;
    a.apn = 0;
    a.amp = 0;
    if (r != r) {  // NaN position: the value stays NaN
        a.c = 0; a.w = r; a.ap = h; a.am = h;
        return a;
    }
    if (i < 0 || (i == 0 && r == (R)0)) {  // x <= 0: node-0 plane
        a.c = 0; a.w = 0; a.am = 0;
        if (i == 0) a.ap = h;                                  // x == 0
        else if (i == -1 && r > (R)1 - h) a.ap = (r - (R)1) + h;  // -h < x < 0
        else a.ap = 0;                                         // x <= -h
        return a;
    }
    if (i >= n - 1) {  // x >= N-1: node-(N-1) plane
        a.c = n - 2; a.w = 1; a.ap = 0;
        const R e = (R)(i - (n - 1)) + r;  // x - (N-1)
        a.am = (e < h) ? (h - e) : (R)0;
        return a;
    }
    a.c = i; a.w = r;  // 0 < x < N-1
    if (r >= h) a.am = h;
    else { a.am = r; if (i >= 1) a.amp = h - r; }
    if (r + h <= (R)1) a.ap = h;
    else { a.ap = (R)1 - r; if (i + 1 <= n - 2) a.apn = (r + h) - (R)1; }
    return a;
}

// axis_setup as selects, no branches: the same values bit for bit (every
// length is the same expression on the same operands), for code where only a
// few lanes of a wavefront need it and a branch tree would cost each of them
// the exec-mask bookkeeping of every arm.
template <typename R>
FQSL_HD Axis<R> axis_setup_sel(int i, R r, int n) {
    const R h = (R)kFdH;
    const bool nan = r != r;
    const bool low = i < 0 || (i == 0 && r == (R)0);  // x <= 0 (NaN positions have i < 0)
    const bool high = i >= n - 1;                      // x >= N-1
    const bool mid = !nan && !low && !high;
    // 0 < x < N-1
    const R am_mid = (r >= h) ? h : r;
    const R amp_mid = (!(r >= h) && i >= 1) ? h - r : (R)0;
    const bool fits = r + h <= (R)1;
    const R ap_mid = fits ? h : (R)1 - r;
    const R apn_mid = (!fits && i + 1 <= n - 2) ? (r + h) - (R)1 : (R)0;
    // x <= 0 and x >= N-1
    const R ap_low = (i == 0) ? h : ((i == -1 && r > (R)1 - h) ? (r - (R)1) + h : (R)0);
    const R e = (R)(i - (n - 1)) + r;
    const R am_high = (e < h) ? (h - e) : (R)0;
    Axis<R> a;
    a.c = (nan || low) ? 0 : (high ? n - 2 : i);
    a.w = (nan || mid) ? r : (low ? (R)0 : (R)1);
    a.ap = nan ? h : (low ? ap_low : (high ? (R)0 : ap_mid));
    a.am = nan ? h : (low ? (R)0 : (high ? am_high : am_mid));
    a.apn = mid ? apn_mid : (R)0;
    a.amp = mid ? amp_mid : (R)0;
    return a;
}

// Trilinear value and the three edge derivatives (B(c+e_j) plane minus B(c)
// plane, bilinear in the other two axes) of a loaded cell.
template <typename R>
FQSL_HD void cell_eval(const R c[3][8], R wx, R wy, R wz, R B[3], R d[3][3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const R dx00 = c[q][1] - c[q][0], dx10 = c[q][3] - c[q][2];
        const R dx01 = c[q][5] - c[q][4], dx11 = c[q][7] - c[q][6];
        const R e00 = c[q][0] + wx * dx00, e10 = c[q][2] + wx * dx10;
        const R e01 = c[q][4] + wx * dx01, e11 = c[q][6] + wx * dx11;
        const R dy0 = e10 - e00, dy1 = e11 - e01;
        const R f0 = e00 + wy * dy0, f1 = e01 + wy * dy1;
        const R gx0 = dx00 + wy * (dx10 - dx00), gx1 = dx01 + wy * (dx11 - dx01);
        const R dz = f1 - f0;
        B[q] = f0 + wz * dz;
        d[2][q] = dz;
        d[1][q] = dy0 + wz * (dy1 - dy0);
        d[0][q] = gx0 + wz * (gx1 - gx0);
    }
}

// The same from trilinear coefficients (FieldCell32): with t = c1 + c4 y,
// u = c6 + c7 y, A = c0 + c2 y + x t, C = c3 + c5 y + x u:
//   B = A + z C,  dB/dz = C,  dB/dx = t + z u,  dB/dy = (c2 + c4 x) + z (c5 + c7 x).
template <typename R>
FQSL_HD void coef_eval(const R c[3][8], R x, R y, R z, R B[3], R d[3][3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const R t = c[q][1] + c[q][4] * y, u = c[q][6] + c[q][7] * y;
        const R A = (c[q][0] + c[q][2] * y) + x * t;
        const R C = (c[q][3] + c[q][5] * y) + x * u;
        B[q] = A + z * C;
        d[2][q] = C;
        d[0][q] = t + z * u;
        d[1][q] = (c[q][2] + c[q][4] * x) + z * (c[q][5] + c[q][7] * x);
    }
}

template <typename R>
FQSL_HD void coef_value(const R c[3][8], R x, R y, R z, R B[3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const R A = (c[q][0] + c[q][2] * y) + x * (c[q][1] + c[q][4] * y);
        const R C = (c[q][3] + c[q][5] * y) + x * (c[q][6] + c[q][7] * y);
        B[q] = A + z * C;
    }
}

// Value of the trilinear interpolant only.
template <typename R>
FQSL_HD void cell_value(const R c[3][8], R wx, R wy, R wz, R B[3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const R e00 = c[q][0] + wx * (c[q][1] - c[q][0]), e10 = c[q][2] + wx * (c[q][3] - c[q][2]);
        const R e01 = c[q][4] + wx * (c[q][5] - c[q][4]), e11 = c[q][6] + wx * (c[q][7] - c[q][6]);
        const R f0 = e00 + wy * (e10 - e00), f1 = e01 + wy * (e11 - e01);
        B[q] = f0 + wz * (f1 - f0);
    }
}

// Edge derivative along axis J only (the d[J] of cell_eval / coef_eval).
template <int J, typename R>
FQSL_HD void coef_edge(const R c[3][8], R x, R y, R z, R d[3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (J == 0) d[q] = (c[q][1] + c[q][4] * y) + z * (c[q][6] + c[q][7] * y);
        else if (J == 1) d[q] = (c[q][2] + c[q][4] * x) + z * (c[q][5] + c[q][7] * x);
        else d[q] = (c[q][3] + c[q][5] * y) + x * (c[q][6] + c[q][7] * y);
    }
}

template <int J, typename R>
FQSL_HD void cell_edge(const R c[3][8], R wx, R wy, R wz, R d[3]) {
    // corner bit of axis J and the two other axes (u, v) with their weights
    constexpr int bj = 1 << J, bu = J == 0 ? 2 : 1, bv = J == 2 ? 2 : 4;
    const R wu = J == 0 ? wy : wx, wv = J == 2 ? wy : wz;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const R e00 = c[q][bj] - c[q][0], e10 = c[q][bj + bu] - c[q][bu];
        const R e01 = c[q][bj + bv] - c[q][bv], e11 = c[q][bj + bu + bv] - c[q][bu + bv];
        const R f0 = e00 + wu * (e10 - e00), f1 = e01 + wu * (e11 - e01);
        d[q] = f0 + wv * (f1 - f0);
    }
}

// Dispatch on the field view's cell representation.
template <class F, typename R>
FQSL_HD void field_eval(const R c[3][8], R wx, R wy, R wz, R B[3], R d[3][3]) {
    if constexpr (F::kCoef) coef_eval<R>(c, wx, wy, wz, B, d);
    else cell_eval<R>(c, wx, wy, wz, B, d);
}

template <class F, typename R>
FQSL_HD void field_value(const R c[3][8], R wx, R wy, R wz, R B[3]) {
    if constexpr (F::kCoef) coef_value<R>(c, wx, wy, wz, B);
    else cell_value<R>(c, wx, wy, wz, B);
}

template <class F, int J, typename R>
FQSL_HD void field_edge(const R c[3][8], R wx, R wy, R wz, R d[3]) {
    if constexpr (F::kCoef) coef_edge<J, R>(c, wx, wy, wz, d);
    else cell_edge<J, R>(c, wx, wy, wz, d);
}

// (b(x+h e_j) - b(x-h e_j)) / 2h from B and the covered increments.
template <typename R>
FQSL_HD void fd_quotient(const R B[3], const R ap[3], const R am[3], R out[3]) {
    R Bp[3], Bm[3], D[3], S[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        Bp[q] = B[q] + ap[q];
        Bm[q] = B[q] - am[q];
        D[q] = ap[q] + am[q];
        S[q] = Bp[q] + Bm[q];
    }
    const R np2 = dot3(Bp, Bp), nm2 = dot3(Bm, Bm);
    const R ip = rsq(np2), im = rsq(nm2);
    const R np = np2 * ip, nm = nm2 * im;
    const R inv2h = (R)(0.5 / kFdH);
    const R t1 = (R)0.5 * (ip + im) * inv2h;
    const R t2 = (R)0.5 * dot3(S, D) * ip * im * rcp(np + nm) * inv2h;
#pragma unroll
    for (int q = 0; q < 3; ++q) out[q] = D[q] * t1 - S[q] * t2;
}

// True when `v` holds on some active lane of the wavefront (on the host: v).
FQSL_HD bool wave_any(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __any(v) != 0;
#else
    return v;
#endif
}

// Raw stage position: x = i + r per axis (i = floor(x)).
template <typename R>
struct Pt {
    int i[3];
    R r[3];
};

template <typename R>
FQSL_HD Pt<R> pt_from_double(const double* x) {
    Pt<R> p;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double f = floor(x[a]);
        p.i[a] = to_int(f);
        p.r[a] = (R)(x[a] - f);
    }
    return p;
}

// Neighbour-cell data fetched together with a stage point's own cell, for
// the coefficient-cell float path: the first axis (lowest index) whose +-h
// sample crosses a grid plane -- the same test axis_setup applies, apn != 0 or
// amp != 0 -- and, per component, the four coefficients that the edge
// derivative along that axis reads in the cell across the plane:
//   d_J = (kA + kB s) + t (kC + kD s)
// with (A, B, C, D; s, t) = (1, 4, 6, 7; y, z), (2, 4, 5, 7; x, z),
// (3, 5, 6, 7; y, x) for J = 0, 1, 2 (coef_edge).  Issued before the own
// cell is consumed, so its latency overlaps that load's instead of following
// it; an axis that crosses when a lower one already does loads in its own
// branch.
struct NbCoef {
    int j;  // axis, or -1
    float k[3][4];
};

// Stand-in for the paths that load their crossing neighbours in the branch.
struct NbNone {
    static constexpr int j = -1;
    float k[3][4];
};

template <class F>
FQSL_HD void nb_fetch(const F& f, const Pt<float>& p, const int* c, NbCoef& nb) {
    const float h = (float)kFdH;
    nb.j = -1;
    int step = 0;
#pragma unroll
    for (int a = 2; a >= 0; --a) {
        const int i = p.i[a], n = f.n[a];
        const float r = p.r[a];
        const bool mid = !(r != r) && !(i < 0 || (i == 0 && r == 0.f)) && !(i >= n - 1);
        const bool up = mid && !(r + h <= 1.f) && i + 1 <= n - 2;
        const bool dn = mid && !(r >= h) && i >= 1;
        if (up || dn) { nb.j = a; step = up ? 1 : -1; }
    }
    if (nb.j >= 0) {
        int ci[3] = {c[0], c[1], c[2]};
        ci[nb.j] += step;
        const FQSL_GLOBAL float* q = f.cell(ci[0], ci[1], ci[2]);
        const int kA = 1 + nb.j, kB = nb.j == 2 ? 5 : 4, kC = nb.j == 1 ? 5 : 6;
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            nb.k[comp][0] = q[8 * comp + kA];
            nb.k[comp][1] = q[8 * comp + kB];
            nb.k[comp][2] = q[8 * comp + kC];
            nb.k[comp][3] = q[8 * comp + 7];
        }
    }
}

// Row J of grad b at p from the cell's B and edge derivative dj along J:
// the closed form on a simple axis, else the exact difference quotient, which
// takes the neighbour cell's edge derivative where the +-h sample crosses a
// grid plane (crossing on both sides at once would need 2h > 1 cell).
template <int J, typename R, class F, class Nb>
FQSL_HD void scott_axis(const F& f, const Pt<R>& p, const int* c, const R* w, bool simple, bool flat,
                        const R B[3], const R b[3], R inb, const R dj[3], const Nb& nb, R out[3]) {
    if (simple) {
        const R bd = dot3(b, dj);
        const R g = flat ? (R)0 : inb;
#pragma unroll
        for (int q = 0; q < 3; ++q) out[q] = (dj[q] - b[q] * bd) * g;
        return;
    }
    constexpr bool kPre = std::is_same<Nb, NbCoef>::value;
    const Axis<R> ax = kPre ? axis_setup_sel<R>(p.i[J], p.r[J], f.n[J]) : axis_setup<R>(p.i[J], p.r[J], f.n[J]);
    R ap[3], am[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ap[q] = ax.ap * dj[q];
        am[q] = ax.am * dj[q];
    }
    if (ax.apn != (R)0 || ax.amp != (R)0) {  // the +-h sample crosses a grid plane
        const bool up = ax.apn != (R)0;
        const R coef = up ? ax.apn : ax.amp;
        R dn[3];
        if (kPre && nb.j == J) {
            const R s = J == 1 ? w[0] : w[1], t = J == 2 ? w[0] : w[2];
#pragma unroll
            for (int q = 0; q < 3; ++q) dn[q] = (nb.k[q][0] + nb.k[q][1] * s) + t * (nb.k[q][2] + nb.k[q][3] * s);
        } else {
            int ci[3] = {c[0], c[1], c[2]};
            ci[J] += up ? 1 : -1;
            R cn[3][8];
            f.template load<R>(ci[0], ci[1], ci[2], cn);
            field_edge<F, J, R>(cn, w[0], w[1], w[2], dn);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (up) ap[q] += coef * dn[q];
            else am[q] += coef * dn[q];
        }
    }
    fd_quotient<R>(B, ap, am, out);
}

// Scott right-hand side at p: b = B/|B| and J[j][q] = (difference quotient of
// b_q along axis j), i.e. grad_unit_vec_B (TraceBlineScott.cu:204-226).
//
// kExact (the FP64 path) evaluates the reference's quotient exactly on every
// axis.  Otherwise an axis whose +-h samples both stay inside the current cell
// uses the closed-form derivative (d - b (b.d)) / |B| of the same interpolant:
// the central quotient differs from it by (h^2/6) times the third derivative
// of b, at most ~2e-7 relative for a field whose scale length is >= 1 cell,
// below float resolution.  Axes whose samples cross a grid plane or the box
// face keep the exact quotient, which there blends two cells or halves the
// one-sided slope; an axis whose samples both lie beyond one face has a zero
// row (the clamped interpolant is constant there).
template <typename R, class F>
FQSL_HD void scott_eval(const F& f, const Pt<R>& p, R b[3], R J[3][3]) {
    constexpr bool kExact = sizeof(R) == 8;
    const R h = (R)kFdH;
    int c[3];
    R w[3];
    bool simple[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int i = p.i[a], n = f.n[a];
        const R r = p.r[a];
        const bool low = i < 0 || (i == 0 && r == (R)0);  // x <= 0
        const bool high = i >= n - 1;                      // x >= N-1
        c[a] = low ? 0 : (high ? n - 2 : i);
        w[a] = (r != r) ? r : (low ? (R)0 : (high ? (R)1 : r));
        simple[a] = !kExact && !low && !high && r >= h && r + h <= (R)1;
    }
    R cc[3][8];
    f.template load<R>(c[0], c[1], c[2], cc);
    constexpr bool kPre = F::kCoef && !kExact;
    typename std::conditional<kPre, NbCoef, NbNone>::type nb;
    if constexpr (kPre) nb.j = -1;
    bool flat[3] = {false, false, false};
    if (!kExact && wave_any(!(simple[0] && simple[1] && simple[2]))) {
        // An axis whose +-h samples both lie beyond one face of the box reads
        // the same clamped value on both sides: its quotient row is zero.
        // Such a lane takes the closed form scaled by zero instead of the
        // quotient (a stage point outside the box, as in a line's last
        // attempt); only the sign of a zero entry can differ.
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int i = p.i[a], n = f.n[a];
            const R r = p.r[a];
            const bool below = i < -1 || (i == -1 && !(r > (R)1 - h));      // x <= -h
            const bool above = i >= n - 1 && !(((R)(i - (n - 1)) + r) < h);  // x >= N-1+h
            flat[a] = r == r && (below || above);
            simple[a] = simple[a] || flat[a];
        }
        // only a lane with a non-simple axis can cross; most stages of a
        // wavefront have none and skip the crossing test
        if constexpr (kPre) nb_fetch<F>(f, p, c, nb);
    }
    R B[3], d[3][3];
    field_eval<F, R>(cc, w[0], w[1], w[2], B, d);
    const R inb = rsq(dot3(B, B));
#pragma unroll
    for (int q = 0; q < 3; ++q) b[q] = B[q] * inb;
    scott_axis<0, R, F>(f, p, c, w, simple[0], flat[0], B, b, inb, d[0], nb, J[0]);
    scott_axis<1, R, F>(f, p, c, w, simple[1], flat[1], B, b, inb, d[1], nb, J[1]);
    scott_axis<2, R, F>(f, p, c, w, simple[2], flat[2], B, b, inb, d[2], nb, J[2]);
}

// The mixed variant's right-hand side at a double stage point: b in FP64
// (the trilinear value from the cell's float data, evaluated in double and
// normalised in double) and grad b in FP32 as scott_eval's float path takes
// it (closed form on simple axes, the exact difference quotient elsewhere).
template <class F>
FQSL_HD void scott_eval_mixed(const F& f, const Pt<double>& pd, double bd[3], float J[3][3]) {
    const double h = kFdH;
    int c[3];
    double wd[3];
    float w[3];
    bool simple[3];
    Pt<float> p;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int i = pd.i[a], n = f.n[a];
        const double r = pd.r[a];
        const bool low = i < 0 || (i == 0 && r == 0.0);  // x <= 0
        const bool high = i >= n - 1;                     // x >= N-1
        c[a] = low ? 0 : (high ? n - 2 : i);
        wd[a] = (r != r) ? r : (low ? 0.0 : (high ? 1.0 : r));
        w[a] = (float)wd[a];
        simple[a] = !low && !high && r >= h && r + h <= 1.0;
        p.i[a] = i;
        p.r[a] = (float)r;
    }
    float cc[3][8];
    f.template load<float>(c[0], c[1], c[2], cc);
    const NbNone nb{};  // fetching the neighbour with the cell costs this variant registers it lacks
    double Bd[3];
    float B[3], d[3][3];
    if constexpr (F::kCoef) {
        const double x = wd[0], y = wd[1], z = wd[2];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double c0 = cc[q][0], c1 = cc[q][1], c2 = cc[q][2], c3 = cc[q][3];
            const double c4 = cc[q][4], c5 = cc[q][5], c6 = cc[q][6], c7 = cc[q][7];
            const double A = (c0 + c2 * y) + x * (c1 + c4 * y);
            const double C = (c3 + c5 * y) + x * (c6 + c7 * y);
            Bd[q] = A + z * C;
            d[2][q] = (float)C;
            d[0][q] = (cc[q][1] + cc[q][4] * w[1]) + w[2] * (cc[q][6] + cc[q][7] * w[1]);
            d[1][q] = (cc[q][2] + cc[q][4] * w[0]) + w[2] * (cc[q][5] + cc[q][7] * w[0]);
        }
    } else {
        double cd[3][8], dd[3][3];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int k = 0; k < 8; ++k) cd[q][k] = cc[q][k];
        cell_eval<double>(cd, wd[0], wd[1], wd[2], Bd, dd);
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int q = 0; q < 3; ++q) d[j][q] = (float)dd[j][q];
    }
    const double inbd = rsq(dot3(Bd, Bd));
    float b[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        bd[q] = Bd[q] * inbd;
        b[q] = (float)bd[q];
        B[q] = (float)Bd[q];
    }
    const float inb = (float)inbd;
    scott_axis<0, float, F>(f, p, c, w, simple[0], false, B, b, inb, d[0], nb, J[0]);
    scott_axis<1, float, F>(f, p, c, w, simple[1], false, B, b, inb, d[1], nb, J[1]);
    scott_axis<2, float, F>(f, p, c, w, simple[2], false, B, b, inb, d[2], nb, J[2]);
}

// Unit vector b only (plain tracer RHS: Interp3dxyzn(..., true)).
template <typename R, class F>
FQSL_HD void unit_b(const F& f, const Pt<R>& p, R b[3]) {
    int ci[3];
    R w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const Axis<R> ax = axis_setup<R>(p.i[a], p.r[a], f.n[a]);
        ci[a] = ax.c; w[a] = ax.w;
    }
    R c[3][8], B[3];
    f.template load<R>(ci[0], ci[1], ci[2], c);
    field_value<F, R>(c, w[0], w[1], w[2], B);
    const R inb = rsq(dot3(B, B));
#pragma unroll
    for (int q = 0; q < 3; ++q) b[q] = B[q] * inb;
}

// Raw B (Interp3dxyzn(..., false)) at a double position.
template <typename R, class F>
FQSL_HD void raw_b(const F& f, const double* x, double B[3]) {
    const Pt<R> p = pt_from_double<R>(x);
    int ci[3];
    R w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const Axis<R> ax = axis_setup<R>(p.i[a], p.r[a], f.n[a]);
        ci[a] = ax.c; w[a] = ax.w;
    }
    R c[3][8], Bv[3];
    f.template load<R>(ci[0], ci[1], ci[2], c);
    field_value<F, R>(c, w[0], w[1], w[2], Bv);
    B[0] = Bv[0]; B[1] = Bv[1]; B[2] = Bv[2];
}

// Interp3d (TraceBlineScott.cu:47-93) for the three components: the point is
// first pulled into [1e-4, N-1-1e-4] (NaN -> 1e-4), in float as the reference.
template <typename R, class F>
FQSL_HD void b_clamped(const F& f, const float* x, float out[3]) {
    double q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v = x[a];
        v = (v > 0.f) ? v : 0.0001f;
        v = (v < (float)(f.n[a] - 1)) ? v : (float)((double)f.n[a] - 1 - 0.0001);
        q[a] = v;
    }
    double B[3];
    raw_b<R>(f, q, B);
    out[0] = (float)B[0]; out[1] = (float)B[1]; out[2] = (float)B[2];
}

// ------------------------------------------------------------------ twist
// Current-density field for the twist number: none (the tracers' default).
struct NoCurrent {
    int n[3];
};

template <class FJ>
constexpr bool has_current() { return !std::is_same<FJ, NoCurrent>::value; }

constexpr double kQuarterOverPi = 0.25 / 3.14159265358979323846;

// J.B / |B|^2 / 4pi at x from the raw interpolants of J and B
// (TraceBlineScott.cu:424-426, TraceBline.cu:326-328); B is returned too.
template <typename R, class F, class FJ>
FQSL_HD double twist_rate(const F& f, const FJ& fj, const double* x, double B[3]) {
    double J[3];
    raw_b<R>(fj, x, J);
    raw_b<R>(f, x, B);
    return dot3(J, B) / dot3(B, B) * kQuarterOverPi;
}

// checkFlag (TraceBlineScott.cu:310-332) in double.
FQSL_HD int check_flag(const int* n, const double* p) {
    const double X = p[0], Y = p[1], Z = p[2];
    const double mx = n[0] - 1., my = n[1] - 1., mz = n[2] - 1.;
    if (X >= 0. && Y > 0. && Z >= 0. && X <= mx && Y <= my && Z <= mz) return 0;
    int fl = kClassOther;
    if (X < 0. && Y >= 0. && Z >= 0. && X < mx && Y < my && Z < mz) fl = 1;
    if (X >= 0. && Y >= 0. && Z >= 0. && X >= mx && Y < my && Z < mz) fl = 2;
    if (X >= 0. && Y < 0. && Z >= 0. && X < mx && Y < my && Z < mz) fl = 3;
    if (X >= 0. && Y >= 0. && Z >= 0. && X < mx && Y >= my && Z < mz) fl = 4;
    if (X >= 0. && Y >= 0. && Z < 0. && X < mx && Y < my && Z < mz) fl = 5;
    if (X >= 0. && Y >= 0. && Z >= 0. && X < mx && Y < my && Z >= mz) fl = 6;
    return fl;
}

// ----------------------------------------------------------------- positions
// Position of a line: compensated float pair or plain double.
template <typename R> struct Pos;

template <>
struct Pos<float> {
    float hi[3], lo[3];
    FQSL_HD void set(const double* x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            hi[a] = (float)x[a];
            lo[a] = (float)(x[a] - (double)hi[a]);
        }
    }
    FQSL_HD void set_f(const float* x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) { hi[a] = x[a]; lo[a] = 0.f; }
    }
    FQSL_HD double get(int a) const { return (double)hi[a] + (double)lo[a]; }
    // stage point x + delta, formed relative to floor(hi) so the weight keeps
    // full float resolution
    FQSL_HD Pt<float> stage(const float* delta) const {
        Pt<float> p;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float fh = floorf(hi[a]);
            const float u = (hi[a] - fh) + (lo[a] + delta[a]);
            const float fu = floorf(u);
            p.i[a] = to_int(fh) + to_int(fu);
            p.r[a] = u - fu;
        }
        return p;
    }
    // x += inc with a compensated (two-sum) update
    FQSL_HD void add(const float* inc) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float t = lo[a] + inc[a];
            const float s = hi[a] + t;
            const float bb = s - hi[a];
            const float err = (hi[a] - (s - bb)) + (t - bb);
            hi[a] = s;
            lo[a] = err;
        }
    }
};

template <>
struct Pos<double> {
    double x[3];
    FQSL_HD void set(const double* v) { x[0] = v[0]; x[1] = v[1]; x[2] = v[2]; }
    FQSL_HD void set_f(const float* v) { x[0] = v[0]; x[1] = v[1]; x[2] = v[2]; }
    FQSL_HD double get(int a) const { return x[a]; }
    FQSL_HD Pt<double> stage(const double* delta) const {
        Pt<double> p;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double v = x[a] + delta[a];
            const double f = floor(v);
            p.i[a] = to_int(f);
            p.r[a] = v - f;
        }
        return p;
    }
    FQSL_HD void add(const double* inc) {
#pragma unroll
        for (int a = 0; a < 3; ++a) x[a] += inc[a];
    }
};

// ------------------------------------------------------------- Scott tracer
struct ScottParams {
    int method;     // kRKF45 or kRK4
    double s0;      // initial (RK4: fixed) step, cells
    double tol;     // RKF45 absolute per-step tolerance, cells
    double ctl;     // (tol / 11.09)^0.2, the tolerance factor of the step controller
                    // (formed once per kernel on the device: scott_ctl)
};

// (tol / 11.09)^0.2: the controller's scale (tol / err / 11.09)^0.2
// (TraceBlineScott.cu:413) is ctl * (err^2)^(-1/10).
FQSL_HD double scott_ctl(double tol) { return tol > 0 ? pow(tol / 11.09, 0.2) : 0.0; }

// e2^(-1/10) to double precision without a double pow: a float estimate
// y0 = 2^(-0.1 log2 e2) (relative error up to ~1e-6, from the float log2 of
// values up to |100|) refined by two Newton steps for y^10 e2 = 1,
// y <- y (1.1 - 0.1 y^10 e2); each step maps a relative error d to 5.5 d^2,
// so the result is exact to double rounding.  Zero, NaN, infinite and extreme
// inputs take the library pow.
// pow(e2, -0.1) out of line: inlined, its polynomial constants held in
// registers cost the stepping kernels about 40 VGPRs for a path that only a
// zero or huge error norm takes.
FQSL_COLD double pow_minus_tenth_wide(double e2) { return pow(e2, -0.1); }

FQSL_HD double pow_minus_tenth(double e2) {
    if (!(e2 > 1e-30 && e2 < 1e30)) return pow_minus_tenth_wide(e2);
    double y = (double)exp2f(-0.1f * log2f((float)e2));
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double y2 = y * y, y4 = y2 * y2, y8 = y4 * y4;
        y = y * (1.1 - 0.1 * (y8 * y2 * e2));
    }
    return y;
}

// Status of a line after init / attempt.
enum : int { kTracing = 0, kFinished = 1, kAborted = 2 };

// Footpoint of one side: position, deviation vectors u, v, class.
template <typename R>
struct ScottEnd {
    R x[3], u[3], v[3];
    int flag;
};

// State of one Scott line (TraceBlineScott, TraceBlineScott.cu:334-519).  Only
// what a line carries from one attempt to the next lives here; the seed frame
// (b0, u0, v0) is recomputed when the second direction starts and the end
// footpoint is handed back to the caller when the line finishes.  X is the
// position path's type (R, or double in the mixed variant).
template <typename R, typename X = R>
struct ScottLine {
    float p0[3];
    Pos<X> x;
    R u[3], v[3];
    double s, len;
    double tw;                 // twist accumulated over both directions
    int steps, attempts, dir;  // dir: -1 backward (first), +1 forward
    uint32_t evals;
    ScottEnd<R> st;            // start (backward) footpoint once known
    const R* fr;               // precomputed seed frame (u0, v0, b0z, |B0|) or nullptr
};

// Seed frame record written by a preparation pass: u0[3], v0[3], b0z, |B0|.
constexpr int kFrameWords = 8;

// Unit field b0 at the seed, |B0| and the initial deviation vectors
// v0 = normalize((b.y, -b.x, 0) - ((b.y,-b.x,0).b) b), u0 = normalize(b x v0)
// (TraceBlineScott.cu:368-386).
template <typename R, class F>
FQSL_HD void seed_frame(const F& f, const float* p0, R u0[3], R v0[3], R* b0z, double* nb0) {
    const double p0d[3] = {p0[0], p0[1], p0[2]};
    double B0[3];
    raw_b<R>(f, p0d, B0);
    if (nb0) *nb0 = sqrt(B0[0] * B0[0] + B0[1] * B0[1] + B0[2] * B0[2]);
    R Bv[3] = {(R)B0[0], (R)B0[1], (R)B0[2]};
    const R inb = rsq(dot3(Bv, Bv));
    const R b0[3] = {Bv[0] * inb, Bv[1] * inb, Bv[2] * inb};
    *b0z = b0[2];
    R v[3] = {b0[1], -b0[0], (R)0};
    const R dv = dot3(v, b0);
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] -= dv * b0[a];
    const R iv = rsq(dot3(v, v));
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] *= iv;
    R u[3] = {b0[1] * v[2] - b0[2] * v[1], b0[2] * v[0] - b0[0] * v[2], b0[0] * v[1] - b0[1] * v[0]};
    const R iu = rsq(dot3(u, u));
#pragma unroll
    for (int a = 0; a < 3; ++a) { u0[a] = u[a] * iu; v0[a] = v[a]; }
}

// Record a footpoint for the active side (start when dir < 0, else `e`).
// (Two explicit branches: selecting a reference to one of the two records
// would force both, and the line state, out of registers into scratch.  For
// the same reason no reference into a line's state, or to the field view,
// is ever passed to an out-of-line (FQSL_COLD) function.)
template <typename R, typename X>
FQSL_HD void scott_record(ScottLine<R, X>& L, ScottEnd<R>& e, const double* x, const R* u, const R* v, int flag) {
    if (L.dir < 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) { L.st.x[a] = (R)x[a]; L.st.u[a] = u[a]; L.st.v[a] = v[a]; }
        L.st.flag = flag;
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) { e.x[a] = (R)x[a]; e.u[a] = u[a]; e.v[a] = v[a]; }
        e.flag = flag;
    }
}

template <typename R, typename X>
FQSL_HD bool scott_can_step(const ScottLine<R, X>& L, int len_lim) {
    return L.len < (double)len_lim && L.steps < kScottStepLim && L.attempts < kAttemptCap;
}

template <class F>
FQSL_HD int scott_len_lim(const F& f) {
    return (int)kScottLenRatio * (f.n[0] + f.n[1] + f.n[2]);
}

// Start direction L.dir from the seed: the z0 shortcut
// (TraceBlineScott.cu:396-401) and directions that cannot take a step because
// the shared length/step budget is spent record their footpoint at once.
// Returns kTracing, or kFinished once the forward side is recorded.
template <typename R, typename X, class F>
FQSL_HD int scott_begin_dir(ScottLine<R, X>& L, const F& f, ScottEnd<R>& e) {
    for (;;) {
        R u0[3], v0[3], b0z;
        if (L.fr) {
#pragma unroll
            for (int a = 0; a < 3; ++a) { u0[a] = L.fr[a]; v0[a] = L.fr[3 + a]; }
            b0z = L.fr[6];
        } else {
            seed_frame<R>(f, L.p0, u0, v0, &b0z, nullptr);
        }
        const bool z0 = L.p0[2] < 1e-6f;
        L.x.set_f(L.p0);
#pragma unroll
        for (int a = 0; a < 3; ++a) { L.u[a] = u0[a]; L.v[a] = v0[a]; }
        const double px[3] = {L.p0[0], L.p0[1], L.p0[2]};
        if (z0 && (double)b0z * L.dir < 0) {
            scott_record(L, e, px, u0, v0, 5);
        } else {
            L.s = fabs(L.s);
            if (scott_can_step(L, scott_len_lim(f))) return kTracing;
            scott_record(L, e, px, u0, v0, kClassLimit);
        }
        if (L.dir < 0) { L.dir = 1; continue; }
        return kFinished;
    }
}

template <typename R, typename X, class F>
FQSL_HD int scott_init(ScottLine<R, X>& L, const F& f, const ScottParams& P, float px, float py, float pz,
                         ScottEnd<R>& e, const R* frame = nullptr) {
    L.p0[0] = px; L.p0[1] = py; L.p0[2] = pz;
    L.fr = frame;
    L.len = 0.0;
    L.tw = 0.0;
    L.steps = 0; L.attempts = 0; L.evals = 0;
    L.s = P.s0;
    L.dir = -1;
    if (pz < -1e-8f) {  // TraceBlineScott.cu:366
#pragma unroll
        for (int a = 0; a < 3; ++a) { L.st.x[a] = (R)L.p0[a]; e.x[a] = (R)L.p0[a]; }
        L.st.flag = e.flag = kClassBelow;
        return kFinished;
    }
    return scott_begin_dir<R>(L, f, e);
}

// Right-hand side for a stage: k = s * f(y + delta) for the 9-vector, the
// position part kx (type X, step scx) and the deviation part ku = (u', v')
// (type R, step sc).
template <typename R, typename X, class F>
FQSL_HD void scott_stage(const F& f, const ScottLine<R, X>& L, const X* dx, const R* du, const R* dv,
                         X scx, R sc, X kx[3], R ku[6]) {
    R J[3][3];
    if constexpr (std::is_same<R, X>::value) {
        const Pt<R> p = L.x.stage(dx);
        R b[3];
        scott_eval<R>(f, p, b, J);
#pragma unroll
        for (int q = 0; q < 3; ++q) kx[q] = sc * b[q];
    } else {
        const Pt<X> p = L.x.stage(dx);
        X b[3];
        scott_eval_mixed<F>(f, p, b, J);
#pragma unroll
        for (int q = 0; q < 3; ++q) kx[q] = scx * b[q];
    }
    R uu[3], vv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) { uu[a] = L.u[a] + du[a]; vv[a] = L.v[a] + dv[a]; }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ku[q] = sc * (uu[0] * J[0][q] + uu[1] * J[1][q] + uu[2] * J[2][q]);
        ku[3 + q] = sc * (vv[0] * J[0][q] + vv[1] * J[1][q] + vv[2] * J[2][q]);
    }
}

// Close the active side; returns the line status after it.
template <typename R, typename X, class F>
FQSL_HD int scott_side_done(ScottLine<R, X>& L, const F& f, ScottEnd<R>& e) {
    if (L.dir < 0) {
        L.dir = 1;
        return scott_begin_dir<R>(L, f, e);
    }
    return kFinished;
}

// The stage sets of one attempt.  Each returns the increment of the state
// (and, for RKF45, the squared position error estimate); the controller and
// the step bookkeeping follow in scott_attempt.  The unrolled forms hold
// every stage's derivative in registers; on the device the FP32 path runs the
// same expressions with one copy of the stage evaluation in a loop and the
// RKF45 stage derivatives k_1..k_5 in LDS, which lets the kernel run three
// waves per SIMD instead of two -- the same bits.
template <typename R, typename X, class F>
FQSL_HD void rkf45_stages_unrolled(const F& f, const ScottLine<R, X>& L, X scx, R sc, X incx[3], R incu[6],
                                   double& e2) {
    // stage k_i = (kx_i, ku_i); t = the stage offset sum_j a_ij k_j
    X x1[3], x2[3], x3[3], x4[3], x5[3], x6[3], tx[3];
    R u1[6], u2[6], u3[6], u4[6], u5[6], u6[6], tu[6];
    const X zx[3] = {0, 0, 0};
    const R zu[6] = {0, 0, 0, 0, 0, 0};
    scott_stage<R, X>(f, L, zx, zu, zu + 3, scx, sc, x1, u1);
#pragma unroll
    for (int c = 0; c < 3; ++c) tx[c] = (X)Tab::a21 * x1[c];
#pragma unroll
    for (int c = 0; c < 6; ++c) tu[c] = (R)Tab::a21 * u1[c];
    scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, x2, u2);
#pragma unroll
    for (int c = 0; c < 3; ++c) tx[c] = (X)Tab::a31 * x1[c] + (X)Tab::a32 * x2[c];
#pragma unroll
    for (int c = 0; c < 6; ++c) tu[c] = (R)Tab::a31 * u1[c] + (R)Tab::a32 * u2[c];
    scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, x3, u3);
#pragma unroll
    for (int c = 0; c < 3; ++c) tx[c] = (X)Tab::a41 * x1[c] + (X)Tab::a42 * x2[c] + (X)Tab::a43 * x3[c];
#pragma unroll
    for (int c = 0; c < 6; ++c) tu[c] = (R)Tab::a41 * u1[c] + (R)Tab::a42 * u2[c] + (R)Tab::a43 * u3[c];
    scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, x4, u4);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        tx[c] = (X)Tab::a51 * x1[c] + (X)Tab::a52 * x2[c] + (X)Tab::a53 * x3[c] + (X)Tab::a54 * x4[c];
#pragma unroll
    for (int c = 0; c < 6; ++c)
        tu[c] = (R)Tab::a51 * u1[c] + (R)Tab::a52 * u2[c] + (R)Tab::a53 * u3[c] + (R)Tab::a54 * u4[c];
    scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, x5, u5);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        tx[c] = (X)Tab::a61 * x1[c] + (X)Tab::a62 * x2[c] + (X)Tab::a63 * x3[c] + (X)Tab::a64 * x4[c] +
                (X)Tab::a65 * x5[c];
#pragma unroll
    for (int c = 0; c < 6; ++c)
        tu[c] = (R)Tab::a61 * u1[c] + (R)Tab::a62 * u2[c] + (R)Tab::a63 * u3[c] + (R)Tab::a64 * u4[c] +
                (R)Tab::a65 * u5[c];
    scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, x6, u6);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        incx[c] = (X)Tab::b1 * x1[c] + (X)Tab::b3 * x3[c] + (X)Tab::b4 * x4[c] + (X)Tab::b5 * x5[c] +
                  (X)Tab::b6 * x6[c];
#pragma unroll
    for (int c = 0; c < 6; ++c)
        incu[c] = (R)Tab::b1 * u1[c] + (R)Tab::b3 * u3[c] + (R)Tab::b4 * u4[c] + (R)Tab::b5 * u5[c] +
                  (R)Tab::b6 * u6[c];
    e2 = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double ec = Tab::e1 * (double)x1[c] + Tab::e3 * (double)x3[c] + Tab::e4 * (double)x4[c] +
                          Tab::e5 * (double)x5[c] + Tab::e6 * (double)x6[c];
        e2 += ec * ec;
    }
}

template <typename R, typename X, class F>
FQSL_HD void rk4_stages_unrolled(const F& f, const ScottLine<R, X>& L, X scx, R sc, X incx[3], R incu[6]) {
    // RK4 (TraceBline.cu:226-234) on the 9-vector, fixed step
    X kx[3], tx[3];
    R ku[6], tu[6];
    const X zx[3] = {0, 0, 0};
    const R zu[6] = {0, 0, 0, 0, 0, 0};
    const X halfx = (X)0.5;
    const R half = (R)0.5;
    scott_stage<R, X>(f, L, zx, zu, zu + 3, scx, sc, kx, ku);
#pragma unroll
    for (int c = 0; c < 3; ++c) { incx[c] = kx[c]; tx[c] = halfx * kx[c]; }
#pragma unroll
    for (int c = 0; c < 6; ++c) { incu[c] = ku[c]; tu[c] = half * ku[c]; }
    scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, kx, ku);
#pragma unroll
    for (int c = 0; c < 3; ++c) { incx[c] += (X)2 * kx[c]; tx[c] = halfx * kx[c]; }
#pragma unroll
    for (int c = 0; c < 6; ++c) { incu[c] += (R)2 * ku[c]; tu[c] = half * ku[c]; }
    scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, kx, ku);
#pragma unroll
    for (int c = 0; c < 3; ++c) { incx[c] += (X)2 * kx[c]; tx[c] = kx[c]; }
#pragma unroll
    for (int c = 0; c < 6; ++c) { incu[c] += (R)2 * ku[c]; tu[c] = ku[c]; }
    scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, kx, ku);
#pragma unroll
    for (int c = 0; c < 3; ++c) incx[c] = (incx[c] + kx[c]) * (X)(1.0 / 6.0);
#pragma unroll
    for (int c = 0; c < 6; ++c) incu[c] = (incu[c] + ku[c]) * (R)(1.0 / 6.0);
}

#if defined(__HIP_DEVICE_COMPILE__)
// Threads per workgroup of every kernel that traces Scott lines (their
// __launch_bounds__): the LDS array of the rolled RKF45 stages holds one
// column per thread.
constexpr int kScottBlock = 256;

template <class F>
__device__ void rkf45_stages_rolled(const F& f, const ScottLine<float, float>& L, float scx, float sc,
                                    float incx[3], float incu[6], double& e2) {
    using X = float;
    using R = float;
    __shared__ float klds[5 * 9 * kScottBlock];
    const int tid = threadIdx.x;
    auto K = [&](int j, int c) -> float& { return klds[(j * 9 + c) * kScottBlock + tid]; };
    float x6[3], u6[6];
#pragma clang loop unroll(disable)
    for (int st = 0; st < 6; ++st) {
        // the stage offset sum_j a_ij k_j, the unrolled form's expressions
        float tx[3], tu[6];
        switch (st) {
        case 0:
#pragma unroll
            for (int c = 0; c < 3; ++c) tx[c] = 0.f;
#pragma unroll
            for (int c = 0; c < 6; ++c) tu[c] = 0.f;
            break;
        case 1:
#pragma unroll
            for (int c = 0; c < 3; ++c) tx[c] = (X)Tab::a21 * K(0, c);
#pragma unroll
            for (int c = 0; c < 6; ++c) tu[c] = (R)Tab::a21 * K(0, 3 + c);
            break;
        case 2:
#pragma unroll
            for (int c = 0; c < 3; ++c) tx[c] = (X)Tab::a31 * K(0, c) + (X)Tab::a32 * K(1, c);
#pragma unroll
            for (int c = 0; c < 6; ++c) tu[c] = (R)Tab::a31 * K(0, 3 + c) + (R)Tab::a32 * K(1, 3 + c);
            break;
        case 3:
#pragma unroll
            for (int c = 0; c < 3; ++c)
                tx[c] = (X)Tab::a41 * K(0, c) + (X)Tab::a42 * K(1, c) + (X)Tab::a43 * K(2, c);
#pragma unroll
            for (int c = 0; c < 6; ++c)
                tu[c] = (R)Tab::a41 * K(0, 3 + c) + (R)Tab::a42 * K(1, 3 + c) + (R)Tab::a43 * K(2, 3 + c);
            break;
        case 4:
#pragma unroll
            for (int c = 0; c < 3; ++c)
                tx[c] = (X)Tab::a51 * K(0, c) + (X)Tab::a52 * K(1, c) + (X)Tab::a53 * K(2, c) +
                        (X)Tab::a54 * K(3, c);
#pragma unroll
            for (int c = 0; c < 6; ++c)
                tu[c] = (R)Tab::a51 * K(0, 3 + c) + (R)Tab::a52 * K(1, 3 + c) + (R)Tab::a53 * K(2, 3 + c) +
                        (R)Tab::a54 * K(3, 3 + c);
            break;
        default:
#pragma unroll
            for (int c = 0; c < 3; ++c)
                tx[c] = (X)Tab::a61 * K(0, c) + (X)Tab::a62 * K(1, c) + (X)Tab::a63 * K(2, c) +
                        (X)Tab::a64 * K(3, c) + (X)Tab::a65 * K(4, c);
#pragma unroll
            for (int c = 0; c < 6; ++c)
                tu[c] = (R)Tab::a61 * K(0, 3 + c) + (R)Tab::a62 * K(1, 3 + c) + (R)Tab::a63 * K(2, 3 + c) +
                        (R)Tab::a64 * K(3, 3 + c) + (R)Tab::a65 * K(4, 3 + c);
            break;
        }
        float xs[3], us[6];
        scott_stage<R, X>(f, L, tx, tu, tu + 3, scx, sc, xs, us);
        if (st < 5) {
#pragma unroll
            for (int c = 0; c < 3; ++c) K(st, c) = xs[c];
#pragma unroll
            for (int c = 0; c < 6; ++c) K(st, 3 + c) = us[c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) x6[c] = xs[c];
#pragma unroll
            for (int c = 0; c < 6; ++c) u6[c] = us[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        incx[c] = (X)Tab::b1 * K(0, c) + (X)Tab::b3 * K(2, c) + (X)Tab::b4 * K(3, c) + (X)Tab::b5 * K(4, c) +
                  (X)Tab::b6 * x6[c];
#pragma unroll
    for (int c = 0; c < 6; ++c)
        incu[c] = (R)Tab::b1 * K(0, 3 + c) + (R)Tab::b3 * K(2, 3 + c) + (R)Tab::b4 * K(3, 3 + c) +
                  (R)Tab::b5 * K(4, 3 + c) + (R)Tab::b6 * u6[c];
    e2 = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double ec = Tab::e1 * (double)K(0, c) + Tab::e3 * (double)K(2, c) + Tab::e4 * (double)K(3, c) +
                          Tab::e5 * (double)K(4, c) + Tab::e6 * (double)x6[c];
        e2 += ec * ec;
    }
}

template <class F>
__device__ void rk4_stages_rolled(const F& f, const ScottLine<float, float>& L, float scx, float sc, float incx[3],
                                  float incu[6]) {
    float tx[3] = {0.f, 0.f, 0.f};
    float tu[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma clang loop unroll(disable)
    for (int st = 0; st < 4; ++st) {
        float kx[3], ku[6];
        scott_stage<float, float>(f, L, tx, tu, tu + 3, scx, sc, kx, ku);
        if (st == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) { incx[c] = kx[c]; tx[c] = 0.5f * kx[c]; }
#pragma unroll
            for (int c = 0; c < 6; ++c) { incu[c] = ku[c]; tu[c] = 0.5f * ku[c]; }
        } else if (st < 3) {
            const float g = st == 1 ? 0.5f : 1.f;
#pragma unroll
            for (int c = 0; c < 3; ++c) { incx[c] += 2.f * kx[c]; tx[c] = g * kx[c]; }
#pragma unroll
            for (int c = 0; c < 6; ++c) { incu[c] += 2.f * ku[c]; tu[c] = g * ku[c]; }
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) incx[c] = (incx[c] + kx[c]) * (float)(1.0 / 6.0);
#pragma unroll
            for (int c = 0; c < 6; ++c) incu[c] = (incu[c] + ku[c]) * (float)(1.0 / 6.0);
        }
    }
}
#endif

// The rolled form serves the three-wave FP32c kernels (kRoll) without the
// twist number (whose extra loads at each accepted step leave no registers
// for a third wave); at two waves the unrolled form is faster.
template <typename R, typename X, class FJ>
constexpr bool rollable_stages() {
    return std::is_same<R, float>::value && std::is_same<X, float>::value && !has_current<FJ>();
}

template <typename R, typename X, class FJ, bool kRoll, class F>
FQSL_HD void rkf45_stages(const F& f, const ScottLine<R, X>& L, X scx, R sc, X incx[3], R incu[6], double& e2) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (kRoll && rollable_stages<R, X, FJ>())
        rkf45_stages_rolled<F>(f, L, scx, sc, incx, incu, e2);
    else
#endif
        rkf45_stages_unrolled<R, X, F>(f, L, scx, sc, incx, incu, e2);
}

template <typename R, typename X, class FJ, bool kRoll, class F>
FQSL_HD void rk4_stages(const F& f, const ScottLine<R, X>& L, X scx, R sc, X incx[3], R incu[6]) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (kRoll && rollable_stages<R, X, FJ>())
        rk4_stages_rolled<F>(f, L, scx, sc, incx, incu);
    else
#endif
        rk4_stages_unrolled<R, X, F>(f, L, scx, sc, incx, incu);
}

// One integrator attempt (one pass of the while loop at
// TraceBlineScott.cu:407-470).  Returns kTracing, kFinished (both footpoints
// known: start in L.st, end in e) or kAborted (class 42: Q = 0).
template <typename R, class F, class FJ = NoCurrent, typename X = R, bool kRoll = false>
FQSL_HD int scott_attempt(ScottLine<R, X>& L, const F& f, const ScottParams& P, ScottEnd<R>& e,
                          const FJ& fj = FJ{}) {
    const R sc = (R)(L.s * L.dir);
    const X scx = (X)(L.s * L.dir);
    X incx[3];
    R incu[6];
    L.attempts += 1;
    if (P.method == kRKF45) {
        double e2;
        rkf45_stages<R, X, FJ, kRoll, F>(f, L, scx, sc, incx, incu, e2);
        L.evals += 6;
        // step controller (TraceBlineScott.cu:413-419), FP64: the FP64 path
        // evaluates the reference's expression as written; the FP32 and mixed
        // paths the same quantity as ctl * e2^(-1/10) (double-accurate, no
        // pow/sqrt/div)
        const double scale = sizeof(R) == 4 ? P.ctl * pow_minus_tenth(e2)
                                            : pow(P.tol / sqrt(e2) / 11.09, 0.2);
        if (scale < 0.618) {
            L.s = L.s * 0.618;
            if (!scott_can_step(L, scott_len_lim(f))) {  // attempt budget spent: limit exit
                const double xa[3] = {L.x.get(0), L.x.get(1), L.x.get(2)};
                scott_record(L, e, xa, L.u, L.v, kClassLimit);
                return scott_side_done<R>(L, f, e);
            }
            return kTracing;
        }
        L.s = L.s * scale;
        if (L.s > 100.) L.s = 100.;
        if (L.s < 1. / 10.) L.s = 1. / 10.;
    } else {
        // RK4 (TraceBline.cu:226-234) on the 9-vector, fixed step
        rk4_stages<R, X, FJ, kRoll, F>(f, L, scx, sc, incx, incu);
        L.evals += 4;
    }

    // accepted step: a = current state, b = a + inc
    const double xa[3] = {L.x.get(0), L.x.get(1), L.x.get(2)};
    Pos<X> xb = L.x;
    xb.add(incx);
    const double xbv[3] = {xb.get(0), xb.get(1), xb.get(2)};
    // step length: the increment is X, so its norm is taken in X; the sum over
    // steps is kept in double
    const double dL = sizeof(X) == 4 ? (double)sqrtf((float)(incx[0] * incx[0] + incx[1] * incx[1] + incx[2] * incx[2]))
                                     : sqrt((double)incx[0] * incx[0] + (double)incx[1] * incx[1] +
                                            (double)incx[2] * incx[2]);
    L.len += dL;
    // twist at the step's end point (TraceBlineScott.cu:423-427)
    double tw_rate = 0.0, Bend[3] = {0.0, 0.0, 0.0};
    if constexpr (has_current<FJ>()) {
        tw_rate = twist_rate<R>(f, fj, xbv, Bend);
        L.tw += tw_rate * dL;
    }
    const int flag = check_flag(f.n, xbv);
    if (flag > 0) {
        L.len -= dL;
        if constexpr (has_current<FJ>()) L.tw -= tw_rate * dL;  // :432-434
        if (flag <= 6) {
            const int dim = (flag - 1) / 2;
            const double p1 = xa[dim], p2 = xbv[dim];
            if (fabs(p1 - p2) > 0) {
                // linear landing of the 9-vector on the face (TraceBlineScott.cu:442-448);
                // max faces land on index N
                const double pm = (flag % 2 == 1) ? 0.0 : (double)f.n[dim];
                const double wa = (p2 - pm) / (p2 - p1), wb = (pm - p1) / (p2 - p1);
                double xl[3];
                R ul[3], vl[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    xl[a] = (xa[a] * (p2 - pm) + xbv[a] * (pm - p1)) / (p2 - p1);
                    ul[a] = (R)((double)L.u[a] * wa + (double)(L.u[a] + incu[a]) * wb);
                    vl[a] = (R)((double)L.v[a] * wa + (double)(L.v[a] + incu[3 + a]) * wb);
                }
                // arc length to the face: with twist on, the reference's B_P1
                // is the raw B at the step's end point (:425) and also scales
                // the landing twist (:455-457); with it off B_P1 is unset and b
                // at the last in-box state is used
                double bdim;
                if constexpr (has_current<FJ>()) {
                    bdim = Bend[dim];
                } else {
                    X ba[3];
                    const X zero3[3] = {0, 0, 0};
                    unit_b<X>(f, L.x.stage(zero3), ba);
                    bdim = (double)ba[dim];
                }
                const double arc = fabs(pm - p1) / (1e-4 + fabs(bdim));
                L.len += arc;
                if constexpr (has_current<FJ>()) L.tw += tw_rate * arc;
                scott_record(L, e, xl, ul, vl, flag);
            } else {
                scott_record(L, e, xa, L.u, L.v, flag);
            }
            L.steps += 1;
            return scott_side_done<R>(L, f, e);
        }
        // class 42 (TraceBlineScott.cu:460-465): q = 0, the line stops here;
        // a forward side never traced keeps the seed
        scott_record(L, e, xa, L.u, L.v, kClassOther);
        if (L.dir < 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) { e.x[a] = (R)L.p0[a]; e.u[a] = 0; e.v[a] = 0; }
            e.flag = kClassOther;
        }
        return kAborted;
    }
    L.x = xb;
#pragma unroll
    for (int a = 0; a < 3; ++a) { L.u[a] += incu[a]; L.v[a] += incu[3 + a]; }
    L.steps += 1;
    if (!scott_can_step(L, scott_len_lim(f))) {  // length / step limit: class 0, last state
        scott_record(L, e, xbv, L.u, L.v, kClassLimit);
        return scott_side_done<R>(L, f, e);
    }
    return kTracing;
}

// Q assembly (TraceBlineScott.cu:475-508) in FP64 from the start footpoint
// st, the end footpoint e and |B0| at the seed.
template <typename R, class F>
FQSL_COLD void scott_q(const ScottEnd<R>& st, const ScottEnd<R>& e, int status, double nb0, const F f,
                       double* q0, double* qperp) {
    if (status == kAborted) { *q0 = 0.0; *qperp = 0.0; return; }
    if (st.flag < 1 || st.flag > 6 || e.flag < 1 || e.flag > 6) {
        *q0 = NAN; *qperp = NAN; return;
    }
    double Bs[3], Be[3];
    const double xs[3] = {(double)st.x[0], (double)st.x[1], (double)st.x[2]};
    const double xe[3] = {(double)e.x[0], (double)e.x[1], (double)e.x[2]};
    raw_b<R>(f, xs, Bs);
    raw_b<R>(f, xe, Be);
    const int ds = (st.flag - 1) / 2, de = (e.flag - 1) / 2;
    double us[3], vs[3], ue[3], ve[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        us[a] = st.u[a]; vs[a] = st.v[a]; ue[a] = e.u[a]; ve[a] = e.v[a];
    }
    double us1[3], vs1[3], ue1[3], ve1[3];
    const double BNs = Bs[ds], BNe = Be[de];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        us1[a] = us[a] - us[ds] / BNs * Bs[a];
        vs1[a] = vs[a] - vs[ds] / BNs * Bs[a];
        ue1[a] = ue[a] - ue[de] / BNe * Be[a];
        ve1[a] = ve[a] - ve[de] / BNe * Be[a];
    }
    const double nb2 = nb0 * nb0;
    double num = dot3(ue1, ue1) * dot3(vs1, vs1) + dot3(us1, us1) * dot3(ve1, ve1) -
                 2.0 * dot3(ue1, ve1) * dot3(us1, vs1);
    *q0 = fabs(num) / fabs(nb2 / (BNs * BNe));
    const double nbs = sqrt(dot3(Bs, Bs)), nbe = sqrt(dot3(Be, Be));
    const double due = dot3(ue, Be) / nbe, dve = dot3(ve, Be) / nbe;
    const double dus = dot3(us, Bs) / nbs, dvs = dot3(vs, Bs) / nbs;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ue1[a] = ue[a] - due * (Be[a] / nbe);
        ve1[a] = ve[a] - dve * (Be[a] / nbe);
        us1[a] = us[a] - dus * (Bs[a] / nbs);
        vs1[a] = vs[a] - dvs * (Bs[a] / nbs);
    }
    num = dot3(ue1, ue1) * dot3(vs1, vs1) + dot3(us1, us1) * dot3(ve1, ve1) -
          2.0 * dot3(ue1, ve1) * dot3(us1, vs1);
    *qperp = fabs(num) / (nb2 / (nbs * nbe));
}

// ------------------------------------------------------------- plain tracer
struct PlainParams {
    int method;        // kRKF45 or kRK4
    double s0;         // initial step (each direction starts from it)
    double tol_coef;   // RKF45 tolerance coefficient
    float ncross[3];   // plane normal for the tolerance rule and B_flag
};

// State of one plain line (TraceBlineAdap / TraceBline, TraceBlineAdaptive.cu:333-426).
template <typename R>
struct PlainLine {
    float p0[3];
    Pos<R> x;
    double s, len, len_lim, tol, len_total;
    int steps, attempts, dir, step_lim;  // dir +1 forward (first), -1 backward
    int flag_e, flag_s;
    double xe[3], xs[3];
    double tw;  // twist, forward plus backward
    bool done;
    uint32_t evals;
};

template <typename R>
FQSL_HD bool plain_can_step(const PlainLine<R>& L) {
    return L.len < L.len_lim && L.steps < L.step_lim && L.attempts < kAttemptCap;
}

template <typename R>
FQSL_HD void plain_record(PlainLine<R>& L, const double* x, int flag) {
    if (L.dir > 0) {
        L.xe[0] = x[0]; L.xe[1] = x[1]; L.xe[2] = x[2];
        L.flag_e = flag;
    } else {
        L.xs[0] = x[0]; L.xs[1] = x[1]; L.xs[2] = x[2];
        L.flag_s = flag;
    }
}

// Begin the direction L.dir from the seed with a fresh step and budget.
template <typename R, class F>
FQSL_HD void plain_begin_dir(PlainLine<R>& L, const F& f, const PlainParams& P) {
    for (;;) {
        L.x.set_f(L.p0);
        L.s = P.s0;
        L.len = 0.0;
        L.steps = 0;
        L.attempts = 0;
        if (plain_can_step(L)) return;
        const double px[3] = {L.p0[0], L.p0[1], L.p0[2]};
        plain_record(L, px, kClassLimit);
        if (L.dir > 0) { L.dir = -1; continue; }
        L.done = true;
        return;
    }
}

template <typename R, class F>
FQSL_HD void plain_init(PlainLine<R>& L, const F& f, const PlainParams& P, float px, float py, float pz) {
    L.p0[0] = px; L.p0[1] = py; L.p0[2] = pz;
    const bool rkf = P.method == kRKF45;
    L.len_lim = (rkf ? kPlainRkf45LenRatio : kPlainRk4LenRatio) * (f.n[0] + f.n[1] + f.n[2]);
    L.step_lim = rkf ? kPlainRkf45StepLim : kPlainRk4StepLim;
    L.len_total = 0.0;
    L.tw = 0.0;
    L.evals = 0;
    L.done = false;
    L.flag_e = L.flag_s = kClassLimit;
    L.tol = 0.0;
    if (rkf) {  // TraceBlineAdaptive.cu:354-359
        const double p0d[3] = {px, py, pz};
        R b0[3];
        unit_b<R>(f, pt_from_double<R>(p0d), b0);
        const double bn = fabs((double)b0[0] * P.ncross[0] + (double)b0[1] * P.ncross[1] +
                               (double)b0[2] * P.ncross[2]);
        L.tol = (bn <= 0.05 ? kTolBase / 8e3 : kTolBase * pow(bn, 1.5)) * P.tol_coef;
    }
    L.dir = 1;
    plain_begin_dir(L, f, P);
}

template <typename R, class F>
FQSL_HD void plain_finish_dir(PlainLine<R>& L, const F& f, const PlainParams& P) {
    L.len_total += L.len;
    if (L.dir > 0) { L.dir = -1; plain_begin_dir(L, f, P); } else L.done = true;
}

// RK4_boundary (TraceBline.cu:253-274): integrate dx/dx_dim = b / b_dim over s.
template <typename R, class F>
FQSL_COLD void rk4_boundary(const F f, const double* p, double s, int dim, double out[3]) {
    double k[4][3], t[3];
    const double cf[4] = {0.0, 0.5, 0.5, 1.0};
#pragma unroll
    for (int st = 0; st < 4; ++st) {
#pragma unroll
        for (int a = 0; a < 3; ++a) t[a] = st == 0 ? p[a] : p[a] + s * k[st - 1][a] * cf[st];
        R b[3];
        unit_b<R>(f, pt_from_double<R>(t), b);
        const double bd = (double)b[dim];
#pragma unroll
        for (int a = 0; a < 3; ++a) k[st][a] = (double)b[a] / bd;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = p[a] + (1. / 6.) * s * (k[0][a] + 2.0 * k[1][a] + 2.0 * k[2][a] + k[3][a]);
}

template <typename R, class F, class FJ = NoCurrent>
FQSL_HD void plain_attempt(PlainLine<R>& L, const F& f, const PlainParams& P, const FJ& fj = FJ{}) {
    const R sc = (R)(L.s * L.dir);
    R inc[3];
    L.attempts += 1;
    if (P.method == kRKF45) {
        R k[6][3], t[3];
        const R zero[3] = {0, 0, 0};
        unit_b<R>(f, L.x.stage(zero), k[0]);
#pragma unroll
        for (int a = 0; a < 3; ++a) k[0][a] *= sc;
#pragma unroll
        for (int a = 0; a < 3; ++a) t[a] = (R)Tab::a21 * k[0][a];
        unit_b<R>(f, L.x.stage(t), k[1]);
#pragma unroll
        for (int a = 0; a < 3; ++a) { k[1][a] *= sc; t[a] = (R)Tab::a31 * k[0][a] + (R)Tab::a32 * k[1][a]; }
        unit_b<R>(f, L.x.stage(t), k[2]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            k[2][a] *= sc;
            t[a] = (R)Tab::a41 * k[0][a] + (R)Tab::a42 * k[1][a] + (R)Tab::a43 * k[2][a];
        }
        unit_b<R>(f, L.x.stage(t), k[3]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            k[3][a] *= sc;
            t[a] = (R)Tab::a51 * k[0][a] + (R)Tab::a52 * k[1][a] + (R)Tab::a53 * k[2][a] + (R)Tab::a54 * k[3][a];
        }
        unit_b<R>(f, L.x.stage(t), k[4]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            k[4][a] *= sc;
            t[a] = (R)Tab::a61 * k[0][a] + (R)Tab::a62 * k[1][a] + (R)Tab::a63 * k[2][a] + (R)Tab::a64 * k[3][a] +
                   (R)Tab::a65 * k[4][a];
        }
        unit_b<R>(f, L.x.stage(t), k[5]);
        double e2 = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            k[5][a] *= sc;
            inc[a] = (R)Tab::b1 * k[0][a] + (R)Tab::b3 * k[2][a] + (R)Tab::b4 * k[3][a] + (R)Tab::b5 * k[4][a] +
                     (R)Tab::b6 * k[5][a];
            const double e = Tab::e1 * (double)k[0][a] + Tab::e3 * (double)k[2][a] + Tab::e4 * (double)k[3][a] +
                             Tab::e5 * (double)k[4][a] + Tab::e6 * (double)k[5][a];
            e2 += e * e;
        }
        L.evals += 6;
        const double scale = pow(L.tol / sqrt(e2) / 11.09, 0.2);  // TraceBlineAdaptive.cu:365-370
        if (scale < 0.618) {
            L.s = L.s * 0.618;
            if (!plain_can_step(L)) {
                const double xa[3] = {L.x.get(0), L.x.get(1), L.x.get(2)};
                plain_record(L, xa, kClassLimit);
                plain_finish_dir(L, f, P);
            }
            return;
        }
        L.s = L.s * scale;
        if (L.s > 100.) L.s = 100.;
        if (L.s < 1. / 10.) L.s = 1. / 10.;
    } else {
        R k[3], t[3];
        const R zero[3] = {0, 0, 0};
        unit_b<R>(f, L.x.stage(zero), k);
#pragma unroll
        for (int a = 0; a < 3; ++a) { inc[a] = k[a]; t[a] = sc * k[a] * (R)0.5; }
        unit_b<R>(f, L.x.stage(t), k);
#pragma unroll
        for (int a = 0; a < 3; ++a) { inc[a] += (R)2 * k[a]; t[a] = sc * k[a] * (R)0.5; }
        unit_b<R>(f, L.x.stage(t), k);
#pragma unroll
        for (int a = 0; a < 3; ++a) { inc[a] += (R)2 * k[a]; t[a] = sc * k[a]; }
        unit_b<R>(f, L.x.stage(t), k);
#pragma unroll
        for (int a = 0; a < 3; ++a) inc[a] = (R)(1.0 / 6.0) * sc * (inc[a] + k[a]);
        L.evals += 4;
    }
    const double xa[3] = {L.x.get(0), L.x.get(1), L.x.get(2)};
    Pos<R> xb = L.x;
    xb.add(inc);
    const double xbv[3] = {xb.get(0), xb.get(1), xb.get(2)};
    const double dl = sqrt((double)inc[0] * inc[0] + (double)inc[1] * inc[1] + (double)inc[2] * inc[2]);
    L.len += dl;
    // twist at the step's start point (TraceBline.cu:325-329,
    // TraceBlineAdaptive.cu:373-377)
    double tw_rate = 0.0, Jp1[3] = {0.0, 0.0, 0.0};
    if constexpr (has_current<FJ>()) {
        double Bp1[3];
        raw_b<R>(fj, xa, Jp1);
        raw_b<R>(f, xa, Bp1);
        tw_rate = dot3(Jp1, Bp1) / dot3(Bp1, Bp1) * kQuarterOverPi;
        L.tw += tw_rate * dl;
    }
    const int flag = check_flag(f.n, xbv);
    if (flag > 0) {
        L.len -= dl;
        if constexpr (has_current<FJ>()) L.tw -= tw_rate * dl;  // TraceBline.cu:334-336
        if (flag <= 6) {
            const int dim = (flag - 1) / 2;
            const double pa = xa[dim], pb = xbv[dim];
            if (fabs(pa - pb) > 1e-3) {  // TraceBlineAdaptive.cu:391-405
                const double pm = (flag % 2 == 1) ? 0.0 : (double)f.n[dim];
                R b1[3], b2[3];
                unit_b<R>(f, pt_from_double<R>(xa), b1);
                unit_b<R>(f, pt_from_double<R>(xbv), b2);
                double out[3];
                if (fabs((double)b1[dim]) < 0.2 || fabs((double)b2[dim]) < 0.2) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) out[a] = (xa[a] * (pb - pm) + xbv[a] * (pm - pa)) / (pb - pa);
                } else {
                    rk4_boundary<R>(f, xa, pm - pa, dim, out);
                    L.evals += 4;
                }
                plain_record(L, out, flag);
                const double arc = fabs(pm - pa) / (1e-4 + fabs((double)b1[dim]));
                L.len += arc;
                // RK4 only (the RKF45 tracer's term is commented out): J.b/|b|^2
                // with b the normalised B at p1, as the reference re-assigns
                // B_P1 before it (TraceBline.cu:362-365)
                if constexpr (has_current<FJ>()) {
                    if (P.method == kRK4) {
                        const double bd[3] = {(double)b1[0], (double)b1[1], (double)b1[2]};
                        L.tw += dot3(Jp1, bd) / dot3(bd, bd) * kQuarterOverPi * arc;
                    }
                }
            } else {
                plain_record(L, xa, flag);
            }
        } else {
            plain_record(L, xbv, flag);  // "ignore": the outside point itself
        }
        L.steps += 1;
        plain_finish_dir(L, f, P);
        return;
    }
    L.x = xb;
    L.steps += 1;
    if (!plain_can_step(L)) {
        plain_record(L, xbv, kClassLimit);
        plain_finish_dir(L, f, P);
    }
}

}  // namespace fqsl
