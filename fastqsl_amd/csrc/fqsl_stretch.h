// fqsl_stretch.h -- stretched (non-uniform) grids for the plain RKF45 tracer.
//
// The reference's stretched tracer (TraceBlineAdaptiveStretch.cu) traces in
// the axes' own coordinates: node (i, j, k) sits at (xa[i], ya[j], za[k]),
// every axis strictly increasing.  Interpolation first pulls the point into
// [a0 + 1e-6, a_{n-1} - 1e-6] per axis (NaN goes to the low bound,
// :102-107), then picks a cell c by one of two rules and weighs with
// r = (p - a[c]) / (a[c+1] - a[c]) (:121-123):
//   kLookupBisect   GetIndexBiSearch (:42-65).  Its bisection stops once the
//                   bracket spans three cells or fewer and then chooses
//                   between the bracket's middle node and the one below, so
//                   a point in the top cell of a three-cell bracket gets the
//                   cell below (r in (1, 2]).  The cell it picks is a function
//                   of the cell that holds the point.
//   kLookupUniform  GetIndexUniStretch (:68-75): floor((p - a0) / (span / n)),
//                   span / n rather than span / (n - 1), clamped to [0, n-2];
//                   r may fall outside [0, 1].  Evaluated as the reference
//                   does (a division by the rounded width) in the path's
//                   precision.
//   kLookupExact    the cell that holds the point (not in the reference; with
//                   it the landing plane is the box face, see below).
// The interpolant is the cell's trilinear polynomial evaluated at r, so a cell
// other than the holding one extrapolates; both reference rules make the
// interpolant jump at some nodes.  The landing plane of a face exit is the
// coordinate 0 (min face) or n_dim (max face) as the reference writes it
// (:385-386), which is the face only when the axis spans [0, n]; the exact
// rule lands on a0 / a_{n-1}.
//
// On the device every rule is one table load: the axis is cut into uniform
// bins narrower than the spacing of the rule's breakpoints (the interior
// nodes for the bisection and exact rules, the points where
// floor((p - a0) / (span / n)) steps for the uniform one), so a bin holds at
// most one breakpoint, and each bin's record carries it (the split) and, for
// each side, the cell the rule interpolates in with that cell's low node and
// inverse width:
//   brec[b] = {split, a[m_lo], 1/(a[m_lo+1]-a[m_lo]), a[m_hi], 1/(...), m_lo, m_hi}
// (split = +inf for a bin without a breakpoint; a point above the split takes
// m_hi).  For the uniform rule the split is the largest value, in the path's
// precision, whose formula cell is still the lower one, so the reference's
// rounding decides ties.  Where the bin count had to be capped and a bin
// holds several breakpoints, the record says so and a search over per-cell
// records (or the formula) takes over:
//   rec[c] = {a[c], a[c+1], a[m], 1 / (a[m+1] - a[m])},  m = ref[c].
// Built on the host (build_axis_tables) once per field.
#pragma once

#include "fqsl_tracer.h"

#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

namespace fqsl {

enum { kLookupBisect = 0, kLookupUniform = 1, kLookupExact = 2 };

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

// One bin of an axis: 32 bytes in float, 48 in double, read with two or
// three 16-byte loads issued together.  m_lo < 0 marks a bin that holds more
// than one breakpoint (the search fallback).
template <typename T>
struct alignas(16) BinRec {
    T split, a_lo, iw_lo, a_hi, iw_hi;
    int m_lo, m_hi;
};

FQSL_HD int bits_of(float f) {
    int i;
    memcpy(&i, &f, 4);
    return i;
}

FQSL_HD BinRec<float> load_binrec(const BinRec<float>* p) {
    const FQSL_GLOBAL float4* q = (const FQSL_GLOBAL float4*)p;
    const float4 v0 = q[0], v1 = q[1];
    BinRec<float> r;
    r.split = v0.x; r.a_lo = v0.y; r.iw_lo = v0.z; r.a_hi = v0.w;
    r.iw_hi = v1.x; r.m_lo = bits_of(v1.y); r.m_hi = bits_of(v1.z);
    return r;
}

FQSL_HD BinRec<double> load_binrec(const BinRec<double>* p) {
    const FQSL_GLOBAL double2* q = (const FQSL_GLOBAL double2*)p;
    const double2 v0 = q[0], v1 = q[1], v2 = q[2];
    BinRec<double> r;
    r.split = v0.x; r.a_lo = v0.y; r.iw_lo = v1.x; r.a_hi = v1.y; r.iw_hi = v2.x;
    long long m;
    memcpy(&m, &v2.y, 8);
    r.m_lo = (int)(m & 0xffffffffll);
    r.m_hi = (int)(m >> 32);
    return r;
}

template <typename T>
struct StretchAxis {
    const BinRec<T>* brec;              // per uniform bin, see above
    const typename Vec4<T>::type* rec;  // per holding cell, see above (multi-breakpoint bins)
    const int* ref;                     // per holding cell: the cell the rule interpolates in
    const int* bins;                    // per uniform bin: lowest cell a point of the bin can lie in
    T lo, hi;                           // a[0], a[n-1]
    T clo, chi;                         // a[0] + 1e-6, a[n-1] - 1e-6
    T bin_scale;                        // nbins / (hi - lo)
    T uni_width;                        // (hi - lo) / n in T, GetIndexUniStretch's cell width
    int n, nbins;
    int lbin, lcell;                    // this axis's first entries in the compact table (below)
};

template <typename T>
struct StretchGrid {
    StretchAxis<T> ax[3];
    int rule;
    double lo[3], hi[3];            // the box, for the exit test
    double land_lo[3], land_hi[3];  // landing planes of min / max face exits
    double len_lim;                 // 2 x (sum of the three spans), :330-332
    // Compact copy of the float bin records for workgroup-local (LDS) lookup,
    // 8 bytes an entry: per bin {split bits, m_lo | m_hi << 16} (m_lo =
    // 0xffff for a bin holding several breakpoints), then per cell m
    // {a[m], 1 / (a[m+1] - a[m])} as float bits -- the a_lo/iw_lo, a_hi/iw_hi
    // of a BinRec are the cell values of its m_lo, m_hi.  compact_n = 0 when
    // an axis has more than 65535 cells.
    const uint2* compact;
    int compact_n;
};

// What the kernels take: the grid lives in device memory beside its tables
// (as a kernel argument its ~80 dwords would crowd the scalar registers).
template <typename T>
struct StretchRef {
    const StretchGrid<T>* g;
};

template <typename T>
struct CellWeight {
    int c;
    T w;
};

// A bin that holds several breakpoints (only where the bin count had to be
// capped): the uniform formula, or a search for the holding cell from the
// bin's lowest possible one followed by the rule's cell.
template <typename T>
FQSL_COLD CellWeight<T> stretch_search(const FQSL_CONST StretchAxis<T>* A, int b, int rule, T p, T phi, T plo) {
    CellWeight<T> o;
    if (rule == kLookupUniform) {  // the reference's division, so ties fall its way
        int k = to_int(flr((p - A->lo) / A->uni_width));
        k = k < 0 ? 0 : (k > A->n - 2 ? A->n - 2 : k);
        const auto r = A->rec[k];
        o.c = k;
        o.w = ((phi - r.x) + plo) * r.w;
        return o;
    }
    int k = A->bins[b];
    auto r = A->rec[k];
    while (k < A->n - 2 && p > r.y) { ++k; r = A->rec[k]; }
    while (k > 0 && !(p > r.x)) { --k; r = A->rec[k]; }
    o.c = (rule == kLookupExact) ? k : A->ref[k];
    o.w = ((phi - r.z) + plo) * r.w;
    return o;
}

// Cell and weight of the stage coordinate phi + plo along one axis (phi the
// leading part, plo a small remainder: the weight is formed as
// ((phi - a[m]) + plo) / width so it keeps the remainder's resolution).
template <typename T>
FQSL_HD void stretch_locate(const FQSL_CONST StretchAxis<T>& A, int rule, T phi, T plo, int& c, T& w) {
    T p = phi + plo;
    if (!(p > A.lo)) { p = A.clo; phi = A.clo; plo = (T)0; }
    if (!(p < A.hi)) { p = A.chi; phi = A.chi; plo = (T)0; }
    int b = to_int((p - A.lo) * A.bin_scale);
    b = b < 0 ? 0 : (b >= A.nbins ? A.nbins - 1 : b);
    const BinRec<T> br = load_binrec(A.brec + b);
    const bool up = p > br.split;  // on a node: the cell below it
    c = up ? br.m_hi : br.m_lo;
    w = ((phi - (up ? br.a_hi : br.a_lo)) + plo) * (up ? br.iw_hi : br.iw_lo);
    if (br.m_lo >= 0) return;
    const CellWeight<T> s = stretch_search<T>(&A, b, rule, p, phi, plo);
    c = s.c;
    w = s.w;
}

FQSL_HD float float_of_bits(unsigned u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// The workgroup's copy of StretchGrid<float>::compact (dynamic LDS), filled
// by the kernels that trace with kL = true before their first lookup.
extern __shared__ uint2 fqsl_stretch_lds[];

// stretch_locate<float> from the LDS copy: the same cell and weight bit for
// bit (the same record fields, the same expressions).
FQSL_HD void stretch_locate_lds(const FQSL_CONST StretchAxis<float>& A, int rule, float phi, float plo, int& c,
                                float& w) {
    float p = phi + plo;
    if (!(p > A.lo)) { p = A.clo; phi = A.clo; plo = 0.f; }
    if (!(p < A.hi)) { p = A.chi; phi = A.chi; plo = 0.f; }
    int b = to_int((p - A.lo) * A.bin_scale);
    b = b < 0 ? 0 : (b >= A.nbins ? A.nbins - 1 : b);
    const uint2 br = fqsl_stretch_lds[A.lbin + b];
    const int m_lo = (int)(br.y & 0xffffu), m_hi = (int)(br.y >> 16);
    const bool multi = m_lo == 0xffff;
    const bool up = p > float_of_bits(br.x);  // on a node: the cell below it
    c = (up || multi) ? m_hi : m_lo;
    const uint2 cw = fqsl_stretch_lds[A.lcell + c];
    w = ((phi - float_of_bits(cw.x)) + plo) * float_of_bits(cw.y);
    if (!multi) return;
    const CellWeight<float> s = stretch_search<float>(&A, b, rule, p, phi, plo);
    c = s.c;
    w = s.w;
}

// Leading part and remainder of a stage point x + delta.
FQSL_HD void stage_parts(const Pos<float>& x, const float* d, float* phi, float* plo) {
#pragma unroll
    for (int a = 0; a < 3; ++a) { phi[a] = x.hi[a]; plo[a] = x.lo[a] + d[a]; }
}
FQSL_HD void stage_parts(const Pos<double>& x, const double* d, double* phi, double* plo) {
#pragma unroll
    for (int a = 0; a < 3; ++a) { phi[a] = x.x[a]; plo[a] = d[a]; }
}
template <typename R>
FQSL_HD void point_parts(const double* x, R* phi, R* plo) {
#pragma unroll
    for (int a = 0; a < 3; ++a) { phi[a] = (R)x[a]; plo[a] = (R)(x[a] - (double)phi[a]); }
}

// Raw B (Interp3dxyzn(..., false), :82-194) at a split stage point.  kL:
// look the cells up in the workgroup's LDS copy of the tables (float only).
template <typename R, class F, bool kL = false>
FQSL_HD void stretch_raw(const F& f, const FQSL_CONST StretchGrid<R>& g, const R* phi, const R* plo, R B[3]) {
    int c[3];
    R w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if constexpr (kL) stretch_locate_lds(g.ax[a], g.rule, phi[a], plo[a], c[a], w[a]);
        else stretch_locate<R>(g.ax[a], g.rule, phi[a], plo[a], c[a], w[a]);
    }
    R cc[3][8];
    f.template load<R>(c[0], c[1], c[2], cc);
    field_value<F, R>(cc, w[0], w[1], w[2], B);
}

template <typename R, class F, bool kL = false>
FQSL_HD void stretch_unit(const F& f, const FQSL_CONST StretchGrid<R>& g, const R* phi, const R* plo, R b[3]) {
    R B[3];
    stretch_raw<R, F, kL>(f, g, phi, plo, B);
    const R inb = rsq(dot3(B, B));
#pragma unroll
    for (int q = 0; q < 3; ++q) b[q] = B[q] * inb;
}

template <typename R, class F, bool kL = false>
FQSL_HD void stretch_unit_at(const F& f, const FQSL_CONST StretchGrid<R>& g, const double* x, R b[3]) {
    R phi[3], plo[3];
    point_parts<R>(x, phi, plo);
    stretch_unit<R, F, kL>(f, g, phi, plo, b);
}

template <typename R, class F, bool kL = false>
FQSL_HD void stretch_raw_at(const F& f, const FQSL_CONST StretchGrid<R>& g, const double* x, double B[3]) {
    R phi[3], plo[3], Bv[3];
    point_parts<R>(x, phi, plo);
    stretch_raw<R, F, kL>(f, g, phi, plo, Bv);
    B[0] = Bv[0]; B[1] = Bv[1]; B[2] = Bv[2];
}

// checkFlag of the stretched tracer (:298-321): the box is [a0, a_{n-1}] per axis.
FQSL_HD int check_flag_box(const double* lo, const double* hi, const double* p) {
    const double X = p[0], Y = p[1], Z = p[2];
    const double x0 = lo[0], y0 = lo[1], z0 = lo[2], mx = hi[0], my = hi[1], mz = hi[2];
    if (X >= x0 && Y > y0 && Z >= z0 && X <= mx && Y <= my && Z <= mz) return 0;
    int fl = kClassOther;
    if (X < x0 && Y >= y0 && Z >= z0 && X < mx && Y < my && Z < mz) fl = 1;
    if (X >= x0 && Y >= y0 && Z >= z0 && X >= mx && Y < my && Z < mz) fl = 2;
    if (X >= x0 && Y < y0 && Z >= z0 && X < mx && Y < my && Z < mz) fl = 3;
    if (X >= x0 && Y >= y0 && Z >= z0 && X < mx && Y >= my && Z < mz) fl = 4;
    if (X >= x0 && Y >= y0 && Z < z0 && X < mx && Y < my && Z < mz) fl = 5;
    if (X >= x0 && Y >= y0 && Z >= z0 && X < mx && Y < my && Z >= mz) fl = 6;
    return fl;
}

// --------------------------------------------------------- line driver
// TraceBlineAdap of the stretched tracer (:323-419), RKF45 only (the
// reference has no other stretched integrator).  PlainLine keeps the state;
// differences from the uniform tracer: the box and landing planes above, no
// step-count limit, len_lim from the spans, and a landing twist term
// J.b/|b|^2 (b the normalised B at p1, :403-406) that the uniform RKF45
// tracer has commented out.
template <typename R, class F, bool kL = false>
FQSL_HD void stretch_init(PlainLine<R>& L, const F& f, const FQSL_CONST StretchGrid<R>& g, const PlainParams& P, float px,
                          float py, float pz) {
    L.p0[0] = px; L.p0[1] = py; L.p0[2] = pz;
    L.len_lim = g.len_lim;
    L.step_lim = INT_MAX;
    L.len_total = 0.0;
    L.tw = 0.0;
    L.evals = 0;
    L.done = false;
    L.flag_e = L.flag_s = kClassLimit;
    const double p0d[3] = {px, py, pz};
    R b0[3];
    stretch_unit_at<R, F, kL>(f, g, p0d, b0);
    const double bn = fabs((double)b0[0] * P.ncross[0] + (double)b0[1] * P.ncross[1] + (double)b0[2] * P.ncross[2]);
    L.tol = (bn <= 0.05 ? kTolBase / 8e3 : kTolBase * pow(bn, 1.5)) * P.tol_coef;  // :345-352
    L.dir = 1;
    plain_begin_dir(L, f, P);
}

// RK4_boundary (:274-296) on the stretched interpolant.
template <typename R, class F, bool kL = false>
FQSL_HD void stretch_rk4_boundary(const F& f, const FQSL_CONST StretchGrid<R>& g, const double* p, double s, int dim,
                                    double out[3]) {
    double k[4][3], t[3];
    const double cf[4] = {0.0, 0.5, 0.5, 1.0};
#pragma unroll
    for (int st = 0; st < 4; ++st) {
#pragma unroll
        for (int a = 0; a < 3; ++a) t[a] = st == 0 ? p[a] : p[a] + s * k[st - 1][a] * cf[st];
        R b[3];
        stretch_unit_at<R, F, kL>(f, g, t, b);
        const double bd = (double)b[dim];
#pragma unroll
        for (int a = 0; a < 3; ++a) k[st][a] = (double)b[a] / bd;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = p[a] + (1. / 6.) * s * (k[0][a] + 2.0 * k[1][a] + 2.0 * k[2][a] + k[3][a]);
}

// Stage offset sum_j a_{st,j} k_j of the Fehlberg tableau (st = 0..5).
template <typename R>
FQSL_HD void rkf45_offset(int st, const R k[6][3], R t[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        switch (st) {
            case 0: t[a] = (R)0; break;
            case 1: t[a] = (R)Tab::a21 * k[0][a]; break;
            case 2: t[a] = (R)Tab::a31 * k[0][a] + (R)Tab::a32 * k[1][a]; break;
            case 3: t[a] = (R)Tab::a41 * k[0][a] + (R)Tab::a42 * k[1][a] + (R)Tab::a43 * k[2][a]; break;
            case 4:
                t[a] = (R)Tab::a51 * k[0][a] + (R)Tab::a52 * k[1][a] + (R)Tab::a53 * k[2][a] + (R)Tab::a54 * k[3][a];
                break;
            default:
                t[a] = (R)Tab::a61 * k[0][a] + (R)Tab::a62 * k[1][a] + (R)Tab::a63 * k[2][a] +
                       (R)Tab::a64 * k[3][a] + (R)Tab::a65 * k[4][a];
        }
    }
}

template <typename R, class F, class FJ = NoCurrent, bool kL = false>
FQSL_HD void stretch_attempt(PlainLine<R>& L, const F& f, const FQSL_CONST StretchGrid<R>& g, const PlainParams& P,
                             const FJ& fj = FJ{}) {
    const R sc = (R)(L.s * L.dir);
    L.attempts += 1;
    // RKF45 (:200-241): stages on the normalised interpolant
    R k[6][3];
#pragma unroll
    for (int st = 0; st < 6; ++st) {
        R t[3];
        rkf45_offset<R>(st, k, t);
        R phi[3], plo[3];
        stage_parts(L.x, t, phi, plo);
        stretch_unit<R, F, kL>(f, g, phi, plo, k[st]);
#pragma unroll
        for (int a = 0; a < 3; ++a) k[st][a] *= sc;
    }
    R inc[3];
    double e2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        inc[a] = (R)Tab::b1 * k[0][a] + (R)Tab::b3 * k[2][a] + (R)Tab::b4 * k[3][a] + (R)Tab::b5 * k[4][a] +
                 (R)Tab::b6 * k[5][a];
        const double e = Tab::e1 * (double)k[0][a] + Tab::e3 * (double)k[2][a] + Tab::e4 * (double)k[3][a] +
                         Tab::e5 * (double)k[4][a] + Tab::e6 * (double)k[5][a];
        e2 += e * e;
    }
    L.evals += 6;
    const double scale = pow(L.tol / sqrt(e2) / 11.09, 0.2);  // :358-363
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
    const double xa[3] = {L.x.get(0), L.x.get(1), L.x.get(2)};
    Pos<R> xb = L.x;
    xb.add(inc);
    const double xbv[3] = {xb.get(0), xb.get(1), xb.get(2)};
    const double dl = sqrt((double)inc[0] * inc[0] + (double)inc[1] * inc[1] + (double)inc[2] * inc[2]);
    L.len += dl;
    // twist at the step's start point (:366-370)
    double tw_rate = 0.0, Jp1[3] = {0.0, 0.0, 0.0};
    if constexpr (has_current<FJ>()) {
        double Bp1[3];
        stretch_raw_at<R, FJ, kL>(fj, g, xa, Jp1);
        stretch_raw_at<R, F, kL>(f, g, xa, Bp1);
        tw_rate = dot3(Jp1, Bp1) / dot3(Bp1, Bp1) * kQuarterOverPi;
        L.tw += tw_rate * dl;
    }
    const int flag = check_flag_box(g.lo, g.hi, xbv);
    if (flag > 0) {
        L.len -= dl;
        if constexpr (has_current<FJ>()) L.tw -= tw_rate * dl;  // :373-377
        if (flag <= 6) {
            const int dim = (flag - 1) / 2;
            const double pa = xa[dim], pb = xbv[dim];
            if (fabs(pa - pb) > 1e-3) {  // :378-406
                const double pm = (flag % 2 == 1) ? g.land_lo[dim] : g.land_hi[dim];
                R b1[3], b2[3];
                stretch_unit_at<R, F, kL>(f, g, xa, b1);
                stretch_unit_at<R, F, kL>(f, g, xbv, b2);
                double out[3];
                if (fabs((double)b1[dim]) < 0.2 || fabs((double)b2[dim]) < 0.2) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) out[a] = (xa[a] * (pb - pm) + xbv[a] * (pm - pa)) / (pb - pa);
                } else {
                    stretch_rk4_boundary<R, F, kL>(f, g, xa, pm - pa, dim, out);
                    L.evals += 4;
                }
                plain_record(L, out, flag);
                const double arc = fabs(pm - pa) / (1e-4 + fabs((double)b1[dim]));
                L.len += arc;
                if constexpr (has_current<FJ>()) {
                    const double bd[3] = {(double)b1[0], (double)b1[1], (double)b1[2]};
                    L.tw += dot3(Jp1, bd) / dot3(bd, bd) * kQuarterOverPi * arc;
                }
            } else {
                plain_record(L, xa, flag);  // no landing segment: p_mid unset, no landing twist
            }
        } else {
            plain_record(L, xbv, flag);
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

// --------------------------------------------------------- host tables
// GetIndexBiSearch (:42-65) on the host, for the cell map of kLookupBisect.
inline int reference_bisect(double p, int n, const double* a) {
    int lo = 0, hi = n - 1, mid = n / 2;
    while (hi - lo > 3) {
        if (p > a[mid]) lo = mid;
        else hi = mid;
        mid = (lo + hi) / 2;
    }
    const int r = (p > a[mid]) ? mid : mid - 1;
    return r > n - 2 ? n - 2 : r;
}

struct AxisTables {
    std::vector<float4> rec32;
    std::vector<double4> rec64;
    std::vector<BinRec<float>> brec32;
    std::vector<BinRec<double>> brec64;
    std::vector<int> ref, bins;
    double lo = 0, hi = 0;
    int n = 0, nbins = 0, multi = 0;  // multi: bins holding more than one interior node
};

constexpr int kMaxStretchBins = 1 << 16;

// The uniform rule's cell of p evaluated in T as the reference does:
// floor((p - a0) / w) with w = (a_{n-1} - a0) / n in T, clamped to [0, n-2].
template <typename T>
inline int uniform_cell(T p, T a0, T w, int n) {
    const T r = std::floor((p - a0) / w);
    if (!(r > (T)0)) return 0;
    return r > (T)(n - 2) ? n - 2 : (int)r;
}

// Finite T values in their numeric order as integers (+0 and -0 both map to
// 0), and back.
template <typename T>
inline long long ordinal_of(T x) {
    if (sizeof(T) == 4) {
        int i;
        memcpy(&i, &x, 4);
        return i >= 0 ? (long long)i : -(long long)(i & 0x7fffffff);
    }
    long long i;
    memcpy(&i, &x, 8);
    return i >= 0 ? i : -(i & 0x7fffffffffffffffLL);
}

template <typename T>
inline T from_ordinal(long long o) {
    T x;
    if (sizeof(T) == 4) {
        const unsigned u = o >= 0 ? (unsigned)o : ((unsigned)(-o) | 0x80000000u);
        memcpy(&x, &u, 4);
    } else {
        const unsigned long long u = o >= 0 ? (unsigned long long)o : ((unsigned long long)(-o) | (1ULL << 63));
        memcpy(&x, &u, 8);
    }
    return x;
}

// The largest T value whose uniform-rule cell is below j (the split of
// breakpoint j): a point above it takes cell j.  The cell is monotone in p,
// so this bisects over the ordered values between a bracket around the
// guess (at most 64 halvings, wherever the split lies, zero included).
template <typename T>
inline T uniform_split(int j, T a0, T w, int n, double guess) {
    const double step = 2.0 * (double)w;
    double dl = guess - step, dh = guess + step;
    while (uniform_cell<T>((T)dl, a0, w, n) >= j) dl -= (guess - dl);
    while (uniform_cell<T>((T)dh, a0, w, n) < j) dh += (dh - guess);
    long long lo = ordinal_of<T>((T)dl), hi = ordinal_of<T>((T)dh);
    // the gap between ordinals of opposite signs can exceed LLONG_MAX
    for (unsigned long long gap; (gap = (unsigned long long)hi - (unsigned long long)lo) > 1;) {
        const long long mid = lo + (long long)(gap / 2);
        if (uniform_cell<T>(from_ordinal<T>(mid), a0, w, n) >= j) hi = mid;
        else lo = mid;
    }
    return from_ordinal<T>(lo);
}

// Bin records of one precision from the breakpoints split[j - 1] (between
// cells j - 1 and j, ascending) and the rule's cell map.
template <typename T>
inline int fill_bins(const std::vector<T>& split, const std::vector<int>& ref, const double* a, int n, double lo,
                     double span, int nbins, std::vector<BinRec<T>>& out) {
    out.resize(nbins);
    const int ns = (int)split.size(), nc = n - 1;
    int below = 0, multi = 0;
    for (int b = 0; b < nbins; ++b) {
        const double e = lo + span * ((double)b / (double)nbins);
        const double e1 = (b + 1 == nbins) ? lo + 2.0 * span : lo + span * ((double)(b + 1) / (double)nbins);
        while (below < ns && (double)split[below] < e) ++below;
        int inside = 0;
        while (below + inside < ns && (double)split[below + inside] < e1) ++inside;
        const int lo_c = below, hi_c = below + 1 < nc ? below + 1 : nc - 1;
        const int m_lo = ref[lo_c], m_hi = ref[hi_c];
        BinRec<T> r;
        r.split = inside == 1 ? split[below] : (T)HUGE_VAL;
        r.a_lo = (T)a[m_lo];
        r.iw_lo = (T)(1.0 / (a[m_lo + 1] - a[m_lo]));
        r.a_hi = (T)a[m_hi];
        r.iw_hi = (T)(1.0 / (a[m_hi + 1] - a[m_hi]));
        r.m_lo = inside > 1 ? -1 : m_lo;  // several breakpoints: search
        r.m_hi = m_hi;
        out[b] = r;
        multi += inside > 1 ? 1 : 0;
    }
    return multi;
}

// Tables of one strictly increasing axis of n >= 2 finite nodes; returns an
// error message or nullptr.
inline const char* build_axis_tables(const double* a, int n, int rule, AxisTables& t) {
    if (n < 2) return "an axis needs at least 2 nodes";
    for (int i = 0; i < n; ++i) {
        if (!(a[i] == a[i]) || a[i] - a[i] != 0.0) return "axis coordinates must be finite";
        if (i > 0 && !(a[i] - a[i - 1] > 0.0)) return "axis coordinates must be strictly increasing";
    }
    t.n = n;
    t.lo = a[0];
    t.hi = a[n - 1];
    const double span = t.hi - t.lo;
    const int nc = n - 1;
    t.ref.resize(nc);
    t.rec32.resize(nc);
    t.rec64.resize(nc);
    for (int c = 0; c < nc; ++c) {
        int m = c;
        if (rule == kLookupBisect) m = reference_bisect(0.5 * (a[c] + a[c + 1]), n, a);
        t.ref[c] = m;
        const double iw = 1.0 / (a[m + 1] - a[m]);
        t.rec64[c] = make_double4(a[c], a[c + 1], a[m], iw);
        t.rec32[c] = make_float4((float)a[c], (float)a[c + 1], (float)a[m], (float)iw);
    }
    // breakpoints: interior nodes, or where the uniform formula steps
    std::vector<double> s64(n - 2);
    std::vector<float> s32(n - 2);
    const double w64 = (a[n - 1] - a[0]) / (double)n;
    const float a0f = (float)a[0], w32 = ((float)a[n - 1] - a0f) / (float)n;
    for (int j = 1; j <= n - 2; ++j) {
        if (rule == kLookupUniform) {
            const double guess = a[0] + j * w64;
            s64[j - 1] = uniform_split<double>(j, a[0], w64, n, guess);
            s32[j - 1] = uniform_split<float>(j, a0f, w32, n, guess);
        } else {
            s64[j - 1] = a[j];
            s32[j - 1] = (float)a[j];
        }
    }
    // bins narrower than the closest two breakpoints
    double dmin = span;
    for (int j = 1; j < n - 2; ++j) {
        const double d64 = s64[j] - s64[j - 1], d32 = (double)s32[j] - (double)s32[j - 1];
        if (d64 > 0.0 && d64 < dmin) dmin = d64;
        if (d32 > 0.0 && d32 < dmin) dmin = d32;
    }
    double want = 1.25 * span / dmin + 2.0;
    if (want < n) want = n;
    t.nbins = want > kMaxStretchBins ? kMaxStretchBins : (int)want;
    // bins[b]: the cell of the bin's lower edge e (largest c with a[c] < e, or 0)
    t.bins.resize(t.nbins);
    int c = 0;
    for (int b = 0; b < t.nbins; ++b) {
        const double e = t.lo + span * ((double)b / (double)t.nbins);
        while (c < nc - 1 && a[c + 1] < e) ++c;
        t.bins[b] = c;
    }
    const int m64 = fill_bins<double>(s64, t.ref, a, n, t.lo, span, t.nbins, t.brec64);
    const int m32 = fill_bins<float>(s32, t.ref, a, n, t.lo, span, t.nbins, t.brec32);
    t.multi = m64 > m32 ? m64 : m32;
    return nullptr;
}

// Entries of the compact table (StretchGrid::compact) for three axes: the
// float bin records, then the float cell records; lbin/lcell get each axis's
// first entry.  Returns false when an axis has more than 65535 cells.
inline bool build_compact(const AxisTables (&t)[3], std::vector<uint2>& out, int (&lbin)[3], int (&lcell)[3]) {
    out.clear();
    for (int q = 0; q < 3; ++q)
        if (t[q].n - 1 > 0xfffe) return false;
    for (int q = 0; q < 3; ++q) {
        lbin[q] = (int)out.size();
        for (const BinRec<float>& r : t[q].brec32) {
            unsigned split;
            memcpy(&split, &r.split, 4);
            const unsigned lo = r.m_lo < 0 ? 0xffffu : (unsigned)r.m_lo;
            out.push_back(make_uint2(split, lo | ((unsigned)r.m_hi << 16)));
        }
    }
    for (int q = 0; q < 3; ++q) {
        lcell[q] = (int)out.size();
        for (int m = 0; m + 1 < t[q].n; ++m) {
            // the float values fill_bins stores as a_lo / iw_lo for cell m
            const float a = (float)t[q].rec64[m].x;
            const double a0 = t[q].rec64[m].x, a1 = t[q].rec64[m].y;
            const float iw = (float)(1.0 / (a1 - a0));
            unsigned ua, uw;
            memcpy(&ua, &a, 4);
            memcpy(&uw, &iw, 4);
            out.push_back(make_uint2(ua, uw));
        }
    }
    return true;
}

// The grid view over tables placed at brec[q] (BinRec<T> per bin of axis q),
// rec[q] (float4 or double4 per cell), ref[q] and bins[q]: clamp bounds, bin and cell scales, the box,
// the landing planes (0 / n for the reference rules, the faces for the exact
// rule) and len_lim.
template <typename T>
inline void fill_grid(const AxisTables (&t)[3], int rule, const void* const (&brec)[3], const void* const (&rec)[3],
                      const int* const (&ref)[3], const int* const (&bins)[3], StretchGrid<T>& g) {
    double span = 0.0;
    for (int q = 0; q < 3; ++q) {
        const double lo = t[q].lo, hi = t[q].hi;
        span += hi - lo;
        StretchAxis<T>& A = g.ax[q];
        A.brec = (const BinRec<T>*)brec[q];
        A.rec = (const typename Vec4<T>::type*)rec[q];
        A.ref = ref[q];
        A.bins = bins[q];
        A.lo = (T)lo;
        A.hi = (T)hi;
        A.clo = (T)(lo + 1e-6);
        A.chi = (T)(hi - 1e-6);
        A.bin_scale = (T)(t[q].nbins / (hi - lo));
        A.uni_width = ((T)hi - (T)lo) / (T)t[q].n;
        A.n = t[q].n;
        A.nbins = t[q].nbins;
        g.lo[q] = lo;
        g.hi[q] = hi;
        g.land_lo[q] = rule == kLookupExact ? lo : 0.0;
        g.land_hi[q] = rule == kLookupExact ? hi : (double)t[q].n;
    }
    g.rule = rule;
    g.len_lim = 2.0 * span;  // MAX_STEP_RATIO, TraceBlineAdaptiveStretch.cu:11,330-332
}

}  // namespace fqsl
