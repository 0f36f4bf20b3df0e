// Test-only CPU build of fastqsl_amd/csrc/fqsl_tracer.h and fqsl_stretch.h: runs the per-line
// tracer code of the HIP kernels on the host so that its logic (the exact
// difference-quotient stencil, the compensated position, the driver) can be
// compared with the FP64 oracle without a GPU.  Never part of the product.
#include <string.h>

#include <vector>

#include "fqsl_stretch.h"
#include "fqsl_tracer.h"

using namespace fqsl;

namespace {
// host copy of the cell layout (trilinear coefficients)
std::vector<float4> build_cells(const float* bx, const float* by, const float* bz, int nx, int ny, int nz) {
    const long cx = nx - 1, cy = ny - 1, cz = nz - 1, nxy = (long)nx * ny;
    std::vector<float4> cells(8 * cx * cy * cz);
    const float* comp[3] = {bx, by, bz};
    for (long t = 0; t < cx * cy * cz; ++t) {
        const long i = t % cx, j = (t / cx) % cy, k = t / (cx * cy);
        const long base = k * nxy + j * nx + i;
        for (int q = 0; q < 3; ++q) {
            float v[8], c8[8];
            for (int c = 0; c < 8; ++c) v[c] = comp[q][base + (c & 1) + ((c >> 1) & 1) * nx + ((c >> 2) & 1) * nxy];
            cell_coefs(v, c8);
            cells[8 * t + 2 * q] = make_float4(c8[0], c8[1], c8[2], c8[3]);
            cells[8 * t + 2 * q + 1] = make_float4(c8[4], c8[5], c8[6], c8[7]);
        }
        cells[8 * t + 6] = make_float4(0, 0, 0, 0);
        cells[8 * t + 7] = make_float4(0, 0, 0, 0);
    }
    return cells;
}

template <typename R, class F, class FJ = NoCurrent, typename X = R>
void run_scott(const F& f, const ScottParams& P, long n, const float* seeds, double* out_start, double* out_end,
               int* fs, int* fe, double* q0, double* qp, double* len, long* evals, const FJ& fj = FJ{},
               double* twist = nullptr) {
    for (long i = 0; i < n; ++i) {
        ScottLine<R, X> L;
        ScottEnd<R> e;
        int st = scott_init<R>(L, f, P, seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2], e);
        while (st == kTracing) st = scott_attempt<R, F, FJ>(L, f, P, e, fj);
        if (twist) twist[i] = L.tw;
        double nb0 = 0.0;
        {
            R u0[3], v0[3], b0z;
            seed_frame<R>(f, L.p0, u0, v0, &b0z, &nb0);
            nb0 = (double)(R)nb0;
        }
        scott_q<R>(L.st, e, st, nb0, f, &q0[i], &qp[i]);
        for (int a = 0; a < 3; ++a) { out_start[3 * i + a] = L.st.x[a]; out_end[3 * i + a] = e.x[a]; }
        fs[i] = L.st.flag; fe[i] = e.flag; len[i] = L.len; evals[i] = L.evals;
    }
}
}  // namespace

extern "C" {
// precision 0: float + cell layout from the float field; 1: double SoA;
// 2: the mixed variant (FP64 position path, FP32 Jacobian) on the cell layout.
// jx/jy/jz (may be NULL): current density for the twist output.
int hc_trace_scott(const double* bx, const double* by, const double* bz, int nx, int ny, int nz, int precision,
                   int method, double s0, double tol, long n, const float* seeds, double* start, double* end,
                   int* fs, int* fe, double* q0, double* qp, double* len, long* evals, const double* jx,
                   const double* jy, const double* jz, double* twist) {
    ScottParams P{method, s0, tol, scott_ctl(tol)};
    const long nn = (long)nx * ny * nz;
    if (precision == 1) {
        FieldSoA<double> f{bx, by, bz, {nx, ny, nz}};
        if (jx) {
            FieldSoA<double> fj{jx, jy, jz, {nx, ny, nz}};
            run_scott<double>(f, P, n, seeds, start, end, fs, fe, q0, qp, len, evals, fj, twist);
        } else {
            run_scott<double>(f, P, n, seeds, start, end, fs, fe, q0, qp, len, evals);
        }
        return 0;
    }
    std::vector<float> fx(bx, bx + nn), fy(by, by + nn), fz(bz, bz + nn);
    std::vector<float4> cells = build_cells(fx.data(), fy.data(), fz.data(), nx, ny, nz);
    FieldCell32 f{cells.data(), {nx, ny, nz}};
    if (jx) {
        std::vector<float> gx(jx, jx + nn), gy(jy, jy + nn), gz(jz, jz + nn);
        FieldSoA<float> fj{gx.data(), gy.data(), gz.data(), {nx, ny, nz}};
        if (precision == 2)
            run_scott<float, FieldCell32, FieldSoA<float>, double>(f, P, n, seeds, start, end, fs, fe, q0, qp, len,
                                                                   evals, fj, twist);
        else
            run_scott<float>(f, P, n, seeds, start, end, fs, fe, q0, qp, len, evals, fj, twist);
    } else if (precision == 2) {
        run_scott<float, FieldCell32, NoCurrent, double>(f, P, n, seeds, start, end, fs, fe, q0, qp, len, evals);
    } else {
        run_scott<float>(f, P, n, seeds, start, end, fs, fe, q0, qp, len, evals);
    }
    return 0;
}

// The FP32 path's step-controller power e2^(-1/10) at n values.
void hc_pow_minus_tenth(long n, const double* e2, double* out) {
    for (long i = 0; i < n; ++i) out[i] = pow_minus_tenth(e2[i]);
}

// grad b by the stencil formula at n points (double SoA); out[9*i + 3*j + c].
void hc_grad_b(const double* bx, const double* by, const double* bz, int nx, int ny, int nz, long n,
               const double* pts, double* out) {
    FieldSoA<double> f{bx, by, bz, {nx, ny, nz}};
    for (long i = 0; i < n; ++i) {
        const Pt<double> p = pt_from_double<double>(pts + 3 * i);
        double b[3], J[3][3];
        scott_eval<double>(f, p, b, J);
        memcpy(out + 9 * i, J, sizeof(J));
    }
}

void hc_trace_plain(const double* bx, const double* by, const double* bz, int nx, int ny, int nz, int precision,
                    int method, double s0, double tol_coef, const float* ncross, long n, const float* seeds,
                    double* start, double* end, int* fs, int* fe, double* len, long* evals) {
    PlainParams P;
    P.method = method; P.s0 = s0; P.tol_coef = tol_coef;
    P.ncross[0] = ncross[0]; P.ncross[1] = ncross[1]; P.ncross[2] = ncross[2];
    const long nn = (long)nx * ny * nz;
    std::vector<float> fx(bx, bx + nn), fy(by, by + nn), fz(bz, bz + nn);
    std::vector<float4> cells = build_cells(fx.data(), fy.data(), fz.data(), nx, ny, nz);
    FieldSoA<double> fd{bx, by, bz, {nx, ny, nz}};
    FieldCell32 fc{cells.data(), {nx, ny, nz}};
    for (long i = 0; i < n; ++i) {
        if (precision == 1) {
            PlainLine<double> L;
            plain_init<double>(L, fd, P, seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]);
            while (!L.done) plain_attempt<double>(L, fd, P);
            for (int a = 0; a < 3; ++a) { start[3 * i + a] = L.xs[a]; end[3 * i + a] = L.xe[a]; }
            fs[i] = L.flag_s; fe[i] = L.flag_e; len[i] = L.len_total; evals[i] = L.evals;
        } else {
            PlainLine<float> L;
            plain_init<float>(L, fc, P, seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]);
            while (!L.done) plain_attempt<float>(L, fc, P);
            for (int a = 0; a < 3; ++a) { start[3 * i + a] = L.xs[a]; end[3 * i + a] = L.xe[a]; }
            fs[i] = L.flag_s; fe[i] = L.flag_e; len[i] = L.len_total; evals[i] = L.evals;
        }
    }
}

// Cell and weight along one stretched axis by the device lookup (bin table,
// cell records, rule map) at npts points; precision 1 = double tables, 0 =
// float tables.
int hc_stretch_cell(const double* a, int n, int rule, int precision, long npts, const double* p, int* cell,
                    double* w) {
    AxisTables t[3];
    for (int q = 0; q < 3; ++q)
        if (build_axis_tables(a, n, rule, t[q])) return -1;
    const void* r64[3] = {t[0].rec64.data(), t[1].rec64.data(), t[2].rec64.data()};
    const void* r32[3] = {t[0].rec32.data(), t[1].rec32.data(), t[2].rec32.data()};
    const int* ref[3] = {t[0].ref.data(), t[1].ref.data(), t[2].ref.data()};
    const int* bins[3] = {t[0].bins.data(), t[1].bins.data(), t[2].bins.data()};
    StretchGrid<double> g64;
    StretchGrid<float> g32;
    const void* b64[3] = {t[0].brec64.data(), t[1].brec64.data(), t[2].brec64.data()};
    const void* b32[3] = {t[0].brec32.data(), t[1].brec32.data(), t[2].brec32.data()};
    fill_grid<double>(t, rule, b64, r64, ref, bins, g64);
    fill_grid<float>(t, rule, b32, r32, ref, bins, g32);
    for (long i = 0; i < npts; ++i) {
        if (precision == 1) {
            double wd;
            stretch_locate<double>(g64.ax[0], rule, p[i], 0.0, cell[i], wd);
            w[i] = wd;
        } else {
            const float hi = (float)p[i], lo = (float)(p[i] - (double)hi);
            float wf;
            stretch_locate<float>(g32.ax[0], rule, hi, lo, cell[i], wf);
            w[i] = wf;
        }
    }
    return 0;
}

// The stretched tracer of the kernels on the host; precision 1 = FP64 over
// the double SoA field, 0 = FP32c over the coefficient cells.
int hc_trace_stretch(const double* bx, const double* by, const double* bz, int nx, int ny, int nz,
                     const double* xa, const double* ya, const double* za, int rule, int precision, double s0,
                     double tol_coef, const float* ncross, long n, const float* seeds, double* start, double* end,
                     int* fs, int* fe, double* len, long* evals, const double* jx, const double* jy,
                     const double* jz, double* twist) {
    const int nn3[3] = {nx, ny, nz};
    const double* ax[3] = {xa, ya, za};
    AxisTables t[3];
    for (int q = 0; q < 3; ++q)
        if (build_axis_tables(ax[q], nn3[q], rule, t[q])) return -1;
    const void* r64[3] = {t[0].rec64.data(), t[1].rec64.data(), t[2].rec64.data()};
    const void* r32[3] = {t[0].rec32.data(), t[1].rec32.data(), t[2].rec32.data()};
    const int* ref[3] = {t[0].ref.data(), t[1].ref.data(), t[2].ref.data()};
    const int* bins[3] = {t[0].bins.data(), t[1].bins.data(), t[2].bins.data()};
    StretchGrid<double> g64;
    StretchGrid<float> g32;
    const void* b64[3] = {t[0].brec64.data(), t[1].brec64.data(), t[2].brec64.data()};
    const void* b32[3] = {t[0].brec32.data(), t[1].brec32.data(), t[2].brec32.data()};
    fill_grid<double>(t, rule, b64, r64, ref, bins, g64);
    fill_grid<float>(t, rule, b32, r32, ref, bins, g32);
    PlainParams P;
    P.method = kRKF45; P.s0 = s0; P.tol_coef = tol_coef;
    P.ncross[0] = ncross[0]; P.ncross[1] = ncross[1]; P.ncross[2] = ncross[2];
    const long nn = (long)nx * ny * nz;
    std::vector<float> fx(bx, bx + nn), fy(by, by + nn), fz(bz, bz + nn);
    std::vector<float4> cells = build_cells(fx.data(), fy.data(), fz.data(), nx, ny, nz);
    FieldSoA<double> fd{bx, by, bz, {nx, ny, nz}};
    FieldCell32 fc{cells.data(), {nx, ny, nz}};
    std::vector<float> cx, cy, cz;
    if (jx) { cx.assign(jx, jx + nn); cy.assign(jy, jy + nn); cz.assign(jz, jz + nn); }
    FieldSoA<double> jd{jx, jy, jz, {nx, ny, nz}};
    FieldSoA<float> jf{cx.data(), cy.data(), cz.data(), {nx, ny, nz}};
    for (long i = 0; i < n; ++i) {
        const float* s = seeds + 3 * i;
        if (precision == 1) {
            PlainLine<double> L;
            stretch_init<double>(L, fd, g64, P, s[0], s[1], s[2]);
            while (!L.done) {
                if (jx) stretch_attempt<double, FieldSoA<double>, FieldSoA<double>>(L, fd, g64, P, jd);
                else stretch_attempt<double>(L, fd, g64, P);
            }
            for (int a = 0; a < 3; ++a) { start[3 * i + a] = L.xs[a]; end[3 * i + a] = L.xe[a]; }
            fs[i] = L.flag_s; fe[i] = L.flag_e; len[i] = L.len_total; evals[i] = L.evals;
            if (twist) twist[i] = L.tw;
        } else {
            PlainLine<float> L;
            stretch_init<float>(L, fc, g32, P, s[0], s[1], s[2]);
            while (!L.done) {
                if (jx) stretch_attempt<float, FieldCell32, FieldSoA<float>>(L, fc, g32, P, jf);
                else stretch_attempt<float>(L, fc, g32, P);
            }
            for (int a = 0; a < 3; ++a) { start[3 * i + a] = L.xs[a]; end[3 * i + a] = L.xe[a]; }
            fs[i] = L.flag_s; fe[i] = L.flag_e; len[i] = L.len_total; evals[i] = L.evals;
            if (twist) twist[i] = L.tw;
        }
    }
    return 0;
}
}
