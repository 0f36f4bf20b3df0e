// fqsl_kernels.hip -- gfx950 kernels and the C ABI of libfqsl.so.
//
// Kernels:
//   k_build_cells          SoA float field -> 128-byte cells of trilinear coefficients
//   k_scott_persistent     Scott tracer, persistent grid, per-lane refill
//                          from a device work queue (one line per lane)
//   k_scott_per_line       Scott tracer, one thread per line
//   k_plain_persistent / k_plain_per_line   position-only tracer (grid-index
//                          coordinates, or a stretched grid's own coordinates)
//   k_qcalc_plane          QCalcPlane (FastQSL.py:65-117)
//   k_pseudo_*             CookPseudoLine compaction (FastQSL.py:4-57)
//   k_qcut_scatter         Qcut + merge (demo/Q_RKF45CuPy.ipynb cells [10]-[11])
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "fqsl.h"
#include "fqsl_stretch.h"
#include "fqsl_tracer.h"

using namespace fqsl;

// ------------------------------------------------------------ error state
static thread_local char g_err[512] = "";
static thread_local hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;  // whole trace
static thread_local hipEvent_t g_evs0 = nullptr, g_evs1 = nullptr;  // stepping kernel only
static thread_local bool g_ev_valid = false;
static thread_local bool g_evs_valid = false;

static int set_err(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_err(FQSL_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

extern "C" const char* fqsl_last_error(void) { return g_err; }
extern "C" const char* fqsl_version(void) { return "fastqsl_amd 0.1.0 (gfx950)"; }

// ------------------------------------------------------------------- field
struct fqsl_field {
    int n[3];
    int device;
    float* soa32[3];
    double* soa64[3];
    float4* cell32;
    float* cur32[3];   // current density J (twist), float and double SoA
    double* cur64[3];
    float4* curcell;   // J in the cell layout (with a cell32 field): the FP32 twist reads it
    size_t bytes;
    // stretched grid: node coordinates per axis and their lookup tables
    void* axes_mem;
    size_t axes_bytes;
    bool stretched;
    const StretchGrid<float>* sg32;   // device copies, beside the tables
    const StretchGrid<double>* sg64;
    int sg_lds_bytes;                 // bytes of sg32's compact table (0: none)
};

__global__ void k_convert_to_f32(const double* in, float* out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}
__global__ void k_convert_to_f64(const float* in, double* out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

// One thread per cell: gather the 8 corners of the 3 components into 8 float4.
__global__ void k_build_cells(const float* __restrict__ bx, const float* __restrict__ by,
                              const float* __restrict__ bz, int nx, int ny, int nz, float4* __restrict__ cells) {
    const long cx = nx - 1, cy = ny - 1, cz = nz - 1;
    const long ncell = cx * cy * cz;
    const long nxy = (long)nx * ny;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
        const long i = t % cx, j = (t / cx) % cy, k = t / (cx * cy);
        const long base = k * nxy + j * nx + i;
        const float* comp[3] = {bx, by, bz};
        float4* o = cells + 8 * t;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float* b = comp[q];
            float v[8], c8[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) v[c] = b[base + (c & 1) + ((c >> 1) & 1) * nx + ((c >> 2) & 1) * nxy];
            cell_coefs(v, c8);
            o[2 * q] = make_float4(c8[0], c8[1], c8[2], c8[3]);
            o[2 * q + 1] = make_float4(c8[4], c8[5], c8[6], c8[7]);
        }
        o[6] = make_float4(0.f, 0.f, 0.f, 0.f);
        o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

static int grid_for(long n, int block) {
    long g = (n + block - 1) / block;
    if (g > 65535L * 16) g = 65535L * 16;
    if (g < 1) g = 1;
    return (int)g;
}

static void field_free(fqsl_field* f) {
    for (int q = 0; q < 3; ++q) {
        if (f->soa32[q]) (void)hipFree(f->soa32[q]);
        if (f->soa64[q]) (void)hipFree(f->soa64[q]);
        if (f->cur32[q]) (void)hipFree(f->cur32[q]);
        if (f->cur64[q]) (void)hipFree(f->cur64[q]);
    }
    if (f->cell32) (void)hipFree(f->cell32);
    if (f->curcell) (void)hipFree(f->curcell);
    if (f->axes_mem) (void)hipFree(f->axes_mem);
    delete f;
}

extern "C" int fqsl_field_create(const void* bx, const void* by, const void* bz, int nx, int ny, int nz,
                                 int dtype, int on_device, int layouts, fqsl_field** out) {
    if (!out) return set_err(FQSL_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!bx || !by || !bz) return set_err(FQSL_ERR_INVALID, "field component pointer is NULL");
    if (nx < 2 || ny < 2 || nz < 2) return set_err(FQSL_ERR_INVALID, "each axis needs >= 2 nodes (%d,%d,%d)", nx, ny, nz);
    if (dtype != FQSL_DTYPE_F32 && dtype != FQSL_DTYPE_F64) return set_err(FQSL_ERR_INVALID, "dtype %d", dtype);
    if ((layouts & FQSL_LAYOUT_ALL) == 0) return set_err(FQSL_ERR_INVALID, "no layout requested");
    const long nn = (long)nx * ny * nz;
    if (nn > (1L << 31) - 1) return set_err(FQSL_ERR_INVALID, "field of %ld nodes exceeds int32 indexing", nn);
    fqsl_field* f = new fqsl_field();
    memset(f, 0, sizeof(*f));
    f->n[0] = nx; f->n[1] = ny; f->n[2] = nz;
    HIP_TRY(hipGetDevice(&f->device));
    const void* src[3] = {bx, by, bz};
    const size_t esz = dtype == FQSL_DTYPE_F64 ? 8 : 4;
    // stage the input on the device in its own dtype
    void* staged[3] = {nullptr, nullptr, nullptr};
    int rc = FQSL_OK;
    auto fail = [&](int code, const char* what, hipError_t e) {
        for (int q = 0; q < 3; ++q)
            if (staged[q] && !on_device) (void)hipFree(staged[q]);
        field_free(f);
        return set_err(code, "%s: %s", what, hipGetErrorString(e));
    };
    for (int q = 0; q < 3; ++q) {
        if (on_device) { staged[q] = const_cast<void*>(src[q]); continue; }
        hipError_t e = hipMalloc(&staged[q], nn * esz);
        if (e != hipSuccess) return fail(FQSL_ERR_NOMEM, "hipMalloc(field staging)", e);
        e = hipMemcpy(staged[q], src[q], nn * esz, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(FQSL_ERR_HIP, "hipMemcpy(field)", e);
    }
    const bool need32 = (layouts & (FQSL_LAYOUT_SOA32 | FQSL_LAYOUT_CELL32)) != 0;
    for (int q = 0; q < 3; ++q) {
        if (need32) {
            hipError_t e = hipMalloc(&f->soa32[q], nn * 4);
            if (e != hipSuccess) return fail(FQSL_ERR_NOMEM, "hipMalloc(soa32)", e);
            if (dtype == FQSL_DTYPE_F32) e = hipMemcpy(f->soa32[q], staged[q], nn * 4, hipMemcpyDeviceToDevice);
            else {
                k_convert_to_f32<<<grid_for(nn, 256), 256>>>((const double*)staged[q], f->soa32[q], nn);
                e = hipGetLastError();
            }
            if (e != hipSuccess) return fail(FQSL_ERR_HIP, "soa32 conversion", e);
            f->bytes += nn * 4;
        }
        if (layouts & FQSL_LAYOUT_SOA64) {
            hipError_t e = hipMalloc(&f->soa64[q], nn * 8);
            if (e != hipSuccess) return fail(FQSL_ERR_NOMEM, "hipMalloc(soa64)", e);
            if (dtype == FQSL_DTYPE_F64) e = hipMemcpy(f->soa64[q], staged[q], nn * 8, hipMemcpyDeviceToDevice);
            else {
                k_convert_to_f64<<<grid_for(nn, 256), 256>>>((const float*)staged[q], f->soa64[q], nn);
                e = hipGetLastError();
            }
            if (e != hipSuccess) return fail(FQSL_ERR_HIP, "soa64 conversion", e);
            f->bytes += nn * 8;
        }
    }
    if (layouts & FQSL_LAYOUT_CELL32) {
        const long ncell = (long)(nx - 1) * (ny - 1) * (nz - 1);
        hipError_t e = hipMalloc(&f->cell32, ncell * 128);
        if (e != hipSuccess) return fail(FQSL_ERR_NOMEM, "hipMalloc(cell32)", e);
        k_build_cells<<<grid_for(ncell, 256), 256>>>(f->soa32[0], f->soa32[1], f->soa32[2], nx, ny, nz, f->cell32);
        e = hipGetLastError();
        if (e != hipSuccess) return fail(FQSL_ERR_HIP, "k_build_cells", e);
        f->bytes += ncell * 128;
    }
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(FQSL_ERR_HIP, "field build", e);
    for (int q = 0; q < 3; ++q)
        if (!on_device && staged[q]) (void)hipFree(staged[q]);
    if (!(layouts & FQSL_LAYOUT_SOA32) && (layouts & FQSL_LAYOUT_CELL32)) {
        // the float SoA copy was only needed to build the cells
        for (int q = 0; q < 3; ++q) { (void)hipFree(f->soa32[q]); f->soa32[q] = nullptr; f->bytes -= nn * 4; }
    }
    (void)rc;
    *out = f;
    return FQSL_OK;
}

extern "C" int fqsl_field_destroy(fqsl_field* f) {
    if (!f) return FQSL_OK;
    field_free(f);
    return FQSL_OK;
}

extern "C" size_t fqsl_field_bytes(const fqsl_field* f) { return f ? f->bytes : 0; }

extern "C" int fqsl_field_set_current(fqsl_field* f, const void* jx, const void* jy, const void* jz, int dtype,
                                      int on_device) {
    if (!f) return set_err(FQSL_ERR_INVALID, "field is NULL");
    if (!jx || !jy || !jz) return set_err(FQSL_ERR_INVALID, "current component pointer is NULL");
    if (dtype != FQSL_DTYPE_F32 && dtype != FQSL_DTYPE_F64) return set_err(FQSL_ERR_INVALID, "dtype %d", dtype);
    const long nn = (long)f->n[0] * f->n[1] * f->n[2];
    const size_t esz = dtype == FQSL_DTYPE_F64 ? 8 : 4;
    const void* src[3] = {jx, jy, jz};
    const long ncell = (long)(f->n[0] - 1) * (f->n[1] - 1) * (f->n[2] - 1);
    for (int q = 0; q < 3; ++q) {
        if (f->cur32[q]) { (void)hipFree(f->cur32[q]); f->cur32[q] = nullptr; f->bytes -= nn * 4; }
        if (f->cur64[q]) { (void)hipFree(f->cur64[q]); f->cur64[q] = nullptr; f->bytes -= nn * 8; }
    }
    if (f->curcell) { (void)hipFree(f->curcell); f->curcell = nullptr; f->bytes -= ncell * 128; }
    for (int q = 0; q < 3; ++q) {
        void* staged = const_cast<void*>(src[q]);
        if (!on_device) {
            HIP_TRY(hipMalloc(&staged, nn * esz));
            hipError_t e = hipMemcpy(staged, src[q], nn * esz, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                (void)hipFree(staged);
                return set_err(FQSL_ERR_HIP, "hipMemcpy(current): %s", hipGetErrorString(e));
            }
        }
        hipError_t e = hipMalloc(&f->cur32[q], nn * 4);
        if (e == hipSuccess) e = hipMalloc(&f->cur64[q], nn * 8);
        if (e == hipSuccess) {
            if (dtype == FQSL_DTYPE_F64) {
                e = hipMemcpy(f->cur64[q], staged, nn * 8, hipMemcpyDeviceToDevice);
                if (e == hipSuccess) k_convert_to_f32<<<grid_for(nn, 256), 256>>>((const double*)staged, f->cur32[q], nn);
            } else {
                e = hipMemcpy(f->cur32[q], staged, nn * 4, hipMemcpyDeviceToDevice);
                if (e == hipSuccess) k_convert_to_f64<<<grid_for(nn, 256), 256>>>((const float*)staged, f->cur64[q], nn);
            }
            if (e == hipSuccess) e = hipGetLastError();
            if (e == hipSuccess) e = hipDeviceSynchronize();
        }
        if (!on_device) (void)hipFree(staged);
        if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "current layout: %s", hipGetErrorString(e));
        f->bytes += nn * 12;
    }
    if (f->cell32) {  // the FP32 tracers read B from cells; J the same way (6 vector loads, not 24 gathers)
        hipError_t e = hipMalloc(&f->curcell, ncell * 128);
        if (e != hipSuccess) return set_err(FQSL_ERR_NOMEM, "hipMalloc(current cells): %s", hipGetErrorString(e));
        k_build_cells<<<grid_for(ncell, 256), 256>>>(f->cur32[0], f->cur32[1], f->cur32[2], f->n[0], f->n[1], f->n[2],
                                                     f->curcell);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "current cells: %s", hipGetErrorString(e));
        f->bytes += ncell * 128;
    }
    return FQSL_OK;
}

extern "C" int fqsl_field_shape(const fqsl_field* f, int* nx, int* ny, int* nz) {
    if (!f) return set_err(FQSL_ERR_INVALID, "field is NULL");
    if (nx) *nx = f->n[0];
    if (ny) *ny = f->n[1];
    if (nz) *nz = f->n[2];
    return FQSL_OK;
}

// -------------------------------------------------------- stretched grids
// Lookup tables of the three axes in one device allocation; fills the float
// and double views.  Landing planes: 0 / n (the reference rules) or the box
// faces (exact rule).
static int build_stretch(const std::vector<double> (&a)[3], const int* n, int rule, void** mem, size_t* bytes,
                         const StretchGrid<float>** g32, const StretchGrid<double>** g64, int* lds_bytes) {
    if (rule != kLookupBisect && rule != kLookupUniform && rule != kLookupExact)
        return set_err(FQSL_ERR_INVALID, "lookup rule %d", rule);
    AxisTables t[3];
    for (int q = 0; q < 3; ++q) {
        if ((int)a[q].size() != n[q]) return set_err(FQSL_ERR_INVALID, "axis %d has %zu nodes, field %d", q, a[q].size(), n[q]);
        const char* why = build_axis_tables(a[q].data(), n[q], rule, t[q]);
        if (why) return set_err(FQSL_ERR_INVALID, "axis %d: %s", q, why);
    }
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t off64[3], off32[3], offr[3], offb[3], offb64[3], offb32[3], total = 0;
    for (int q = 0; q < 3; ++q) { offb64[q] = total; total = up(total + t[q].brec64.size() * sizeof(BinRec<double>)); }
    for (int q = 0; q < 3; ++q) { offb32[q] = total; total = up(total + t[q].brec32.size() * sizeof(BinRec<float>)); }
    for (int q = 0; q < 3; ++q) { off64[q] = total; total = up(total + t[q].rec64.size() * sizeof(double4)); }
    for (int q = 0; q < 3; ++q) { off32[q] = total; total = up(total + t[q].rec32.size() * sizeof(float4)); }
    for (int q = 0; q < 3; ++q) { offr[q] = total; total = up(total + t[q].ref.size() * sizeof(int)); }
    for (int q = 0; q < 3; ++q) { offb[q] = total; total = up(total + t[q].bins.size() * sizeof(int)); }
    std::vector<uint2> compact;
    int lbin[3] = {0, 0, 0}, lcell[3] = {0, 0, 0};
    if (!build_compact(t, compact, lbin, lcell)) compact.clear();
    const size_t offc = total;
    total = up(total + compact.size() * sizeof(uint2));
    const size_t offg32 = total;
    total = up(total + sizeof(StretchGrid<float>));
    const size_t offg64 = total;
    total = up(total + sizeof(StretchGrid<double>));
    std::vector<char> host(total, 0);
    for (int q = 0; q < 3; ++q) {
        memcpy(host.data() + offb64[q], t[q].brec64.data(), t[q].brec64.size() * sizeof(BinRec<double>));
        memcpy(host.data() + offb32[q], t[q].brec32.data(), t[q].brec32.size() * sizeof(BinRec<float>));
        memcpy(host.data() + off64[q], t[q].rec64.data(), t[q].rec64.size() * sizeof(double4));
        memcpy(host.data() + off32[q], t[q].rec32.data(), t[q].rec32.size() * sizeof(float4));
        memcpy(host.data() + offr[q], t[q].ref.data(), t[q].ref.size() * sizeof(int));
        memcpy(host.data() + offb[q], t[q].bins.data(), t[q].bins.size() * sizeof(int));
    }
    memcpy(host.data() + offc, compact.data(), compact.size() * sizeof(uint2));
    char* d = nullptr;
    HIP_TRY(hipMalloc(&d, total));
    const void* rec64[3] = {d + off64[0], d + off64[1], d + off64[2]};
    const void* rec32[3] = {d + off32[0], d + off32[1], d + off32[2]};
    const int* ref[3] = {(const int*)(d + offr[0]), (const int*)(d + offr[1]), (const int*)(d + offr[2])};
    const int* bins[3] = {(const int*)(d + offb[0]), (const int*)(d + offb[1]), (const int*)(d + offb[2])};
    const void* br64[3] = {d + offb64[0], d + offb64[1], d + offb64[2]};
    const void* br32[3] = {d + offb32[0], d + offb32[1], d + offb32[2]};
    StretchGrid<double> h64;
    StretchGrid<float> h32;
    fill_grid<double>(t, rule, br64, rec64, ref, bins, h64);
    fill_grid<float>(t, rule, br32, rec32, ref, bins, h32);
    h64.compact = nullptr;
    h64.compact_n = 0;
    h32.compact = compact.empty() ? nullptr : (const uint2*)(d + offc);
    h32.compact_n = (int)compact.size();
    for (int q = 0; q < 3; ++q) {
        h32.ax[q].lbin = lbin[q]; h32.ax[q].lcell = lcell[q];
        h64.ax[q].lbin = 0; h64.ax[q].lcell = 0;
    }
    memcpy(host.data() + offg32, &h32, sizeof(h32));
    memcpy(host.data() + offg64, &h64, sizeof(h64));
    hipError_t e = hipMemcpy(d, host.data(), total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return set_err(FQSL_ERR_HIP, "hipMemcpy(axis tables): %s", hipGetErrorString(e));
    }
    *g32 = (const StretchGrid<float>*)(d + offg32);
    *g64 = (const StretchGrid<double>*)(d + offg64);
    *lds_bytes = (int)(compact.size() * sizeof(uint2));
    *mem = d;
    *bytes = total;
    return FQSL_OK;
}

// Host copy (double) of n values of an axis array in dtype.
static int axis_to_host(const void* src, int n, int dtype, int on_device, std::vector<double>& out) {
    out.resize(n);
    const size_t esz = dtype == FQSL_DTYPE_F64 ? 8 : 4;
    std::vector<char> raw(n * esz);
    if (on_device) HIP_TRY(hipMemcpy(raw.data(), src, n * esz, hipMemcpyDeviceToHost));
    else memcpy(raw.data(), src, n * esz);
    for (int i = 0; i < n; ++i)
        out[i] = dtype == FQSL_DTYPE_F64 ? ((const double*)raw.data())[i] : (double)((const float*)raw.data())[i];
    return FQSL_OK;
}

extern "C" int fqsl_field_set_axes(fqsl_field* f, const void* xa, const void* ya, const void* za, int dtype,
                                   int on_device, int lookup) {
    if (!f) return set_err(FQSL_ERR_INVALID, "field is NULL");
    if (!xa && !ya && !za) {  // back to grid-index coordinates
        if (f->axes_mem) { (void)hipFree(f->axes_mem); f->bytes -= f->axes_bytes; }
        f->axes_mem = nullptr; f->axes_bytes = 0; f->stretched = false;
        return FQSL_OK;
    }
    if (!xa || !ya || !za) return set_err(FQSL_ERR_INVALID, "axis pointer is NULL");
    if (dtype != FQSL_DTYPE_F32 && dtype != FQSL_DTYPE_F64) return set_err(FQSL_ERR_INVALID, "dtype %d", dtype);
    std::vector<double> a[3];
    const void* src[3] = {xa, ya, za};
    for (int q = 0; q < 3; ++q) {
        int rc = axis_to_host(src[q], f->n[q], dtype, on_device, a[q]);
        if (rc) return rc;
    }
    void* mem = nullptr;
    size_t bytes = 0;
    const StretchGrid<float>* g32 = nullptr;
    const StretchGrid<double>* g64 = nullptr;
    int lds_bytes = 0;
    int rc = build_stretch(a, f->n, lookup, &mem, &bytes, &g32, &g64, &lds_bytes);
    if (rc) return rc;
    if (f->axes_mem) { (void)hipFree(f->axes_mem); f->bytes -= f->axes_bytes; }
    f->axes_mem = mem; f->axes_bytes = bytes; f->bytes += bytes;
    f->sg32 = g32; f->sg64 = g64; f->sg_lds_bytes = lds_bytes; f->stretched = true;
    return FQSL_OK;
}

extern "C" int fqsl_field_stretched(const fqsl_field* f) { return f && f->stretched ? 1 : 0; }

// ------------------------------------------------------------- tuning knobs
// Read by every launch, written by fqsl_tune_set from any host thread.
static std::atomic<int> g_tune[FQSL_TUNE_COUNT] = {3, 100, 1, 8, 1, 64, 64, 1};

extern "C" int fqsl_tune_set(int knob, int value) {
    if (knob < 0 || knob >= FQSL_TUNE_COUNT) return set_err(FQSL_ERR_INVALID, "tuning knob %d", knob);
    if (knob == FQSL_TUNE_SCOTT_WAVES && value != 0 && value != 2 && value != 3)
        return set_err(FQSL_ERR_INVALID, "scott waves/SIMD target %d not in {0, 2, 3}", value);
    if (knob == FQSL_TUNE_QUEUES && value != 1 && value != 8)
        return set_err(FQSL_ERR_INVALID, "queue count %d not in {1, 8}", value);
    if ((knob == FQSL_TUNE_REFILL_MIN || knob == FQSL_TUNE_PLAIN_REFILL) && (value < 1 || value > 64))
        return set_err(FQSL_ERR_INVALID, "refill group %d not in 1..64", value);
    if (knob == FQSL_TUNE_ORDER && (value < 0 || value > 2))
        return set_err(FQSL_ERR_INVALID, "order knob must be 0, 1 or 2 (got %d)", value);
    if (knob == FQSL_TUNE_ORDER_BUCKETS && (value < 2 || value > 256 || (value & (value - 1))))
        return set_err(FQSL_ERR_INVALID, "order buckets must be a power of two in [2, 256] (got %d)", value);
    if (knob == FQSL_TUNE_GRID_PCT && (value < 1 || value > 1000))
        return set_err(FQSL_ERR_INVALID, "grid percentage %d not in 1..1000", value);
    if (knob == FQSL_TUNE_STRETCH_LDS && value != 0 && value != 1)
        return set_err(FQSL_ERR_INVALID, "stretch_lds must be 0 or 1 (got %d)", value);
    g_tune[knob] = value;
    return FQSL_OK;
}

extern "C" int fqsl_tune_get(int knob, int* value) {
    if (knob < 0 || knob >= FQSL_TUNE_COUNT || !value) return set_err(FQSL_ERR_INVALID, "tuning knob %d", knob);
    *value = g_tune[knob];
    return FQSL_OK;
}

// ------------------------------------------------------- device work queue
// A ring of 64-bit queue counters; each trace call takes the next slot and
// clears it with a fill kernel on its stream before the persistent launch.
// Calls that are more than kQueueSlots launches apart never share a slot.
static constexpr int kQueueSlots = 4096;
static constexpr int kQueueSlotWidth = 8;  // counters per slot (one per XCD queue)
struct DeviceQueues {
    unsigned long long* slots = nullptr;
    unsigned next = 0;
};
static std::mutex g_q_mu;
static std::vector<DeviceQueues> g_queues;
static std::vector<int> g_cu_count;

static int queue_slot(unsigned long long** slot) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_q_mu);
    if ((int)g_queues.size() <= dev) g_queues.resize(dev + 1);
    DeviceQueues& d = g_queues[dev];
    if (!d.slots) {
        HIP_TRY(hipMalloc(&d.slots, (size_t)kQueueSlots * kQueueSlotWidth * sizeof(unsigned long long)));
        HIP_TRY(hipMemset(d.slots, 0, (size_t)kQueueSlots * kQueueSlotWidth * sizeof(unsigned long long)));
    }
    *slot = d.slots + (size_t)kQueueSlotWidth * (d.next++ % kQueueSlots);
    return FQSL_OK;
}

static int cu_count() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lk(g_q_mu);
    if ((int)g_cu_count.size() <= dev) g_cu_count.resize(dev + 1, 0);
    if (!g_cu_count[dev]) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        g_cu_count[dev] = c;
    }
    return g_cu_count[dev];
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
    if ((int)threadIdx.x < count) p[threadIdx.x] = 0ull;
}

// Scalars the reference passes as one-element device arrays.  When `shape`
// is non-NULL the kernel reads them on the device.
struct DevScalars {
    const int* shape;
    const float* s_len;
    const unsigned long long* N;
    const float* tol_coef;
    const bool* twist_flag;
    double* twist;
    // plain tracers: line count written on the device by an earlier kernel of
    // the stream (the pseudo-line compaction); the kernel traces
    // min(n, *n_dev) lines, n being the capacity of the seed and output arrays
    const unsigned long long* n_dev;
};

struct Seeds {
    const float* x;
    const float* y;
    const float* z;
};

struct ScottOut {
    float *sx, *sy, *sz, *ex, *ey, *ez;
    int32_t *fs, *fe;
    float *btx, *bty, *btz, *q0, *qp;
    double* len;
    uint32_t* evals;
    double* tw;
};

struct PlainOut {
    float *sx, *sy, *sz, *ex, *ey, *ez;
    int32_t *fs, *fe;
    float *btx, *bty, *btz;
    int32_t* bflag;
    float *bsx, *bsy, *bsz, *bex, *bey, *bez;
    double* len;
    uint32_t* evals;
    double* tw;
};

template <typename R, typename X, class F>
__device__ __forceinline__ void scott_write(const ScottLine<R, X>& L, const ScottEnd<R>& e, int status,
                                            const F& f, unsigned long long i, const ScottOut& o,
                                            const DevScalars& ds) {
    double q0, qp, nb0 = 0.0;
    if (status == kFinished) {
        R u0[3], v0[3], b0z;
        seed_frame<R>(f, L.p0, u0, v0, &b0z, &nb0);
        nb0 = (double)(R)nb0;  // |B0| as the persistent path's seed frame stores it
    }
    const ScottEnd<R> st = L.st, en = e;  // copies: scott_q is out of line
    scott_q<R>(st, en, status, nb0, f, &q0, &qp);
    if (o.sx) { o.sx[i] = (float)L.st.x[0]; o.sy[i] = (float)L.st.x[1]; o.sz[i] = (float)L.st.x[2]; }
    if (o.ex) { o.ex[i] = (float)e.x[0]; o.ey[i] = (float)e.x[1]; o.ez[i] = (float)e.x[2]; }
    if (o.fs) o.fs[i] = L.st.flag;
    if (o.fe) o.fe[i] = e.flag;
    if (o.q0) o.q0[i] = (float)q0;
    if (o.qp) o.qp[i] = (float)qp;
    if (o.len) o.len[i] = L.len;
    if (o.evals) o.evals[i] = L.evals;
    if (o.tw) o.tw[i] = L.tw;
    if (o.btx) {
        float bt[3];
        b_clamped<R>(f, L.p0, bt);
        o.btx[i] = bt[0]; o.bty[i] = bt[1]; o.btz[i] = bt[2];
    }
}

template <class F>
__device__ __forceinline__ void apply_scalars(F& f, ScottParams& P, unsigned long long& n, const DevScalars& ds) {
    if (!ds.shape) return;
    f.n[0] = ds.shape[0]; f.n[1] = ds.shape[1]; f.n[2] = ds.shape[2];
    P.s0 = (double)ds.s_len[0];
    P.tol = 1e-3 * (double)ds.tol_coef[0];
    n = ds.N[0];
}

__device__ __forceinline__ unsigned long long bcast64(unsigned long long v, int src) {
    const int lo = __shfl((int)(unsigned)(v & 0xffffffffull), src);
    const int hi = __shfl((int)(unsigned)(v >> 32), src);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// Finished line as handed from the stepping kernel to the Q-assembly kernel.
template <typename R>
struct ScottRec {
    ScottEnd<R> st, en;
    double len;
    int32_t status;
    uint32_t evals;
    double tw;
};

// Work queues of the persistent kernels: kQueues counters, queue q owning
// seeds [q*n/kQueues, (q+1)*n/kQueues).  A workgroup starts on queue
// blockIdx % kQueues -- consecutive workgroups land on different XCDs, so each
// XCD first drains its own contiguous stretch of seeds (its L2 then holds that
// region of the field) -- and moves on to the other queues when its own is
// empty.  With one queue all workgroups share a single counter.
constexpr int kQueues = 8;

struct QueueCursor {
    int cur, tried;
};

// Claim `cnt` consecutive seed indices for a wavefront; returns false once
// every queue is empty.  base/avail describe the claimed range.
__device__ __forceinline__ bool claim(unsigned long long* queues, int nq, unsigned long long n, QueueCursor& qc,
                                      int cnt, int leader, unsigned long long& base, int& avail) {
    const int lane = __lane_id();
    while (qc.tried < nq) {
        const unsigned long long lo = n * (unsigned long long)qc.cur / nq;
        const unsigned long long hi = n * (unsigned long long)(qc.cur + 1) / nq;
        unsigned long long got = 0;
        if (lane == leader) got = atomicAdd(queues + qc.cur, (unsigned long long)cnt);
        got = bcast64(got, leader);
        if (lo + got < hi) {
            base = lo + got;
            const unsigned long long left = hi - base;
            avail = left < (unsigned long long)cnt ? (int)left : cnt;
            if (left <= (unsigned long long)cnt) { qc.cur = (qc.cur + 1) % nq; ++qc.tried; }
            return true;
        }
        qc.cur = (qc.cur + 1) % nq;
        ++qc.tried;
    }
    return false;
}

// ------------------------------------------------------ longest-first order
// Lines take 10 to 10^4 steps.  A persistent launch that hands them out in
// seed order ends with a tail of long lines running in nearly empty
// wavefronts (about a third of the launch on the 1024^2 multipole map).  A
// coarse estimate of each line's length -- a plain RK4 trace of b with 8-cell
// steps, at most kPredSteps per direction, so lines up to 256 cells are told
// apart -- sorts the queue longest first (counting sort into 2..256 length
// buckets; within a bucket lines keep roughly their seed order, so fewer
// buckets keep more of the field locality of neighbouring seeds), which
// removes most of that tail for ~3 % more work.  (4-cell steps, twice as
// many, ordered no better and cost 0.25 ms more per 1024^2 map; 16-cell
// steps, half as many, save 0.15 ms and order worse by as much.)
// With many lines per lane the tail is short anyway and the lost locality
// costs more than it saves, so by default lines are only ordered up to 64
// lines per persistent lane.  The order only decides which lane traces a line
// and when; the line's result is the same.
constexpr int kPredSteps = 16;
constexpr float kPredStep = 8.0f;
constexpr int kOrderBuckets = 256;

template <typename R>
__device__ __forceinline__ Pt<R> pt_from_real(const R* x) {
    Pt<R> p;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const R fl = flr(x[a]);
        p.i[a] = to_int(fl);
        p.r[a] = x[a] - fl;
    }
    return p;
}

// Length (cells) a coarse RK4 trace of +-b from p0 covers inside the box,
// capped at 2 * kPredSteps * kPredStep.
template <typename R, class F>
__device__ float predict_length(const F& f, const float* p0) {
    const R h = (R)kPredStep;
    float total = 0.f;
    for (int dir = -1; dir <= 1; dir += 2) {
        R x[3] = {(R)p0[0], (R)p0[1], (R)p0[2]};
        const R hs = h * (R)dir;
        for (int st = 0; st < kPredSteps; ++st) {
            R k1[3], k2[3], k3[3], k4[3], y[3];
            unit_b<R>(f, pt_from_real<R>(x), k1);
#pragma unroll
            for (int a = 0; a < 3; ++a) y[a] = x[a] + (R)0.5 * hs * k1[a];
            unit_b<R>(f, pt_from_real<R>(y), k2);
#pragma unroll
            for (int a = 0; a < 3; ++a) y[a] = x[a] + (R)0.5 * hs * k2[a];
            unit_b<R>(f, pt_from_real<R>(y), k3);
#pragma unroll
            for (int a = 0; a < 3; ++a) y[a] = x[a] + hs * k3[a];
            unit_b<R>(f, pt_from_real<R>(y), k4);
            bool inside = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                x[a] += hs / (R)6 * (k1[a] + (R)2 * k2[a] + (R)2 * k3[a] + k4[a]);
                inside = inside && x[a] >= (R)0 && x[a] <= (R)(f.n[a] - 1);  // false for NaN
            }
            total += kPredStep;
            if (!inside) break;
        }
    }
    return total;
}

__global__ void k_zero_u32(unsigned* p, int count) {
    for (int t = threadIdx.x; t < count; t += blockDim.x) p[t] = 0u;
}

// Bucket key per line (0 = longest) and the bucket counts.
template <typename R, class F>
__global__ void __launch_bounds__(256) k_order_predict(F f, Seeds s, unsigned long long n, int nbuckets,
                                                       unsigned char* __restrict__ key, unsigned* counts) {
    __shared__ unsigned hist[kOrderBuckets];
    for (int t = threadIdx.x; t < kOrderBuckets; t += blockDim.x) hist[t] = 0u;
    __syncthreads();
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const float p0[3] = {s.x[i], s.y[i], s.z[i]};
        const float len = predict_length<R, F>(f, p0);
        int b = (int)(len * ((float)nbuckets / (2.f * kPredSteps * kPredStep)));
        b = b < 0 ? 0 : (b > nbuckets - 1 ? nbuckets - 1 : b);
        const int kb = nbuckets - 1 - b;  // bucket 0 holds the longest lines
        key[i] = (unsigned char)kb;
        atomicAdd(&hist[kb], 1u);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kOrderBuckets; t += blockDim.x)
        if (hist[t]) atomicAdd(&counts[t], hist[t]);
}

// Exclusive scan of the bucket counts (one block).
__global__ void __launch_bounds__(kOrderBuckets) k_order_scan(unsigned* counts) {
    __shared__ unsigned v[kOrderBuckets];
    v[threadIdx.x] = counts[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned acc = 0;
        for (int t = 0; t < kOrderBuckets; ++t) {
            const unsigned c = v[t];
            v[t] = acc;
            acc += c;
        }
    }
    __syncthreads();
    counts[threadIdx.x] = v[threadIdx.x];
}

// order[offset(bucket) + rank] = line, 256 lines per block step: one global
// atomic per (block step, bucket) reserves the block's range in a bucket.
__global__ void __launch_bounds__(256) k_order_scatter(const unsigned char* __restrict__ key, unsigned long long n,
                                                       unsigned* offs, unsigned* __restrict__ order) {
    __shared__ unsigned hist[kOrderBuckets], base[kOrderBuckets];
    for (unsigned long long c0 = blockIdx.x * 256ull; c0 < n; c0 += gridDim.x * 256ull) {
        hist[threadIdx.x] = 0u;
        __syncthreads();
        const unsigned long long i = c0 + threadIdx.x;
        const bool valid = i < n;
        const int kb = valid ? (int)key[i] : 0;
        const unsigned rank = valid ? atomicAdd(&hist[kb], 1u) : 0u;
        __syncthreads();
        if (hist[threadIdx.x]) base[threadIdx.x] = atomicAdd(&offs[threadIdx.x], hist[threadIdx.x]);
        __syncthreads();
        if (valid) order[base[kb] + rank] = (unsigned)i;
        __syncthreads();
    }
}

// Seed preparation, one thread per seed: frame (u0, v0, b0z, |B0|) for the
// stepping kernel and B at the seed (Interp3d rule) straight to the output.
template <typename R, class F>
__global__ void __launch_bounds__(256) k_scott_prep(F f, ScottParams P, Seeds s, unsigned long long n,
                                                    R* __restrict__ frames, ScottOut o, DevScalars ds) {
    apply_scalars(f, P, n, ds);
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const float p0[3] = {s.x[i], s.y[i], s.z[i]};
        R u0[3], v0[3], b0z;
        double nb0;
        seed_frame<R>(f, p0, u0, v0, &b0z, &nb0);
        R* fr = frames + (unsigned long long)kFrameWords * i;
#pragma unroll
        for (int a = 0; a < 3; ++a) { fr[a] = u0[a]; fr[3 + a] = v0[a]; }
        fr[6] = b0z;
        fr[7] = (R)nb0;
        if (o.btx) {
            float bt[3];
            b_clamped<R>(f, p0, bt);
            o.btx[i] = bt[0]; o.bty[i] = bt[1]; o.btz[i] = bt[2];
        }
    }
}

// Persistent stepping: every lane owns one line; idle lanes take the next
// queue entries once at least `refill_min` lanes of the wavefront are idle (or
// all are), so the setup of refilled lines runs once for the group.  A wave
// exits once every queue is empty and all its lanes are idle.
template <typename R, class F, class FJ, typename X, bool kRoll = false>
__device__ __forceinline__ void scott_persistent_body(F f, ScottParams P, Seeds s, unsigned long long n,
                                                      const R* __restrict__ frames, ScottRec<R>* __restrict__ recs,
                                                      DevScalars ds, unsigned long long* queues, int nq,
                                                      int refill_min, const unsigned* __restrict__ order, FJ fj) {
    apply_scalars(f, P, n, ds);
    fj.n[0] = f.n[0]; fj.n[1] = f.n[1]; fj.n[2] = f.n[2];  // J lives on B's grid
    P.ctl = scott_ctl(P.tol);  // on the device, so every path forms the same bits
    ScottLine<R, X> L;
    bool active = false;
    bool drained = false;
    unsigned long long idx = 0;
    const int lane = __lane_id();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    QueueCursor qc{(int)(blockIdx.x % (unsigned)nq), 0};
    for (;;) {
        if (!drained) {
            const unsigned long long need = __ballot(!active);
            const int cnt = __popcll(need);
            if (cnt >= refill_min || need == ~0ull) {
                const int leader = __ffsll((long long)need) - 1;
                unsigned long long base = 0;
                int avail = 0;
                if (!claim(queues, nq, n, qc, cnt, leader, base, avail)) drained = true;
                const int rank = __popcll(need & below);
                if (!active && rank < avail) {
                    idx = base + (unsigned long long)rank;
                    if (order) idx = order[idx];
                    ScottEnd<R> e;
                    const int st = scott_init<R>(L, f, P, s.x[idx], s.y[idx], s.z[idx], e,
                                                 frames + (unsigned long long)kFrameWords * idx);
                    if (st != kTracing) {
                        recs[idx] = ScottRec<R>{L.st, e, L.len, st, L.evals, L.tw};
                    } else {
                        active = true;
                    }
                }
            }
        }
        if (!__any(active)) {
            if (drained) break;
            continue;
        }
        if (active) {
            ScottEnd<R> e;
            const int st = scott_attempt<R, F, FJ, X, kRoll>(L, f, P, e, fj);
            if (st != kTracing) {
                recs[idx] = ScottRec<R>{L.st, e, L.len, st, L.evals, L.tw};
                active = false;
            }
        }
    }
}

// W = waves per SIMD the register allocation targets.
#define FQSL_PERSISTENT_KERNEL(W)                                                                     \
    template <typename R, class F, class FJ, typename X>                                              \
    __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))                \
    k_scott_persistent_w##W(F f, ScottParams P, Seeds s, unsigned long long n, const R* frames,       \
                            ScottRec<R>* recs, DevScalars ds, unsigned long long* queues, int nq,     \
                            int refill_min, const unsigned* order, FJ fj) {                           \
        scott_persistent_body<R, F, FJ, X, W == 3>(f, P, s, n, frames, recs, ds, queues, nq, refill_min, order, \
                                                  fj);                                                \
    }

template <typename R, class F, class FJ, typename X>
__global__ void __launch_bounds__(256) k_scott_persistent(F f, ScottParams P, Seeds s, unsigned long long n,
                                                          const R* frames, ScottRec<R>* recs, DevScalars ds,
                                                          unsigned long long* queues, int nq, int refill_min,
                                                          const unsigned* order, FJ fj) {
    scott_persistent_body<R, F, FJ, X>(f, P, s, n, frames, recs, ds, queues, nq, refill_min, order, fj);
}
FQSL_PERSISTENT_KERNEL(2)
FQSL_PERSISTENT_KERNEL(3)

// Q assembly and output, one thread per line (TraceBlineScott.cu:475-508,
// outputs as TraceAllBlineScott :565-582).
template <typename R, class F>
__global__ void __launch_bounds__(256) k_scott_finish(F f, ScottParams P, unsigned long long n,
                                                      const R* __restrict__ frames,
                                                      const ScottRec<R>* __restrict__ recs, ScottOut o,
                                                      DevScalars ds) {
    apply_scalars(f, P, n, ds);
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const ScottRec<R> r = recs[i];
        double q0, qp;
        scott_q<R>(r.st, r.en, r.status, (double)frames[(unsigned long long)kFrameWords * i + 7], f, &q0, &qp);
        if (o.sx) { o.sx[i] = (float)r.st.x[0]; o.sy[i] = (float)r.st.x[1]; o.sz[i] = (float)r.st.x[2]; }
        if (o.ex) { o.ex[i] = (float)r.en.x[0]; o.ey[i] = (float)r.en.x[1]; o.ez[i] = (float)r.en.x[2]; }
        if (o.fs) o.fs[i] = r.st.flag;
        if (o.fe) o.fe[i] = r.en.flag;
        if (o.q0) o.q0[i] = (float)q0;
        if (o.qp) o.qp[i] = (float)qp;
        if (o.len) o.len[i] = r.len;
        if (o.evals) o.evals[i] = r.evals;
        if (o.tw) o.tw[i] = r.tw;
    }
}

// One thread per line (grid-stride), the reference's mapping
// (TraceBlineScott.cu:530-553).
template <typename R, class F, class FJ, typename X, bool kRoll = false>
__device__ __forceinline__ void scott_per_line_body(F f, ScottParams P, Seeds s, unsigned long long n,
                                                    ScottOut out, DevScalars ds, FJ fj) {
    apply_scalars(f, P, n, ds);
    fj.n[0] = f.n[0]; fj.n[1] = f.n[1]; fj.n[2] = f.n[2];  // J lives on B's grid
    P.ctl = scott_ctl(P.tol);  // on the device, so every path forms the same bits
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n; i += stride) {
        ScottLine<R, X> L;
        ScottEnd<R> e;
        int st = scott_init<R>(L, f, P, s.x[i], s.y[i], s.z[i], e);
        while (st == kTracing) st = scott_attempt<R, F, FJ, X, kRoll>(L, f, P, e, fj);
        scott_write<R>(L, e, st, f, i, out, ds);
    }
}

template <typename R, class F, class FJ, typename X>
__global__ void __launch_bounds__(256) k_scott_per_line(F f, ScottParams P, Seeds s, unsigned long long n,
                                                        ScottOut out, DevScalars ds, FJ fj) {
    scott_per_line_body<R, F, FJ, X>(f, P, s, n, out, ds, fj);
}

// FP32 variants compiled for two or three waves per SIMD (like the
// persistent kernel).
template <typename R, class F, class FJ, typename X>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_scott_per_line_w2(F f, ScottParams P, Seeds s, unsigned long long n, ScottOut out, DevScalars ds, FJ fj) {
    scott_per_line_body<R, F, FJ, X>(f, P, s, n, out, ds, fj);
}

template <typename R, class F, class FJ, typename X>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_scott_per_line_w3(F f, ScottParams P, Seeds s, unsigned long long n, ScottOut out, DevScalars ds, FJ fj) {
    scott_per_line_body<R, F, FJ, X, true>(f, P, s, n, out, ds, fj);
}

// ------------------------------------------------------------ plain kernels
template <typename R, class F>
__device__ __forceinline__ void plain_write(const PlainLine<R>& L, const F& f, const PlainParams& P,
                                            unsigned long long i, const PlainOut& o, const DevScalars& ds) {
    const float xs[3] = {(float)L.xs[0], (float)L.xs[1], (float)L.xs[2]};
    const float xe[3] = {(float)L.xe[0], (float)L.xe[1], (float)L.xe[2]};
    if (o.sx) { o.sx[i] = xs[0]; o.sy[i] = xs[1]; o.sz[i] = xs[2]; }
    if (o.ex) { o.ex[i] = xe[0]; o.ey[i] = xe[1]; o.ez[i] = xe[2]; }
    if (o.fs) o.fs[i] = L.flag_s;
    if (o.fe) o.fe[i] = L.flag_e;
    if (o.len) o.len[i] = L.len_total;
    if (o.evals) o.evals[i] = L.evals;
    if (o.tw) o.tw[i] = L.tw;
    float b[3];
    if (o.bsx) { b_clamped<R>(f, xs, b); o.bsx[i] = b[0]; o.bsy[i] = b[1]; o.bsz[i] = b[2]; }
    if (o.bex) { b_clamped<R>(f, xe, b); o.bex[i] = b[0]; o.bey[i] = b[1]; o.bez[i] = b[2]; }
    b_clamped<R>(f, L.p0, b);
    if (o.btx) { o.btx[i] = b[0]; o.bty[i] = b[1]; o.btz[i] = b[2]; }
    if (o.bflag) {
        // TraceBlineAdaptive.cu:498-501, in float as the reference
        const float bn = fabsf(b[0] * P.ncross[0] + b[1] * P.ncross[1] + b[2] * P.ncross[2]) * 100.f;
        o.bflag[i] = bn < sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) ? 1 : 0;
    }
}

// The stretched tracer writes B_start / B_end / B_this by the stretched
// Interp3dxyzn (raw, TraceBlineAdaptiveStretch.cu:463-496).
template <typename R, class F, bool kL = false>
__device__ __forceinline__ void stretch_write(const PlainLine<R>& L, const F& f, const FQSL_CONST StretchGrid<R>& g,
                                              const PlainParams& P, unsigned long long i, const PlainOut& o) {
    const float xs[3] = {(float)L.xs[0], (float)L.xs[1], (float)L.xs[2]};
    const float xe[3] = {(float)L.xe[0], (float)L.xe[1], (float)L.xe[2]};
    if (o.sx) { o.sx[i] = xs[0]; o.sy[i] = xs[1]; o.sz[i] = xs[2]; }
    if (o.ex) { o.ex[i] = xe[0]; o.ey[i] = xe[1]; o.ez[i] = xe[2]; }
    if (o.fs) o.fs[i] = L.flag_s;
    if (o.fe) o.fe[i] = L.flag_e;
    if (o.len) o.len[i] = L.len_total;
    if (o.evals) o.evals[i] = L.evals;
    if (o.tw) o.tw[i] = L.tw;
    double B[3];
    if (o.bsx) {
        const double p[3] = {xs[0], xs[1], xs[2]};
        stretch_raw_at<R, F, kL>(f, g, p, B);
        o.bsx[i] = (float)B[0]; o.bsy[i] = (float)B[1]; o.bsz[i] = (float)B[2];
    }
    if (o.bex) {
        const double p[3] = {xe[0], xe[1], xe[2]};
        stretch_raw_at<R, F, kL>(f, g, p, B);
        o.bex[i] = (float)B[0]; o.bey[i] = (float)B[1]; o.bez[i] = (float)B[2];
    }
    const double p0[3] = {L.p0[0], L.p0[1], L.p0[2]};
    stretch_raw_at<R, F, kL>(f, g, p0, B);
    const float b[3] = {(float)B[0], (float)B[1], (float)B[2]};
    if (o.btx) { o.btx[i] = b[0]; o.bty[i] = b[1]; o.btz[i] = b[2]; }
    if (o.bflag) {
        const float bn = fabsf(b[0] * P.ncross[0] + b[1] * P.ncross[1] + b[2] * P.ncross[2]) * 100.f;
        o.bflag[i] = bn < sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) ? 1 : 0;
    }
}

// Geometry of the plain tracer: grid-index coordinates (the uniform tracers)
// or a stretched grid; the overloads pick the line driver.
struct UniformGrid {};

template <typename R, class F>
__device__ __forceinline__ void line_init(PlainLine<R>& L, const F& f, const UniformGrid&, const PlainParams& P,
                                          float x, float y, float z) {
    plain_init<R>(L, f, P, x, y, z);
}
template <typename R>
__device__ __forceinline__ const FQSL_CONST StretchGrid<R>& grid_of(const StretchRef<R>& g) {
    return *(const FQSL_CONST StretchGrid<R>*)g.g;
}

template <typename R, class F>
__device__ __forceinline__ void line_init(PlainLine<R>& L, const F& f, const StretchRef<R>& g, const PlainParams& P,
                                          float x, float y, float z) {
    stretch_init<R, F>(L, f, grid_of(g), P, x, y, z);
}
template <typename R, class F, class FJ>
__device__ __forceinline__ void line_attempt(PlainLine<R>& L, const F& f, const UniformGrid&, const PlainParams& P,
                                             const FJ& fj) {
    plain_attempt<R, F, FJ>(L, f, P, fj);
}
template <typename R, class F, class FJ>
__device__ __forceinline__ void line_attempt(PlainLine<R>& L, const F& f, const StretchRef<R>& g,
                                             const PlainParams& P, const FJ& fj) {
    stretch_attempt<R, F, FJ>(L, f, grid_of(g), P, fj);
}
template <typename R, class F>
__device__ __forceinline__ void line_write(const PlainLine<R>& L, const F& f, const UniformGrid&, const PlainParams& P,
                                           unsigned long long i, const PlainOut& o, const DevScalars& ds) {
    plain_write<R>(L, f, P, i, o, ds);
}
template <typename R, class F>
__device__ __forceinline__ void line_write(const PlainLine<R>& L, const F& f, const StretchRef<R>& g,
                                           const PlainParams& P, unsigned long long i, const PlainOut& o,
                                           const DevScalars&) {
    stretch_write<R>(L, f, grid_of(g), P, i, o);
}

// A float stretched grid whose compact lookup table each workgroup copies to
// LDS at the start (`lds_bytes` of dynamic shared memory): the bin record of
// a stage point is then an LDS read instead of a global load that the cell
// load has to wait for.
struct StretchLds {
    const StretchGrid<float>* g;
    int lds_bytes;
};
__device__ __forceinline__ const FQSL_CONST StretchGrid<float>& grid_of(const StretchLds& g) {
    return *(const FQSL_CONST StretchGrid<float>*)g.g;
}
template <typename R, class F>
__device__ __forceinline__ void line_init(PlainLine<R>& L, const F& f, const StretchLds& g, const PlainParams& P,
                                          float x, float y, float z) {
    stretch_init<R, F, true>(L, f, grid_of(g), P, x, y, z);
}
template <typename R, class F, class FJ>
__device__ __forceinline__ void line_attempt(PlainLine<R>& L, const F& f, const StretchLds& g,
                                             const PlainParams& P, const FJ& fj) {
    stretch_attempt<R, F, FJ, true>(L, f, grid_of(g), P, fj);
}
template <typename R, class F>
__device__ __forceinline__ void line_write(const PlainLine<R>& L, const F& f, const StretchLds& g,
                                           const PlainParams& P, unsigned long long i, const PlainOut& o,
                                           const DevScalars&) {
    stretch_write<R, F, true>(L, f, grid_of(g), P, i, o);
}

// Kernel prologue per geometry: the LDS grid copies its table (every thread
// of the block reaches it before any lookup).
template <class G>
__device__ __forceinline__ void grid_prologue(const G&) {}
__device__ __forceinline__ void grid_prologue(const StretchLds& g) {
    const FQSL_CONST StretchGrid<float>& sg = grid_of(g);
    const int m = sg.compact_n;
    for (int i = threadIdx.x; i < m; i += blockDim.x) fqsl_stretch_lds[i] = sg.compact[i];
    __syncthreads();
}
static int grid_lds_bytes(const UniformGrid&) { return 0; }
template <typename R>
static int grid_lds_bytes(const StretchRef<R>&) { return 0; }
static int grid_lds_bytes(const StretchLds& g) { return g.lds_bytes; }

template <class F>
__device__ __forceinline__ void apply_plain_scalars(F& f, PlainParams& P, unsigned long long& n,
                                                    const DevScalars& ds, const float* ncross) {
    if (ds.n_dev) {
        const unsigned long long m = ds.n_dev[0];
        if (m < n) n = m;
    }
    if (!ds.shape) return;
    f.n[0] = ds.shape[0]; f.n[1] = ds.shape[1]; f.n[2] = ds.shape[2];
    P.s0 = (double)ds.s_len[0];
    P.tol_coef = ds.tol_coef ? (double)ds.tol_coef[0] : 1.0;
    n = ds.N[0];
    P.ncross[0] = ncross[0]; P.ncross[1] = ncross[1]; P.ncross[2] = ncross[2];
}

// Persistent plain tracing: each lane owns one line (forward, then backward)
// and writes its outputs when it ends; idle lanes take the next queue
// entries once at least `refill_min` of the wavefront's lanes are idle (or
// all are), so refills and their setup run together.
template <typename R, class F, class FJ, class G>
__global__ void __launch_bounds__(256) k_plain_persistent(F f, PlainParams P, Seeds s, unsigned long long n,
                                                          PlainOut out, DevScalars ds, const float* ncross,
                                                          unsigned long long* queue, int refill_min, FJ fj, G g) {
    apply_plain_scalars(f, P, n, ds, ncross);
    fj.n[0] = f.n[0]; fj.n[1] = f.n[1]; fj.n[2] = f.n[2];  // J lives on B's grid
    grid_prologue(g);
    PlainLine<R> L;
    bool active = false;
    bool drained = false;
    unsigned long long idx = 0;
    const int lane = __lane_id();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (;;) {
        if (!drained) {
            const unsigned long long need = __ballot(!active);
            const int cnt = __popcll(need);
            if (cnt >= refill_min || need == ~0ull) {
                const int leader = __ffsll((long long)need) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(queue, (unsigned long long)cnt);
                base = bcast64(base, leader);
                if (base + (unsigned long long)cnt >= n) drained = true;
                if (!active) {
                    idx = base + (unsigned long long)__popcll(need & below);
                    if (idx < n) {
                        line_init<R>(L, f, g, P, s.x[idx], s.y[idx], s.z[idx]);
                        if (L.done) line_write<R>(L, f, g, P, idx, out, ds);
                        else active = true;
                    }
                }
            }
        }
        if (!__any(active)) {
            if (drained) break;
            continue;
        }
        if (active) {
            line_attempt<R, F, FJ>(L, f, g, P, fj);
            if (L.done) {
                line_write<R>(L, f, g, P, idx, out, ds);
                active = false;
            }
        }
    }
}

// One thread per line.  Two waves per SIMD as the register target: the FP64
// variants otherwise take 256 VGPRs plus AGPRs and run one wave per SIMD
// (1.2-1.3x slower); the FP32 variants fit two waves either way.
template <typename R, class F, class FJ, class G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_plain_per_line(F f, PlainParams P, Seeds s, unsigned long long n,
                                                        PlainOut out, DevScalars ds, const float* ncross, FJ fj,
                                                        G g) {
    apply_plain_scalars(f, P, n, ds, ncross);
    fj.n[0] = f.n[0]; fj.n[1] = f.n[1]; fj.n[2] = f.n[2];  // J lives on B's grid
    grid_prologue(g);
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n; i += stride) {
        PlainLine<R> L;
        line_init<R>(L, f, g, P, s.x[i], s.y[i], s.z[i]);
        while (!L.done) line_attempt<R, F, FJ>(L, f, g, P, fj);
        line_write<R>(L, f, g, P, i, out, ds);
    }
}

// ------------------------------------------------------------- launching
static int record_start(hipStream_t st) {
    if (!g_ev0) {
        HIP_TRY(hipEventCreate(&g_ev0));
        HIP_TRY(hipEventCreate(&g_ev1));
        HIP_TRY(hipEventCreate(&g_evs0));
        HIP_TRY(hipEventCreate(&g_evs1));
    }
    g_evs_valid = false;
    HIP_TRY(hipEventRecord(g_ev0, st));
    return FQSL_OK;
}
static int record_end(hipStream_t st) {
    HIP_TRY(hipEventRecord(g_ev1, st));
    g_ev_valid = true;
    return FQSL_OK;
}

extern "C" int fqsl_last_step_ms(float* ms) {
    if (!ms) return set_err(FQSL_ERR_INVALID, "ms is NULL");
    if (!g_evs_valid) return set_err(FQSL_ERR_INVALID, "the last trace on this thread had no separate stepping kernel");
    HIP_TRY(hipEventSynchronize(g_evs1));
    HIP_TRY(hipEventElapsedTime(ms, g_evs0, g_evs1));
    return FQSL_OK;
}

extern "C" int fqsl_last_kernel_ms(float* ms) {
    if (!ms) return set_err(FQSL_ERR_INVALID, "ms is NULL");
    if (!g_ev_valid) return set_err(FQSL_ERR_INVALID, "no trace kernel has been launched on this thread");
    HIP_TRY(hipEventSynchronize(g_ev1));
    HIP_TRY(hipEventElapsedTime(ms, g_ev0, g_ev1));
    return FQSL_OK;
}

// Grid of a one-thread-per-line launch.  When the line count lives on the
// device (the reference-shaped entry points, the pseudo-line traces) n_hint
// is only a cap, so the grid
// is bounded by the chip instead: 4096 threads per CU (eight occupancy waves
// at two waves per SIMD), the grid-stride loop covering larger counts.
static int per_line_grid(unsigned long long n_hint, int block, const DevScalars& ds) {
    int g = grid_for((long)n_hint, block);
    if (ds.N || ds.n_dev) {
        const int cap = cu_count() * (4096 / block);
        if (g > cap) g = cap;
    }
    return g;
}

template <typename K>
static int persistent_grid(K kernel, int* grid, int lds_bytes = 0) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds_bytes));
    if (per_cu < 1) per_cu = 1;
    *grid = per_cu * cu_count();
    return FQSL_OK;
}

// Workspace of one persistent Scott trace: seed frames, finished-line
// records and, when the queue is ordered, the order, the bucket keys and
// counts; each part 256-byte aligned.
struct ScottWs {
    size_t frames, recs, order, key, counts, bytes;
};

constexpr size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

template <typename R>
constexpr ScottWs scott_ws_layout(unsigned long long n, bool ordered) {
    constexpr auto up = align256;
    ScottWs w{};
    w.frames = 0;
    w.recs = up((size_t)kFrameWords * sizeof(R) * n);
    w.order = up(w.recs + sizeof(ScottRec<R>) * n);
    w.key = w.order + (ordered ? up(sizeof(unsigned) * n) : 0);
    w.counts = w.key + (ordered ? up(n) : 0);
    w.bytes = w.counts + (ordered ? sizeof(unsigned) * kOrderBuckets : 0);
    return w;
}

// Lines up to which a launch runs the two-wave FP32c kernel whatever the
// knob: one two-wave grid's lanes (2,048 waves of 64 on 256 CUs).  Such a
// launch leaves SIMD slots idle anyway, and its time is the latency of its
// longest lines, which the unrolled two-wave kernel runs faster.
static unsigned long long small_launch_lines() { return (unsigned long long)cu_count() * 4 * 2 * 64; }

// Waves per SIMD a float Scott kernel runs with, for a launch of n lines:
// the knob for FP32c (three: the rolled stages; two: unrolled), two for
// small launches; the mixed variant and the twist tracer (stages unrolled)
// run two; FP64 is left to the compiler.
template <typename R, typename X, class FJ>
static int scott_waves(unsigned long long n) {
    if (sizeof(R) != 4) return 0;
    if (!rollable_stages<R, X, FJ>()) return 2;
    const int w = g_tune[FQSL_TUNE_SCOTT_WAVES].load();
    return (w == 3 && n <= small_launch_lines()) ? 2 : w;
}

// Stepping kernel and grid of the persistent Scott schedule.
template <typename R, class F, class FJ, typename X>
static int scott_persistent_kernel(decltype(&k_scott_persistent<R, F, FJ, X>)* kern, int* grid,
                                   unsigned long long n) {
    const int waves = scott_waves<R, X, FJ>(n);
    *kern = waves == 2 ? k_scott_persistent_w2<R, F, FJ, X>
          : waves == 3 ? k_scott_persistent_w3<R, F, FJ, X>
                       : k_scott_persistent<R, F, FJ, X>;
    int rc = persistent_grid(*kern, grid);
    if (rc) return rc;
    *grid = (int)((long)*grid * g_tune[FQSL_TUNE_GRID_PCT] / 100);
    if (*grid < 1) *grid = 1;
    return FQSL_OK;
}

template <typename R>
static bool scott_ordered(const DevScalars& ds, unsigned long long n_hint, int grid) {
    const int om = g_tune[FQSL_TUNE_ORDER];
    return !ds.shape && n_hint < (1ull << 32) &&
           (om == 2 || (om == 1 && n_hint <= 64ull * 256ull * (unsigned long long)grid));
}

// The persistent schedule's launches on `st` over a workspace `ws` laid out
// by scott_ws_layout and the queue counters `q` (kQueueSlotWidth words).
// `timed` records the trace and stepping-kernel events of this thread.
template <typename R, class F, class FJ, typename X>
static hipError_t scott_persistent_launches(const F& f, const ScottParams& P, const Seeds& s, unsigned long long n,
                                            const ScottOut& o, const DevScalars& ds, hipStream_t st,
                                            unsigned long long n_hint, const FJ& fj, char* ws, const ScottWs& w,
                                            bool ordered, unsigned long long* q,
                                            decltype(&k_scott_persistent<R, F, FJ, X>) kern, int grid, bool timed) {
    R* frames = (R*)(ws + w.frames);
    ScottRec<R>* recs = (ScottRec<R>*)(ws + w.recs);
    unsigned* order = ordered ? (unsigned*)(ws + w.order) : nullptr;
    unsigned char* key = ordered ? (unsigned char*)ws + w.key : nullptr;
    unsigned* counts = ordered ? (unsigned*)(ws + w.counts) : nullptr;
    const int nq = g_tune[FQSL_TUNE_QUEUES];
    const int refill_min = g_tune[FQSL_TUNE_REFILL_MIN];
    k_zero_u64<<<1, 64, 0, st>>>(q, kQueueSlotWidth);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (timed && record_start(st)) return hipErrorUnknown;
    if (ordered) {
        k_zero_u32<<<1, 256, 0, st>>>(counts, kOrderBuckets);
        k_order_predict<R, F><<<grid_for((long)n_hint, 256), 256, 0, st>>>(
            f, s, n, g_tune[FQSL_TUNE_ORDER_BUCKETS], key, counts);
        k_order_scan<<<1, kOrderBuckets, 0, st>>>(counts);
        const int sg = grid_for((long)n_hint, 256) < 4096 ? grid_for((long)n_hint, 256) : 4096;
        k_order_scatter<<<sg, 256, 0, st>>>(key, n, counts, order);
    }
    k_scott_prep<R, F><<<grid_for((long)n_hint, 256), 256, 0, st>>>(f, P, s, n, frames, o, ds);
    if (timed) e = hipEventRecord(g_evs0, st);
    kern<<<grid, 256, 0, st>>>(f, P, s, n, frames, recs, ds, q, nq, refill_min, order, fj);
    if (timed && e == hipSuccess) e = hipEventRecord(g_evs1, st);
    if (timed && e == hipSuccess) g_evs_valid = true;
    k_scott_finish<R, F><<<grid_for((long)n_hint, 256), 256, 0, st>>>(f, P, n, frames, recs, o, ds);
    if (e == hipSuccess) e = hipGetLastError();
    if (timed && e == hipSuccess && record_end(st)) return hipErrorUnknown;
    return e;
}

// ---------------------------------------------- scratch of the no-alloc entry
// fqsl_trace_scott_noalloc runs the persistent schedule over this array
// instead of allocated memory: the queue counters in its first 256 bytes,
// then the workspace of up to 2^20 FP32C or 2^19 FP64 lines (ordered queue
// included).  Each device holds its own copy, at its own address, from the
// library's load until the process ends.
constexpr unsigned long long kNoAllocLinesFp32c = 1ull << 20;
constexpr unsigned long long kNoAllocLinesFp64 = 1ull << 19;
constexpr size_t kNoAllocQueueBytes = 256;
constexpr size_t kNoAllocWsBytes =
    scott_ws_layout<float>(kNoAllocLinesFp32c, true).bytes > scott_ws_layout<double>(kNoAllocLinesFp64, true).bytes
        ? scott_ws_layout<float>(kNoAllocLinesFp32c, true).bytes
        : scott_ws_layout<double>(kNoAllocLinesFp64, true).bytes;
constexpr size_t kNoAllocScratchBytes = kNoAllocQueueBytes + align256(kNoAllocWsBytes);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// X: the position path's type (R, or double for the mixed variant).
template <typename R, class F, class FJ = NoCurrent, typename X = R>
static int launch_scott(const F& f, const ScottParams& P, const Seeds& s, unsigned long long n, const ScottOut& o,
                        const DevScalars& ds, int schedule, hipStream_t st, unsigned long long n_hint,
                        char* noalloc, const FJ& fj = FJ{}) {
    if (schedule == FQSL_SCHED_PERSISTENT) {
        // n_hint: number of lines, known on the host (sizes the workspace)
        if (n_hint == 0) return FQSL_OK;
        decltype(&k_scott_persistent<R, F, FJ, X>) kern = nullptr;
        int grid = 0;
        int rc = scott_persistent_kernel<R, F, FJ, X>(&kern, &grid, n_hint);
        if (rc) return rc;
        const bool ordered = scott_ordered<R>(ds, n_hint, grid);
        const ScottWs w = scott_ws_layout<R>(n_hint, ordered);
        if (noalloc) {  // device address of this device's scratch array: no allocation, no events
            if (w.bytes > kNoAllocScratchBytes - kNoAllocQueueBytes)
                return set_err(FQSL_ERR_INVALID, "%llu lines need %zu scratch bytes, the array holds %zu",
                               n_hint, w.bytes, kNoAllocScratchBytes - kNoAllocQueueBytes);
            g_ev_valid = false;
            g_evs_valid = false;
            const hipError_t e = scott_persistent_launches<R, F, FJ, X>(
                f, P, s, n, o, ds, st, n_hint, fj, noalloc + kNoAllocQueueBytes, w, ordered,
                (unsigned long long*)noalloc, kern, grid, false);
            if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "Scott trace launch: %s", hipGetErrorString(e));
            return FQSL_OK;
        }
        unsigned long long* q = nullptr;
        rc = queue_slot(&q);
        if (rc) return rc;
        void* ws = nullptr;
        HIP_TRY(hipMallocAsync(&ws, w.bytes, st));
        const hipError_t e = scott_persistent_launches<R, F, FJ, X>(f, P, s, n, o, ds, st, n_hint, fj, (char*)ws,
                                                                    w, ordered, q, kern, grid, true);
        const hipError_t ef = hipFreeAsync(ws, st);
        if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "Scott trace launch: %s", hipGetErrorString(e));
        if (ef != hipSuccess) return set_err(FQSL_ERR_HIP, "hipFreeAsync: %s", hipGetErrorString(ef));
        return FQSL_OK;
    }
    if (schedule == FQSL_SCHED_PER_LINE) {
        int rc = record_start(st);
        if (rc) return rc;
        if (noalloc) return set_err(FQSL_ERR_INVALID, "the no-alloc entry runs the persistent schedule only");
        const int waves = scott_waves<R, X, FJ>(n_hint);
        auto kern = waves == 2 ? k_scott_per_line_w2<R, F, FJ, X>
                  : waves == 3 ? k_scott_per_line_w3<R, F, FJ, X>
                               : k_scott_per_line<R, F, FJ, X>;
        kern<<<per_line_grid(n_hint, 256, ds), 256, 0, st>>>(f, P, s, n, o, ds, fj);
        HIP_TRY(hipGetLastError());
        return record_end(st);
    }
    return set_err(FQSL_ERR_INVALID, "schedule %d", schedule);
}

// Workgroup size of the one-thread-per-line plain kernel: one wavefront on a
// uniform grid (a finished wavefront frees its slot at once: 14.8 -> 12.7 ms
// on 1024^2 lines of ~470 evaluations), four on a stretched grid, where the
// four wavefronts of adjacent seeds sharing a CU's L1 outweigh that (16.5
// against 16.7 ms; FP64 27.0 against 30.3 ms).
static int per_line_block(const UniformGrid&) { return 64; }
template <typename R>
static int per_line_block(const StretchRef<R>&) { return 256; }
static int per_line_block(const StretchLds&) { return 256; }

template <typename R, class F, class FJ = NoCurrent, class G = UniformGrid>
static int launch_plain(const F& f, const PlainParams& P, const Seeds& s, unsigned long long n, const PlainOut& o,
                        const DevScalars& ds, const float* ncross, int schedule, hipStream_t st,
                        unsigned long long n_hint, const FJ& fj = FJ{}, const G& g = G{}) {
    if (schedule == FQSL_SCHED_PERSISTENT) {
        unsigned long long* q = nullptr;
        int rc = queue_slot(&q);
        if (rc) return rc;
        int grid = 0;
        const int lds = grid_lds_bytes(g);
        rc = persistent_grid(k_plain_persistent<R, F, FJ, G>, &grid, lds);
        if (rc) return rc;
        k_zero_u64<<<1, 64, 0, st>>>(q, kQueueSlotWidth);
        HIP_TRY(hipGetLastError());
        rc = record_start(st);
        if (rc) return rc;
        k_plain_persistent<R, F, FJ, G><<<grid, 256, lds, st>>>(f, P, s, n, o, ds, ncross, q,
                                                             g_tune[FQSL_TUNE_PLAIN_REFILL], fj, g);
        HIP_TRY(hipGetLastError());
        return record_end(st);
    }
    if (schedule == FQSL_SCHED_PER_LINE) {
        int rc = record_start(st);
        if (rc) return rc;
        const int block = per_line_block(g);
        k_plain_per_line<R, F, FJ, G><<<per_line_grid(n_hint, block, ds), block, grid_lds_bytes(g), st>>>(
            f, P, s, n, o, ds, ncross, fj, g);
        HIP_TRY(hipGetLastError());
        return record_end(st);
    }
    return set_err(FQSL_ERR_INVALID, "schedule %d", schedule);
}

static int check_common(int method, int precision, int schedule, bool scott = false) {
    if (method != FQSL_METHOD_RKF45 && method != FQSL_METHOD_RK4) return set_err(FQSL_ERR_INVALID, "method %d", method);
    if (precision == FQSL_PREC_MIXED && !scott)
        return set_err(FQSL_ERR_UNSUPPORTED, "FQSL_PREC_MIXED is a Scott-tracer precision");
    if (precision != FQSL_PREC_FP32C && precision != FQSL_PREC_FP64 && precision != FQSL_PREC_MIXED)
        return set_err(FQSL_ERR_INVALID, "precision %d", precision);
    if (schedule != FQSL_SCHED_PERSISTENT && schedule != FQSL_SCHED_PER_LINE)
        return set_err(FQSL_ERR_INVALID, "schedule %d", schedule);
    return FQSL_OK;
}

static int trace_scott_impl(const fqsl_field* field, const fqsl_scott_params* p, uint64_t n, const float* x,
                            const float* y, const float* z, const fqsl_scott_out* out, void* stream, char* noalloc) {
    if (!field || !p || !out) return set_err(FQSL_ERR_INVALID, "NULL field/params/out");
    int rc = check_common(p->method, p->precision, p->schedule, true);
    if (rc) return rc;
    if (n == 0) return FQSL_OK;
    if (!x || !y || !z) return set_err(FQSL_ERR_INVALID, "NULL seed array");
    if (!(p->step > 0) || !(p->step <= 100.0)) return set_err(FQSL_ERR_INVALID, "step %g outside (0, 100]", p->step);
    if (p->method == FQSL_METHOD_RKF45 && !(p->tol > 0)) return set_err(FQSL_ERR_INVALID, "tol %g must be > 0", p->tol);
    if (p->method == FQSL_METHOD_RKF45 && p->precision == FQSL_PREC_FP32C && p->tol < kFp32cMinTol)
        return set_err(FQSL_ERR_INVALID,
                       "tol %g is below what the FP32 error norm resolves (%g cells); use FQSL_PREC_MIXED or "
                       "FQSL_PREC_FP64", p->tol, kFp32cMinTol);
    ScottParams P{p->method, p->step, p->tol, 0.0};  // ctl is formed on the device
    Seeds s{x, y, z};
    ScottOut o{out->start_x, out->start_y, out->start_z, out->end_x, out->end_y, out->end_z,
               out->flag_start, out->flag_end, out->b_this_x, out->b_this_y, out->b_this_z,
               out->q_0, out->q_perp, out->line_len, out->evals, out->twist};
    if ((o.sx && (!o.sy || !o.sz)) || (o.ex && (!o.ey || !o.ez)) || (o.btx && (!o.bty || !o.btz)))
        return set_err(FQSL_ERR_INVALID, "x/y/z output arrays must be all set or all NULL");
    DevScalars ds{};
    hipStream_t st = (hipStream_t)stream;
    const int* nn = field->n;
    if (o.tw && !field->cur32[0]) return set_err(FQSL_ERR_INVALID, "twist requested but the field has no current density");
    if (field->stretched)
        return set_err(FQSL_ERR_UNSUPPORTED, "the Scott tracer runs on grid-index coordinates; this field has stretched axes");
    if (p->precision == FQSL_PREC_FP64) {
        if (!field->soa64[0]) return set_err(FQSL_ERR_INVALID, "field has no FP64 layout");
        FieldSoA<double> f{field->soa64[0], field->soa64[1], field->soa64[2], {nn[0], nn[1], nn[2]}};
        if (o.tw) {
            FieldSoA<double> fj{field->cur64[0], field->cur64[1], field->cur64[2], {nn[0], nn[1], nn[2]}};
            return launch_scott<double>(f, P, s, n, o, ds, p->schedule, st, n, noalloc, fj);
        }
        return launch_scott<double>(f, P, s, n, o, ds, p->schedule, st, n, noalloc);
    }
    const FieldSoA<float> fj{field->cur32[0], field->cur32[1], field->cur32[2], {nn[0], nn[1], nn[2]}};
    const bool mixed = p->precision == FQSL_PREC_MIXED;
    if (field->cell32) {
        FieldCell32 f{field->cell32, {nn[0], nn[1], nn[2]}};
        if (o.tw) {
            if (!field->curcell) return set_err(FQSL_ERR_INVALID, "current density has no cell layout");
            const FieldCell32 fjc{field->curcell, {nn[0], nn[1], nn[2]}};
            if (mixed)
                return launch_scott<float, FieldCell32, FieldCell32, double>(f, P, s, n, o, ds, p->schedule, st, n,
                                                                             noalloc, fjc);
            return launch_scott<float>(f, P, s, n, o, ds, p->schedule, st, n, noalloc, fjc);
        }
        if (mixed)
            return launch_scott<float, FieldCell32, NoCurrent, double>(f, P, s, n, o, ds, p->schedule, st, n, noalloc);
        return launch_scott<float>(f, P, s, n, o, ds, p->schedule, st, n, noalloc);
    }
    if (field->soa32[0]) {
        FieldSoA<float> f{field->soa32[0], field->soa32[1], field->soa32[2], {nn[0], nn[1], nn[2]}};
        if (mixed) {
            if (o.tw)
                return launch_scott<float, FieldSoA<float>, FieldSoA<float>, double>(f, P, s, n, o, ds, p->schedule,
                                                                                     st, n, noalloc, fj);
            return launch_scott<float, FieldSoA<float>, NoCurrent, double>(f, P, s, n, o, ds, p->schedule, st, n,
                                                                           noalloc);
        }
        if (o.tw) return launch_scott<float>(f, P, s, n, o, ds, p->schedule, st, n, noalloc, fj);
        return launch_scott<float>(f, P, s, n, o, ds, p->schedule, st, n, noalloc);
    }
    return set_err(FQSL_ERR_INVALID, "field has no FP32 layout");
}

extern "C" int fqsl_trace_scott(const fqsl_field* field, const fqsl_scott_params* p, uint64_t n, const float* x,
                                const float* y, const float* z, const fqsl_scott_out* out, void* stream) {
    return trace_scott_impl(field, p, n, x, y, z, out, stream, nullptr);
}

// Per device: the address of its scratch array, looked up once with that
// device current (the array's host-side name is not a device address).
static std::mutex g_noalloc_mu;
static char* g_noalloc_ptr[64];

extern "C" int fqsl_trace_scott_noalloc(const fqsl_field* field, const fqsl_scott_params* p, uint64_t n,
                                        const float* x, const float* y, const float* z, const fqsl_scott_out* out,
                                        void* stream) {
    if (!p) return set_err(FQSL_ERR_INVALID, "NULL params");
    if (p->schedule != FQSL_SCHED_PERSISTENT)
        return set_err(FQSL_ERR_INVALID, "the no-alloc entry runs the persistent schedule only");
    const unsigned long long cap = p->precision == FQSL_PREC_FP64 ? kNoAllocLinesFp64 : kNoAllocLinesFp32c;  // mixed: float records
    if (n > cap)
        return set_err(FQSL_ERR_INVALID, "%llu lines exceed the no-alloc maximum of %llu for this precision",
                       (unsigned long long)n, cap);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_err(FQSL_ERR_UNSUPPORTED, "device index %d", dev);
    char* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (!g_noalloc_ptr[dev]) {
            void* a = nullptr;
            const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !a)
                return set_err(FQSL_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            (void)cu_count();  // cache the CU count now: later calls only read it
            g_noalloc_ptr[dev] = (char*)a;
        }
        ptr = g_noalloc_ptr[dev];
    }
    return trace_scott_impl(field, p, n, x, y, z, out, stream, ptr);
}

// ---------------------------------------------------- current density (curl)
// curl B on the node grid of non-uniform axes (FastQSL.curl_of_B_grid,
// fastqslpy/FastQSL.py:120-163): at an interior node the derivative along an
// axis is the three-point rule
//   ((c[+1] - c[0]) / dr * dl / (dr + dl) + (c[0] - c[-1]) / dl * dr / (dr + dl))
// with dl, dr the spacings to the neighbours; boundary nodes take the value
// of the nearest interior node (copy_boundary, :133-141, whose six plane
// copies leave every boundary node with the interior value at the clamped
// index).  The differences of B are taken in B's own type and everything
// after in double, in the reference's operation order, so the results are
// the same bits as its array arithmetic.
#pragma clang fp contract(off)
template <typename T>
__device__ __forceinline__ double diff3(const T* __restrict__ c, long idx, long stride, const double* __restrict__ ax,
                                        int i) {
    const double dl = ax[i] - ax[i - 1], dr = ax[i + 1] - ax[i];
    const T dp = c[idx + stride] - c[idx];
    const T dm = c[idx] - c[idx - stride];
    return (double)dp / dr * dl / (dr + dl) + (double)dm / dl * dr / (dr + dl);
}

template <typename T>
__global__ void __launch_bounds__(256) k_curl(const T* __restrict__ bx, const T* __restrict__ by,
                                              const T* __restrict__ bz, int nx, int ny, int nz,
                                              const double* __restrict__ ax, const double* __restrict__ ay,
                                              const double* __restrict__ az, double* __restrict__ cx,
                                              double* __restrict__ cy, double* __restrict__ cz) {
    const long nxy = (long)nx * ny, nn = nxy * nz;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nn; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / nxy);
        const int ci = i < 1 ? 1 : (i > nx - 2 ? nx - 2 : i);
        const int cj = j < 1 ? 1 : (j > ny - 2 ? ny - 2 : j);
        const int ck = k < 1 ? 1 : (k > nz - 2 ? nz - 2 : k);
        const long c = (long)ck * nxy + (long)cj * nx + ci;
        cx[t] = diff3(bz, c, nx, ay, cj) - diff3(by, c, nxy, az, ck);
        cy[t] = diff3(bx, c, nxy, az, ck) - diff3(bz, c, 1, ax, ci);
        cz[t] = diff3(by, c, 1, ax, ci) - diff3(bx, c, nx, ay, cj);
    }
}

// diff_3point (FastQSL.py:143-163): the derivative along `axis` at every
// interior node, out[(k-1)][(j-1)][(i-1)] (x fastest, (nx-2)(ny-2)(nz-2) values).
template <typename T>
__global__ void __launch_bounds__(256) k_diff3(const T* __restrict__ cube, int nx, int ny, int nz,
                                               const double* __restrict__ ax, int axis, double* __restrict__ out) {
    const long mx = nx - 2, my = ny - 2, mxy = mx * my, nn = mxy * (nz - 2), nxy = (long)nx * ny;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nn; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % mx) + 1, j = (int)((t / mx) % my) + 1, k = (int)(t / mxy) + 1;
        const long c = (long)k * nxy + (long)j * nx + i;
        out[t] = axis == 0 ? diff3(cube, c, 1, ax, i) : (axis == 1 ? diff3(cube, c, nx, ax, j) : diff3(cube, c, nxy, ax, k));
    }
}
#pragma clang fp contract(on)

// copy_boundary in place: every boundary node takes the clamped interior value.
__global__ void __launch_bounds__(256) k_copy_boundary(double* __restrict__ a, int nx, int ny, int nz) {
    const long nxy = (long)nx * ny, nn = nxy * nz;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nn; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / nxy);
        const int ci = i < 1 ? 1 : (i > nx - 2 ? nx - 2 : i);
        const int cj = j < 1 ? 1 : (j > ny - 2 ? ny - 2 : j);
        const int ck = k < 1 ? 1 : (k > nz - 2 ? nz - 2 : k);
        if (ci != i || cj != j || ck != k) a[t] = a[(long)ck * nxy + (long)cj * nx + ci];
    }
}

static int check_grid3(int nx, int ny, int nz, int dtype) {
    if (nx < 3 || ny < 3 || nz < 3) return set_err(FQSL_ERR_INVALID, "each axis needs >= 3 nodes (%d,%d,%d)", nx, ny, nz);
    if (dtype != FQSL_DTYPE_F32 && dtype != FQSL_DTYPE_F64) return set_err(FQSL_ERR_INVALID, "dtype %d", dtype);
    if ((long)nx * ny * nz > (1L << 31) - 1) return set_err(FQSL_ERR_INVALID, "grid exceeds int32 indexing");
    return FQSL_OK;
}

extern "C" int fqsl_curl(const void* bx, const void* by, const void* bz, int nx, int ny, int nz, int dtype,
                         const double* ax, const double* ay, const double* az, double* cx, double* cy, double* cz,
                         void* stream) {
    int rc = check_grid3(nx, ny, nz, dtype);
    if (rc) return rc;
    if (!bx || !by || !bz || !ax || !ay || !az || !cx || !cy || !cz) return set_err(FQSL_ERR_INVALID, "NULL array");
    hipStream_t st = (hipStream_t)stream;
    const int g = grid_for((long)nx * ny * nz, 256);
    if (dtype == FQSL_DTYPE_F32)
        k_curl<float><<<g, 256, 0, st>>>((const float*)bx, (const float*)by, (const float*)bz, nx, ny, nz, ax, ay, az, cx,
                                         cy, cz);
    else
        k_curl<double><<<g, 256, 0, st>>>((const double*)bx, (const double*)by, (const double*)bz, nx, ny, nz, ax, ay,
                                          az, cx, cy, cz);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}

extern "C" int fqsl_diff_3point(const void* cube, int nx, int ny, int nz, int dtype, const double* ax, int axis,
                                double* out, void* stream) {
    int rc = check_grid3(nx, ny, nz, dtype);
    if (rc) return rc;
    if (axis < 0 || axis > 2) return set_err(FQSL_ERR_INVALID, "axis %d", axis);
    if (!cube || !ax || !out) return set_err(FQSL_ERR_INVALID, "NULL array");
    hipStream_t st = (hipStream_t)stream;
    const int g = grid_for((long)(nx - 2) * (ny - 2) * (nz - 2), 256);
    if (dtype == FQSL_DTYPE_F32) k_diff3<float><<<g, 256, 0, st>>>((const float*)cube, nx, ny, nz, ax, axis, out);
    else k_diff3<double><<<g, 256, 0, st>>>((const double*)cube, nx, ny, nz, ax, axis, out);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}

extern "C" int fqsl_copy_boundary(double* a, int nx, int ny, int nz, void* stream) {
    int rc = check_grid3(nx, ny, nz, FQSL_DTYPE_F64);
    if (rc) return rc;
    if (!a) return set_err(FQSL_ERR_INVALID, "NULL array");
    k_copy_boundary<<<grid_for((long)nx * ny * nz, 256), 256, 0, (hipStream_t)stream>>>(a, nx, ny, nz);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}

// curl B in grid-index units as demo/Q_RK4.ipynb cell [11] (and
// demo/npzinputRKF45CuPyTwist.ipynb cell [9]) computes it for the twist: on a
// zero array, central differences (c[+1] - c[-1]) / 2 on the index ranges the
// notebook slices (curl_x on [:, 1:-1, 1:-1], curl_y on [1:-1, :, 1:-1],
// curl_z on [1:-1, 1:-1, :]), then at z = 0 on [1:-1, 1:-1] the one-sided
// derivative (-3 c[0] + 4 c[1] - c[2]) / 2 along z for curl_x and curl_y.
// In B's type and in the notebook's operation order, without contraction,
// so the result is the same bits as its array arithmetic.
#pragma clang fp contract(off)
template <typename T>
__global__ void __launch_bounds__(256) k_curl_index(const T* __restrict__ bx, const T* __restrict__ by,
                                                    const T* __restrict__ bz, int nx, int ny, int nz,
                                                    T* __restrict__ cx, T* __restrict__ cy, T* __restrict__ cz) {
    const long nxy = (long)nx * ny, nn = nxy * nz;
    const T two = (T)2, three = (T)3, four = (T)4;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nn; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / nxy);
        const bool ii = i >= 1 && i <= nx - 2, jj = j >= 1 && j <= ny - 2, kk = k >= 1 && k <= nz - 2;
        T vx = (T)0, vy = (T)0, vz = (T)0;
        if (jj && kk) vx = (bz[t + nx] - bz[t - nx]) / two - (by[t + nxy] - by[t - nxy]) / two;
        if (ii && kk) vy = (bx[t + nxy] - bx[t - nxy]) / two - (bz[t + 1] - bz[t - 1]) / two;
        if (ii && jj) vz = (by[t + 1] - by[t - 1]) / two - (bx[t + nx] - bx[t - nx]) / two;
        if (k == 0 && ii && jj) {
            vx = (bz[t + nx] - bz[t - nx]) / two - ((-three * by[t] + four * by[t + nxy]) - by[t + 2 * nxy]) / two;
            vy = ((-three * bx[t] + four * bx[t + nxy]) - bx[t + 2 * nxy]) / two - (bz[t + 1] - bz[t - 1]) / two;
        }
        cx[t] = vx;
        cy[t] = vy;
        cz[t] = vz;
    }
}

// slog10 Q = log10(Q) * sign(Bn) [* (flag <= 6)] (demo/Q_RKF45CuPy.ipynb
// cells [13]-[14], demo/Q_RKF45MixCuPyDiffScott.ipynb cell [8]), with
// NumPy's sign (0 at 0, NaN at NaN) and its left-to-right products.
__global__ void __launch_bounds__(256) k_slog10q(unsigned long long n, const float* __restrict__ q,
                                                 const float* __restrict__ bn, const int32_t* __restrict__ flag,
                                                 float* __restrict__ out) {
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const float b = bn[i];
        const float sg = b > 0.f ? 1.f : (b < 0.f ? -1.f : (b == 0.f ? 0.f : b));
        float r = log10f(q[i]) * sg;
        if (flag) r = r * (flag[i] <= 6 ? 1.f : 0.f);
        out[i] = r;
    }
}
#pragma clang fp contract(on)

extern "C" int fqsl_curl_index(const void* bx, const void* by, const void* bz, int nx, int ny, int nz, int dtype,
                               void* cx, void* cy, void* cz, void* stream) {
    int rc = check_grid3(nx, ny, nz, dtype);
    if (rc) return rc;
    if (!bx || !by || !bz || !cx || !cy || !cz) return set_err(FQSL_ERR_INVALID, "NULL array");
    hipStream_t st = (hipStream_t)stream;
    const int g = grid_for((long)nx * ny * nz, 256);
    if (dtype == FQSL_DTYPE_F32)
        k_curl_index<float><<<g, 256, 0, st>>>((const float*)bx, (const float*)by, (const float*)bz, nx, ny, nz,
                                               (float*)cx, (float*)cy, (float*)cz);
    else
        k_curl_index<double><<<g, 256, 0, st>>>((const double*)bx, (const double*)by, (const double*)bz, nx, ny, nz,
                                                (double*)cx, (double*)cy, (double*)cz);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}

extern "C" int fqsl_slog10q(uint64_t n, const float* q, const float* bn, const int32_t* flag, float* out,
                            void* stream) {
    if (n == 0) return FQSL_OK;
    if (!q || !bn || !out) return set_err(FQSL_ERR_INVALID, "NULL array");
    k_slog10q<<<grid_for((long)((n + 255) / 256 * 256), 256), 256, 0, (hipStream_t)stream>>>(n, q, bn, flag, out);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}

// ----------------------------------------------------------------- batches
// A batch of Scott requests, each with its own field, method, step,
// tolerance, seeds and outputs.  Requests are grouped by (field, method); the
// lines of a group form one contiguous range [base, base + n) of the batch's
// line index and are drained by one persistent launch.  A line finds its
// request by binary search over the requests' first-line offsets and takes
// that request's step and tolerance, so one launch mixes requests of any
// size.  All state lives in a workspace laid out at plan time, and a run only
// launches kernels (the counters are cleared by a kernel), so a run can be
// captured in a graph and replayed.
struct BatchReq {
    Seeds s;
    ScottOut o;
    double step, tol;
    unsigned long long n;
};

struct BatchCtl {
    unsigned long long* queues;     // one counter per group
    unsigned long long* remaining;  // lines of each request not yet finished
    unsigned long long* stamps;     // [r]: device wall clock when request r's last line
                                    // finished stepping; [nreq]: start of the run
};

// Largest r in [r0, r1) with offs[r] <= i (requests without lines share the
// offset of the next one and are never returned for a line).
__device__ __forceinline__ int batch_find(const unsigned long long* __restrict__ offs, int r0, int r1,
                                          unsigned long long i) {
    int lo = r0, hi = r1 - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void batch_line_done(const BatchCtl& c, int r) {
    if (atomicSub(c.remaining + r, 1ull) == 1ull) c.stamps[r] = wall_clock64();
}

__global__ void k_batch_reset(BatchCtl c, const BatchReq* __restrict__ reqs, int nreq, int ngroups) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ngroups) c.queues[t] = 0ull;
    if (t < nreq) {
        c.remaining[t] = reqs[t].n;
        c.stamps[t] = 0ull;
    }
    if (t == 0) c.stamps[nreq] = wall_clock64();
}

struct BatchGroupArgs {
    const BatchReq* reqs;
    const unsigned long long* offs;
    int r0, r1;
    unsigned long long base, n;
};

template <typename R, class F>
__global__ void __launch_bounds__(256) k_batch_prep(F f, BatchGroupArgs g, R* __restrict__ frames) {
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < g.n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long gi = g.base + i;
        const int r = batch_find(g.offs, g.r0, g.r1, gi);
        const unsigned long long li = gi - g.offs[r];
        const Seeds s = g.reqs[r].s;
        const float p0[3] = {s.x[li], s.y[li], s.z[li]};
        R u0[3], v0[3], b0z;
        double nb0;
        seed_frame<R>(f, p0, u0, v0, &b0z, &nb0);
        R* fr = frames + (unsigned long long)kFrameWords * gi;
#pragma unroll
        for (int a = 0; a < 3; ++a) { fr[a] = u0[a]; fr[3 + a] = v0[a]; }
        fr[6] = b0z;
        fr[7] = (R)nb0;
        const ScottOut o = g.reqs[r].o;
        if (o.btx) {
            float bt[3];
            b_clamped<R>(f, p0, bt);
            o.btx[li] = bt[0]; o.bty[li] = bt[1]; o.btz[li] = bt[2];
        }
    }
}

// Persistent stepping over one group: as scott_persistent_body, with the
// step and tolerance of each lane's request.
template <typename R, class F, typename X, bool kRoll = false>
__device__ __forceinline__ void scott_batch_body(F f, int method, BatchGroupArgs g, const R* __restrict__ frames,
                                                 ScottRec<R>* __restrict__ recs, BatchCtl c, int qi,
                                                 int refill_min) {
    ScottLine<R, X> L;
    ScottParams P{method, 0.0, 0.0, 0.0};
    bool active = false;
    bool drained = false;
    unsigned long long idx = 0;
    const int lane = __lane_id();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    QueueCursor qc{0, 0};
    for (;;) {
        if (!drained) {
            const unsigned long long need = __ballot(!active);
            const int cnt = __popcll(need);
            if (cnt >= refill_min || need == ~0ull) {
                const int leader = __ffsll((long long)need) - 1;
                unsigned long long base = 0;
                int avail = 0;
                if (!claim(c.queues + qi, 1, g.n, qc, cnt, leader, base, avail)) drained = true;
                const int rank = __popcll(need & below);
                if (!active && rank < avail) {
                    idx = g.base + base + (unsigned long long)rank;
                    const int r = batch_find(g.offs, g.r0, g.r1, idx);
                    const unsigned long long li = idx - g.offs[r];
                    P.s0 = g.reqs[r].step;
                    P.tol = g.reqs[r].tol;
                    P.ctl = scott_ctl(P.tol);
                    const Seeds s = g.reqs[r].s;
                    ScottEnd<R> e;
                    const int st = scott_init<R>(L, f, P, s.x[li], s.y[li], s.z[li], e,
                                                 frames + (unsigned long long)kFrameWords * idx);
                    if (st != kTracing) {
                        recs[idx] = ScottRec<R>{L.st, e, L.len, st, L.evals, L.tw};
                        batch_line_done(c, r);
                    } else {
                        active = true;
                    }
                }
            }
        }
        if (!__any(active)) {
            if (drained) break;
            continue;
        }
        if (active) {
            ScottEnd<R> e;
            const int st = scott_attempt<R, F, NoCurrent, X, kRoll>(L, f, P, e);
            if (st != kTracing) {
                recs[idx] = ScottRec<R>{L.st, e, L.len, st, L.evals, L.tw};
                active = false;
                batch_line_done(c, batch_find(g.offs, g.r0, g.r1, idx));
            }
        }
    }
}

template <typename R, class F, typename X>
__global__ void __launch_bounds__(256) k_batch_step(F f, int method, BatchGroupArgs g, const R* frames,
                                                    ScottRec<R>* recs, BatchCtl c, int qi, int refill_min) {
    scott_batch_body<R, F, X>(f, method, g, frames, recs, c, qi, refill_min);
}

template <typename R, class F, typename X>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_batch_step_w2(F f, int method, BatchGroupArgs g, const R* frames, ScottRec<R>* recs, BatchCtl c, int qi,
                int refill_min) {
    scott_batch_body<R, F, X>(f, method, g, frames, recs, c, qi, refill_min);
}

template <typename R, class F, typename X>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_batch_step_w3(F f, int method, BatchGroupArgs g, const R* frames, ScottRec<R>* recs, BatchCtl c, int qi,
                int refill_min) {
    scott_batch_body<R, F, X, true>(f, method, g, frames, recs, c, qi, refill_min);
}

template <typename R, class F>
__global__ void __launch_bounds__(256) k_batch_finish(F f, BatchGroupArgs g, const R* __restrict__ frames,
                                                      const ScottRec<R>* __restrict__ recs) {
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < g.n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long gi = g.base + i;
        const int r = batch_find(g.offs, g.r0, g.r1, gi);
        const unsigned long long li = gi - g.offs[r];
        const ScottOut o = g.reqs[r].o;
        const ScottRec<R> rec = recs[gi];
        double q0, qp;
        scott_q<R>(rec.st, rec.en, rec.status, (double)frames[(unsigned long long)kFrameWords * gi + 7], f, &q0,
                   &qp);
        if (o.sx) { o.sx[li] = (float)rec.st.x[0]; o.sy[li] = (float)rec.st.x[1]; o.sz[li] = (float)rec.st.x[2]; }
        if (o.ex) { o.ex[li] = (float)rec.en.x[0]; o.ey[li] = (float)rec.en.x[1]; o.ez[li] = (float)rec.en.x[2]; }
        if (o.fs) o.fs[li] = rec.st.flag;
        if (o.fe) o.fe[li] = rec.en.flag;
        if (o.q0) o.q0[li] = (float)q0;
        if (o.qp) o.qp[li] = (float)qp;
        if (o.len) o.len[li] = rec.len;
        if (o.evals) o.evals[li] = rec.evals;
        if (o.tw) o.tw[li] = rec.tw;
    }
}

struct BatchGroup {
    const fqsl_field* field;
    int method;
    int r0, r1;
    unsigned long long base, n;
};

struct fqsl_scott_batch {
    int device;
    int precision;
    int nreq;
    int refill_min;
    int waves;
    std::vector<BatchGroup> groups;
    std::vector<int> order;  // batch position -> caller's request index
    char* ws;                // caller's workspace
    BatchReq* reqs;
    unsigned long long* offs;
    BatchCtl ctl;
    void* frames;
    void* recs;
};

namespace {
// Workspace layout of a batch: [queues | remaining | stamps | reqs | offs | frames | recs]
struct BatchLayout {
    size_t queues, remaining, stamps, reqs, offs, frames, recs, total;
};

BatchLayout batch_layout(int nreq, int ngroups, unsigned long long lines, int precision) {
    const size_t rsz = precision == FQSL_PREC_FP64 ? sizeof(double) : sizeof(float);
    const size_t recsz = precision == FQSL_PREC_FP64 ? sizeof(ScottRec<double>) : sizeof(ScottRec<float>);
    BatchLayout L;
    size_t o = 0;
    L.queues = o; o = align256(o + sizeof(unsigned long long) * (size_t)(ngroups > 0 ? ngroups : 1));
    L.remaining = o; o = align256(o + sizeof(unsigned long long) * (size_t)(nreq > 0 ? nreq : 1));
    L.stamps = o; o = align256(o + sizeof(unsigned long long) * (size_t)(nreq + 1));
    L.reqs = o; o = align256(o + sizeof(BatchReq) * (size_t)(nreq > 0 ? nreq : 1));
    L.offs = o; o = align256(o + sizeof(unsigned long long) * (size_t)(nreq > 0 ? nreq : 1));
    L.frames = o; o = align256(o + (size_t)kFrameWords * rsz * lines);
    L.recs = o; o = align256(o + recsz * lines);
    L.total = o;
    return L;
}

int check_request(const fqsl_scott_request& q, int precision, int k) {
    if (!q.field) return set_err(FQSL_ERR_INVALID, "request %d: NULL field", k);
    if (q.field->stretched) return set_err(FQSL_ERR_UNSUPPORTED, "request %d: field has stretched axes", k);
    if (q.method != FQSL_METHOD_RKF45 && q.method != FQSL_METHOD_RK4)
        return set_err(FQSL_ERR_INVALID, "request %d: method %d", k, q.method);
    if (!(q.step > 0) || !(q.step <= 100.0)) return set_err(FQSL_ERR_INVALID, "request %d: step %g outside (0, 100]", k, q.step);
    if (q.method == FQSL_METHOD_RKF45 && !(q.tol > 0)) return set_err(FQSL_ERR_INVALID, "request %d: tol %g must be > 0", k, q.tol);
    if (q.method == FQSL_METHOD_RKF45 && precision == FQSL_PREC_FP32C && q.tol < kFp32cMinTol)
        return set_err(FQSL_ERR_INVALID, "request %d: tol %g is below what the FP32 error norm resolves (%g); use "
                       "FQSL_PREC_MIXED or FQSL_PREC_FP64", k, q.tol, kFp32cMinTol);
    if (q.n && (!q.x || !q.y || !q.z)) return set_err(FQSL_ERR_INVALID, "request %d: NULL seed array", k);
    if (q.out.twist) return set_err(FQSL_ERR_UNSUPPORTED, "request %d: twist is not available in batches", k);
    const fqsl_scott_out& o = q.out;
    if ((o.start_x && (!o.start_y || !o.start_z)) || (o.end_x && (!o.end_y || !o.end_z)) ||
        (o.b_this_x && (!o.b_this_y || !o.b_this_z)))
        return set_err(FQSL_ERR_INVALID, "request %d: x/y/z output arrays must be all set or all NULL", k);
    if (precision == FQSL_PREC_FP64 && !q.field->soa64[0])
        return set_err(FQSL_ERR_INVALID, "request %d: field has no FP64 layout", k);
    if (precision != FQSL_PREC_FP64 && !q.field->cell32 && !q.field->soa32[0])
        return set_err(FQSL_ERR_INVALID, "request %d: field has no FP32 layout", k);
    return FQSL_OK;
}

// Requests in batch order: grouped by (field, method), stable within a group.
std::vector<int> batch_order(const fqsl_scott_request* reqs, int nreq) {
    std::vector<int> ord(nreq);
    for (int k = 0; k < nreq; ++k) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        if (reqs[a].field != reqs[b].field) return reqs[a].field < reqs[b].field;
        return reqs[a].method < reqs[b].method;
    });
    return ord;
}

template <typename R, class F, typename X = R>
int batch_launch_group(const fqsl_scott_batch* b, const BatchGroup& G, int gi, const F& f, hipStream_t st) {
    BatchGroupArgs g{b->reqs, b->offs, G.r0, G.r1, G.base, G.n};
    R* frames = (R*)b->frames;
    ScottRec<R>* recs = (ScottRec<R>*)b->recs;
    const int knob = sizeof(R) != 4 ? 0 : (sizeof(X) == 8 ? 2 : b->waves);
    const int waves = (knob == 3 && G.n <= small_launch_lines()) ? 2 : knob;
    auto kern = waves == 2 ? k_batch_step_w2<R, F, X> : waves == 3 ? k_batch_step_w3<R, F, X> : k_batch_step<R, F, X>;
    int grid = 0;
    int rc = persistent_grid(kern, &grid);
    if (rc) return rc;
    grid = (int)((long)grid * g_tune[FQSL_TUNE_GRID_PCT] / 100);
    if (grid < 1) grid = 1;
    k_batch_prep<R, F><<<grid_for((long)G.n, 256), 256, 0, st>>>(f, g, frames);
    kern<<<grid, 256, 0, st>>>(f, G.method, g, frames, recs, b->ctl, gi, b->refill_min);
    k_batch_finish<R, F><<<grid_for((long)G.n, 256), 256, 0, st>>>(f, g, frames, recs);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}
}  // namespace

extern "C" size_t fqsl_scott_batch_workspace_bytes(const fqsl_scott_request* reqs, int nreq, int precision) {
    if (!reqs || nreq <= 0) return 0;
    unsigned long long lines = 0;
    for (int k = 0; k < nreq; ++k) lines += reqs[k].n;
    std::vector<int> ord = batch_order(reqs, nreq);
    int ngroups = 0;
    for (int k = 0; k < nreq; ++k)
        if (k == 0 || reqs[ord[k]].field != reqs[ord[k - 1]].field || reqs[ord[k]].method != reqs[ord[k - 1]].method)
            ++ngroups;
    return batch_layout(nreq, ngroups, lines, precision).total;
}

extern "C" int fqsl_scott_batch_create(const fqsl_scott_request* reqs, int nreq, int precision, void* workspace,
                                       size_t workspace_bytes, void* stream, fqsl_scott_batch** out) {
    if (!out) return set_err(FQSL_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!reqs || nreq <= 0) return set_err(FQSL_ERR_INVALID, "empty batch");
    if (precision != FQSL_PREC_FP32C && precision != FQSL_PREC_FP64 && precision != FQSL_PREC_MIXED)
        return set_err(FQSL_ERR_INVALID, "precision %d", precision);
    for (int k = 0; k < nreq; ++k) {
        int rc = check_request(reqs[k], precision, k);
        if (rc) return rc;
    }
    const size_t need = fqsl_scott_batch_workspace_bytes(reqs, nreq, precision);
    if (!workspace || workspace_bytes < need)
        return set_err(FQSL_ERR_INVALID, "workspace of %zu bytes given, %zu needed", workspace_bytes, need);
    if (((uintptr_t)workspace & 255) != 0) return set_err(FQSL_ERR_INVALID, "workspace must be 256-byte aligned");
    fqsl_scott_batch* b = new fqsl_scott_batch();
    HIP_TRY(hipGetDevice(&b->device));
    b->precision = precision;
    b->nreq = nreq;
    b->refill_min = g_tune[FQSL_TUNE_REFILL_MIN];
    b->waves = g_tune[FQSL_TUNE_SCOTT_WAVES];
    b->order = batch_order(reqs, nreq);
    std::vector<BatchReq> table(nreq);
    std::vector<unsigned long long> offs(nreq);
    unsigned long long lines = 0;
    for (int k = 0; k < nreq; ++k) {
        const fqsl_scott_request& q = reqs[b->order[k]];
        if (k == 0 || q.field != b->groups.back().field || q.method != b->groups.back().method)
            b->groups.push_back(BatchGroup{q.field, q.method, k, k, lines, 0});
        BatchGroup& G = b->groups.back();
        G.r1 = k + 1;
        G.n += q.n;
        offs[k] = lines;
        const fqsl_scott_out& o = q.out;
        table[k] = BatchReq{Seeds{q.x, q.y, q.z},
                            ScottOut{o.start_x, o.start_y, o.start_z, o.end_x, o.end_y, o.end_z, o.flag_start,
                                     o.flag_end, o.b_this_x, o.b_this_y, o.b_this_z, o.q_0, o.q_perp, o.line_len,
                                     o.evals, nullptr},
                            q.step, q.method == FQSL_METHOD_RKF45 ? q.tol : 0.0, q.n};
        lines += q.n;
    }
    const BatchLayout L = batch_layout(nreq, (int)b->groups.size(), lines, precision);
    char* ws = (char*)workspace;
    b->ws = ws;
    b->reqs = (BatchReq*)(ws + L.reqs);
    b->offs = (unsigned long long*)(ws + L.offs);
    b->ctl = BatchCtl{(unsigned long long*)(ws + L.queues), (unsigned long long*)(ws + L.remaining),
                      (unsigned long long*)(ws + L.stamps)};
    b->frames = ws + L.frames;
    b->recs = ws + L.recs;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(b->reqs, table.data(), sizeof(BatchReq) * nreq, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->offs, offs.data(), sizeof(unsigned long long) * nreq, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        delete b;
        return set_err(FQSL_ERR_HIP, "batch table upload: %s", hipGetErrorString(e));
    }
    *out = b;
    return FQSL_OK;
}

extern "C" int fqsl_scott_batch_run(const fqsl_scott_batch* b, void* stream) {
    if (!b) return set_err(FQSL_ERR_INVALID, "batch is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(st, &cap));
    const bool timed = cap == hipStreamCaptureStatusNone;
    if (timed) {
        int rc = record_start(st);
        if (rc) return rc;
    }
    const int ng = (int)b->groups.size();
    k_batch_reset<<<grid_for((long)(b->nreq + 1 > ng ? b->nreq + 1 : ng), 256), 256, 0, st>>>(b->ctl, b->reqs, b->nreq,
                                                                                             ng);
    HIP_TRY(hipGetLastError());
    for (int gi = 0; gi < ng; ++gi) {
        const BatchGroup& G = b->groups[gi];
        if (G.n == 0) continue;
        const fqsl_field* fd = G.field;
        int rc;
        if (b->precision == FQSL_PREC_FP64) {
            FieldSoA<double> f{fd->soa64[0], fd->soa64[1], fd->soa64[2], {fd->n[0], fd->n[1], fd->n[2]}};
            rc = batch_launch_group<double>(b, G, gi, f, st);
        } else if (fd->cell32) {
            FieldCell32 f{fd->cell32, {fd->n[0], fd->n[1], fd->n[2]}};
            rc = b->precision == FQSL_PREC_MIXED ? batch_launch_group<float, FieldCell32, double>(b, G, gi, f, st)
                                                 : batch_launch_group<float>(b, G, gi, f, st);
        } else {
            FieldSoA<float> f{fd->soa32[0], fd->soa32[1], fd->soa32[2], {fd->n[0], fd->n[1], fd->n[2]}};
            rc = b->precision == FQSL_PREC_MIXED ? batch_launch_group<float, FieldSoA<float>, double>(b, G, gi, f, st)
                                                 : batch_launch_group<float>(b, G, gi, f, st);
        }
        if (rc) return rc;
    }
    if (timed) return record_end(st);
    return FQSL_OK;
}

extern "C" int fqsl_scott_batch_times(const fqsl_scott_batch* b, double* ms, void* stream) {
    if (!b || !ms) return set_err(FQSL_ERR_INVALID, "NULL batch or ms");
    std::vector<unsigned long long> stamps(b->nreq + 1);
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(stamps.data(), b->ctl.stamps, sizeof(unsigned long long) * (b->nreq + 1),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
    if (khz <= 0) return set_err(FQSL_ERR_HIP, "device wall clock rate unknown");
    const unsigned long long t0 = stamps[b->nreq];
    for (int k = 0; k < b->nreq; ++k) {
        const unsigned long long t = stamps[k];
        ms[b->order[k]] = t ? (double)(long long)(t - t0) / (double)khz : 0.0;
    }
    return FQSL_OK;
}

extern "C" int fqsl_scott_batch_destroy(fqsl_scott_batch* b) {
    delete b;
    return FQSL_OK;
}

// ------------------------------------------------------------- plane maps
// Seeds of a plane (fqsl_plane), one thread per pixel, row-major.  NumPy's
// linspace(a, b, n)[i] is i * ((b - a) / (n - 1)) + a in double with the last
// point b, and plane_seeds_abc forms A + (B - A) u + (C - A) v in double; both
// without fused multiply-adds, so the rounding to float gives NumPy's seeds.
#pragma clang fp contract(off)
__device__ __forceinline__ double np_linspace(double a, double b, int n, int i) {
    if (n <= 1) return 0.0 * (b - a) + a;  // one point: arange(1) * delta + a
    if (i == n - 1) return b;
    const double step = (b - a) / (double)(n - 1);
    return (double)i * step + a;
}

__global__ void __launch_bounds__(256) k_plane_seeds(fqsl_plane pl, float* __restrict__ x, float* __restrict__ y,
                                                     float* __restrict__ z) {
    const long n = (long)pl.n_u * pl.n_v;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % pl.n_u), j = (int)(t / pl.n_u);
        const double* q = pl.p;
        if (pl.kind == FQSL_PLANE_Z) {
            x[t] = (float)np_linspace(q[0], q[1], pl.n_u, i);
            y[t] = (float)np_linspace(q[2], q[3], pl.n_v, j);
            z[t] = (float)q[4];
        } else {
            const double u = np_linspace(0.0, 1.0, pl.n_u, i), v = np_linspace(0.0, 1.0, pl.n_v, j);
            float* o[3] = {x, y, z};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double a = q[c], ba = q[3 + c] - a, ca = q[6 + c] - a;
                o[c][t] = (float)((a + ba * u) + ca * v);
            }
        }
    }
}
#pragma clang fp contract(on)

extern "C" int fqsl_qmap_plane(const fqsl_field* field, const fqsl_plane* plane, const fqsl_scott_params* params,
                               const fqsl_scott_out* out, void* stream) {
    if (!field || !plane || !params || !out) return set_err(FQSL_ERR_INVALID, "NULL field/plane/params/out");
    if (plane->kind != FQSL_PLANE_Z && plane->kind != FQSL_PLANE_ABC)
        return set_err(FQSL_ERR_INVALID, "plane kind %d", plane->kind);
    if (plane->n_u < 0 || plane->n_v < 0) return set_err(FQSL_ERR_INVALID, "plane of %d x %d pixels", plane->n_u, plane->n_v);
    for (int c = 0; c < 9; ++c)
        if (!(plane->p[c] - plane->p[c] == 0.0)) return set_err(FQSL_ERR_INVALID, "plane coordinate %d is not finite", c);
    const unsigned long long n = (unsigned long long)plane->n_u * (unsigned long long)plane->n_v;
    if (n == 0) return FQSL_OK;
    hipStream_t st = (hipStream_t)stream;
    float* seeds = nullptr;
    HIP_TRY(hipMallocAsync((void**)&seeds, 3 * sizeof(float) * n, st));
    k_plane_seeds<<<grid_for((long)n, 256), 256, 0, st>>>(*plane, seeds, seeds + n, seeds + 2 * n);
    hipError_t e = hipGetLastError();
    int rc = e == hipSuccess ? fqsl_trace_scott(field, params, n, seeds, seeds + n, seeds + 2 * n, out, stream)
                             : set_err(FQSL_ERR_HIP, "plane seeds: %s", hipGetErrorString(e));
    e = hipFreeAsync(seeds, st);
    if (rc) return rc;
    if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "hipFreeAsync: %s", hipGetErrorString(e));
    return FQSL_OK;
}

// The neighbour-difference map of a plane, every step on the stream.
extern "C" int fqsl_qmap_plane_neighbour(const fqsl_field* field, const fqsl_plane* plane,
                                         const fqsl_neighbour_params* p, float* q_all, uint64_t* n_pseudo_dev,
                                         void* stream) {
    if (!field || !plane || !p || !q_all) return set_err(FQSL_ERR_INVALID, "NULL field/plane/params/q_all");
    if (plane->kind != FQSL_PLANE_Z && plane->kind != FQSL_PLANE_ABC)
        return set_err(FQSL_ERR_INVALID, "plane kind %d", plane->kind);
    if (plane->n_u < 3 || plane->n_v < 3)
        return set_err(FQSL_ERR_INVALID, "a neighbour-difference map needs at least 3 x 3 pixels (got %d x %d)",
                       plane->n_u, plane->n_v);
    for (int c = 0; c < 9; ++c)
        if (!(plane->p[c] - plane->p[c] == 0.0)) return set_err(FQSL_ERR_INVALID, "plane coordinate %d is not finite", c);
    const int nu = plane->n_u, nv = plane->n_v;
    const uint64_t n = (uint64_t)nu * (uint64_t)nv;
    const size_t pwb = fqsl_pseudo_workspace_bytes(n);
    // scratch: seeds, main traces, q_inner, pseudo seeds / Bz / pixels / counts, pseudo traces
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off = align256(off + bytes); return o; };
    const size_t o_seed = take(3 * sizeof(float) * n);
    const size_t o_main = take(18 * sizeof(float) * n);
    const size_t o_qin = take(sizeof(float) * (size_t)(nu - 2) * (nv - 2));
    const size_t o_pseed = take(12 * sizeof(float) * n);
    const size_t o_bz0 = take(2 * sizeof(float) * n);
    const size_t o_pix = take(sizeof(uint64_t) * n);
    const size_t o_cnt = take(2 * sizeof(uint64_t));
    const size_t o_pws = take(pwb);
    const size_t o_pout = take(32 * sizeof(float) * n);
    hipStream_t st = (hipStream_t)stream;
    char* ws = nullptr;
    HIP_TRY(hipMallocAsync((void**)&ws, off, st));
    float* sd = (float*)(ws + o_seed);
    float* m = (float*)(ws + o_main);  // 18 arrays of n: start xyz, end xyz, flags, b_this xyz, b_flag, b_start xyz, b_end xyz
    float *sx = m, *sy = m + n, *sz = m + 2 * n, *ex = m + 3 * n, *ey = m + 4 * n, *ez = m + 5 * n;
    int32_t *fs = (int32_t*)(m + 6 * n), *fe = (int32_t*)(m + 7 * n);
    float *btx = m + 8 * n, *bty = m + 9 * n, *btz = m + 10 * n;
    int32_t* bflag = (int32_t*)(m + 11 * n);
    float *bsx = m + 12 * n, *bsy = m + 13 * n, *bsz = m + 14 * n, *bex = m + 15 * n, *bey = m + 16 * n,
          *bez = m + 17 * n;
    float* qin = (float*)(ws + o_qin);
    float* ps = (float*)(ws + o_pseed);
    float* bz0 = (float*)(ws + o_bz0);
    uint64_t* pix = (uint64_t*)(ws + o_pix);
    uint64_t* cnt = (uint64_t*)(ws + o_cnt);
    float* po = (float*)(ws + o_pout);  // 4n-line traces: start xyz, end xyz, flags
    const uint64_t n4 = 4 * n;
    int rc = FQSL_OK;
    k_plane_seeds<<<grid_for((long)n, 256), 256, 0, st>>>(*plane, sd, sd + n, sd + 2 * n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = set_err(FQSL_ERR_HIP, "plane seeds: %s", hipGetErrorString(e));
    if (!rc) {
        fqsl_plain_params mp{p->method, p->precision, p->schedule, 0, p->step, p->tol_coef,
                             {p->normal[0], p->normal[1], p->normal[2]}, 0.f};
        fqsl_plain_out mo{sx, sy, sz, ex, ey, ez, fs, fe, btx, bty, btz, bflag, bsx, bsy, bsz, bex, bey, bez,
                          nullptr, nullptr, nullptr};
        rc = fqsl_trace_plain(field, &mp, n, sd, sd + n, sd + 2 * n, &mo, stream);
    }
    if (!rc)
        rc = fqsl_qcalc_plane(nv, nu, ex, ey, ez, fe, sx, sy, sz, fs, bsz, bez, btx, bty, btz, p->stride, p->normal,
                              qin, stream);
    if (!rc) {
        e = hipMemsetAsync(q_all, 0, sizeof(float) * n, st);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(q_all + nu + 1, sizeof(float) * nu, qin, sizeof(float) * (nu - 2),
                                 sizeof(float) * (nu - 2), nv - 2, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) rc = set_err(FQSL_ERR_HIP, "q_all interior: %s", hipGetErrorString(e));
    }
    const float pstride = (float)p->pseudo_stride;
    if (!rc)
        rc = fqsl_pseudo_seeds_async(n, bflag, sx, sy, sz, fs, bsz, bez, pstride, n, ps, ps + n4, ps + 2 * n4, bz0,
                                     bz0 + n, pix, cnt, ws + o_pws, pwb, stream);
    float *qsx = po, *qsy = po + n4, *qsz = po + 2 * n4, *qex = po + 3 * n4, *qey = po + 4 * n4, *qez = po + 5 * n4;
    int32_t *qfs = (int32_t*)(po + 6 * n4), *qfe = (int32_t*)(po + 7 * n4);
    if (!rc) {
        fqsl_plain_params pp{p->method, p->precision, p->schedule, 0, p->step, p->tol_coef * p->pseudo_tol_scale,
                             {p->pseudo_normal[0], p->pseudo_normal[1], p->pseudo_normal[2]}, 0.f};
        fqsl_plain_out pout{qsx, qsy, qsz, qex, qey, qez, qfs, qfe, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        rc = fqsl_trace_plain_n(field, &pp, n4, cnt + 1, ps, ps + n4, ps + 2 * n4, &pout, stream);
    }
    if (!rc)
        rc = fqsl_qcut_scatter_n(n, cnt, qsx, qsy, qsz, qfs, qex, qey, qez, qfe, bz0, bz0 + n, pstride, pix, q_all,
                                 stream);
    if (!rc && n_pseudo_dev) {
        e = hipMemcpyAsync(n_pseudo_dev, cnt, sizeof(uint64_t), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) rc = set_err(FQSL_ERR_HIP, "pseudo count: %s", hipGetErrorString(e));
    }
    e = hipFreeAsync(ws, st);
    if (rc) return rc;
    if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "hipFreeAsync: %s", hipGetErrorString(e));
    return FQSL_OK;
}

// ---------------------------------------------------------------- serving
// One consumer thread per server drains the queue in batches: it takes the
// waiting requests in submission order up to max_lines lines, traces them as
// one fqsl_scott_batch on its stream, waits for the stream and marks them
// done with the batch's status.  Its workspace grows to the largest batch.
struct ServerReq {
    uint64_t id;
    fqsl_scott_request req;
    hipEvent_t ready;  // recorded on the submitter's stream
};

struct ServerDone {
    int status;
    std::string err;
};

struct fqsl_server {
    int device = 0;
    int precision = FQSL_PREC_FP32C;
    uint64_t max_lines = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerReq> queue;
    std::unordered_map<uint64_t, ServerDone> done;   // finished, not yet waited for
    std::unordered_set<uint64_t> pending;            // queued or running
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    hipStream_t stream = nullptr;
    void* ws = nullptr;
    size_t ws_bytes = 0;
};

static int server_trace(fqsl_server* S, const std::vector<ServerReq>& take) {
    std::vector<fqsl_scott_request> reqs;
    for (const ServerReq& r : take) {
        reqs.push_back(r.req);
        HIP_TRY(hipStreamWaitEvent(S->stream, r.ready, 0));
    }
    const size_t need = fqsl_scott_batch_workspace_bytes(reqs.data(), (int)reqs.size(), S->precision);
    if (need > S->ws_bytes) {
        if (S->ws) HIP_TRY(hipFree(S->ws));
        S->ws = nullptr;
        S->ws_bytes = 0;
        HIP_TRY(hipMalloc(&S->ws, need));
        S->ws_bytes = need;
    }
    fqsl_scott_batch* b = nullptr;
    int rc = fqsl_scott_batch_create(reqs.data(), (int)reqs.size(), S->precision, S->ws, S->ws_bytes, S->stream, &b);
    if (rc) return rc;
    rc = fqsl_scott_batch_run(b, S->stream);
    const hipError_t e = hipStreamSynchronize(S->stream);
    fqsl_scott_batch_destroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "server batch: %s", hipGetErrorString(e));
    return FQSL_OK;
}

static void server_loop(fqsl_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        std::vector<ServerReq> take;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [S] { return S->stop || !S->queue.empty(); });
            if (S->queue.empty()) return;  // stopping and drained
            uint64_t lines = 0;
            while (!S->queue.empty() && (take.empty() || lines + S->queue.front().req.n <= S->max_lines)) {
                lines += S->queue.front().req.n;
                take.push_back(S->queue.front());
                S->queue.pop_front();
            }
        }
        const int rc = server_trace(S, take);
        const std::string err = rc ? std::string(g_err) : std::string();
        if (rc) (void)hipStreamSynchronize(S->stream);  // the waits on the events have ended
        for (const ServerReq& r : take) (void)hipEventDestroy(r.ready);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (const ServerReq& r : take) {
                S->pending.erase(r.id);
                S->done[r.id] = ServerDone{rc, err};
            }
        }
        S->cv_done.notify_all();
    }
}

extern "C" int fqsl_server_create(int device, int precision, uint64_t max_lines, fqsl_server** server) {
    if (!server) return set_err(FQSL_ERR_INVALID, "server is NULL");
    *server = nullptr;
    if (precision != FQSL_PREC_FP32C && precision != FQSL_PREC_FP64 && precision != FQSL_PREC_MIXED)
        return set_err(FQSL_ERR_INVALID, "precision %d", precision);
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return set_err(FQSL_ERR_INVALID, "device %d of %d", device, count);
    fqsl_server* S = new fqsl_server();
    S->device = device;
    S->precision = precision;
    S->max_lines = max_lines ? max_lines : (1ull << 24);
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess) e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        delete S;
        return set_err(FQSL_ERR_HIP, "server stream: %s", hipGetErrorString(e));
    }
    S->worker = std::thread(server_loop, S);
    *server = S;
    return FQSL_OK;
}

extern "C" int fqsl_submit(fqsl_server* S, const fqsl_scott_request* req, void* stream, uint64_t* id) {
    if (!S || !req || !id) return set_err(FQSL_ERR_INVALID, "NULL server/request/id");
    int rc = check_request(*req, S->precision, 0);
    if (rc) return rc;
    hipEvent_t ready = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    const hipError_t e = hipEventRecord(ready, (hipStream_t)stream);
    if (e != hipSuccess) {
        (void)hipEventDestroy(ready);
        return set_err(FQSL_ERR_HIP, "hipEventRecord: %s", hipGetErrorString(e));
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            (void)hipEventDestroy(ready);
            return set_err(FQSL_ERR_INVALID, "server is stopping");
        }
        *id = S->next_id++;
        S->queue.push_back(ServerReq{*id, *req, ready});
        S->pending.insert(*id);
    }
    S->cv_work.notify_one();
    return FQSL_OK;
}

extern "C" int fqsl_poll(fqsl_server* S, uint64_t id, int* done) {
    if (!S || !done) return set_err(FQSL_ERR_INVALID, "NULL server/done");
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->done.count(id)) { *done = 1; return FQSL_OK; }
    if (S->pending.count(id)) { *done = 0; return FQSL_OK; }
    return set_err(FQSL_ERR_INVALID, "unknown request id %llu", (unsigned long long)id);
}

extern "C" int fqsl_wait(fqsl_server* S, uint64_t id) {
    if (!S) return set_err(FQSL_ERR_INVALID, "server is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    if (!S->done.count(id) && !S->pending.count(id))
        return set_err(FQSL_ERR_INVALID, "unknown request id %llu", (unsigned long long)id);
    // a request leaves `pending` when it enters `done`; if it is in neither
    // on wake-up, another thread's wait took it first
    S->cv_done.wait(lk, [S, id] { return S->done.count(id) > 0 || !S->pending.count(id); });
    if (!S->done.count(id))
        return set_err(FQSL_ERR_INVALID, "request id %llu was waited for by another call",
                       (unsigned long long)id);
    const ServerDone d = S->done[id];
    S->done.erase(id);
    if (d.status) return set_err(d.status, "%s", d.err.c_str());
    return FQSL_OK;
}

extern "C" int fqsl_server_destroy(fqsl_server* S) {
    if (!S) return FQSL_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(S->device);
    if (S->ws) (void)hipFree(S->ws);
    if (S->stream) (void)hipStreamDestroy(S->stream);
    (void)hipSetDevice(prev);
    delete S;
    return FQSL_OK;
}

// Largest LDS table the stretched kernels copy per workgroup: two 256-thread
// workgroups (two waves per SIMD) still fit a CU's 160 KiB.
constexpr int kStretchLdsMax = 64 * 1024;
static bool stretch_lds_fits(int bytes) {
    return g_tune[FQSL_TUNE_STRETCH_LDS] && bytes > 0 && bytes <= kStretchLdsMax;
}

// FP32 stretched launches over geometry G (StretchRef<float> or StretchLds).
template <class G>
static int trace_stretched_f32(const fqsl_field* field, const PlainParams& P, const Seeds& s, unsigned long long n,
                               const PlainOut& o, int schedule, hipStream_t st, const DevScalars& ds, const G& g) {
    const int* nn = field->n;
    const FieldSoA<float> fj{field->cur32[0], field->cur32[1], field->cur32[2], {nn[0], nn[1], nn[2]}};
    if (field->cell32) {
        FieldCell32 f{field->cell32, {nn[0], nn[1], nn[2]}};
        if (o.tw) {
            if (!field->curcell) return set_err(FQSL_ERR_INVALID, "current density has no cell layout");
            const FieldCell32 fjc{field->curcell, {nn[0], nn[1], nn[2]}};
            return launch_plain<float>(f, P, s, n, o, ds, nullptr, schedule, st, n, fjc, g);
        }
        return launch_plain<float>(f, P, s, n, o, ds, nullptr, schedule, st, n, NoCurrent{}, g);
    }
    if (field->soa32[0]) {
        FieldSoA<float> f{field->soa32[0], field->soa32[1], field->soa32[2], {nn[0], nn[1], nn[2]}};
        if (o.tw) return launch_plain<float>(f, P, s, n, o, ds, nullptr, schedule, st, n, fj, g);
        return launch_plain<float>(f, P, s, n, o, ds, nullptr, schedule, st, n, NoCurrent{}, g);
    }
    return set_err(FQSL_ERR_INVALID, "field has no FP32 layout");
}

// The stretched tracer on a resident field (RKF45 only, as the reference).
static int trace_plain_stretched(const fqsl_field* field, const PlainParams& P, const Seeds& s, unsigned long long n,
                                 const PlainOut& o, int precision, int schedule, hipStream_t st,
                                 const DevScalars& ds) {
    if (P.method != FQSL_METHOD_RKF45)
        return set_err(FQSL_ERR_UNSUPPORTED, "a stretched grid is traced with RKF45 only (TraceBlineAdaptiveStretch.cu)");
    const int* nn = field->n;
    if (precision == FQSL_PREC_FP64) {
        if (!field->soa64[0]) return set_err(FQSL_ERR_INVALID, "field has no FP64 layout");
        const StretchRef<double> g{field->sg64};
        FieldSoA<double> f{field->soa64[0], field->soa64[1], field->soa64[2], {nn[0], nn[1], nn[2]}};
        if (o.tw) {
            FieldSoA<double> fj{field->cur64[0], field->cur64[1], field->cur64[2], {nn[0], nn[1], nn[2]}};
            return launch_plain<double>(f, P, s, n, o, ds, nullptr, schedule, st, n, fj, g);
        }
        return launch_plain<double>(f, P, s, n, o, ds, nullptr, schedule, st, n, NoCurrent{}, g);
    }
    if (stretch_lds_fits(field->sg_lds_bytes))
        return trace_stretched_f32(field, P, s, n, o, schedule, st, ds, StretchLds{field->sg32, field->sg_lds_bytes});
    return trace_stretched_f32(field, P, s, n, o, schedule, st, ds, StretchRef<float>{field->sg32});
}

static int trace_plain_impl(const fqsl_field* field, const fqsl_plain_params* p, uint64_t n,
                            const unsigned long long* n_dev, const float* x, const float* y, const float* z,
                            const fqsl_plain_out* out, void* stream) {
    if (!field || !p || !out) return set_err(FQSL_ERR_INVALID, "NULL field/params/out");
    int rc = check_common(p->method, p->precision, p->schedule);
    if (rc) return rc;
    if (n == 0) return FQSL_OK;
    if (!x || !y || !z) return set_err(FQSL_ERR_INVALID, "NULL seed array");
    if (!(p->step > 0) || !(p->step <= 100.0)) return set_err(FQSL_ERR_INVALID, "step %g outside (0, 100]", p->step);
    PlainParams P;
    P.method = p->method; P.s0 = p->step; P.tol_coef = p->tol_coef;
    P.ncross[0] = p->ncross[0]; P.ncross[1] = p->ncross[1]; P.ncross[2] = p->ncross[2];
    Seeds s{x, y, z};
    PlainOut o{out->start_x, out->start_y, out->start_z, out->end_x, out->end_y, out->end_z,
               out->flag_start, out->flag_end, out->b_this_x, out->b_this_y, out->b_this_z, out->b_flag,
               out->b_start_x, out->b_start_y, out->b_start_z, out->b_end_x, out->b_end_y, out->b_end_z,
               out->line_len, out->evals, out->twist};
    if ((o.sx && (!o.sy || !o.sz)) || (o.ex && (!o.ey || !o.ez)) || (o.btx && (!o.bty || !o.btz)) ||
        (o.bsx && (!o.bsy || !o.bsz)) || (o.bex && (!o.bey || !o.bez)))
        return set_err(FQSL_ERR_INVALID, "x/y/z output arrays must be all set or all NULL");
    DevScalars ds{};
    ds.n_dev = n_dev;
    hipStream_t st = (hipStream_t)stream;
    const int* nn = field->n;
    if (o.tw && !field->cur32[0]) return set_err(FQSL_ERR_INVALID, "twist requested but the field has no current density");
    if (field->stretched) return trace_plain_stretched(field, P, s, n, o, p->precision, p->schedule, st, ds);
    if (p->precision == FQSL_PREC_FP64) {
        if (!field->soa64[0]) return set_err(FQSL_ERR_INVALID, "field has no FP64 layout");
        FieldSoA<double> f{field->soa64[0], field->soa64[1], field->soa64[2], {nn[0], nn[1], nn[2]}};
        if (o.tw) {
            FieldSoA<double> fj{field->cur64[0], field->cur64[1], field->cur64[2], {nn[0], nn[1], nn[2]}};
            return launch_plain<double>(f, P, s, n, o, ds, nullptr, p->schedule, st, n, fj);
        }
        return launch_plain<double>(f, P, s, n, o, ds, nullptr, p->schedule, st, n);
    }
    const FieldSoA<float> fj{field->cur32[0], field->cur32[1], field->cur32[2], {nn[0], nn[1], nn[2]}};
    if (field->cell32) {
        FieldCell32 f{field->cell32, {nn[0], nn[1], nn[2]}};
        if (o.tw) {
            if (!field->curcell) return set_err(FQSL_ERR_INVALID, "current density has no cell layout");
            const FieldCell32 fjc{field->curcell, {nn[0], nn[1], nn[2]}};
            return launch_plain<float>(f, P, s, n, o, ds, nullptr, p->schedule, st, n, fjc);
        }
        return launch_plain<float>(f, P, s, n, o, ds, nullptr, p->schedule, st, n);
    }
    if (field->soa32[0]) {
        FieldSoA<float> f{field->soa32[0], field->soa32[1], field->soa32[2], {nn[0], nn[1], nn[2]}};
        if (o.tw) return launch_plain<float>(f, P, s, n, o, ds, nullptr, p->schedule, st, n, fj);
        return launch_plain<float>(f, P, s, n, o, ds, nullptr, p->schedule, st, n);
    }
    return set_err(FQSL_ERR_INVALID, "field has no FP32 layout");
}

extern "C" int fqsl_trace_plain(const fqsl_field* field, const fqsl_plain_params* p, uint64_t n, const float* x,
                                const float* y, const float* z, const fqsl_plain_out* out, void* stream) {
    return trace_plain_impl(field, p, n, nullptr, x, y, z, out, stream);
}

extern "C" int fqsl_trace_plain_n(const fqsl_field* field, const fqsl_plain_params* p, uint64_t n_cap,
                                  const uint64_t* n_dev, const float* x, const float* y, const float* z,
                                  const fqsl_plain_out* out, void* stream) {
    if (!n_dev) return set_err(FQSL_ERR_INVALID, "n_dev is NULL");
    return trace_plain_impl(field, p, n_cap, (const unsigned long long*)n_dev, x, y, z, out, stream);
}

// ------------------------------------------- reference-shaped entry points
// The seed count N lives on the device; the per-line launch needs a grid on
// the host, so it runs a chip-sized grid (per_line_grid) whose grid-stride
// loop covers any N, as the reference's fixed grids do.  kPerLineMaxBlocks *
// 256 is only the cap passed as the line-count hint.
static constexpr int kPerLineMaxBlocks = 65536;

extern "C" int fqsl_TraceAllBlineScott(float* Bx, float* By, float* Bz, int* BshapeN, float* curB_x, float* curB_y,
                                       float* curB_z, double* twist, bool* curB_flag, float* inp_x, float* inp_y,
                                       float* inp_z, float* inp_cross_dir, float* start_x, float* start_y,
                                       float* start_z, int* flag_start, float* end_x, float* end_y, float* end_z,
                                       int* flag_end, float* B_this_x, float* B_this_y, float* B_this_z, float* q_0,
                                       float* q_perp, float* s_len, unsigned long long* N, double* LineLen,
                                       float* tol_coef, int precision, int schedule, void* stream) {
    (void)inp_cross_dir;
    if (!Bx || !By || !Bz || !BshapeN || !inp_x || !inp_y || !inp_z || !s_len || !N || !tol_coef)
        return set_err(FQSL_ERR_INVALID, "NULL required argument");
    bool twist_on = false;
    if (curB_flag) {  // the switch is a device scalar: read it (synchronises the stream)
        HIP_TRY(hipMemcpyAsync(&twist_on, curB_flag, sizeof(bool), hipMemcpyDeviceToHost, (hipStream_t)stream));
        HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    }
    if (twist_on && (!curB_x || !curB_y || !curB_z || !twist))
        return set_err(FQSL_ERR_INVALID, "curB_flag is set: curB_x/y/z and twist must be device arrays");
    int rc = check_common(FQSL_METHOD_RKF45, precision, schedule, true);
    if (rc) return rc;
    if (precision == FQSL_PREC_FP64)
        return set_err(FQSL_ERR_UNSUPPORTED, "the in-place float field supports FQSL_PREC_FP32C and FQSL_PREC_MIXED");
    ScottParams P{FQSL_METHOD_RKF45, 0.0, 0.0, 0.0};
    Seeds s{inp_x, inp_y, inp_z};
    ScottOut o{start_x, start_y, start_z, end_x, end_y, end_z, flag_start, flag_end, B_this_x, B_this_y, B_this_z,
               q_0, q_perp, LineLen, nullptr, twist_on ? twist : nullptr};
    DevScalars ds{BshapeN, s_len, N, tol_coef, curB_flag, twist};
    FieldSoA<float> f{Bx, By, Bz, {2, 2, 2}};
    unsigned long long n_hint = (unsigned long long)kPerLineMaxBlocks * 256;
    if (schedule == FQSL_SCHED_PERSISTENT) {
        // the persistent path sizes its workspace from N: read it (synchronises the stream)
        HIP_TRY(hipMemcpyAsync(&n_hint, N, sizeof(n_hint), hipMemcpyDeviceToHost, (hipStream_t)stream));
        HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    }
    const hipStream_t st = (hipStream_t)stream;
    if (twist_on) {
        const FieldSoA<float> fj{curB_x, curB_y, curB_z, {2, 2, 2}};  // shape taken from B's on the device
        if (precision == FQSL_PREC_MIXED)
            return launch_scott<float, FieldSoA<float>, FieldSoA<float>, double>(f, P, s, 0ull, o, ds, schedule, st,
                                                                                 n_hint, nullptr, fj);
        return launch_scott<float>(f, P, s, 0ull, o, ds, schedule, st, n_hint, nullptr, fj);
    }
    if (precision == FQSL_PREC_MIXED)
        return launch_scott<float, FieldSoA<float>, NoCurrent, double>(f, P, s, 0ull, o, ds, schedule, st, n_hint,
                                                                       nullptr);
    return launch_scott<float>(f, P, s, 0ull, o, ds, schedule, st, n_hint, nullptr);
}

extern "C" int fqsl_TraceAllBline(int method, float* Bx, float* By, float* Bz, int* BshapeN, float* curB_x,
                                  float* curB_y, float* curB_z, double* twist, bool* curB_flag, float* inp_x,
                                  float* inp_y, float* inp_z, float* inp_cross_dir, float* start_x, float* start_y,
                                  float* start_z, int* flag_start, float* end_x, float* end_y, float* end_z,
                                  int* flag_end, float* B_this_x, float* B_this_y, float* B_this_z, int* B_flag,
                                  float* B_start_x, float* B_start_y, float* B_start_z, float* B_end_x,
                                  float* B_end_y, float* B_end_z, float* s_len, unsigned long long* N,
                                  double* LineLen, float* tol_coef, int precision, int schedule, void* stream) {
    if (!Bx || !By || !Bz || !BshapeN || !inp_x || !inp_y || !inp_z || !inp_cross_dir || !s_len || !N)
        return set_err(FQSL_ERR_INVALID, "NULL required argument");
    bool twist_on = false;
    if (curB_flag) {  // the switch is a device scalar: read it (synchronises the stream)
        HIP_TRY(hipMemcpyAsync(&twist_on, curB_flag, sizeof(bool), hipMemcpyDeviceToHost, (hipStream_t)stream));
        HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    }
    if (twist_on && (!curB_x || !curB_y || !curB_z || !twist))
        return set_err(FQSL_ERR_INVALID, "curB_flag is set: curB_x/y/z and twist must be device arrays");
    if (method == FQSL_METHOD_RKF45 && !tol_coef) return set_err(FQSL_ERR_INVALID, "RKF45 needs tol_coef");
    int rc = check_common(method, precision, schedule);
    if (rc) return rc;
    if (precision != FQSL_PREC_FP32C)
        return set_err(FQSL_ERR_UNSUPPORTED, "the in-place float field supports FQSL_PREC_FP32C only");
    PlainParams P;
    P.method = method; P.s0 = 0.0; P.tol_coef = 1.0;
    P.ncross[0] = P.ncross[1] = P.ncross[2] = 0.f;
    Seeds s{inp_x, inp_y, inp_z};
    PlainOut o{start_x, start_y, start_z, end_x, end_y, end_z, flag_start, flag_end, B_this_x, B_this_y, B_this_z,
               B_flag, B_start_x, B_start_y, B_start_z, B_end_x, B_end_y, B_end_z, LineLen, nullptr,
               twist_on ? twist : nullptr};
    DevScalars ds{BshapeN, s_len, N, method == FQSL_METHOD_RKF45 ? tol_coef : nullptr, curB_flag, twist};
    FieldSoA<float> f{Bx, By, Bz, {2, 2, 2}};
    const unsigned long long n_hint = (unsigned long long)kPerLineMaxBlocks * 256;
    if (twist_on) {
        const FieldSoA<float> fj{curB_x, curB_y, curB_z, {2, 2, 2}};  // shape taken from B's on the device
        return launch_plain<float>(f, P, s, 0ull, o, ds, inp_cross_dir, schedule, (hipStream_t)stream, n_hint, fj);
    }
    return launch_plain<float>(f, P, s, 0ull, o, ds, inp_cross_dir, schedule, (hipStream_t)stream, n_hint);
}

// Lookup tables of the stretched drop-in, per device, kept while the axes
// and the rule of successive calls are unchanged.
struct DropinAxes {
    std::vector<double> a[3];
    int rule = -1;
    void* mem = nullptr;
    const StretchGrid<float>* g = nullptr;
    int lds_bytes = 0;
};
static std::mutex g_dropin_mu;
static DropinAxes g_dropin_axes[64];

extern "C" int fqsl_TraceAllBlineStretch(float* Bx, float* By, float* Bz, int* BshapeN, float* x_arr, float* y_arr,
                                         float* z_arr, bool* flag_uni, float* curB_x, float* curB_y, float* curB_z,
                                         double* twist, bool* curB_flag, float* inp_x, float* inp_y, float* inp_z,
                                         float* inp_cross_dir, float* start_x, float* start_y, float* start_z,
                                         int* flag_start, float* end_x, float* end_y, float* end_z, int* flag_end,
                                         float* B_this_x, float* B_this_y, float* B_this_z, int* B_flag,
                                         float* B_start_x, float* B_start_y, float* B_start_z, float* B_end_x,
                                         float* B_end_y, float* B_end_z, float* s_len, unsigned long long* N,
                                         double* LineLen, float* tol_coef, int precision, int schedule,
                                         void* stream) {
    if (!Bx || !By || !Bz || !BshapeN || !x_arr || !y_arr || !z_arr || !inp_x || !inp_y || !inp_z ||
        !inp_cross_dir || !s_len || !N || !tol_coef)
        return set_err(FQSL_ERR_INVALID, "NULL required argument");
    int rc = check_common(FQSL_METHOD_RKF45, precision, schedule);
    if (rc) return rc;
    if (precision != FQSL_PREC_FP32C)
        return set_err(FQSL_ERR_UNSUPPORTED, "the in-place float field supports FQSL_PREC_FP32C only");
    hipStream_t st = (hipStream_t)stream;
    // the shape and the two switches are device scalars: read them (synchronises the stream)
    int shape[3] = {0, 0, 0};
    bool uni = false, twist_on = false;
    HIP_TRY(hipMemcpyAsync(shape, BshapeN, sizeof(shape), hipMemcpyDeviceToHost, st));
    if (flag_uni) HIP_TRY(hipMemcpyAsync(&uni, flag_uni, sizeof(bool), hipMemcpyDeviceToHost, st));
    if (curB_flag) HIP_TRY(hipMemcpyAsync(&twist_on, curB_flag, sizeof(bool), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (shape[0] < 2 || shape[1] < 2 || shape[2] < 2)
        return set_err(FQSL_ERR_INVALID, "BshapeN (%d,%d,%d): each axis needs >= 2 nodes", shape[0], shape[1], shape[2]);
    if (twist_on && (!curB_x || !curB_y || !curB_z || !twist))
        return set_err(FQSL_ERR_INVALID, "curB_flag is set: curB_x/y/z and twist must be device arrays");
    std::vector<double> a[3];
    const float* src[3] = {x_arr, y_arr, z_arr};
    for (int q = 0; q < 3; ++q) {
        rc = axis_to_host(src[q], shape[q], FQSL_DTYPE_F32, 1, a[q]);
        if (rc) return rc;
    }
    const int rule = uni ? kLookupUniform : kLookupBisect;  // GetIndexUniStretch / GetIndexBiSearch
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_err(FQSL_ERR_UNSUPPORTED, "device index %d", dev);
    // The lock is held until the launch is enqueued: a later call from another
    // thread that replaces the tables first synchronises the device, which
    // then covers this launch, before it frees them.
    std::lock_guard<std::mutex> lock(g_dropin_mu);
    StretchRef<float> g{nullptr};
    int lds_bytes = 0;
    {
        DropinAxes& c = g_dropin_axes[dev];
        const bool same = c.mem && c.rule == rule && c.a[0] == a[0] && c.a[1] == a[1] && c.a[2] == a[2];
        if (!same) {
            void* mem = nullptr;
            size_t bytes = 0;
            const StretchGrid<double>* g64 = nullptr;
            const StretchGrid<float>* g32 = nullptr;
            int lb = 0;
            rc = build_stretch(a, shape, rule, &mem, &bytes, &g32, &g64, &lb);
            if (rc) return rc;
            if (c.mem) {  // launches enqueued earlier, on any stream, may still read the old tables
                HIP_TRY(hipDeviceSynchronize());
                (void)hipFree(c.mem);
            }
            c.mem = mem;
            c.g = g32;
            c.lds_bytes = lb;
            c.rule = rule;
            for (int q = 0; q < 3; ++q) c.a[q] = a[q];
        }
        g.g = c.g;
        lds_bytes = c.lds_bytes;
    }
    PlainParams P;
    P.method = FQSL_METHOD_RKF45; P.s0 = 0.0; P.tol_coef = 1.0;
    P.ncross[0] = P.ncross[1] = P.ncross[2] = 0.f;
    Seeds s{inp_x, inp_y, inp_z};
    PlainOut o{start_x, start_y, start_z, end_x, end_y, end_z, flag_start, flag_end, B_this_x, B_this_y, B_this_z,
               B_flag, B_start_x, B_start_y, B_start_z, B_end_x, B_end_y, B_end_z, LineLen, nullptr,
               twist_on ? twist : nullptr};
    DevScalars ds{BshapeN, s_len, N, tol_coef, curB_flag, twist};
    FieldSoA<float> f{Bx, By, Bz, {shape[0], shape[1], shape[2]}};
    const unsigned long long n_hint = (unsigned long long)kPerLineMaxBlocks * 256;
    const FieldSoA<float> fj{curB_x, curB_y, curB_z, {shape[0], shape[1], shape[2]}};
    if (stretch_lds_fits(lds_bytes)) {
        const StretchLds gl{g.g, lds_bytes};
        if (twist_on) return launch_plain<float>(f, P, s, 0ull, o, ds, inp_cross_dir, schedule, st, n_hint, fj, gl);
        return launch_plain<float>(f, P, s, 0ull, o, ds, inp_cross_dir, schedule, st, n_hint, NoCurrent{}, gl);
    }
    if (twist_on) return launch_plain<float>(f, P, s, 0ull, o, ds, inp_cross_dir, schedule, st, n_hint, fj, g);
    return launch_plain<float>(f, P, s, 0ull, o, ds, inp_cross_dir, schedule, st, n_hint, NoCurrent{}, g);
}

// ---------------------------------------------- neighbour-difference assembly
// Exit-face -> 2-D footpoint coordinates (FastQSL.py:76-84).
__device__ __forceinline__ void footpoint2d(int flag, float x, float y, float z, float* X, float* Y) {
    const int dim = (flag - 1) >= 0 ? (flag - 1) / 2 : -1;  // Python floor division
    if (dim == 2) { *X = x; *Y = y; }
    else if (dim == 1) { *X = z; *Y = x; }
    else if (dim == 0) { *X = y; *Y = z; }
    else { *X = 0.f; *Y = 0.f; }
}

#pragma clang fp contract(off)
__global__ void k_qcalc_plane(int ny, int nx, const float* __restrict__ xe, const float* __restrict__ ye,
                              const float* __restrict__ ze, const int32_t* __restrict__ fe,
                              const float* __restrict__ xs, const float* __restrict__ ys,
                              const float* __restrict__ zs, const int32_t* __restrict__ fs,
                              const float* __restrict__ bzin, const float* __restrict__ bzout,
                              const float* __restrict__ b0x, const float* __restrict__ b0y,
                              const float* __restrict__ b0z, float cd0, float cd1, float cd2, float scale,
                              float* __restrict__ q) {
    const long inner = (long)(ny - 2) * (nx - 2);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < inner; t += (long)gridDim.x * blockDim.x) {
        const int r = (int)(t / (nx - 2)) + 1, c = (int)(t % (nx - 2)) + 1;
        const long p = (long)r * nx + c;
        const long pn[4] = {p + nx, p - nx, p + 1, p - 1};  // rows +-1 (axis 0), columns +-1
        float X1[4], Y1[4], X2[4], Y2[4];
        bool diff = false;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const long o = pn[m];
            footpoint2d(fs[o], xs[o], ys[o], zs[o], &X1[m], &Y1[m]);
            footpoint2d(fe[o], xe[o], ye[o], ze[o], &X2[m], &Y2[m]);
            diff |= (fs[o] != fs[p]) | (fe[o] != fe[p]);
        }
        const float dx2xc = X2[0] - X2[1], dx2yc = X2[2] - X2[3];
        const float dy2xc = Y2[0] - Y2[1], dy2yc = Y2[2] - Y2[3];
        const float dx1xc = X1[0] - X1[1], dx1yc = X1[2] - X1[3];
        const float dy1xc = Y1[0] - Y1[1], dy1yc = Y1[2] - Y1[3];
        const float a = dx2xc * dy1yc - dx2yc * dy1xc;
        const float b = dx2yc * dx1xc - dx2xc * dx1yc;
        const float cc = dy2xc * dy1yc - dy2yc * dy1xc;
        const float d = dy2yc * dx1xc - dy2xc * dx1yc;
        const float bnorm = b0x[p] * cd0 + b0y[p] * cd1 + b0z[p] * cd2;
        const float bn = fabsf(bnorm);
        const float bnr = fabsf(bzin[p]) / (bn * bn) * fabsf(bzout[p]) * scale;
        float Q = (a * a + b * b + cc * cc + d * d) * bnr;
        if (Q < 1.0f) Q = 1.0f;
        if (diff) Q = NAN;
        q[t] = Q;
    }
}

__global__ void k_qcut_scatter(unsigned long long m, const float* __restrict__ sx, const float* __restrict__ sy,
                               const float* __restrict__ sz, const int32_t* __restrict__ fs,
                               const float* __restrict__ ex, const float* __restrict__ ey,
                               const float* __restrict__ ez, const int32_t* __restrict__ fe,
                               const float* __restrict__ bzs, const float* __restrict__ bze, float scale,
                               const uint64_t* __restrict__ pixel, float* __restrict__ qall,
                               const unsigned long long* __restrict__ m_dev) {
    if (m_dev && m_dev[0] < m) m = m_dev[0];
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < m;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        float X1[4], Y1[4], X2[4], Y2[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const unsigned long long o = 4 * i + k;
            footpoint2d(fs[o], sx[o], sy[o], sz[o], &X1[k], &Y1[k]);
            footpoint2d(fe[o], ex[o], ey[o], ez[o], &X2[k], &Y2[k]);
        }
        const float dx2xc = X2[0] - X2[2], dx2yc = X2[1] - X2[3];
        const float dy2xc = Y2[0] - Y2[2], dy2yc = Y2[1] - Y2[3];
        const float dx1xc = X1[0] - X1[2], dx1yc = X1[1] - X1[3];
        const float dy1xc = Y1[0] - Y1[2], dy1yc = Y1[1] - Y1[3];
        const float a = dx2xc * dy1yc - dx2yc * dy1xc;
        const float b = dx2yc * dx1xc - dx2xc * dx1yc;
        const float c = dy2xc * dy1yc - dy2yc * dy1xc;
        const float d = dy2yc * dx1xc - dy2xc * dx1yc;
        const float bnr = fabsf(bze[i]) / fabsf(bzs[i]) * scale;
        float Q = (a * a + b * b + c * c + d * d) * bnr;
        if (Q < 1.0f) Q = 1.0f;
        qall[pixel[i]] = Q;
    }
}
#pragma clang fp contract(on)

// CookPseudoLine compaction, order-preserving: per-block counts, one-block
// exclusive scan, per-block write.
static constexpr int kCompactBlock = 256;

__global__ void k_pseudo_count(unsigned long long n, const int32_t* __restrict__ bflag, unsigned* __restrict__ counts) {
    __shared__ unsigned s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const unsigned long long i = blockIdx.x * (unsigned long long)kCompactBlock + threadIdx.x;
    const bool sel = i < n && bflag[i] == 1;
    const unsigned long long m = __ballot(sel);
    if ((threadIdx.x & 63) == 0) atomicAdd(&s, (unsigned)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s;
}

// total[0] = selected pixels, total[1] = their pseudo lines (four each).
__global__ void k_pseudo_scan(unsigned nblocks, unsigned* __restrict__ counts, unsigned long long* __restrict__ total) {
    // single block of 1024 threads, each scans a contiguous chunk
    __shared__ unsigned long long part[1024];
    const unsigned per = (nblocks + 1023) / 1024;
    const unsigned b0 = threadIdx.x * per;
    unsigned long long sum = 0;
    for (unsigned b = b0; b < b0 + per && b < nblocks; ++b) sum += counts[b];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int t = 0; t < 1024; ++t) { const unsigned long long v = part[t]; part[t] = run; run += v; }
        total[0] = run;
        total[1] = 4 * run;
    }
    __syncthreads();
    unsigned long long run = part[threadIdx.x];
    for (unsigned b = b0; b < b0 + per && b < nblocks; ++b) {
        const unsigned v = counts[b];
        counts[b] = (unsigned)run;  // exclusive offset (fits: offsets < 2^32 checked on the host)
        run += v;
    }
}

__global__ void k_pseudo_write(unsigned long long n, const int32_t* __restrict__ bflag, const float* __restrict__ xs,
                               const float* __restrict__ ys, const float* __restrict__ zs,
                               const int32_t* __restrict__ fls, const float* __restrict__ bzs,
                               const float* __restrict__ bze, float st, const unsigned* __restrict__ offs,
                               unsigned long long capacity, float* __restrict__ ox, float* __restrict__ oy,
                               float* __restrict__ oz, float* __restrict__ obzs, float* __restrict__ obze,
                               uint64_t* __restrict__ pix) {
    __shared__ unsigned wave_base[kCompactBlock / 64];
    const unsigned long long i = blockIdx.x * (unsigned long long)kCompactBlock + threadIdx.x;
    const bool sel = i < n && bflag[i] == 1;
    const unsigned long long m = __ballot(sel);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wave_base[w] = (unsigned)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned run = 0;
        for (int k = 0; k < kCompactBlock / 64; ++k) { const unsigned v = wave_base[k]; wave_base[k] = run; run += v; }
    }
    __syncthreads();
    if (!sel) return;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const unsigned long long j = (unsigned long long)offs[blockIdx.x] + wave_base[w] + __popcll(m & below);
    if (j >= capacity) return;
    const float x = xs[i], y = ys[i], z = zs[i];
    float cx[4] = {x, x, x, x}, cy[4] = {y, y, y, y}, cz[4] = {z, z, z, z};
    const int f = fls[i];
    const int dim = (f - 1) >= 0 ? (f - 1) / 2 : -1;
    if (dim == 2) { cx[0] = x + st; cx[2] = x - st; cy[1] = y + st; cy[3] = y - st; }   // FastQSL.py:38-41
    else if (dim == 1) { cx[0] = x + st; cx[2] = x - st; cz[1] = z + st; cz[3] = z - st; }  // :42-45
    else if (dim == 0) { cy[0] = y + st; cy[2] = y - st; cz[1] = z + st; cz[3] = z - st; }  // :46-49
#pragma unroll
    for (int k = 0; k < 4; ++k) { ox[4 * j + k] = cx[k]; oy[4 * j + k] = cy[k]; oz[4 * j + k] = cz[k]; }
    obzs[j] = bzs[i];
    obze[j] = bze[i];
    pix[j] = i;
}

extern "C" int fqsl_qcalc_plane(int ny, int nx, const float* x_end, const float* y_end, const float* z_end,
                                const int32_t* flag_end, const float* x_start, const float* y_start,
                                const float* z_start, const int32_t* flag_start, const float* b_in_z,
                                const float* b_out_z, const float* b0_x, const float* b0_y, const float* b0_z,
                                double stride_step, const float* cross_dir_host, float* q, void* stream) {
    if (ny < 3 || nx < 3) return set_err(FQSL_ERR_INVALID, "QCalcPlane needs at least 3x3 pixels (%d x %d)", ny, nx);
    if (!x_end || !y_end || !z_end || !flag_end || !x_start || !y_start || !z_start || !flag_start || !b_in_z ||
        !b_out_z || !b0_x || !b0_y || !b0_z || !q)
        return set_err(FQSL_ERR_INVALID, "NULL array argument");
    if (!(stride_step > 0)) return set_err(FQSL_ERR_INVALID, "stride_step must be > 0");
    const float cd[3] = {cross_dir_host ? cross_dir_host[0] : 0.f, cross_dir_host ? cross_dir_host[1] : 0.f,
                         cross_dir_host ? cross_dir_host[2] : 1.f};
    // ((1./(2.*stride_step))**4.) in double, then float32 (FastQSL.py:113)
    const float scale = (float)pow(1.0 / (2.0 * stride_step), 4.0);
    const long inner = (long)(ny - 2) * (nx - 2);
    k_qcalc_plane<<<grid_for(inner, 256), 256, 0, (hipStream_t)stream>>>(
        ny, nx, x_end, y_end, z_end, flag_end, x_start, y_start, z_start, flag_start, b_in_z, b_out_z, b0_x, b0_y,
        b0_z, cd[0], cd[1], cd[2], scale, q);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}

// Scratch of the compaction: per-block counts (offsets after the scan).
static size_t pseudo_ws_bytes(uint64_t n) {
    const size_t nblocks = (size_t)((n + kCompactBlock - 1) / kCompactBlock);
    return ((nblocks * sizeof(unsigned) + 255) & ~(size_t)255) + 256;
}

extern "C" size_t fqsl_pseudo_workspace_bytes(uint64_t n) { return pseudo_ws_bytes(n); }

static int pseudo_check(uint64_t n, const int32_t* b_flag, const float* x_start, const float* y_start,
                        const float* z_start, const int32_t* flag_start, const float* bz_start, const float* bz_end,
                        uint64_t capacity, float* seed_x, float* seed_y, float* seed_z, float* bz0_start,
                        float* bz0_end, uint64_t* pixel) {
    if (!b_flag || !x_start || !y_start || !z_start || !flag_start || !bz_start || !bz_end)
        return set_err(FQSL_ERR_INVALID, "NULL input array");
    if (capacity && (!seed_x || !seed_y || !seed_z || !bz0_start || !bz0_end || !pixel))
        return set_err(FQSL_ERR_INVALID, "NULL output array");
    if (n > 0xffffffffull) return set_err(FQSL_ERR_INVALID, "pseudo-line compaction supports < 2^32 pixels");
    return FQSL_OK;
}

// count, scan, write on the stream; counts_dev[0] = selected pixels,
// counts_dev[1] = 4 x that (the pseudo lines)
static int pseudo_launch(uint64_t n, const int32_t* b_flag, const float* x_start, const float* y_start,
                         const float* z_start, const int32_t* flag_start, const float* bz_start, const float* bz_end,
                         float stride, uint64_t capacity, float* seed_x, float* seed_y, float* seed_z,
                         float* bz0_start, float* bz0_end, uint64_t* pixel, unsigned long long* counts_dev,
                         void* workspace, hipStream_t st) {
    const unsigned nblocks = (unsigned)((n + kCompactBlock - 1) / kCompactBlock);
    unsigned* counts = (unsigned*)workspace;
    k_pseudo_count<<<nblocks, kCompactBlock, 0, st>>>(n, b_flag, counts);
    k_pseudo_scan<<<1, 1024, 0, st>>>(nblocks, counts, counts_dev);
    k_pseudo_write<<<nblocks, kCompactBlock, 0, st>>>(n, b_flag, x_start, y_start, z_start, flag_start, bz_start,
                                                      bz_end, stride, counts, capacity, seed_x, seed_y, seed_z,
                                                      bz0_start, bz0_end, pixel);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "pseudo-line compaction: %s", hipGetErrorString(e));
    return FQSL_OK;
}

extern "C" int fqsl_pseudo_seeds_async(uint64_t n, const int32_t* b_flag, const float* x_start, const float* y_start,
                                       const float* z_start, const int32_t* flag_start, const float* bz_start,
                                       const float* bz_end, float stride, uint64_t capacity, float* seed_x,
                                       float* seed_y, float* seed_z, float* bz0_start, float* bz0_end,
                                       uint64_t* pixel, uint64_t* counts_dev, void* workspace,
                                       size_t workspace_bytes, void* stream) {
    if (!counts_dev) return set_err(FQSL_ERR_INVALID, "counts_dev is NULL");
    hipStream_t st = (hipStream_t)stream;
    if (n == 0) {  // nothing selected: clear the two counts on the stream
        k_zero_u64<<<1, 64, 0, st>>>((unsigned long long*)counts_dev, 2);
        HIP_TRY(hipGetLastError());
        return FQSL_OK;
    }
    int rc = pseudo_check(n, b_flag, x_start, y_start, z_start, flag_start, bz_start, bz_end, capacity, seed_x, seed_y,
                          seed_z, bz0_start, bz0_end, pixel);
    if (rc) return rc;
    if (!workspace || workspace_bytes < pseudo_ws_bytes(n))
        return set_err(FQSL_ERR_INVALID, "workspace of %zu bytes, %zu needed (fqsl_pseudo_workspace_bytes)",
                       workspace_bytes, pseudo_ws_bytes(n));
    return pseudo_launch(n, b_flag, x_start, y_start, z_start, flag_start, bz_start, bz_end, stride, capacity, seed_x,
                         seed_y, seed_z, bz0_start, bz0_end, pixel, (unsigned long long*)counts_dev, workspace, st);
}

// Pinned read-back word per host thread (allocated once, kept until exit).
static thread_local unsigned long long* g_pinned_count = nullptr;

extern "C" int fqsl_pseudo_seeds(uint64_t n, const int32_t* b_flag, const float* x_start, const float* y_start,
                                 const float* z_start, const int32_t* flag_start, const float* bz_start,
                                 const float* bz_end, float stride, uint64_t capacity, float* seed_x, float* seed_y,
                                 float* seed_z, float* bz0_start, float* bz0_end, uint64_t* pixel,
                                 uint64_t* count_host, void* stream) {
    if (!count_host) return set_err(FQSL_ERR_INVALID, "count_host is NULL");
    *count_host = 0;
    if (n == 0) return FQSL_OK;
    int rc = pseudo_check(n, b_flag, x_start, y_start, z_start, flag_start, bz_start, bz_end, capacity, seed_x, seed_y,
                          seed_z, bz0_start, bz0_end, pixel);
    if (rc) return rc;
    if (!g_pinned_count) HIP_TRY(hipHostMalloc((void**)&g_pinned_count, 2 * sizeof(unsigned long long)));
    hipStream_t st = (hipStream_t)stream;
    const size_t wsb = pseudo_ws_bytes(n);
    void* ws = nullptr;
    HIP_TRY(hipMallocAsync(&ws, wsb + 2 * sizeof(unsigned long long), st));
    unsigned long long* total = (unsigned long long*)((char*)ws + wsb);
    rc = pseudo_launch(n, b_flag, x_start, y_start, z_start, flag_start, bz_start, bz_end, stride, capacity, seed_x,
                       seed_y, seed_z, bz0_start, bz0_end, pixel, total, ws, st);
    hipError_t e = rc ? hipSuccess : hipMemcpyAsync(g_pinned_count, total, sizeof(unsigned long long),
                                                     hipMemcpyDeviceToHost, st);
    (void)hipFreeAsync(ws, st);
    if (rc) return rc;
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "pseudo-line compaction: %s", hipGetErrorString(e));
    const unsigned long long cnt = g_pinned_count[0];
    *count_host = cnt;
    if (capacity == 0) return FQSL_OK;  // count only
    if (cnt > capacity)
        return set_err(FQSL_ERR_INVALID, "%llu pseudo pixels exceed capacity %llu (outputs hold the first %llu)",
                       (unsigned long long)cnt, (unsigned long long)capacity, (unsigned long long)capacity);
    return FQSL_OK;
}

static int qcut_impl(uint64_t m, const unsigned long long* m_dev, const float* start_x, const float* start_y,
                     const float* start_z, const int32_t* flag_start, const float* end_x, const float* end_y,
                     const float* end_z, const int32_t* flag_end, const float* bz0_start, const float* bz0_end,
                     float stride, const uint64_t* pixel, float* q_all, void* stream) {
    if (m == 0) return FQSL_OK;
    if (!start_x || !start_y || !start_z || !flag_start || !end_x || !end_y || !end_z || !flag_end || !bz0_start ||
        !bz0_end || !pixel || !q_all)
        return set_err(FQSL_ERR_INVALID, "NULL array argument");
    if (!(stride > 0)) return set_err(FQSL_ERR_INVALID, "stride must be > 0");
    // (1/stride/2)**4 in float32 (demo/Q_RKF45CuPy.ipynb cell [10])
    const float scale = powf((1.0f / stride) / 2.0f, 4.0f);
    k_qcut_scatter<<<grid_for((long)m, 256), 256, 0, (hipStream_t)stream>>>(
        m, start_x, start_y, start_z, flag_start, end_x, end_y, end_z, flag_end, bz0_start, bz0_end, scale, pixel,
        q_all, m_dev);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}

extern "C" int fqsl_qcut_scatter(uint64_t m, const float* start_x, const float* start_y, const float* start_z,
                                 const int32_t* flag_start, const float* end_x, const float* end_y,
                                 const float* end_z, const int32_t* flag_end, const float* bz0_start,
                                 const float* bz0_end, float stride, const uint64_t* pixel, float* q_all,
                                 void* stream) {
    return qcut_impl(m, nullptr, start_x, start_y, start_z, flag_start, end_x, end_y, end_z, flag_end, bz0_start,
                     bz0_end, stride, pixel, q_all, stream);
}

extern "C" int fqsl_qcut_scatter_n(uint64_t m_cap, const uint64_t* m_dev, const float* start_x, const float* start_y,
                                   const float* start_z, const int32_t* flag_start, const float* end_x,
                                   const float* end_y, const float* end_z, const int32_t* flag_end,
                                   const float* bz0_start, const float* bz0_end, float stride, const uint64_t* pixel,
                                   float* q_all, void* stream) {
    if (!m_dev) return set_err(FQSL_ERR_INVALID, "m_dev is NULL");
    return qcut_impl(m_cap, (const unsigned long long*)m_dev, start_x, start_y, start_z, flag_start, end_x, end_y,
                     end_z, flag_end, bz0_start, bz0_end, stride, pixel, q_all, stream);
}

// Pixels of a neighbour-difference map that the Scott method re-traces
// (demo/Q_RKF45MixCuPyDiffScott.ipynb cell [9]): B_flag + isnan(Q) > 0.5.
__device__ __forceinline__ bool bad_pixel(const int32_t* __restrict__ bflag, const float* __restrict__ q,
                                          unsigned long long i) {
    const float v = q[i];
    return bflag[i] + (v != v ? 1 : 0) > 0;
}

__global__ void k_bad_count(unsigned long long n, const int32_t* __restrict__ bflag, const float* __restrict__ q,
                            unsigned* __restrict__ counts) {
    __shared__ unsigned s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const unsigned long long i = blockIdx.x * (unsigned long long)kCompactBlock + threadIdx.x;
    const bool sel = i < n && bad_pixel(bflag, q, i);
    const unsigned long long m = __ballot(sel);
    if ((threadIdx.x & 63) == 0) atomicAdd(&s, (unsigned)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s;
}

__global__ void k_bad_write(unsigned long long n, const int32_t* __restrict__ bflag, const float* __restrict__ q,
                            const float* __restrict__ xs, const float* __restrict__ ys, const float* __restrict__ zs,
                            const unsigned* __restrict__ offs, unsigned long long capacity, float* __restrict__ ox,
                            float* __restrict__ oy, float* __restrict__ oz, uint64_t* __restrict__ pix) {
    __shared__ unsigned wave_base[kCompactBlock / 64];
    const unsigned long long i = blockIdx.x * (unsigned long long)kCompactBlock + threadIdx.x;
    const bool sel = i < n && bad_pixel(bflag, q, i);
    const unsigned long long m = __ballot(sel);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wave_base[w] = (unsigned)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned run = 0;
        for (int k = 0; k < kCompactBlock / 64; ++k) { const unsigned v = wave_base[k]; wave_base[k] = run; run += v; }
    }
    __syncthreads();
    if (!sel) return;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const unsigned long long j = (unsigned long long)offs[blockIdx.x] + wave_base[w] + __popcll(m & below);
    if (j >= capacity) return;
    ox[j] = xs[i];
    oy[j] = ys[i];
    oz[j] = zs[i];
    pix[j] = i;
}

__global__ void k_scatter_f32(unsigned long long m, const uint64_t* __restrict__ pix, const float* __restrict__ v,
                              float* __restrict__ target) {
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < m;
         i += (unsigned long long)gridDim.x * blockDim.x)
        target[pix[i]] = v[i];
}

extern "C" int fqsl_bad_pixels(uint64_t n, const int32_t* b_flag, const float* q, const float* x, const float* y,
                               const float* z, uint64_t capacity, float* seed_x, float* seed_y, float* seed_z,
                               uint64_t* pixel, uint64_t* count_host, void* stream) {
    if (!count_host) return set_err(FQSL_ERR_INVALID, "count_host is NULL");
    *count_host = 0;
    if (n == 0) return FQSL_OK;
    if (!b_flag || !q || !x || !y || !z) return set_err(FQSL_ERR_INVALID, "NULL input array");
    if (capacity && (!seed_x || !seed_y || !seed_z || !pixel)) return set_err(FQSL_ERR_INVALID, "NULL output array");
    if (n > 0xffffffffull) return set_err(FQSL_ERR_INVALID, "bad-pixel compaction supports < 2^32 pixels");
    hipStream_t st = (hipStream_t)stream;
    const unsigned nblocks = (unsigned)((n + kCompactBlock - 1) / kCompactBlock);
    // block counts and the total in one stream-ordered allocation
    const size_t counts_bytes = align256((size_t)nblocks * sizeof(unsigned) + 16);
    unsigned* counts = nullptr;
    HIP_TRY(hipMallocAsync((void**)&counts, counts_bytes + sizeof(unsigned long long), st));
    unsigned long long* total = (unsigned long long*)((char*)counts + counts_bytes);
    k_bad_count<<<nblocks, kCompactBlock, 0, st>>>(n, b_flag, q, counts);
    k_pseudo_scan<<<1, 1024, 0, st>>>(nblocks, counts, total);
    k_bad_write<<<nblocks, kCompactBlock, 0, st>>>(n, b_flag, q, x, y, z, counts, capacity, seed_x, seed_y, seed_z,
                                                   pixel);
    hipError_t e = hipGetLastError();
    unsigned long long cnt = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt, total, sizeof(cnt), hipMemcpyDeviceToHost, st);
    (void)hipFreeAsync(counts, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_err(FQSL_ERR_HIP, "bad-pixel compaction: %s", hipGetErrorString(e));
    *count_host = cnt;
    if (capacity == 0) return FQSL_OK;
    if (cnt > capacity)
        return set_err(FQSL_ERR_INVALID, "%llu bad pixels exceed capacity %llu (outputs hold the first %llu)",
                       (unsigned long long)cnt, (unsigned long long)capacity, (unsigned long long)capacity);
    return FQSL_OK;
}

extern "C" int fqsl_scatter_f32(uint64_t m, const uint64_t* pixel, const float* values, float* target, void* stream) {
    if (m == 0) return FQSL_OK;
    if (!pixel || !values || !target) return set_err(FQSL_ERR_INVALID, "NULL array");
    k_scatter_f32<<<grid_for((long)m, 256), 256, 0, (hipStream_t)stream>>>(m, pixel, values, target);
    HIP_TRY(hipGetLastError());
    return FQSL_OK;
}

