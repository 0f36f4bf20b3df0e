/*
 * fqsl.h -- C ABI of libfqsl.so, the MI355X (gfx950) squashing-factor engine.
 *
 * Plain pointers, sizes and status codes; no C++ or torch types.  Every entry
 * point returns FQSL_OK (0) or an error code and leaves a message for
 * fqsl_last_error() (thread-local).  Array arguments are device pointers
 * unless a parameter says otherwise; all work is enqueued on the given HIP
 * stream (NULL = the default stream) and nothing synchronises unless stated.
 *
 * Grid-index coordinates are used throughout: node (i, j, k) sits at
 * (x, y, z) = (i, j, k); fields are three arrays Bx, By, Bz of nx*ny*nz values
 * with x fastest, element (i, j, k) at k*ny*nx + j*nx + i
 * (fastqslpy/cu/TraceBlineScott.cu:36-39).
 *
 * Line classes written to flag_start / flag_end (TraceBlineScott.cu:310-332):
 *   1..6  left through x min, x max, y min, y max, z min, z max
 *   0     stopped on the length / step limit (its side's endpoint is the last
 *         state reached; Q = NaN)
 *   42    any other exit (edge, corner, NaN); the line stops, Q = 0
 *   -1    seed below z = -1e-8 (Scott tracer only; Q = NaN)
 */
#ifndef FQSL_H
#define FQSL_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define FQSL_OK 0
#define FQSL_ERR_INVALID 1     /* bad argument */
#define FQSL_ERR_HIP 2         /* a HIP runtime call failed */
#define FQSL_ERR_NOMEM 3       /* device allocation failed */
#define FQSL_ERR_UNSUPPORTED 4 /* valid request the library does not implement */

#define FQSL_DTYPE_F32 0
#define FQSL_DTYPE_F64 1

#define FQSL_METHOD_RKF45 0
#define FQSL_METHOD_RK4 1

/* precision of the tracing arithmetic */
#define FQSL_PREC_FP32C 0 /* FP32 field/stages, compensated position, FP64 controller and Q */
#define FQSL_PREC_FP64 1  /* FP64 throughout */
#define FQSL_PREC_MIXED 2 /* Scott tracer: FP32 Jacobian (grad b, u, v); FP64 position path (b at the
                             stage points, stage positions, increments, error norm), controller and Q */

/* how lines are mapped to lanes */
#define FQSL_SCHED_PERSISTENT 0 /* persistent grid, each lane refills from a device queue */
#define FQSL_SCHED_PER_LINE 1   /* one thread per line, one launch */

/* field layouts held on the device (bit mask for fqsl_field_create) */
#define FQSL_LAYOUT_SOA32 1  /* three float arrays, x fastest */
#define FQSL_LAYOUT_SOA64 2  /* three double arrays, x fastest */
#define FQSL_LAYOUT_CELL32 4 /* 128-byte cells: 8 trilinear coefficients x 3 components, float */
#define FQSL_LAYOUT_ALL 7

const char *fqsl_last_error(void);
const char *fqsl_version(void);

/* ------------------------------------------------------------------ fields */
typedef struct fqsl_field fqsl_field;

/* Copy a field to the current device and build the requested layouts; the
 * library owns the result.  bx/by/bz are host pointers when on_device == 0,
 * device pointers otherwise.  Synchronous.  Replaces the notebooks' cupy.array
 * upload of Bx/By/Bz (demo/Q_RKF45CuPyScott.ipynb cell [4]). */
int fqsl_field_create(const void *bx, const void *by, const void *bz, int nx, int ny, int nz,
                      int dtype, int on_device, int layouts, fqsl_field **out);
int fqsl_field_destroy(fqsl_field *field);
/* Device bytes held by the field's layouts. */
size_t fqsl_field_bytes(const fqsl_field *field);
int fqsl_field_shape(const fqsl_field *field, int *nx, int *ny, int *nz);
/* Attach the current density J (three arrays on the field's grid, x fastest)
 * for the twist number: the tracers' curB_x/y/z (TraceBline.cu:301).  Kept in
 * float and double.  Synchronous; replaces an earlier one. */
int fqsl_field_set_current(fqsl_field *field, const void *jx, const void *jy, const void *jz, int dtype,
                           int on_device);

/* Give the field stretched (non-uniform) axes: node (i, j, k) at
 * (xa[i], ya[j], za[k]), each axis strictly increasing with the field's
 * node count.  Seeds and results of fqsl_trace_plain are then in these
 * coordinates, and it runs the stretched RKF45 tracer of
 * TraceBlineAdaptiveStretch.cu:323-508 (the reference passes x_arr / y_arr /
 * z_arr and flag_uni per launch, :422-431).  lookup picks the interpolation
 * cell: FQSL_LOOKUP_BISECT (GetIndexBiSearch, :42-65, the reference with
 * flag_uni false), FQSL_LOOKUP_UNIFORM (GetIndexUniStretch, :68-75, flag_uni
 * true) -- both land face exits on the plane 0 / n_dim as the reference does
 * (:385-386) -- or FQSL_LOOKUP_EXACT (the cell holding the point; landing on
 * the box faces).  All three axis pointers NULL returns the field to
 * grid-index coordinates.  Scott tracing and batches reject a stretched
 * field.  Synchronous. */
#define FQSL_LOOKUP_BISECT 0
#define FQSL_LOOKUP_UNIFORM 1
#define FQSL_LOOKUP_EXACT 2
int fqsl_field_set_axes(fqsl_field *field, const void *xa, const void *ya, const void *za, int dtype,
                        int on_device, int lookup);
/* 1 if the field has stretched axes, else 0. */
int fqsl_field_stretched(const fqsl_field *field);

/* ----------------------------------------------------- current density */
/* curl B on the node grid of axes ax/ay/az (device double arrays of nx, ny,
 * nz coordinates), FastQSL.curl_of_B_grid (fastqslpy/FastQSL.py:120-131):
 * three-point differences on non-uniform spacing at interior nodes, boundary
 * nodes copied from the nearest interior node.  B in dtype, curl in double;
 * all arrays x fastest, nx, ny, nz >= 3. */
int fqsl_curl(const void *bx, const void *by, const void *bz, int nx, int ny, int nz, int dtype,
              const double *ax, const double *ay, const double *az, double *cx, double *cy, double *cz,
              void *stream);
/* diff_3point (FastQSL.py:143-163): derivative along axis (0, 1, 2) at the
 * interior nodes, out of (nx-2)(ny-2)(nz-2) doubles, x fastest. */
int fqsl_diff_3point(const void *cube, int nx, int ny, int nz, int dtype, const double *ax, int axis,
                     double *out, void *stream);
/* copy_boundary (FastQSL.py:133-141) in place on a double array. */
int fqsl_copy_boundary(double *a, int nx, int ny, int nz, void *stream);

/* ------------------------------------------------------------ Scott tracer */
typedef struct {
    int method;     /* FQSL_METHOD_RKF45 (TraceBlineScott.cu:256) or FQSL_METHOD_RK4 */
    int precision;  /* FQSL_PREC_* */
    int schedule;   /* FQSL_SCHED_* */
    int reserved;
    double step;    /* initial step (RKF45) or fixed step (RK4), cells */
    double tol;     /* RKF45 absolute per-step tolerance, cells */
} fqsl_scott_params;

/* Per-line outputs (device arrays of n elements).  Any pointer may be NULL
 * to skip that output.  Mirrors the outputs of TraceAllBlineScott
 * (TraceBlineScott.cu:521-528) plus a stage-evaluation count. */
typedef struct {
    float *start_x, *start_y, *start_z; /* backward footpoint */
    float *end_x, *end_y, *end_z;       /* forward footpoint */
    int32_t *flag_start, *flag_end;
    float *b_this_x, *b_this_y, *b_this_z; /* B at the seed, Interp3d rule */
    float *q_0, *q_perp;                /* TraceBlineScott.cu:495-508 */
    double *line_len;
    uint32_t *evals;                    /* RHS evaluations spent on the line */
    double *twist;                      /* twist number (needs the field's current density,
                                           fqsl_field_set_current; TraceBlineScott.cu:423-457) */
} fqsl_scott_out;

/* Trace n seeds (device float arrays x, y, z) with the Scott method and
 * assemble q_0 and q_perp per line.  Replaces a launch of
 * TraceAllBlineScott (fastqslpy/cu/TraceBlineScott.cu:521-595).
 * FQSL_PREC_FP32C accepts RKF45 tolerances down to 1e-9 cells (its error
 * norm is formed from float stages, with a round-off floor near 1e-8 x the
 * step); a smaller tol fails with FQSL_ERR_INVALID -- FQSL_PREC_MIXED and
 * FQSL_PREC_FP64 resolve any tolerance. */
int fqsl_trace_scott(const fqsl_field *field, const fqsl_scott_params *params, uint64_t n,
                     const float *x, const float *y, const float *z, const fqsl_scott_out *out,
                     void *stream);

/* fqsl_trace_scott that allocates nothing.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The persistent schedule's workspace (queue counters, seed
 * frames, finished-line records, queue order) lives in a fixed array the
 * library holds per device from its load until the process ends (about
 * 148 MiB): n is at most 2^20 lines with FQSL_PREC_FP32C or FQSL_PREC_MIXED
 * and 2^19 with FQSL_PREC_FP64, larger n fails with FQSL_ERR_INVALID, and params->schedule
 * must be FQSL_SCHED_PERSISTENT.  Results are the same bits as
 * fqsl_trace_scott's.  All calls on one device share that array: calls that
 * overlap on one device (other streams or host threads) race and must be
 * serialised by the caller.  It records no timing events
 * (fqsl_last_kernel_ms / fqsl_last_step_ms then fail). */
int fqsl_trace_scott_noalloc(const fqsl_field *field, const fqsl_scott_params *params, uint64_t n,
                             const float *x, const float *y, const float *z,
                             const fqsl_scott_out *out, void *stream);

/* ----------------------------------------------------------------- batches */
/* Many Scott requests -- each with its own field, method, step, tolerance,
 * seeds and outputs -- traced together.  The reference has no batch call:
 * its notebooks launch TraceAllBlineScott once per map
 * (demo/Q_RKF45CuPy3DScott.ipynb cell [7] loops over slices); a batch replaces
 * such a loop of launches.  Requests are grouped by (field, method) and each
 * group is drained by one persistent launch. */
typedef struct {
    const fqsl_field *field;
    int method;       /* FQSL_METHOD_* */
    int reserved;
    double step;      /* initial (RKF45) or fixed (RK4) step, cells */
    double tol;       /* RKF45 absolute per-step tolerance, cells */
    uint64_t n;       /* number of seeds (0 allowed) */
    const float *x, *y, *z;
    fqsl_scott_out out;
} fqsl_scott_request;

typedef struct fqsl_scott_batch fqsl_scott_batch;

/* Device workspace (bytes, 256-byte aligned) a batch of these requests needs:
 * request table, counters and per-line scratch (8 words + one record per line). */
size_t fqsl_scott_batch_workspace_bytes(const fqsl_scott_request *reqs, int nreq, int precision);

/* Plans a batch in `workspace` (device memory the caller owns and keeps alive,
 * with the fields, seeds and outputs, for the batch's lifetime): uploads the
 * request table and synchronises `stream` once. */
int fqsl_scott_batch_create(const fqsl_scott_request *reqs, int nreq, int precision, void *workspace,
                            size_t workspace_bytes, void *stream, fqsl_scott_batch **batch);

/* Traces every request of the batch on `stream`.  Only launches kernels (no
 * allocation, no memset, no synchronisation), so it may be captured in a HIP
 * graph and replayed; each run overwrites the outputs.  Runs of one batch must
 * not overlap (they share the workspace). */
int fqsl_scott_batch_run(const fqsl_scott_batch *batch, void *stream);

/* Per request (caller's order): milliseconds from the start of the last run
 * to the moment its last line finished stepping, from the device clock.
 * Synchronises `stream`. */
int fqsl_scott_batch_times(const fqsl_scott_batch *batch, double *ms, void *stream);

int fqsl_scott_batch_destroy(fqsl_scott_batch *batch);

/* -------------------------------------------------------------- plane maps */
#define FQSL_PLANE_Z 0   /* x = linspace(x0, x1, n_u), y = linspace(y0, y1, n_v), z = z0: the
                            notebooks' constant-z grid (demo/Q_RKF45CuPy.ipynb cell [6]) */
#define FQSL_PLANE_ABC 1 /* A + u (B - A) + v (C - A), u = linspace(0, 1, n_u),
                            v = linspace(0, 1, n_v): a plane through three points
                            (demo/Q_RKF45CuPyPlane.ipynb cell [4]) */
typedef struct {
    int kind;      /* FQSL_PLANE_* */
    int n_u, n_v;  /* pixels per row and rows; maps are n_v x n_u, row-major */
    int reserved;
    double p[9];   /* FQSL_PLANE_Z: x0, x1, y0, y1, z0 (p[5..8] unused, finite, e.g. 0);
                      FQSL_PLANE_ABC: A, B, C.  A non-finite entry is rejected. */
} fqsl_plane;

/* Scott Q map of a plane of seeds, formed on the device: linspace as NumPy
 * evaluates it (i * ((b - a) / (n - 1)) + a in double, the last point exactly
 * b) and, for FQSL_PLANE_ABC, A + (B - A) u + (C - A) v in double without
 * fused multiply-adds, each coordinate then rounded to float -- the seeds of
 * fastqsl_amd.synth.plane_seeds_z / plane_seeds_abc bit for bit -- traced as
 * fqsl_trace_scott traces them.  Every output array holds n_v * n_u values,
 * row-major (rows follow y, or C - A).  Replaces the notebook host code
 * around one TraceAllBlineScott launch (demo/Q_RKF45CuPyScott.ipynb cells
 * [5]-[7]); the seeds live in stream-ordered scratch freed on `stream`. */
int fqsl_qmap_plane(const fqsl_field *field, const fqsl_plane *plane, const fqsl_scott_params *params,
                    const fqsl_scott_out *out, void *stream);

/* Neighbour-difference Q map of a plane of seeds (the pipeline of
 * demo/Q_RKF45CuPy.ipynb cells [6]-[11]): the plane's seeds as in
 * fqsl_qmap_plane, position-only traces (fqsl_trace_plain), QCalcPlane
 * (FastQSL.py:65-117) into the interior of q_all, CookPseudoLine
 * (FastQSL.py:4-57) for the pixels with B_flag set, their four pseudo lines
 * traced with tol_coef x pseudo_tol_scale, and Qcut scattered into q_all --
 * all on `stream`, the pseudo-line count kept on the device.  q_all holds
 * n_v * n_u floats (row-major, border pixels 0); *n_pseudo_dev (a device
 * word, may be NULL) receives the number of pseudo-line pixels.  Plain
 * methods and precisions only (FQSL_PREC_FP32C or FQSL_PREC_FP64); the plane
 * needs at least 3 x 3 pixels.  Its scratch is stream-ordered and freed on
 * `stream`. */
typedef struct {
    int method, precision, schedule, reserved;  /* as fqsl_plain_params */
    double step;             /* s_len, cells */
    double tol_coef;         /* tolerance coefficient of the main lines */
    double pseudo_tol_scale; /* the pseudo lines use tol_coef * pseudo_tol_scale (notebooks: 0.1) */
    double stride;           /* pixel spacing in cells (QCalcPlane's stride_step) */
    double pseudo_stride;    /* offset of the four pseudo seeds (CookPseudoLine's stride) */
    float normal[3];         /* plane normal: tolerance rule, B_flag and QCalcPlane */
    float pseudo_normal[3];  /* normal for the pseudo lines' tolerance rule */
} fqsl_neighbour_params;

int fqsl_qmap_plane_neighbour(const fqsl_field *field, const fqsl_plane *plane,
                              const fqsl_neighbour_params *params, float *q_all, uint64_t *n_pseudo_dev,
                              void *stream);

/* ----------------------------------------------------------------- serving */
/* A request queue with one consumer thread per server that traces what has
 * arrived as batches (continuous batching): whenever the device is free, the
 * consumer takes the waiting requests in submission order, up to max_lines
 * lines (at least one request), traces them as one fqsl_scott_batch on its
 * own stream and marks them done.  The reference has no serving layer; its
 * notebooks launch one kernel per map.  Submission is thread-safe.  The
 * caller keeps a request's field, seeds and outputs alive until fqsl_poll
 * reports it done or fqsl_wait returns; each request's outputs are the bits
 * fqsl_trace_scott gives for it alone. */
typedef struct fqsl_server fqsl_server;

/* A server on `device` tracing with `precision` (FQSL_PREC_*); max_lines
 * bounds one batch (0: 2^24 lines). */
int fqsl_server_create(int device, int precision, uint64_t max_lines, fqsl_server **server);

/* Queue one request (copied); *id identifies it (ids start at 1).  The
 * request is ordered after the work enqueued on `stream` (a stream of the
 * server's device, NULL the null stream) up to this call, so its seeds may
 * still be in flight there. */
int fqsl_submit(fqsl_server *server, const fqsl_scott_request *req, void *stream, uint64_t *id);

/* *done = 1 once request `id` has been traced (its status is then kept for
 * fqsl_wait), 0 while it is queued or running.  Unknown ids fail. */
int fqsl_poll(fqsl_server *server, uint64_t id, int *done);

/* Blocks until request `id` is done, returns its status (FQSL_OK or the
 * batch's error, with fqsl_last_error set on this thread) and forgets it.
 * Of several calls waiting on one id, one gets the status and the others
 * fail with FQSL_ERR_INVALID. */
int fqsl_wait(fqsl_server *server, uint64_t id);

/* Traces what is still queued, stops the consumer and frees the server;
 * requests not yet waited for are forgotten.  No other call on the server
 * may be in progress or follow. */
int fqsl_server_destroy(fqsl_server *server);

/* ------------------------------------------------------------ plain tracer */
typedef struct {
    int method;      /* FQSL_METHOD_RKF45 (TraceBlineAdaptive.cu) or FQSL_METHOD_RK4 (TraceBline.cu) */
    int precision;   /* FQSL_PREC_* */
    int schedule;    /* FQSL_SCHED_* */
    int reserved;
    double step;     /* initial step, cells (each direction restarts from it) */
    double tol_coef; /* tolerance coefficient, TraceBlineAdaptive.cu:354-359 */
    float ncross[3]; /* plane normal: tolerance rule and B_flag */
    float reserved2;
} fqsl_plain_params;

/* Mirrors the outputs of TraceAllBline (TraceBlineAdaptive.cu:429-437). */
typedef struct {
    float *start_x, *start_y, *start_z;
    float *end_x, *end_y, *end_z;
    int32_t *flag_start, *flag_end;
    float *b_this_x, *b_this_y, *b_this_z;
    int32_t *b_flag;
    float *b_start_x, *b_start_y, *b_start_z;
    float *b_end_x, *b_end_y, *b_end_z;
    double *line_len;
    uint32_t *evals;
    double *twist;   /* twist number, forward plus backward (TraceBline.cu:325-365 for RK4,
                        TraceBlineAdaptive.cu:373-384 for RKF45); needs the current density */
} fqsl_plain_out;

/* Position-only tracing of n seeds, forward then backward.  Replaces a launch
 * of TraceAllBline (TraceBlineAdaptive.cu:429-510, TraceBline.cu:383-464). */
int fqsl_trace_plain(const fqsl_field *field, const fqsl_plain_params *params, uint64_t n,
                     const float *x, const float *y, const float *z, const fqsl_plain_out *out,
                     void *stream);

/* fqsl_trace_plain of min(n_cap, *n_dev) seeds, the count read on the device
 * when the kernel runs (n_dev: one device word written earlier on the stream,
 * e.g. counts_dev + 1 of fqsl_pseudo_seeds_async); seed and output arrays
 * hold n_cap entries.  Nothing is read back to the host.  Replaces the second
 * TraceAllBline launch of the pseudo lines (demo/Q_RKF45CuPy.ipynb cell [9]),
 * whose N the notebook sizes on the host. */
int fqsl_trace_plain_n(const fqsl_field *field, const fqsl_plain_params *params, uint64_t n_cap,
                       const uint64_t *n_dev, const float *x, const float *y, const float *z,
                       const fqsl_plain_out *out, void *stream);

/* ------------------------------------------- neighbour-difference assembly */
/* QCalcPlane (fastqslpy/FastQSL.py:65-117) on a (ny, nx) row-major pixel grid:
 * writes q of (ny-2)*(nx-2) floats.  Arithmetic follows the reference's
 * float32 expression tree operation by operation.  b_in_z / b_out_z are Bz at
 * the start / end footpoints; b0_* is B at the seed. */
int fqsl_qcalc_plane(int ny, int nx, const float *x_end, const float *y_end, const float *z_end,
                     const int32_t *flag_end, const float *x_start, const float *y_start,
                     const float *z_start, const int32_t *flag_start, const float *b_in_z,
                     const float *b_out_z, const float *b0_x, const float *b0_y, const float *b0_z,
                     double stride_step, const float *cross_dir_host, float *q, void *stream);

/* CookPseudoLine (FastQSL.py:4-57): for every pixel with b_flag == 1 (in
 * pixel order) write four seeds at its start footpoint +/- stride along the
 * two in-face axes, plus Bz at both footpoints and the pixel index.  The
 * output arrays hold up to `capacity` pixels (4*capacity seeds); the number of
 * selected pixels is written to *count_host after a stream synchronisation.
 * capacity == 0 only counts (output pointers may then be NULL). */
int fqsl_pseudo_seeds(uint64_t n, const int32_t *b_flag, const float *x_start, const float *y_start,
                      const float *z_start, const int32_t *flag_start, const float *bz_start,
                      const float *bz_end, float stride, uint64_t capacity, float *seed_x,
                      float *seed_y, float *seed_z, float *bz0_start, float *bz0_end,
                      uint64_t *pixel, uint64_t *count_host, void *stream);

/* Device bytes of the workspace fqsl_pseudo_seeds_async needs for n pixels. */
size_t fqsl_pseudo_workspace_bytes(uint64_t n);

/* fqsl_pseudo_seeds without a host round trip (FastQSL.py:4-57 as the first
 * half of a pipeline that stays on the stream): the counts go to two device
 * words, counts_dev[0] = selected pixels and counts_dev[1] = their 4x pseudo
 * lines; outputs hold the first min(count, capacity) pixels.  The caller
 * passes a device workspace of fqsl_pseudo_workspace_bytes(n); nothing is
 * allocated and the stream is not synchronised. */
int fqsl_pseudo_seeds_async(uint64_t n, const int32_t *b_flag, const float *x_start,
                            const float *y_start, const float *z_start, const int32_t *flag_start,
                            const float *bz_start, const float *bz_end, float stride, uint64_t capacity,
                            float *seed_x, float *seed_y, float *seed_z, float *bz0_start,
                            float *bz0_end, uint64_t *pixel, uint64_t *counts_dev, void *workspace,
                            size_t workspace_bytes, void *stream);

/* Qcut (demo/Q_RKF45CuPy.ipynb cell [10]) from the traces of the 4*m pseudo
 * seeds, scattered into q_all[pixel[i]] (cell [11]). */
int fqsl_qcut_scatter(uint64_t m, const float *start_x, const float *start_y, const float *start_z,
                      const int32_t *flag_start, const float *end_x, const float *end_y,
                      const float *end_z, const int32_t *flag_end, const float *bz0_start,
                      const float *bz0_end, float stride, const uint64_t *pixel, float *q_all,
                      void *stream);

/* fqsl_qcut_scatter over min(m_cap, *m_dev) pseudo pixels, the count read on
 * the device (counts_dev of fqsl_pseudo_seeds_async); same cells. */
int fqsl_qcut_scatter_n(uint64_t m_cap, const uint64_t *m_dev, const float *start_x,
                        const float *start_y, const float *start_z, const int32_t *flag_start,
                        const float *end_x, const float *end_y, const float *end_z,
                        const int32_t *flag_end, const float *bz0_start, const float *bz0_end,
                        float stride, const uint64_t *pixel, float *q_all, void *stream);

/* ------------------------------------------------- around the Q map */
/* slog10 Q = log10(q) * sign(bn) [* (flag <= 6) when flag is not NULL]
 * (demo/Q_RKF45CuPy.ipynb cells [13]-[14], demo/Q_RKF45MixCuPyDiffScott.ipynb
 * cell [8]); n device floats, NumPy's sign and operation order. */
int fqsl_slog10q(uint64_t n, const float *q, const float *bn, const int32_t *flag, float *out, void *stream);

/* The pixels the Scott method re-traces in the mixed method
 * (demo/Q_RKF45MixCuPyDiffScott.ipynb cell [9]): b_flag + isnan(q) > 0.5,
 * in pixel order.  Writes up to capacity seeds (x, y, z of those pixels) and
 * pixel indices; count_host receives the number of such pixels (capacity 0:
 * count only).  Synchronises the stream. */
int fqsl_bad_pixels(uint64_t n, const int32_t *b_flag, const float *q, const float *x, const float *y,
                    const float *z, uint64_t capacity, float *seed_x, float *seed_y, float *seed_z,
                    uint64_t *pixel, uint64_t *count_host, void *stream);

/* target[pixel[i]] = values[i] for i < m (the fix-up's
 * Qall_1d[idx_bad] = q_perp_bad, same cell). */
int fqsl_scatter_f32(uint64_t m, const uint64_t *pixel, const float *values, float *target, void *stream);

/* curl B in grid-index units as the twist notebooks compute it
 * (demo/Q_RK4.ipynb cell [11], demo/npzinputRKF45CuPyTwist.ipynb cell [9]):
 * central differences on the notebook's index ranges, a one-sided
 * second-order z derivative at z = 0 for curl_x and curl_y, zero elsewhere;
 * arrays x fastest in dtype, curl in the same dtype. */
int fqsl_curl_index(const void *bx, const void *by, const void *bz, int nx, int ny, int nz, int dtype, void *cx,
                    void *cy, void *cz, void *stream);

/* ------------------------------------------- reference-shaped entry points */
/* Same parameters, order and meaning as the CUDA kernels the fastqslpy host
 * code launches (all device pointers; scalars as one-element device arrays),
 * plus the stream.  The field arrays are read in place (SoA float layout).
 * With curB_flag set (read from the device: one stream synchronisation) the
 * twist number is written to twist[] from the current density curB_x/y/z
 * (float, same layout as B); with it clear, twist[] and curB are not read. */

/* TraceAllBlineScott (TraceBlineScott.cu:521-528).  tol_coef scales the
 * per-step tolerance 1e-3 * tol_coef.  precision (FQSL_PREC_FP32C or
 * FQSL_PREC_MIXED) / schedule select the path; the persistent schedule reads
 * N back to the host (one stream synchronisation) to size its workspace. */
int fqsl_TraceAllBlineScott(float *Bx, float *By, float *Bz, int *BshapeN, float *curB_x,
                            float *curB_y, float *curB_z, double *twist, bool *curB_flag,
                            float *inp_x, float *inp_y, float *inp_z, float *inp_cross_dir,
                            float *start_x, float *start_y, float *start_z, int *flag_start,
                            float *end_x, float *end_y, float *end_z, int *flag_end,
                            float *B_this_x, float *B_this_y, float *B_this_z, float *q_0,
                            float *q_perp, float *s_len, unsigned long long *N, double *LineLen,
                            float *tol_coef, int precision, int schedule, void *stream);

/* TraceAllBline of TraceBlineAdaptive.cu:429-437 (method RKF45) or of
 * TraceBline.cu:383-391 (method RK4; tol_coef may be NULL). */
int fqsl_TraceAllBline(int method, float *Bx, float *By, float *Bz, int *BshapeN, float *curB_x,
                       float *curB_y, float *curB_z, double *twist, bool *curB_flag, float *inp_x,
                       float *inp_y, float *inp_z, float *inp_cross_dir, float *start_x,
                       float *start_y, float *start_z, int *flag_start, float *end_x, float *end_y,
                       float *end_z, int *flag_end, float *B_this_x, float *B_this_y,
                       float *B_this_z, int *B_flag, float *B_start_x, float *B_start_y,
                       float *B_start_z, float *B_end_x, float *B_end_y, float *B_end_z,
                       float *s_len, unsigned long long *N, double *LineLen, float *tol_coef,
                       int precision, int schedule, void *stream);

/* TraceAllBline of TraceBlineAdaptiveStretch.cu:422-508 with its 39
 * arguments in the reference's order (device pointers; BshapeN, flag_uni,
 * curB_flag, s_len, N and tol_coef are one-element device arrays).  Reads
 * BshapeN, flag_uni, curB_flag and the three axes to the host (one stream
 * synchronisation), builds the axes' lookup tables (kept for the next call
 * while axes and flag_uni are unchanged), then traces the in-place float
 * field.  precision must be FQSL_PREC_FP32C. */
int fqsl_TraceAllBlineStretch(float *Bx, float *By, float *Bz, int *BshapeN, float *x_arr, float *y_arr,
                              float *z_arr, bool *flag_uni, float *curB_x, float *curB_y, float *curB_z,
                              double *twist, bool *curB_flag, float *inp_x, float *inp_y, float *inp_z,
                              float *inp_cross_dir, float *start_x, float *start_y, float *start_z,
                              int *flag_start, float *end_x, float *end_y, float *end_z, int *flag_end,
                              float *B_this_x, float *B_this_y, float *B_this_z, int *B_flag,
                              float *B_start_x, float *B_start_y, float *B_start_z, float *B_end_x,
                              float *B_end_y, float *B_end_z, float *s_len, unsigned long long *N,
                              double *LineLen, float *tol_coef, int precision, int schedule, void *stream);

/* ---------------------------------------------------------- tuning knobs */
/* Process-wide launch parameters (fastqsl_amd._lib.tune_set / tune_get).
 * They change speed only, never results. */
#define FQSL_TUNE_SCOTT_WAVES 0 /* waves/SIMD of the FP32c Scott kernels: 3 (default: the
                                 stages rolled into one loop), 2 (unrolled) or 0
                                 (compiler's choice); a launch of at most one two-wave
                                 grid's lanes (131,072 lines on 256 CUs) runs two; the
                                 mixed variant and the twist tracer always run two */
#define FQSL_TUNE_GRID_PCT 1    /* persistent grid size in % of one full occupancy wave (default 100) */
#define FQSL_TUNE_QUEUES 2      /* seed queues of the persistent Scott kernel: 1 (default) or 8 (one per XCD) */
#define FQSL_TUNE_REFILL_MIN 3  /* idle lanes a wavefront gathers before it refills them (1..64, default 8) */
#define FQSL_TUNE_ORDER 4       /* persistent Scott queue order: 0 = seed order, 2 = longest first by a
                                   coarse length estimate, 1 (default) = longest first up to 64 lines per
                                   persistent lane (ordering changes timing, never results) */
#define FQSL_TUNE_ORDER_BUCKETS 5 /* length buckets of that order: power of two in [2, 256], default 64 */
#define FQSL_TUNE_PLAIN_REFILL 6  /* idle lanes a wavefront of the persistent plain tracer gathers before it
                                     refills them (1..64, default 64: a wavefront claims new lines
                                     once all its lanes are idle) */
#define FQSL_TUNE_STRETCH_LDS 7   /* FP32 stretched tracer: 1 (default) = each workgroup copies the axis
                                     lookup table to LDS when it fits 64 KiB; 0 = global-memory tables */
#define FQSL_TUNE_COUNT 8
int fqsl_tune_set(int knob, int value);
int fqsl_tune_get(int knob, int *value);

/* -------------------------------------------------------------- timing */
/* Duration in milliseconds of the most recent trace kernel launched by
 * fqsl_trace_scott / fqsl_trace_plain on this thread, measured with HIP
 * events recorded around it on its stream (synchronises on the end event). */
int fqsl_last_kernel_ms(float *ms);

/* Duration in milliseconds of the stepping kernel alone of the most recent
 * persistent fqsl_trace_scott on this thread (HIP events around that launch;
 * synchronises on its end event).  FQSL_ERR_INVALID if that trace had no
 * separate stepping kernel (per-line schedule). */
int fqsl_last_step_ms(float *ms);

#ifdef __cplusplus
}
#endif

#endif /* FQSL_H */
