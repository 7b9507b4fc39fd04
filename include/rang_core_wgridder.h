/*
 * C ABI of the exact w-gridder (vis2dirty / dirty2vis), MI355X build.
 *
 * The rang_core_wgridder_* symbols are the drop-in replacement of rang-core's
 * core/src/ffi_wgridder.rs: identical names, argument order, struct layout
 * and status codes (0 ok, 1 invalid arguments with the message kept per
 * thread for rang_core_wgridder_last_error, 2 a failure inside the library).
 * All buffers are host memory owned by the caller and borrowed for the call;
 * the work runs on the GPU and outputs are complete on return.
 *
 * The rang_amd_* symbols have no counterpart in the reference: the same
 * transforms on device-resident buffers and a caller's HIP stream, for
 * callers that keep visibilities in HBM (solvers, benchmarks).
 *
 * Conventions (core/src/wgridder.rs:30-36): pixel (i, j) of an nx * ny image
 * is at i * ny + j, l = lshift + (i - nx/2) dx, m = mshift + (j - ny/2) dy;
 * vis2dirty = sum_k wgt_k Re[V_k exp(+2 pi i (u l + v m + w (n - 1)))],
 * dirty2vis its adjoint V_k = wgt_k sum I exp(-2 pi i (...)); uvw in metres
 * (nrow * 3 doubles), freq in Hz ascending (nchan doubles), visibilities
 * nrow * nchan interleaved complex values, row-major.
 */
#ifndef RANG_CORE_WGRIDDER_H
#define RANG_CORE_WGRIDDER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* core/src/ffi_wgridder.rs:22-43 (#[repr(C)] WgridderOptions);
 * kernel_index < 0 lets the cost model choose kernel and grid. */
typedef struct WgridderOptions {
    size_t nx;
    size_t ny;
    double pixsize_x;
    double pixsize_y;
    double epsilon;
    double center_x;
    double center_y;
    double sigma_min;
    double sigma_max;
    int32_t do_wgridding;
    int32_t divide_by_n;
    int32_t flip_u;
    int32_t flip_v;
    int32_t flip_w;
    int32_t allow_nshift;
    int64_t kernel_index;
    size_t nu;
    size_t nv;
} WgridderOptions;

/* Number of doubles written to `info` (core/src/wgridder.rs:490): kernel
 * index, support, nu, nv, nplanes, nvis, wmin_d, wmax_d, nm1min, nm1max,
 * nshift, dw, wmin, ofactor. */
#define RANG_WGRIDDER_N_INFO 14

/* vis_kind / wgt_kind (core/src/ffi_wgridder.rs:296-310) */
#define WGRIDDER_VIS_UNIT 0
#define WGRIDDER_VIS_C64 1
#define WGRIDDER_VIS_C128 2
#define WGRIDDER_WGT_NONE 0
#define WGRIDDER_WGT_F32 1
#define WGRIDDER_WGT_F64 2
#define WGRIDDER_WGT_F32_PRODUCT 3

/* core/src/ffi_wgridder.rs:155 */
int32_t rang_core_wgridder_vis2dirty_f64(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                         const double *uvw, const double *freq, const double *vis,
                                         const double *wgt, const uint8_t *mask, double *dirty,
                                         double *info);
/* core/src/ffi_wgridder.rs:191 */
int32_t rang_core_wgridder_vis2dirty_f32(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                         const double *uvw, const double *freq, const float *vis,
                                         const float *wgt, const uint8_t *mask, float *dirty,
                                         double *info);
/* core/src/ffi_wgridder.rs:228 */
int32_t rang_core_wgridder_dirty2vis_f64(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                         const double *uvw, const double *freq, const double *dirty,
                                         const double *wgt, const uint8_t *mask, double *vis,
                                         double *info);
/* core/src/ffi_wgridder.rs:264 */
int32_t rang_core_wgridder_dirty2vis_f32(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                         const double *uvw, const double *freq, const float *dirty,
                                         const float *wgt, const uint8_t *mask, float *vis,
                                         double *info);
/* core/src/ffi_wgridder.rs:454 */
int32_t rang_core_wgridder_vis2dirty_ex_f64(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                            const double *uvw, const double *freq, const void *vis,
                                            int32_t vis_kind, const void *wgt, int32_t wgt_kind,
                                            const uint8_t *mask, double *dirty, double *psf,
                                            double *info);
/* core/src/ffi_wgridder.rs:496 */
int32_t rang_core_wgridder_vis2dirty_ex_f32(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                            const double *uvw, const double *freq, const void *vis,
                                            int32_t vis_kind, const void *wgt, int32_t wgt_kind,
                                            const uint8_t *mask, float *dirty, float *psf,
                                            double *info);
/* core/src/ffi_wgridder.rs:542 */
int32_t rang_core_wgridder_vis2dirty_ex2_f64(const WgridderOptions *opts, size_t nrow,
                                             size_t nchan, const double *uvw, const double *freq,
                                             const void *vis, int32_t vis_kind, const void *wgt,
                                             const void *wgt2, int32_t wgt_kind,
                                             const uint8_t *mask, double *dirty, double *psf,
                                             double *info);
/* core/src/ffi_wgridder.rs:573 */
int32_t rang_core_wgridder_vis2dirty_ex2_f32(const WgridderOptions *opts, size_t nrow,
                                             size_t nchan, const double *uvw, const double *freq,
                                             const void *vis, int32_t vis_kind, const void *wgt,
                                             const void *wgt2, int32_t wgt_kind,
                                             const uint8_t *mask, float *dirty, float *psf,
                                             double *info);
/* core/src/ffi_wgridder.rs:668 */
int32_t rang_core_wgridder_dirty2vis_ex_f64(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                            const double *uvw, const double *freq,
                                            const double *dirty, const void *wgt, int32_t wgt_kind,
                                            const uint8_t *mask, double *vis, double *info);
/* core/src/ffi_wgridder.rs:695 */
int32_t rang_core_wgridder_dirty2vis_ex_f32(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                            const double *uvw, const double *freq,
                                            const float *dirty, const void *wgt, int32_t wgt_kind,
                                            const uint8_t *mask, float *vis, double *info);
/* core/src/ffi_wgridder.rs:722: seven doubles (W, ofactor, epsilon, beta,
 * e0, ndim, singleprec); 1 past the end of the table. */
int32_t rang_core_wgridder_kernel_row(size_t index, double *out);
/* core/src/ffi_wgridder.rs:749: this thread's last status-1 message into
 * buf (at most len bytes, not terminated); returns its full length. */
size_t rang_core_wgridder_last_error(uint8_t *buf, size_t len);

/* --- No counterpart in the reference: device-resident buffers. ----------
 * As the _ex2 / dirty2vis_ex entry points, with every array pointer a
 * device pointer of the current HIP device, run on `stream` (a
 * hipStream_t, NULL for the null stream); `info` is host memory. The call
 * synchronises `stream` before returning. */
int32_t rang_amd_wgridder_vis2dirty_dev_f64(const WgridderOptions *opts, size_t nrow,
                                            size_t nchan, const double *uvw, const double *freq,
                                            const void *vis, int32_t vis_kind, const void *wgt,
                                            const void *wgt2, int32_t wgt_kind,
                                            const uint8_t *mask, double *dirty, double *psf,
                                            double *info, void *stream);
int32_t rang_amd_wgridder_vis2dirty_dev_f32(const WgridderOptions *opts, size_t nrow,
                                            size_t nchan, const double *uvw, const double *freq,
                                            const void *vis, int32_t vis_kind, const void *wgt,
                                            const void *wgt2, int32_t wgt_kind,
                                            const uint8_t *mask, float *dirty, float *psf,
                                            double *info, void *stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): as
 * rang_amd_wgridder_vis2dirty_dev_f32, but every device buffer of the call
 * (sorted records, bucket counts, tables, w-planes, transforms) is carved
 * from one fixed block of RANG_WGRIDDER_NOALLOC_MAX_BYTES per GPU, held from
 * the library's load until the process ends; the call allocates no device
 * or pinned host memory. A call that needs more returns 1 with a message.
 * Each call rewrites the block (nothing carries over); calls on one GPU are
 * serialised by a lock and each returns after its stream is synchronised. */
#define RANG_WGRIDDER_NOALLOC_MAX_BYTES (256u << 20)
int32_t rang_amd_wgridder_vis2dirty_noalloc_f32(const WgridderOptions *opts, size_t nrow,
                                                size_t nchan, const double *uvw,
                                                const double *freq, const void *vis,
                                                int32_t vis_kind, const void *wgt,
                                                const void *wgt2, int32_t wgt_kind,
                                                const uint8_t *mask, float *dirty, float *psf,
                                                double *info, void *stream);
int32_t rang_amd_wgridder_dirty2vis_dev_f64(const WgridderOptions *opts, size_t nrow,
                                            size_t nchan, const double *uvw, const double *freq,
                                            const double *dirty, const void *wgt,
                                            int32_t wgt_kind, const uint8_t *mask, double *vis,
                                            double *info, void *stream);
int32_t rang_amd_wgridder_dirty2vis_dev_f32(const WgridderOptions *opts, size_t nrow,
                                            size_t nchan, const double *uvw, const double *freq,
                                            const float *dirty, const void *wgt, int32_t wgt_kind,
                                            const uint8_t *mask, float *vis, double *info,
                                            void *stream);
/* --- Plans without gridding: the reference's Rust-side wgridder::plan and
 * gridding_footprint (core/src/wgridder.rs:3525-3586), which have no C entry
 * point there, on host buffers. The scan runs on the GPU. ------------------
 * plan: the 14 info doubles the call with these arrays would report (all zero
 * with no active visibility); vis kind unit plans the PSF / degridding pass.
 * footprint: the plan of the gridding pass on float64 weights for nout (1 or
 * 2) images and the device bytes that pass holds in this implementation. */
int32_t rang_amd_wgridder_plan_f64(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                   const double *uvw, const double *freq, const void *vis,
                                   int32_t vis_kind, const void *wgt, int32_t wgt_kind,
                                   const uint8_t *mask, double *info);
int32_t rang_amd_wgridder_plan_f32(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                   const double *uvw, const double *freq, const void *vis,
                                   int32_t vis_kind, const void *wgt, int32_t wgt_kind,
                                   const uint8_t *mask, double *info);
int32_t rang_amd_wgridder_footprint_f64(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                        const double *uvw, const double *freq, const double *wgt,
                                        size_t nout, double *info, size_t *bytes);
int32_t rang_amd_wgridder_footprint_f32(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                        const double *uvw, const double *freq, const double *wgt,
                                        size_t nout, double *info, size_t *bytes);

/* --- Prepared plans: the reference's Rust-side Prepared (core/src/wgridder.rs:
 * 3165-3499), which has no C entry point there, on device buffers. ---------
 * A plan and both sorted layouts of one set of visibilities (uvw nrow x 3,
 * weights nrow x nchan of kind none / f32 / f64, all device pointers; freq a
 * device pointer too), kept on the GPU until _free. The active set is the
 * nonzero weights (Prepared::new, :3182-3228). Returns NULL on failure, with
 * the message for rang_core_wgridder_last_error. Calls on one plan are
 * serialised. */
typedef struct RangPrepared RangPrepared;
RangPrepared *rang_amd_wgridder_prepare_dev_f64(const WgridderOptions *opts, size_t nrow,
                                                size_t nchan, const double *uvw,
                                                const double *freq, const void *wgt,
                                                int32_t wgt_kind, void *stream);
RangPrepared *rang_amd_wgridder_prepare_dev_f32(const WgridderOptions *opts, size_t nrow,
                                                size_t nchan, const double *uvw,
                                                const double *freq, const void *wgt,
                                                int32_t wgt_kind, void *stream);
/* Prepared::vis2dirty (:3432-3456): the image of `vis` (kind c64 / c128, nrow x
 * nchan) with the prepared weights; info as the gridder's. */
int32_t rang_amd_wgridder_prepared_vis2dirty_f64(RangPrepared *p, const void *vis,
                                                 int32_t vis_kind, double *dirty, double *info,
                                                 void *stream);
int32_t rang_amd_wgridder_prepared_vis2dirty_f32(RangPrepared *p, const void *vis,
                                                 int32_t vis_kind, float *dirty, double *info,
                                                 void *stream);
/* Prepared::dirty2vis (:3458-3486): complex visibilities (interleaved re, im)
 * of the image, times the prepared weights when `weighted`; zero where the
 * weight is zero. */
int32_t rang_amd_wgridder_prepared_dirty2vis_f64(RangPrepared *p, const double *dirty,
                                                 int32_t weighted, double *vis, void *stream);
int32_t rang_amd_wgridder_prepared_dirty2vis_f32(RangPrepared *p, const float *dirty,
                                                 int32_t weighted, float *vis, void *stream);
/* Prepared::residual (:3230-3255, 3300-3383): the image of vis - A model (model
 * NULL: of vis) with the prepared weights, and misfit = sum w |vis - A model|^2
 * (a parallel sum in double, not the reference's serial order). The
 * prediction and residual are formed in the degridder and handed to the
 * gridder in one value per active visibility; no visibility-sized buffer. */
int32_t rang_amd_wgridder_prepared_residual_f64(RangPrepared *p, const void *vis,
                                                int32_t vis_kind, const double *model,
                                                double *dirty, double *misfit, void *stream);
int32_t rang_amd_wgridder_prepared_residual_f32(RangPrepared *p, const void *vis,
                                                int32_t vis_kind, const float *model,
                                                float *dirty, double *misfit, void *stream);
/* Prepared::residual_sorted (:3257-3267): as _residual with the data held only
 * for the active visibilities, in the order _sorted gives (the prepared
 * plan's gridding order); no visibility-sized array is read or held. */
int32_t rang_amd_wgridder_prepared_residual_sorted_f64(RangPrepared *p, const void *vis,
                                                       int32_t vis_kind, const double *model,
                                                       double *dirty, double *misfit,
                                                       void *stream);
int32_t rang_amd_wgridder_prepared_residual_sorted_f32(RangPrepared *p, const void *vis,
                                                       int32_t vis_kind, const float *model,
                                                       float *dirty, double *misfit, void *stream);
/* Prepared::sorted (:3269-3286): `dense` (nrow x nchan elements of elem_bytes
 * = 4, 8 or 16, device memory) at the active visibilities, in that order,
 * into `out` (active elements). */
int32_t rang_amd_wgridder_prepared_sorted_f64(RangPrepared *p, const void *dense, size_t elem_bytes,
                                              void *out, void *stream);
int32_t rang_amd_wgridder_prepared_sorted_f32(RangPrepared *p, const void *dense, size_t elem_bytes,
                                              void *out, void *stream);
/* The normal operator A^H W A x of the prepared plan (WgridMeasurement::
 * gram_apply, core/src/clean_certified.rs:233-236): out = vis2dirty(dirty2vis(x)). */
int32_t rang_amd_wgridder_prepared_gram_f64(RangPrepared *p, const double *x, double *out,
                                            void *stream);
int32_t rang_amd_wgridder_prepared_gram_f32(RangPrepared *p, const float *x, float *out,
                                            void *stream);
/* clean::conjugate_gradients (core/src/clean.rs:263-306) on A^H W A x = rhs from
 * zero: at most `iterations` Gram applications, stop at |r| <= tolerance |rhs|
 * or <p, A p> <= 0; *used applications, *rel = |r| / |rhs|. */
int32_t rang_amd_wgridder_prepared_cg_f64(RangPrepared *p, const double *rhs, double *x,
                                          int32_t iterations, double tolerance, int32_t *used,
                                          double *rel, void *stream);
int32_t rang_amd_wgridder_prepared_cg_f32(RangPrepared *p, const float *rhs, float *x,
                                          int32_t iterations, double tolerance, int32_t *used,
                                          double *rel, void *stream);
/* prox_solver::solve (core/src/prox_solver.rs:283-445), FISTA or forward-
 * backward for 1/2 |y - Phi m|^2_W + R(m) on the prepared plan's normal
 * operator (weights as prepared, normally normalised to sum to one), with
 * the L1 prior R(x) = sum lambda factor_i |x_i| (L1Prior, :55-132; factors
 * NULL: all one) and x >= 0 unless `signed_`. `dirty` = Phi^H W y and q0 =
 * 1/2 |y|^2_W (Problem, :167-172); x holds the start on entry and the result
 * on return, gx (may be NULL) G x. Settings as the reference's (:175-208;
 * a NaN tolerance is None; objective_every does not apply to this prior).
 * One Gram application per iteration; records[k] (up to nrecords) is the
 * reference's Record of iterate k. Image vectors nx x ny on the device. */
typedef struct RangFistaSettings {
    double step;
    uint64_t max_iterations;
    int32_t momentum;
    int32_t restart;
    double gap_tolerance;
    double kkt_tolerance;
    double step_tolerance;
} RangFistaSettings;
typedef struct RangFistaRecord {
    uint64_t iteration;
    double objective;
    double misfit;
    double penalty;
    double gap;
    double kkt;
    uint64_t nonzero;
    double flux;
    double curvature;
    double relative_step;
    int32_t restarted;
    int32_t reserved;
} RangFistaRecord;
typedef struct RangFistaOutcome {
    uint64_t iterations;
    int32_t converged; /* stop: 1 "converged", 0 "iterations" */
    int32_t reserved;
    uint64_t restarts;
    uint64_t objective_rises;
    double largest_rise;
    uint64_t negative_curvatures;
    double min_curvature;
    double max_curvature;
    RangFistaRecord last;
} RangFistaOutcome;
int32_t rang_amd_wgridder_prepared_fista_f64(RangPrepared *p, const double *dirty, double q0,
                                             double lambda, const double *factors, int32_t signed_,
                                             const RangFistaSettings *settings, double *x,
                                             double *gx, RangFistaRecord *records,
                                             size_t nrecords, RangFistaOutcome *out, void *stream);
int32_t rang_amd_wgridder_prepared_fista_f32(RangPrepared *p, const float *dirty, double q0,
                                             double lambda, const float *factors, int32_t signed_,
                                             const RangFistaSettings *settings, float *x,
                                             float *gx, RangFistaRecord *records, size_t nrecords,
                                             RangFistaOutcome *out, void *stream);
/* prox_solver::OuterSchedule (:760-772): at most `cycles` major cycles, stop
 * when the exact optimality residual at `lambda` is <= kkt; a cycle solves at
 * max(lambda, gamma * peak residual). */
typedef struct RangOuterSchedule {
    uint64_t cycles;
    double kkt;
    double lambda;
    double gamma;
} RangOuterSchedule;
/* prox_solver::OuterRecord (:447-462): one major cycle, judged exactly. */
typedef struct RangOuterRecord {
    uint64_t outer;
    double objective;
    double misfit;
    double gap;
    double kkt;
    uint64_t inner_iterations; /* inner iterations summed before this cycle */
    double defect_sup;         /* max |G x - G~ x| */
} RangOuterRecord;
/* How defect_correction ended (prox_solver.rs:778-875): `converged` 1 for
 * "converged", 0 for "cycles"; the inner iterations summed; the exact and
 * cheap operator applications made. */
typedef struct RangDefectOutcome {
    uint64_t iterations;
    int32_t converged;
    int32_t reserved;
    uint64_t cycles;
    uint64_t restarts;
    uint64_t negative_curvatures;
    uint64_t exact_applies;
    uint64_t cheap_applies;
    RangOuterRecord last;
} RangDefectOutcome;
/* prox_solver::power_iteration (:886-908): the largest eigenvalue of the normal
 * operator from the reference's seeded start (xorshift64, :871-882), stopping
 * at a relative change <= tolerance; *used applications. */
int32_t rang_amd_wgridder_prepared_power_f64(RangPrepared *p, uint64_t seed, double tolerance,
                                             uint64_t max_iterations, double *value,
                                             uint64_t *used, void *stream);
int32_t rang_amd_wgridder_prepared_power_f32(RangPrepared *p, uint64_t seed, double tolerance,
                                             uint64_t max_iterations, double *value,
                                             uint64_t *used, void *stream);
/* Prepared::active (:3257-3262), memory (:3386-3418, total device bytes),
 * release_grids (:3420-3429), and the plan's release. */
size_t rang_amd_wgridder_prepared_active(const RangPrepared *p);
size_t rang_amd_wgridder_prepared_memory(const RangPrepared *p);
void rang_amd_wgridder_prepared_release_grids(RangPrepared *p);
void rang_amd_wgridder_prepared_free(RangPrepared *p);

/* No counterpart in the reference: frees the device memory the library
 * keeps between calls on every device (its per-device workspaces and
 * sub-band batch pools; not the fixed block of the allocation-free entry
 * point, nor the buffers of live prepared plans). Waits for the devices'
 * work first; the next call allocates again. */
int32_t rang_amd_release_workspaces(void);

/* Per-stage timing (HIP events on the call's stream) of this thread's later
 * calls: on != 0 enables it; count, names and milliseconds of the last call. */
void rang_amd_stage_timing(int32_t on);
int32_t rang_amd_stage_count(void);
const char *rang_amd_stage_name(int32_t i);
double rang_amd_stage_ms(int32_t i);

#ifdef __cplusplus
}
#endif

#endif
