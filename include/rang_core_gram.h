/*
 * The normal operator without visibilities, MI355X build: rang-core's
 * Rust-side gridder_gram::{WExpansion, GramOperator} (core/src/gridder_gram.rs),
 * which has no C entry point there, on device buffers, and the preconditioned
 * conjugate gradients that use its circulant inverse.
 *
 * Status codes and messages as include/rang_core_wgridder.h (0 ok, 1 invalid
 * arguments, 2 device failure; rang_core_wgridder_last_error holds the
 * message). Image vectors are n x n device arrays, pixel (i, j) at i * n + j.
 */
#ifndef RANG_CORE_GRAM_H
#define RANG_CORE_GRAM_H

#include <stddef.h>
#include <stdint.h>

#include "rang_core_wgridder.h"

#ifdef __cplusplus
extern "C" {
#endif

/* WExpansion::new (core/src/gridder_gram.rs:84-141) on the host: the fit of
 * `stencil` (even) nodes at phase step theta, regularised by ridge. Writes
 * its measured sup_error and lebesgue constant (:143-163) and, when `table`
 * is not NULL, the (8192 + 1) * stencil coefficient table. No GPU needed. */
#define RANG_WEXP_BINS 8192
int32_t rang_amd_wexpansion(int32_t stencil, double theta, double ridge, double *sup_error,
                            double *lebesgue, double *table);

typedef struct RangGram RangGram;
/* GramOperator::new (:226-407): the w-aware operator of the measurement with
 * weights `wgt` (nrow x nchan, kind f32 or f64, used as given) from the
 * expansion (stencil, theta, ridge), each node gridded at `epsilon`. uvw,
 * freq and wgt are device pointers; the options give the square even image
 * (no shift, no divide_by_n), the flips and the cell. NULL on failure. */
RangGram *rang_amd_gram_new_f64(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                const double *uvw, const double *freq, const void *wgt,
                                int32_t wgt_kind, int32_t stencil, double theta, double ridge,
                                double epsilon, void *stream);
RangGram *rang_amd_gram_new_f32(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                const double *uvw, const double *freq, const void *wgt,
                                int32_t wgt_kind, int32_t stencil, double theta, double ridge,
                                double epsilon, void *stream);
/* GramOperator::no_w (:414-478): one node, the translation-invariant PSF. */
RangGram *rang_amd_gram_no_w_f64(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                 const double *uvw, const double *freq, const void *wgt,
                                 int32_t wgt_kind, double epsilon, void *stream);
RangGram *rang_amd_gram_no_w_f32(const WgridderOptions *opts, size_t nrow, size_t nchan,
                                 const double *uvw, const double *freq, const void *wgt,
                                 int32_t wgt_kind, double epsilon, void *stream);
/* GramOperator::inverse_filter (:488-509): the circulant operator whose
 * spectrum is 1 / (max(K, 0) + rho) of a one-node operator's K; a new
 * operator of the same precision. */
RangGram *rang_amd_gram_inverse_filter(const RangGram *g, double rho, void *stream);
/* GramOperator::apply (:519-647): out = G x. */
int32_t rang_amd_gram_apply_f64(RangGram *g, const double *x, double *out, void *stream);
int32_t rang_amd_gram_apply_f32(RangGram *g, const float *x, float *out, void *stream);
/* Stats (:189-214): planes, stencil, theta, wmin, wmax, x_range, sup_error,
 * lebesgue, active_pairs, cube_bytes, hermitian_check, spectral_sparsity;
 * then the largest |K^| over the nodes (the scale for an inverse filter's
 * rho). */
#define RANG_GRAM_N_STATS 13
int32_t rang_amd_gram_stats(const RangGram *g, double *out);
/* The operator's image side (0 for NULL). */
size_t rang_amd_gram_side(const RangGram *g);
void rang_amd_gram_free(RangGram *g);

/* Conjugate gradients on the prepared plan's exact normal operator (as
 * rang_amd_wgridder_prepared_cg_*) preconditioned by the operator `m`
 * (a symmetric positive definite one, such as an inverse filter): the same
 * stopping rule on |r| / |rhs|, so the Gram applications compare. */
int32_t rang_amd_wgridder_prepared_pcg_f64(RangPrepared *p, RangGram *m, const double *rhs,
                                           double *x, int32_t iterations, double tolerance,
                                           int32_t *used, double *rel, void *stream);
int32_t rang_amd_wgridder_prepared_pcg_f32(RangPrepared *p, RangGram *m, const float *rhs,
                                           float *x, int32_t iterations, double tolerance,
                                           int32_t *used, double *rel, void *stream);

/* prox_solver::defect_correction (core/src/prox_solver.rs:778-875): major
 * cycles of the L1-regularised problem (prior lambda * factor_i, x >= 0
 * unless signed_) on the prepared plan's exact normal operator, each inner
 * FISTA solve (settings `inner`; its gap tolerance is not used) on the cheap
 * operator `cheap` with the exact operator's defect at the cycle's iterate
 * moved into the data. dirty = Phi^H W y and q0 = 1/2 |y|^2_W as for
 * rang_amd_wgridder_prepared_fista_*; x holds the start and receives the last
 * judged iterate, gx (may be NULL) its exact G x; records (may be NULL) one
 * per cycle, at most nrecords. The cheap operator's side must equal the
 * plan's square image. */
int32_t rang_amd_wgridder_prepared_defect_f64(RangPrepared *p, RangGram *cheap, const double *dirty,
                                              double q0, const double *factors, int32_t signed_,
                                              const RangFistaSettings *inner,
                                              const RangOuterSchedule *schedule, double *x, double *gx,
                                              RangOuterRecord *records, size_t nrecords,
                                              RangDefectOutcome *out, void *stream);
int32_t rang_amd_wgridder_prepared_defect_f32(RangPrepared *p, RangGram *cheap, const float *dirty,
                                              double q0, const float *factors, int32_t signed_,
                                              const RangFistaSettings *inner,
                                              const RangOuterSchedule *schedule, float *x, float *gx,
                                              RangOuterRecord *records, size_t nrecords,
                                              RangDefectOutcome *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif
