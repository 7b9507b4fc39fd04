/*
 * Sub-band batches of the exact w-gridder, MI355X build (no counterpart in
 * the reference, which grids one image per call: core/src/wgridder.rs:2430-2696
 * per image, core/src/ffi_wgridder.rs:155-227 per call).
 *
 * A batch is a list of independent images, each with its own options, rows,
 * channels, weights and flags; vis2dirty of all of them runs with the
 * gridding of every image in one persistent launch (per group of images
 * whose buffers fit `budget_bytes`; the largest images go first), each
 * image's result equal to what rang_amd_wgridder_vis2dirty_dev_* returns for
 * it alone. All arrays are device pointers of the current HIP device except
 * `freq` (host, nchan doubles: the plan is made on the host). Status codes
 * and messages as include/rang_core_wgridder.h; a refusal names the image.
 *
 * The tuner times one image's vis2dirty over the cost model's cheapest
 * kernel rows (one per support, each reaching the image's epsilon) and over
 * gridding tile sides, and writes the fastest into the image (kernel_index,
 * nu, nv in its options, tile); choices are kept per signature for the
 * process and, when `cache_path` is not NULL, appended to that text file
 * (one line per signature) and read back by later processes.
 */
#ifndef RANG_CORE_BATCH_H
#define RANG_CORE_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "rang_core_wgridder.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RangBatchImage {
    WgridderOptions opts;
    size_t nrow;
    size_t nchan;
    const double *uvw;   /* nrow * 3 */
    const double *freq;  /* host */
    const void *vis;     /* nrow * nchan complex of vis_kind */
    const void *wgt;     /* nrow * nchan of wgt_kind, or NULL */
    const uint8_t *mask; /* nrow * nchan, or NULL */
    void *dirty;         /* nx * ny of the precision */
    int32_t vis_kind;    /* WGRIDDER_VIS_C64 / C128 */
    int32_t wgt_kind;    /* WGRIDDER_WGT_* */
    int32_t tile;        /* gridding tile side to aim for; 0 the default */
    int32_t reserved;
    double info[RANG_WGRIDDER_N_INFO]; /* out: the image's plan */
} RangBatchImage;

int32_t rang_amd_wgridder_batch_vis2dirty_f32(RangBatchImage *images, size_t n, size_t budget_bytes,
                                              void *stream);
int32_t rang_amd_wgridder_batch_vis2dirty_f64(RangBatchImage *images, size_t n, size_t budget_bytes,
                                              void *stream);

typedef struct RangTuneReport {
    double ms;         /* the chosen plan's vis2dirty time */
    double default_ms; /* the cost model's own choice's time (0 when cached) */
    int32_t cached;    /* 1 when the signature was already known */
    int32_t reserved;
} RangTuneReport;

/* `kernels`: how many of the cheapest modelled kernel rows to time (>= 1). */
int32_t rang_amd_wgridder_tune_f32(RangBatchImage *image, const char *cache_path, int32_t kernels,
                                   RangTuneReport *report, void *stream);
int32_t rang_amd_wgridder_tune_f64(RangBatchImage *image, const char *cache_path, int32_t kernels,
                                   RangTuneReport *report, void *stream);

#ifdef __cplusplus
}
#endif

#endif
