/*
 * C ABI of the MI355X w-stacking gridder: the entry points of rang-core's
 * core/src/ffi_grid.rs and core/src/gpu/ffi.rs (WSClean 3.6's inversion,
 * core/src/grid.rs), same structs, same status codes, run on the GPU.
 */
#ifndef RANG_CORE_GRID_H
#define RANG_CORE_GRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* core/src/ffi_grid.rs:7-29: borrowed host inputs, lengths explicit. */
typedef struct {
    size_t nrow, nchan, width, height, kernel_size, oversampling, n_w_layers;
    double pixel_size_x, pixel_size_y; /* radians */
    double padding;                    /* 1.2 as WSClean's default */
    const double *uvw;                 /* nrow * 3 metres */
    const double *frequencies;         /* nchan Hz */
    const float *data;                 /* nrow * nchan * 2 interleaved complex */
    const float *weights;              /* nrow * nchan visibility weights */
    const float *imaging_weights;      /* nrow * nchan imaging weights */
} GridInputs;

/* core/src/ffi_grid.rs:31-45: caller-owned outputs; images width * height
 * doubles in [x + y * width] order; stats eight doubles: image weight,
 * gridded count, visibility weight sum, largest gridded weight, min w, max w,
 * PSF normalisation, number of w-layers. */
typedef struct {
    double *dirty;
    double *psf;
    double *stats;
} GridOutputs;

#define RANG_GRID_N_STATS 8
#define RANG_GRID_N_TIMINGS 14

/* core/src/ffi_grid.rs:54-145: invert to a dirty image and PSF; 0 ok, 1
 * invalid arguments, 2 a failure inside. Runs the double-precision GPU path,
 * with transforms of any padded side as the reference's rustfft takes them
 * (sides with prime factors above 11 by Bluestein's chirp transform, sides
 * beyond one LDS in four steps). */
int32_t rang_core_grid_invert(const GridInputs *inputs, GridOutputs *outputs);

/* core/src/ffi_grid.rs:154-170: the oversampled Kaiser-Bessel kernel,
 * kernel_size * oversampling doubles; 0 ok, 1 invalid arguments. */
int32_t rang_core_grid_kernel(size_t kernel_size, size_t oversampling, double *kernel);

/* core/src/gpu/ffi.rs:104-137: precision 0 double, 1 single; timings 14
 * doubles in the order of Timings::as_array (core/src/gpu/invert.rs:57-74):
 * context, compile, prepare, upload, compute, download, finish seconds;
 * device bytes, batches, layers per batch; grid, fft, project, bucket
 * seconds. 0 ok, 1 invalid arguments, 2 a failure inside, 3 the GPU path
 * failed. As the reference's GPU path, padded sides with a prime factor
 * above 11 are refused (3). */
int32_t rang_core_grid_invert_gpu(const GridInputs *inputs, GridOutputs *outputs, int32_t precision,
                                  double *timings);

/* No counterpart in the reference: the message of this thread's last
 * nonzero status into buf (at most len bytes, not terminated); returns its
 * full length. */
size_t rang_amd_grid_last_error(uint8_t *buf, size_t len);
/* No counterpart in the reference (grid_gpu_adapter's string): the current
 * HIP device's name and architecture into buf; returns the full length, 0
 * without a device. */
size_t rang_amd_device_name(uint8_t *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif
