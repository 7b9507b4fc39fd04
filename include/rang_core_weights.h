/*
 * C ABI of the WSClean imaging weights and Stokes I formation, MI355X build:
 * the drop-in replacement of rang-core's core/src/ffi_weights.rs (identical
 * names, struct layouts and status codes: 0 ok, 1 invalid arguments, 2 a
 * failure inside the library). The rang_core_* entry points take host
 * buffers, stage them to the GPU and return complete outputs; the rang_amd_*
 * variants take device pointers of the current HIP device and run on the
 * caller's stream (synchronised before returning).
 *
 * On the GPU the weight grid is summed with double atomics, so a cell's sum
 * may differ from the reference's serial sum in its last bits; the cells
 * each sample reaches, the finish formulas and the lookup are the
 * reference's.
 */
#ifndef RANG_CORE_WEIGHTS_H
#define RANG_CORE_WEIGHTS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* core/src/ffi_weights.rs:9-27 */
typedef struct StokesIInputs {
    size_t ncell;
    const float *data;     /* ncell * 4 interleaved complex floats (XX, XY, YX, YY) */
    const float *weights;  /* ncell * 4 */
    const uint8_t *flags;  /* ncell * 4, non-zero when flagged */
} StokesIInputs;
typedef struct StokesIOutputs {
    float *vis;    /* ncell interleaved complex floats */
    float *weight; /* ncell */
} StokesIOutputs;

/* core/src/ffi_weights.rs:29-69; class 0 natural, 1 uniform, 2 Briggs. */
typedef struct ImagingWeightInputs {
    size_t nrow;
    size_t nchan;
    const double *uvw;        /* nrow * 3 metres */
    const double *wavelength; /* nchan metres */
    const float *weights;     /* nrow * nchan Stokes I visibility weights */
    const uint8_t *row_selected; /* nrow */
    uint32_t class_;
    double robust;
    size_t width;
    size_t height;
    double pixel_scale_x;
    double pixel_scale_y;
    double super_weight;
    double rank_filter_level;
    size_t rank_filter_size;
} ImagingWeightInputs;
typedef struct ImagingWeightOutputs {
    float *weights; /* nrow * nchan */
    double *grid;   /* grid_width * grid_height / 2, or NULL */
    double *sums;   /* visibility weight sum, total weight, max gridded weight, count; or NULL */
} ImagingWeightOutputs;

/* core/src/ffi_weights.rs:97-117 */
int32_t rang_core_weight_grid_shape(size_t width, size_t height, double super_weight, size_t *shape);
/* core/src/ffi_weights.rs:119-124 */
size_t rang_core_padded_size(size_t trimmed, double padding);
/* core/src/ffi_weights.rs:126-158 */
int32_t rang_core_stokes_i(const StokesIInputs *inputs, StokesIOutputs *outputs);
/* core/src/ffi_weights.rs:160-230 */
int32_t rang_core_imaging_weights(const ImagingWeightInputs *inputs, ImagingWeightOutputs *outputs);

/* No counterpart in the reference: the same on device buffers and `stream`. */
int32_t rang_amd_stokes_i_dev(const StokesIInputs *inputs, StokesIOutputs *outputs, void *stream);
int32_t rang_amd_imaging_weights_dev(const ImagingWeightInputs *inputs, ImagingWeightOutputs *outputs,
                                     void *stream);

#ifdef __cplusplus
}
#endif

#endif
