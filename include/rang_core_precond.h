/* The learned preconditioner's kernels: a BF16 U-Net (FP32 accumulation on
 * the matrix cores) for inference and training, random training images and
 * the vector kernels of flexible conjugate gradients, all on device buffers.
 *
 * None of these has a counterpart in the reference (rang-core has no learned
 * preconditioner); they serve the preconditioned solve of the normal
 * equations A^H W A x = A^H W v, which the reference solves with plain
 * conjugate gradients (core/src/clean.rs:263-306) and whose residual its exact
 * Gram operator judges (core/src/clean_certified.rs:233-236).
 *
 * Tensors are NHWC: B images of H x W pixels, pixel p = (n H + h) W + w,
 * channel c at p * ld + off + c. BF16 values are uint16_t bit patterns. Channel
 * counts of convolutions are multiples of 16; strides and offsets of BF16
 * inputs multiples of 4. Weights are [cout][k * k][cin]. All calls are
 * asynchronous on `stream` (a hipStream_t); status 0 ok, 1 bad arguments, 2 a
 * HIP failure, with the message for rang_amd_nn_last_error. */
#ifndef RANG_CORE_PRECOND_H
#define RANG_CORE_PRECOND_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

size_t rang_amd_nn_last_error(uint8_t *buf, size_t len);

/* y = conv_k(x, w) (k = ks, 1 or 3, zero padding): FP32 into y and BF16
 * relu(y) (or y when relu is 0) into a; either may be NULL. */
int32_t rang_amd_nn_conv_fwd(const uint16_t *x, int32_t ldx, int32_t xoff, const uint16_t *w, float *y,
                             uint16_t *a, int32_t ldy, int32_t yoff, int32_t B, int32_t H, int32_t W,
                             int32_t cin, int32_t cout, int32_t ks, int32_t relu, void *stream);
/* dw += the weight gradient of conv_k for output gradient dy (FP32 dw). */
int32_t rang_amd_nn_conv_wgrad(const uint16_t *x, int32_t ldx, int32_t xoff, const uint16_t *dy, int32_t lddy,
                               int32_t dyoff, float *dw, int32_t B, int32_t H, int32_t W, int32_t cin,
                               int32_t cout, int32_t ks, void *stream);
/* The data-gradient weights: wt[ci][k*k-1-tap][co] = bf16(w[co][tap][ci]), so
 * that conv_fwd(dy, wt) (cin and cout swapped) is the input gradient. */
int32_t rang_amd_nn_flip(const float *w, uint16_t *wt, int32_t cin, int32_t cout, int32_t ks, void *stream);
int32_t rang_amd_nn_to_bf16(const float *s, uint16_t *d, int64_t n, void *stream);
/* g = bf16(d * (pre > 0)) (pre NULL: no mask) for C channels of npix pixels. */
int32_t rang_amd_nn_relu_bwd(const float *d, int32_t ldd, int32_t doff, const float *pre, int32_t ldp,
                             int32_t poff, uint16_t *g, int32_t ldg, int32_t goff, int64_t npix, int32_t C,
                             void *stream);
/* 2 x 2 average pooling from H x W (even) and its gradient (added to gx). */
int32_t rang_amd_nn_pool(const uint16_t *x, int32_t ldx, int32_t xoff, uint16_t *y, int32_t ldy, int32_t yoff,
                         int32_t B, int32_t H, int32_t W, int32_t C, void *stream);
int32_t rang_amd_nn_pool_bwd(const float *gy, int32_t ldy, int32_t yoff, float *gx, int32_t ldx, int32_t xoff,
                             int32_t B, int32_t H, int32_t W, int32_t C, void *stream);
/* Nearest 2x upsampling to H x W and its gradient (written to gx). */
int32_t rang_amd_nn_up(const uint16_t *x, int32_t ldx, int32_t xoff, uint16_t *y, int32_t ldy, int32_t yoff,
                       int32_t B, int32_t H, int32_t W, int32_t C, void *stream);
int32_t rang_amd_nn_up_bwd(const float *gy, int32_t ldy, int32_t yoff, float *gx, int32_t ldx, int32_t xoff,
                           int32_t B, int32_t H, int32_t W, int32_t C, void *stream);
/* Adam step `step` (1, 2, ...) on FP32 weights, refreshing the BF16 copy wb
 * (may be NULL). */
int32_t rang_amd_nn_adam(float *w, const float *g, float *m, float *v, uint16_t *wb, int64_t n, double lr,
                         double b1, double b2, double eps, int64_t step, void *stream);
/* B images of hw values, times s[n] (s NULL: 1), into channel c of a BF16
 * tensor of ldx channels; and channel c of an FP32 tensor, times s[n], out. */
int32_t rang_amd_nn_pack_f64(const double *img, const double *s, uint16_t *x, int32_t ldx, int32_t c,
                             int32_t B, int64_t hw, void *stream);
int32_t rang_amd_nn_pack_f32(const float *img, const double *s, uint16_t *x, int32_t ldx, int32_t c,
                             int32_t B, int64_t hw, void *stream);
int32_t rang_amd_nn_unpack_f64(const float *y, int32_t ldy, int32_t c, const double *s, double *out,
                               int32_t B, int64_t hw, void *stream);
int32_t rang_amd_nn_unpack_f32(const float *y, int32_t ldy, int32_t c, const double *s, float *out,
                               int32_t B, int64_t hw, void *stream);
/* s[n] = sum a b over image n (b NULL: sum a^2), in double. */
int32_t rang_amd_nn_image_dot_f64(const double *a, const double *b, double *s, int32_t B, int64_t hw,
                                  void *stream);
int32_t rang_amd_nn_image_dot_f32(const float *a, const float *b, double *s, int32_t B, int64_t hw,
                                  void *stream);
/* loss[0] = sum_n inv[n] |y_n - t_n|^2 / 2 on channel c of y (ldy channels);
 * g (ldg channels) = the gradient on channel c, zero elsewhere. */
int32_t rang_amd_nn_loss(const float *y, int32_t ldy, int32_t c, const float *t, const double *inv, float *g,
                         int32_t ldg, int32_t B, int64_t hw, double *loss, void *stream);
/* B random images: K Gaussians each, log-uniform widths in [wmin, wmax]
 * pixels, amplitudes in [-1, 1] (positive: [0, 1]), from `seed`. */
int32_t rang_amd_nn_sky_f64(double *out, int32_t B, int32_t H, int32_t W, int32_t K, double wmin, double wmax,
                            int32_t positive, uint64_t seed, void *stream);
int32_t rang_amd_nn_sky_f32(float *out, int32_t B, int32_t H, int32_t W, int32_t K, double wmin, double wmax,
                            int32_t positive, uint64_t seed, void *stream);
/* out[n] = s[n] x[n]: B FP64 images of hw values scaled into FP32. */
int32_t rang_amd_nn_scale_images(const double *x, const double *s, float *out, int32_t B, int64_t hw,
                                 void *stream);
/* An output gradient from B FP64 images: channel c of g (ldg channels) =
 * s[n] img[n], the other channels zero. */
int32_t rang_amd_nn_grad_images(const double *img, const double *s, float *g, int32_t ldg, int32_t c,
                                int32_t B, int64_t hw, void *stream);
/* y = a x + b y. */
int32_t rang_amd_nn_axpby_f64(const double *x, double *y, double a, double b, int64_t n, void *stream);
int32_t rang_amd_nn_axpby_f32(const float *x, float *y, double a, double b, int64_t n, void *stream);

#ifdef __cplusplus
}
#endif

#endif
