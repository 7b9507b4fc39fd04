// Kernels of the learned preconditioner: a small U-Net in BF16 with FP32
// accumulation on the matrix cores (gfx950), for inference and training, and
// the vector kernels of flexible conjugate gradients. The reference has no
// counterpart; the network's graph is assembled by the caller over the C ABI
// of include/rang_core_precond.h.
//
// Tensors are NHWC: pixel p = (n H + h) W + w, channel c at p * ld + off + c,
// so a concatenation along channels is two writers into one buffer with
// different offsets. Activations and gradients that feed the matrix cores
// are BF16 (round to nearest even); pre-activations, gradients being summed
// and the master weights are FP32.
//
// A 3 x 3 (or 1 x 1) convolution is an implicit GEMM: pixels x (taps x
// input channels) times (taps x input channels) x output channels. One wave
// computes 16 pixels x 32 output channels with v_mfma_f32_16x16x16_bf16:
// A row = pixel (lane mod 16), k = 4 consecutive input channels per lane
// group (lane / 16); B column = output channel, the same k split; D column =
// output channel (lane mod 16), rows 4 (lane / 16) + t. Channel counts are
// multiples of 16 (callers pad the first and last layers).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/rang_core_precond.h"

namespace rang {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

thread_local std::string g_nn_err;

int nn_fail(hipError_t e) {
    if (e == hipSuccess) return 0;
    g_nn_err = hipGetErrorString(e);
    return 2;
}

__device__ __forceinline__ std::uint16_t to_bf16(float f) {
    std::uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return std::uint16_t((u >> 16) | 0x40u);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return std::uint16_t(u >> 16);
}

__device__ __forceinline__ float from_bf16(std::uint16_t b) { return __uint_as_float(std::uint32_t(b) << 16); }

// Four consecutive BF16 values (8 bytes, aligned when the channel offset and
// stride are multiples of 4) or zeros.
__device__ __forceinline__ bf16x4 load4(const std::uint16_t* p, bool ok) {
    if (!ok) return bf16x4{0, 0, 0, 0};
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = short(v.x & 0xffffu);
    r[1] = short(v.x >> 16);
    r[2] = short(v.y & 0xffffu);
    r[3] = short(v.y >> 16);
    return r;
}

// y = conv(x, w) over B images of H x W: w as [cout][ks * ks][cin] BF16; the
// FP32 result into y (pre-activation, may be null) and BF16 relu(y) (or y
// when relu is 0) into a (may be null). Block: 4 waves, 64 pixels x 32
// output channels.
template <int KS>
__global__ __launch_bounds__(256) void k_conv_fwd(const std::uint16_t* __restrict__ x, int ldx, int xoff,
                                                  const std::uint16_t* __restrict__ w, float* __restrict__ y,
                                                  std::uint16_t* __restrict__ a, int ldy, int yoff, int B,
                                                  int H, int W, int cin, int cout, int relu) {
    const long long npix = (long long)B * H * W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long p0 = (long long)blockIdx.x * 64 + wave * 16;
    const int co0 = blockIdx.y * 32;
    const int row = lane & 15, q = lane >> 4;
    // this lane's A pixel
    const long long pa = p0 + row;
    int n = 0, h = 0, wc = 0;
    if (pa < npix) {
        n = int(pa / ((long long)H * W));
        const int r = int(pa - (long long)n * H * W);
        h = r / W;
        wc = r - h * W;
    }
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    constexpr int T = KS * KS;
    for (int tap = 0; tap < T; ++tap) {
        const int dy = KS == 3 ? tap / 3 - 1 : 0, dx = KS == 3 ? tap % 3 - 1 : 0;
        const int hh = h + dy, ww = wc + dx;
        const bool in = pa < npix && hh >= 0 && hh < H && ww >= 0 && ww < W;
        const long long src = ((long long)n * H + hh) * W + ww;
        const std::uint16_t* xa = x + src * ldx + xoff + 4 * q;
        const std::uint16_t* wb0 = w + ((long long)(co0 + row) * T + tap) * cin + 4 * q;
        const std::uint16_t* wb1 = w + ((long long)(co0 + 16 + row) * T + tap) * cin + 4 * q;
        for (int c0 = 0; c0 < cin; c0 += 16) {
            const bf16x4 av = load4(xa + c0, in);
            const bf16x4 b0 = load4(wb0 + c0, true);
            const bf16x4 b1 = load4(wb1 + c0, co0 + 16 < cout);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, b1, acc1, 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const long long p = p0 + 4 * q + t;
        if (p >= npix) continue;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int co = co0 + 16 * half + row;
            if (co >= cout) continue;
            const float v = half ? acc1[t] : acc0[t];
            const long long at = p * ldy + yoff + co;
            if (y) y[at] = v;
            if (a) a[at] = to_bf16(relu ? fmaxf(v, 0.f) : v);
        }
    }
}

// dw += sum over pixels of dy (x) the shifted input, dw as [cout][ks * ks][cin]
// FP32: a GEMM over pixels (16 per MFMA) for one (16 cout, tap, 16 cin)
// tile per wave; each block sums a chunk of pixels and adds its tile.
template <int KS>
__global__ __launch_bounds__(256) void k_conv_wgrad(const std::uint16_t* __restrict__ x, int ldx, int xoff,
                                                    const std::uint16_t* __restrict__ dy, int lddy, int dyoff,
                                                    float* __restrict__ dw, int B, int H, int W, int cin,
                                                    int cout, long long chunk) {
    constexpr int T = KS * KS;
    const long long npix = (long long)B * H * W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nct = cin / 16, ncot = cout / 16;
    const int tile = blockIdx.x * 4 + wave;  // (cout tile, tap, cin tile)
    if (tile >= ncot * T * nct) return;
    const int cot = tile / (T * nct), rem = tile - cot * T * nct;
    const int tap = rem / nct, ct = rem - tap * nct;
    const int ddy = KS == 3 ? tap / 3 - 1 : 0, ddx = KS == 3 ? tap % 3 - 1 : 0;
    const int row = lane & 15, q = lane >> 4;
    const long long pbeg = (long long)blockIdx.y * chunk, pend = pbeg + chunk < npix ? pbeg + chunk : npix;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long long pb = pbeg; pb < pend; pb += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const long long p = pb + 4 * q + t;
            short va = 0, vb = 0;
            if (p < pend) {
                va = short(dy[p * lddy + dyoff + cot * 16 + row]);
                const int n = int(p / ((long long)H * W));
                const int r = int(p - (long long)n * H * W);
                const int hh = r / W + ddy, ww = r % W + ddx;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W)
                    vb = short(x[(((long long)n * H + hh) * W + ww) * ldx + xoff + ct * 16 + row]);
            }
            av[t] = va;
            bv[t] = vb;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    // D row = cout within the tile (4 q + t), column = cin within the tile (row)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int co = cot * 16 + 4 * q + t, ci = ct * 16 + row;
        atomicAdd(&dw[((long long)co * T + tap) * cin + ci], acc[t]);
    }
}

// The weights of the data gradient: wt[ci][T - 1 - tap][co] = w[co][tap][ci],
// BF16 from the FP32 master weights.
__global__ __launch_bounds__(256) void k_flip(const float* __restrict__ w, std::uint16_t* __restrict__ wt,
                                              int cin, int cout, int T) {
    const long long n = (long long)cin * cout * T;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int co = int(e / ((long long)T * cin));
    const int rem = int(e - (long long)co * T * cin);
    const int tap = rem / cin, ci = rem - tap * cin;
    wt[((long long)ci * T + (T - 1 - tap)) * cout + co] = to_bf16(w[e]);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ s, std::uint16_t* __restrict__ d,
                                                 long long n) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
        d[e] = to_bf16(s[e]);
}

// g (BF16) = d (FP32) times (pre > 0) when pre is given, for C channels at
// stride ld / offset off of each of n pixels; accumulate adds into gacc
// (FP32) first when it is given.
__global__ __launch_bounds__(256) void k_relu_bwd(const float* __restrict__ d, int ldd, int doff,
                                                  const float* __restrict__ pre, int ldp, int poff,
                                                  std::uint16_t* __restrict__ g, int ldg, int goff, long long n,
                                                  int C) {
    const long long tot = n * C;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const long long p = e / C;
        const int c = int(e - p * C);
        float v = d[p * ldd + doff + c];
        if (pre && !(pre[p * ldp + poff + c] > 0.f)) v = 0.f;
        g[p * ldg + goff + c] = to_bf16(v);
    }
}

// 2 x 2 average pooling (H, W even) of C channels: BF16 in, BF16 out.
__global__ __launch_bounds__(256) void k_pool(const std::uint16_t* __restrict__ x, int ldx, int xoff,
                                              std::uint16_t* __restrict__ y, int ldy, int yoff, int B, int H,
                                              int W, int C) {
    const int Ho = H / 2, Wo = W / 2;
    const long long tot = (long long)B * Ho * Wo * C;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const long long po = e / C;
        const int c = int(e - po * C);
        const int n = int(po / ((long long)Ho * Wo));
        const int r = int(po - (long long)n * Ho * Wo);
        const int ho = r / Wo, wo = r - ho * Wo;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long long pi = ((long long)n * H + 2 * ho + (k >> 1)) * W + 2 * wo + (k & 1);
            s += from_bf16(x[pi * ldx + xoff + c]);
        }
        y[po * ldy + yoff + c] = to_bf16(0.25f * s);
    }
}

// The gradient of the pooling: each input pixel gets a quarter of its
// output's gradient, added to g (FP32) at full resolution.
__global__ __launch_bounds__(256) void k_pool_bwd(const float* __restrict__ gy, int ldy, int yoff,
                                                  float* __restrict__ gx, int ldx, int xoff, int B, int H, int W,
                                                  int C) {
    const long long tot = (long long)B * H * W * C;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const long long p = e / C;
        const int c = int(e - p * C);
        const int n = int(p / ((long long)H * W));
        const int r = int(p - (long long)n * H * W);
        const int h = r / W, w = r - h * W;
        const long long po = ((long long)n * (H / 2) + h / 2) * (W / 2) + w / 2;
        gx[p * ldx + xoff + c] += 0.25f * gy[po * ldy + yoff + c];
    }
}

// Nearest 2x upsampling of C channels (BF16) from H/2 x W/2 to H x W.
__global__ __launch_bounds__(256) void k_up(const std::uint16_t* __restrict__ x, int ldx, int xoff,
                                            std::uint16_t* __restrict__ y, int ldy, int yoff, int B, int H, int W,
                                            int C) {
    const long long tot = (long long)B * H * W * C;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const long long p = e / C;
        const int c = int(e - p * C);
        const int n = int(p / ((long long)H * W));
        const int r = int(p - (long long)n * H * W);
        const int h = r / W, w = r - h * W;
        const long long pi = ((long long)n * (H / 2) + h / 2) * (W / 2) + w / 2;
        y[p * ldy + yoff + c] = x[pi * ldx + xoff + c];
    }
}

// The gradient of the upsampling: each coarse pixel sums its four fine
// pixels' gradients (FP32), written (not added) to gx.
__global__ __launch_bounds__(256) void k_up_bwd(const float* __restrict__ gy, int ldy, int yoff,
                                                float* __restrict__ gx, int ldx, int xoff, int B, int H, int W,
                                                int C) {
    const int Ho = H / 2, Wo = W / 2;
    const long long tot = (long long)B * Ho * Wo * C;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const long long po = e / C;
        const int c = int(e - po * C);
        const int n = int(po / ((long long)Ho * Wo));
        const int r = int(po - (long long)n * Ho * Wo);
        const int ho = r / Wo, wo = r - ho * Wo;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long long pi = ((long long)n * H + 2 * ho + (k >> 1)) * W + 2 * wo + (k & 1);
            s += gy[pi * ldy + yoff + c];
        }
        gx[po * ldx + xoff + c] = s;
    }
}

// Adam on FP32 master weights with the BF16 copy refreshed
// (Kingma and Ba 2015, bias-corrected moments).
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              std::uint16_t* __restrict__ wb, long long n, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.f - b1) * gi;
        const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[e] = wi;
        if (wb) wb[e] = to_bf16(wi);
    }
}

// Images (B of H x W, FP32 or FP64) as channel `c` of a BF16 tensor scaled
// by s[n] per image, and back: out[n] = scale[n] * channel c.
template <class T>
__global__ __launch_bounds__(256) void k_pack(const T* __restrict__ img, const double* __restrict__ s,
                                              std::uint16_t* __restrict__ x, int ldx, int c, int B, long long hw) {
    const long long tot = (long long)B * hw;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int n = int(e / hw);
        x[e * ldx + c] = to_bf16(float(double(img[e]) * (s ? s[n] : 1.0)));
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_unpack(const float* __restrict__ y, int ldy, int c,
                                                const double* __restrict__ s, T* __restrict__ out, int B,
                                                long long hw) {
    const long long tot = (long long)B * hw;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int n = int(e / hw);
        out[e] = T(double(y[e * ldy + c]) * (s ? s[n] : 1.0));
    }
}

// Per-image sums over hw values: s[n] = sum a b (b null: a a), in double.
template <class T>
__global__ __launch_bounds__(256) void k_image_dot(const T* __restrict__ a, const T* __restrict__ b,
                                                   double* __restrict__ s, long long hw) {
    __shared__ double red[256];
    const int n = blockIdx.y;
    double q = 0.0;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < hw; e += (long long)gridDim.x * 256) {
        const double av = double(a[n * hw + e]);
        q += av * (b ? double(b[n * hw + e]) : av);
    }
    red[threadIdx.x] = q;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&s[n], red[0]);
}

// The loss's gradient on the output channel: g = (y - t) * inv[n] (FP32, at
// channel c of a tensor of ld channels; the other channels zero) and the
// partial sums of inv[n] (y - t)^2 into loss[0].
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, int ldy, int c,
                                              const float* __restrict__ t, const double* __restrict__ inv,
                                              float* __restrict__ g, int ldg, int B, long long hw,
                                              double* __restrict__ loss) {
    __shared__ double red[256];
    double q = 0.0;
    const long long tot = (long long)B * hw;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int n = int(e / hw);
        const double d = double(y[e * ldy + c]) - double(t[e]);
        for (int k = 0; k < ldg; ++k) g[e * ldg + k] = 0.f;
        g[e * ldg + c] = float(d * inv[n]);
        q += 0.5 * d * d * inv[n];
    }
    red[threadIdx.x] = q;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Random images for training: per image K Gaussian components of random
// centre, log-uniform width in [wmin, wmax] pixels and amplitude uniform in
// [-1, 1] (or [0, 1] when positive), from a counter-based hash of (seed,
// image, component).
__device__ __forceinline__ std::uint64_t mix64(std::uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unit(std::uint64_t seed, std::uint64_t a, std::uint64_t b, int k) {
    const std::uint64_t h = mix64(seed ^ mix64(a * 0x100000001b3ull + b * 0x2545f4914f6cdd1dull + std::uint64_t(k)));
    return double(h >> 11) * (1.0 / 9007199254740992.0);
}

template <class T>
__global__ __launch_bounds__(256) void k_sky(T* __restrict__ out, int B, int H, int W, int K, double wmin,
                                             double wmax, int positive, std::uint64_t seed) {
    const long long hw = (long long)H * W;
    const long long tot = (long long)B * hw;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int n = int(e / hw);
        const int r = int(e - (long long)n * hw);
        const int h = r / W, w = r - h * W;
        double v = 0.0;
        for (int k = 0; k < K; ++k) {
            const double ch = unit(seed, n, k, 0) * H, cw = unit(seed, n, k, 1) * W;
            const double sw = wmin * exp(unit(seed, n, k, 2) * log(wmax / wmin));
            const double amp = positive ? unit(seed, n, k, 3) : 2.0 * unit(seed, n, k, 3) - 1.0;
            const double dh = h - ch, dw = w - cw;
            const double q = (dh * dh + dw * dw) / (2.0 * sw * sw);
            if (q < 40.0) v += amp * exp(-q);
        }
        out[e] = T(v);
    }
}

// An output gradient from images: g[e * ldg + c] = s[n] img[e] (FP64 to
// FP32), the other channels zero.
__global__ __launch_bounds__(256) void k_grad_images(const double* __restrict__ img, const double* __restrict__ s,
                                                     float* __restrict__ g, int ldg, int c, int B, long long hw) {
    const long long tot = (long long)B * hw;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        for (int k = 0; k < ldg; ++k) g[e * ldg + k] = 0.f;
        g[e * ldg + c] = float(img[e] * s[e / hw]);
    }
}

// out[n] = s[n] x[n] (FP64 images to FP32).
__global__ __launch_bounds__(256) void k_scale_images(const double* __restrict__ x, const double* __restrict__ s,
                                                      float* __restrict__ out, int B, long long hw) {
    const long long tot = (long long)B * hw;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long long)gridDim.x * 256)
        out[e] = float(x[e] * s[e / hw]);
}

// y = a x + b y.
template <class T>
__global__ __launch_bounds__(256) void k_axpby(const T* __restrict__ x, T* __restrict__ y, double a, double b,
                                               long long n) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
        y[e] = T(a * double(x[e]) + b * double(y[e]));
}

int grid1(long long n) {
    const long long b = (n + 255) / 256;
    return int(b < 1 ? 1 : b > 65536 ? 65536 : b);
}

}  // namespace
}  // namespace rang

using namespace rang;

extern "C" {

size_t rang_amd_nn_last_error(uint8_t* buf, size_t len) {
    const std::string& e = g_nn_err;
    if (!buf || len == 0) return e.size();
    const size_t n = e.size() < len ? e.size() : len;
    for (size_t i = 0; i < n; ++i) buf[i] = uint8_t(e[i]);
    return n;
}

static int conv_args_ok(int ldx, int xoff, int ldy, int yoff, int cin, int cout, int ks) {
    if (ks != 1 && ks != 3) return g_nn_err = "kernel size must be 1 or 3", 1;
    if (cin <= 0 || cout <= 0 || cin % 16 || cout % 16) return g_nn_err = "channel counts must be multiples of 16", 1;
    if (ldx % 4 || xoff % 4 || ldx < xoff + cin || ldy < yoff + cout)
        return g_nn_err = "channel strides must hold the channels, in multiples of 4", 1;
    return 0;
}

int32_t rang_amd_nn_conv_fwd(const uint16_t* x, int32_t ldx, int32_t xoff, const uint16_t* w, float* y,
                             uint16_t* a, int32_t ldy, int32_t yoff, int32_t B, int32_t H, int32_t W,
                             int32_t cin, int32_t cout, int32_t ks, int32_t relu, void* stream) {
    if (const int e = conv_args_ok(ldx, xoff, ldy, yoff, cin, cout, ks)) return e;
    if (!x || !w || (!y && !a) || B <= 0 || H <= 0 || W <= 0) return g_nn_err = "null array or empty image", 1;
    const long long npix = (long long)B * H * W;
    const dim3 grid(unsigned((npix + 63) / 64), unsigned((cout + 31) / 32));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (ks == 3)
        k_conv_fwd<3><<<grid, 256, 0, st>>>(x, ldx, xoff, w, y, a, ldy, yoff, B, H, W, cin, cout, relu);
    else
        k_conv_fwd<1><<<grid, 256, 0, st>>>(x, ldx, xoff, w, y, a, ldy, yoff, B, H, W, cin, cout, relu);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_conv_wgrad(const uint16_t* x, int32_t ldx, int32_t xoff, const uint16_t* dy, int32_t lddy,
                               int32_t dyoff, float* dw, int32_t B, int32_t H, int32_t W, int32_t cin,
                               int32_t cout, int32_t ks, void* stream) {
    if (const int e = conv_args_ok(ldx, xoff, lddy, dyoff, cin, cout, ks)) return e;
    if (!x || !dy || !dw || B <= 0 || H <= 0 || W <= 0) return g_nn_err = "null array or empty image", 1;
    const long long npix = (long long)B * H * W;
    const int tiles = (cout / 16) * ks * ks * (cin / 16);
    // pixel chunks: enough blocks to fill the chip, chunks a multiple of 16
    long long nchunk = (2048 + (tiles + 3) / 4 - 1) / ((tiles + 3) / 4);
    if (nchunk < 1) nchunk = 1;
    long long chunk = (npix + nchunk - 1) / nchunk;
    chunk = (chunk + 15) / 16 * 16;
    nchunk = (npix + chunk - 1) / chunk;
    const dim3 grid(unsigned((tiles + 3) / 4), unsigned(nchunk));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (ks == 3)
        k_conv_wgrad<3><<<grid, 256, 0, st>>>(x, ldx, xoff, dy, lddy, dyoff, dw, B, H, W, cin, cout, chunk);
    else
        k_conv_wgrad<1><<<grid, 256, 0, st>>>(x, ldx, xoff, dy, lddy, dyoff, dw, B, H, W, cin, cout, chunk);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_flip(const float* w, uint16_t* wt, int32_t cin, int32_t cout, int32_t ks, void* stream) {
    if (!w || !wt || cin <= 0 || cout <= 0 || (ks != 1 && ks != 3)) return g_nn_err = "bad weights", 1;
    const long long n = (long long)cin * cout * ks * ks;
    k_flip<<<unsigned((n + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>(w, wt, cin, cout, ks * ks);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_to_bf16(const float* s, uint16_t* d, int64_t n, void* stream) {
    if (!s || !d || n < 0) return g_nn_err = "null array", 1;
    if (n == 0) return 0;
    k_to_bf16<<<grid1(n), 256, 0, static_cast<hipStream_t>(stream)>>>(s, d, n);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_relu_bwd(const float* d, int32_t ldd, int32_t doff, const float* pre, int32_t ldp, int32_t poff,
                             uint16_t* g, int32_t ldg, int32_t goff, int64_t npix, int32_t C, void* stream) {
    if (!d || !g || npix < 0 || C <= 0) return g_nn_err = "null array", 1;
    if (npix == 0) return 0;
    k_relu_bwd<<<grid1(npix * C), 256, 0, static_cast<hipStream_t>(stream)>>>(d, ldd, doff, pre, ldp, poff, g, ldg,
                                                                               goff, npix, C);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_pool(const uint16_t* x, int32_t ldx, int32_t xoff, uint16_t* y, int32_t ldy, int32_t yoff,
                         int32_t B, int32_t H, int32_t W, int32_t C, void* stream) {
    if (!x || !y || H % 2 || W % 2) return g_nn_err = "pooling needs even sides", 1;
    k_pool<<<grid1((long long)B * H * W * C / 4), 256, 0, static_cast<hipStream_t>(stream)>>>(x, ldx, xoff, y, ldy,
                                                                                              yoff, B, H, W, C);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_pool_bwd(const float* gy, int32_t ldy, int32_t yoff, float* gx, int32_t ldx, int32_t xoff,
                             int32_t B, int32_t H, int32_t W, int32_t C, void* stream) {
    if (!gy || !gx || H % 2 || W % 2) return g_nn_err = "pooling needs even sides", 1;
    k_pool_bwd<<<grid1((long long)B * H * W * C), 256, 0, static_cast<hipStream_t>(stream)>>>(gy, ldy, yoff, gx, ldx,
                                                                                              xoff, B, H, W, C);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_up(const uint16_t* x, int32_t ldx, int32_t xoff, uint16_t* y, int32_t ldy, int32_t yoff,
                       int32_t B, int32_t H, int32_t W, int32_t C, void* stream) {
    if (!x || !y || H % 2 || W % 2) return g_nn_err = "upsampling needs even sides", 1;
    k_up<<<grid1((long long)B * H * W * C), 256, 0, static_cast<hipStream_t>(stream)>>>(x, ldx, xoff, y, ldy, yoff, B,
                                                                                        H, W, C);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_up_bwd(const float* gy, int32_t ldy, int32_t yoff, float* gx, int32_t ldx, int32_t xoff,
                           int32_t B, int32_t H, int32_t W, int32_t C, void* stream) {
    if (!gy || !gx || H % 2 || W % 2) return g_nn_err = "upsampling needs even sides", 1;
    k_up_bwd<<<grid1((long long)B * H * W * C / 4), 256, 0, static_cast<hipStream_t>(stream)>>>(gy, ldy, yoff, gx, ldx,
                                                                                                xoff, B, H, W, C);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, double lr, double b1,
                         double b2, double eps, int64_t step, void* stream) {
    if (!w || !g || !m || !v || n < 0 || step < 1) return g_nn_err = "bad optimiser arguments", 1;
    if (n == 0) return 0;
    const double c1 = 1.0 - std::pow(b1, double(step)), c2 = 1.0 - std::pow(b2, double(step));
    k_adam<<<grid1(n), 256, 0, static_cast<hipStream_t>(stream)>>>(w, g, m, v, wb, n, float(lr), float(b1), float(b2),
                                                                   float(eps), float(c1), float(c2));
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_pack_f64(const double* img, const double* s, uint16_t* x, int32_t ldx, int32_t c, int32_t B,
                             int64_t hw, void* stream) {
    if (!img || !x || c < 0 || c >= ldx) return g_nn_err = "bad pack arguments", 1;
    k_pack<double><<<grid1(B * hw), 256, 0, static_cast<hipStream_t>(stream)>>>(img, s, x, ldx, c, B, hw);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_pack_f32(const float* img, const double* s, uint16_t* x, int32_t ldx, int32_t c, int32_t B,
                             int64_t hw, void* stream) {
    if (!img || !x || c < 0 || c >= ldx) return g_nn_err = "bad pack arguments", 1;
    k_pack<float><<<grid1(B * hw), 256, 0, static_cast<hipStream_t>(stream)>>>(img, s, x, ldx, c, B, hw);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_unpack_f64(const float* y, int32_t ldy, int32_t c, const double* s, double* out, int32_t B,
                               int64_t hw, void* stream) {
    if (!y || !out || c < 0 || c >= ldy) return g_nn_err = "bad unpack arguments", 1;
    k_unpack<double><<<grid1(B * hw), 256, 0, static_cast<hipStream_t>(stream)>>>(y, ldy, c, s, out, B, hw);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_unpack_f32(const float* y, int32_t ldy, int32_t c, const double* s, float* out, int32_t B,
                               int64_t hw, void* stream) {
    if (!y || !out || c < 0 || c >= ldy) return g_nn_err = "bad unpack arguments", 1;
    k_unpack<float><<<grid1(B * hw), 256, 0, static_cast<hipStream_t>(stream)>>>(y, ldy, c, s, out, B, hw);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_image_dot_f64(const double* a, const double* b, double* s, int32_t B, int64_t hw, void* stream) {
    if (!a || !s || B <= 0) return g_nn_err = "bad dot arguments", 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (const int e = nn_fail(hipMemsetAsync(s, 0, sizeof(double) * B, st))) return e;
    const dim3 grid(unsigned(grid1(hw) < 256 ? grid1(hw) : 256), unsigned(B));
    k_image_dot<double><<<grid, 256, 0, st>>>(a, b, s, hw);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_image_dot_f32(const float* a, const float* b, double* s, int32_t B, int64_t hw, void* stream) {
    if (!a || !s || B <= 0) return g_nn_err = "bad dot arguments", 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (const int e = nn_fail(hipMemsetAsync(s, 0, sizeof(double) * B, st))) return e;
    const dim3 grid(unsigned(grid1(hw) < 256 ? grid1(hw) : 256), unsigned(B));
    k_image_dot<float><<<grid, 256, 0, st>>>(a, b, s, hw);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_loss(const float* y, int32_t ldy, int32_t c, const float* t, const double* inv, float* g,
                         int32_t ldg, int32_t B, int64_t hw, double* loss, void* stream) {
    if (!y || !t || !inv || !g || !loss || c < 0 || c >= ldy || c >= ldg) return g_nn_err = "bad loss arguments", 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (const int e = nn_fail(hipMemsetAsync(loss, 0, sizeof(double), st))) return e;
    k_loss<<<unsigned(grid1(B * hw) < 1024 ? grid1(B * hw) : 1024), 256, 0, st>>>(y, ldy, c, t, inv, g, ldg, B, hw,
                                                                                   loss);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_sky_f64(double* out, int32_t B, int32_t H, int32_t W, int32_t K, double wmin, double wmax,
                            int32_t positive, uint64_t seed, void* stream) {
    if (!out || B <= 0 || H <= 0 || W <= 0 || K < 0 || !(wmin > 0.0) || !(wmax >= wmin))
        return g_nn_err = "bad sky arguments", 1;
    k_sky<double><<<grid1((long long)B * H * W), 256, 0, static_cast<hipStream_t>(stream)>>>(out, B, H, W, K, wmin,
                                                                                             wmax, positive, seed);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_sky_f32(float* out, int32_t B, int32_t H, int32_t W, int32_t K, double wmin, double wmax,
                            int32_t positive, uint64_t seed, void* stream) {
    if (!out || B <= 0 || H <= 0 || W <= 0 || K < 0 || !(wmin > 0.0) || !(wmax >= wmin))
        return g_nn_err = "bad sky arguments", 1;
    k_sky<float><<<grid1((long long)B * H * W), 256, 0, static_cast<hipStream_t>(stream)>>>(out, B, H, W, K, wmin,
                                                                                            wmax, positive, seed);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_scale_images(const double* x, const double* s, float* out, int32_t B, int64_t hw,
                                 void* stream) {
    if (!x || !s || !out || B <= 0 || hw < 0) return g_nn_err = "bad scale arguments", 1;
    k_scale_images<<<grid1(B * hw), 256, 0, static_cast<hipStream_t>(stream)>>>(x, s, out, B, hw);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_grad_images(const double* img, const double* s, float* g, int32_t ldg, int32_t c, int32_t B,
                                int64_t hw, void* stream) {
    if (!img || !s || !g || c < 0 || c >= ldg || B <= 0 || hw < 0) return g_nn_err = "bad gradient arguments", 1;
    k_grad_images<<<grid1(B * hw), 256, 0, static_cast<hipStream_t>(stream)>>>(img, s, g, ldg, c, B, hw);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_axpby_f64(const double* x, double* y, double a, double b, int64_t n, void* stream) {
    if (!x || !y || n < 0) return g_nn_err = "null array", 1;
    if (n == 0) return 0;
    k_axpby<double><<<grid1(n), 256, 0, static_cast<hipStream_t>(stream)>>>(x, y, a, b, n);
    return nn_fail(hipGetLastError());
}

int32_t rang_amd_nn_axpby_f32(const float* x, float* y, double a, double b, int64_t n, void* stream) {
    if (!x || !y || n < 0) return g_nn_err = "null array", 1;
    if (n == 0) return 0;
    k_axpby<float><<<grid1(n), 256, 0, static_cast<hipStream_t>(stream)>>>(x, y, a, b, n);
    return nn_fail(hipGetLastError());
}

}  // extern "C"
