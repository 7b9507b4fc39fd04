// Transforms of any length for the w-stacking gridder's grids (rang-core's
// CPU path, core/src/grid.rs:206-259, transforms any side through rustfft).
//
// LineFft<T> transforms contiguous lines of length N in place, unnormalised,
// X[k] = sum_j x[j] exp(sgn 2 pi i j k / N):
//  * N a product of 2, 3, 5, 7 and 11 within one LDS: the in-LDS passes of
//    wg_fft.h, one line per workgroup;
//  * N such a product beyond one LDS: N = M1 M2 in four steps (Bailey):
//    transforms along M1 at stride M2 with the twiddles w_N^(j2 k1), then along
//    M2, which leaves X[k1 + M1 k2] at k1 M2 + k2; a transpose of each
//    M1 x M2 line puts it in natural order;
//  * any other N by Bluestein's chirp z-transform: j k = (j^2 + k^2 -
//    (k - j)^2) / 2 makes the transform the convolution of x_j c_j with
//    b_m = conj(c_m), c_k = exp(sgn pi i k^2 / N), done by transforms of a
//    good length M >= 2N - 1 (in LDS, or in four steps; the product with the
//    filter's spectrum is taken in whatever order the forward transform
//    leaves, and the four-step inverse reads that order back to natural).
// Chirps and twiddles are formed in double from exact integer reductions
// (k^2 mod 2N, j k mod M).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "wg_fft.h"

namespace rang {
namespace ws {

template <class T>
__device__ __forceinline__ cx<T> unit_root(long long num, long long den, int sgn) {
    // exp(sgn 2 pi i num / den) with num reduced mod den exactly
    long long r = num % den;
    if (r < 0) r += den;
    double s, c;
    sincos(6.283185307179586476925286766559 * double(r) / double(den), &s, &c);
    return cx<T>{T(c), T(sgn > 0 ? s : -s)};
}

// One line per workgroup in LDS.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_line_lds(cx<T>* __restrict__ data, FftDesc d, const cx<T>* __restrict__ tw,
                                                 int sgn) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    cx<T>* line = data + (long long)blockIdx.x * d.n;
    for (int e = threadIdx.x; e < d.n; e += NT) buf[e] = line[e];
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, 1, tw, sgn);
    for (int e = threadIdx.x; e < d.n; e += NT) line[e] = buf[e];
}

// Bluestein in LDS, one line of length N per workgroup: x c padded to M,
// forward, times the filter's spectrum (natural order, scaled by 1 / M),
// inverse, times c.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_blue_lds(cx<T>* __restrict__ data, int N, FftDesc dm, const cx<T>* __restrict__ tw,
                                                 const cx<T>* __restrict__ chirp, const cx<T>* __restrict__ bhat) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    cx<T>* line = data + (long long)blockIdx.x * N;
    for (int e = threadIdx.x; e < dm.n; e += NT) buf[e] = e < N ? cmul(line[e], chirp[e]) : cx<T>{T(0), T(0)};
    __syncthreads();
    fft_lds<T, NT, MT>(buf, dm, 1, tw, -1);
    for (int e = threadIdx.x; e < dm.n; e += NT) buf[e] = cmul(buf[e], bhat[e]);
    __syncthreads();
    fft_lds<T, NT, MT>(buf, dm, 1, tw, +1);
    for (int e = threadIdx.x; e < N; e += NT) line[e] = cmul(buf[e], chirp[e]);
}

// Four-step pass over the columns of lines viewed as M1 x M2 matrices
// (element [a][b] at a M2 + b): C columns per workgroup transformed along a
// (length M1), then, with `twid`, element [k1][j2] times w_M^(sgn j2 k1).
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_fs_cols(cx<T>* __restrict__ buf, int M1, int M2, int C, FftDesc d1,
                                                const cx<T>* __restrict__ tw1, int sgn, int twid) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* s = reinterpret_cast<cx<T>*>(smem);
    const int ncb = (M2 + C - 1) / C;
    const long long line = blockIdx.x / ncb;
    const int j0 = int(blockIdx.x - line * ncb) * C, nc = min(C, M2 - j0);
    const long long M = (long long)M1 * M2;
    cx<T>* x = buf + line * M;
    for (int e = threadIdx.x; e < nc * M1; e += NT) {
        const int a = e / nc, c = e - a * nc;
        s[c * M1 + a] = x[(long long)a * M2 + j0 + c];
    }
    __syncthreads();
    fft_lds<T, NT, MT>(s, d1, nc, tw1, sgn);
    for (int e = threadIdx.x; e < nc * M1; e += NT) {
        const int a = e / nc, c = e - a * nc;
        cx<T> v = s[c * M1 + a];
        if (twid) v = cmul(v, unit_root<T>((long long)(j0 + c) * a, M, sgn));
        x[(long long)a * M2 + j0 + c] = v;
    }
}

// Four-step pass over the rows: K rows of length M2 per workgroup, then,
// with `twid`, element [k1][j2] times w_M^(sgn j2 k1).
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_fs_rows(cx<T>* __restrict__ buf, int M1, int M2, int K, FftDesc d2,
                                                const cx<T>* __restrict__ tw2, int sgn, int twid) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* s = reinterpret_cast<cx<T>*>(smem);
    const int nrb = (M1 + K - 1) / K;
    const long long line = blockIdx.x / nrb;
    const int r0 = int(blockIdx.x - line * nrb) * K, nr = min(K, M1 - r0);
    const long long M = (long long)M1 * M2;
    cx<T>* x = buf + line * M + (long long)r0 * M2;
    for (int e = threadIdx.x; e < nr * M2; e += NT) s[e] = x[e];
    __syncthreads();
    fft_lds<T, NT, MT>(s, d2, nr, tw2, sgn);
    for (int e = threadIdx.x; e < nr * M2; e += NT) {
        cx<T> v = s[e];
        if (twid) {
            const int r = e / M2, b = e - r * M2;
            v = cmul(v, unit_root<T>((long long)b * (r0 + r), M, sgn));
        }
        x[e] = v;
    }
}

// out[l][b][a] = in[l][a][b] for lines of A x B, 32 x 32 tiles through LDS.
template <class T>
__global__ __launch_bounds__(256) void k_transpose(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, int A, int B) {
    __shared__ cx<T> tile[32][33];
    const long long line = blockIdx.z;
    const int a0 = blockIdx.y * 32, b0 = blockIdx.x * 32;
    const cx<T>* src = in + line * (long long)A * B;
    cx<T>* dst = out + line * (long long)A * B;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int a = a0 + r, b = b0 + tx;
        if (a < A && b < B) tile[r][tx] = src[(long long)a * B + b];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int b = b0 + r, a = a0 + tx;
        if (a < A && b < B) dst[(long long)b * A + a] = tile[tx][r];
    }
}

// lines of N -> lines of M: x c, zero-padded.
template <class T>
__global__ __launch_bounds__(256) void k_blue_in(const cx<T>* __restrict__ data, int N, long long M, long long lines,
                                                 const cx<T>* __restrict__ chirp, cx<T>* __restrict__ buf) {
    const long long total = lines * M;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const long long l = e / M;
        const int j = int(e - l * M);
        buf[e] = j < N ? cmul(data[l * N + j], chirp[j]) : cx<T>{T(0), T(0)};
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_mul_spec(cx<T>* __restrict__ buf, long long M, long long lines,
                                                  const cx<T>* __restrict__ bhat) {
    const long long total = lines * M;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256)
        buf[e] = cmul(buf[e], bhat[e % M]);
}

template <class T>
__global__ __launch_bounds__(256) void k_blue_out(const cx<T>* __restrict__ buf, int N, long long M, long long lines,
                                                  const cx<T>* __restrict__ chirp, cx<T>* __restrict__ data) {
    const long long total = lines * N;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const long long l = e / N;
        const int k = int(e - l * N);
        data[e] = cmul(buf[l * M + k], chirp[k]);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_scale(cx<T>* __restrict__ v, long long n, double s) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) v[e] = cx<T>{T(double(v[e].re) * s), T(double(v[e].im) * s)};
}

// Largest in-LDS transform in T (values held); RANG_WS_FFT_LDS_MAX lowers it
// (a test knob that sends small sides through the four-step and chirp
// paths; it never raises it).
template <class T>
int lds_values() {
    int lim = sizeof(T) == 4 ? 16384 : 8192;
    if (const char* e = std::getenv("RANG_WS_FFT_LDS_MAX")) {
        const long v = std::atol(e);
        if (v >= 16 && v < lim) lim = int(v);
    }
    return lim;
}

// Smallest product of 2, 3, 5, 7 and 11 at least n.
inline long long good_length(long long n) {
    for (long long m = std::max<long long>(n, 1);; ++m) {
        long long r = m;
        for (long long p : {2, 3, 5, 7, 11})
            while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
}

inline bool is_good(long long n) { return n >= 1 && good_length(n) == n; }

// n = n1 n2, both within `lim`, n1 nearest sqrt(n); false when none.
inline bool split_two(long long n, int lim, int& n1, int& n2) {
    n1 = n2 = 0;
    double best = 1e300;
    for (long long d = 2; d <= lim && d < n; ++d) {
        if (n % d || n / d > lim) continue;
        const double dist = std::fabs(double(d) - std::sqrt(double(n)));
        if (dist < best) {
            best = dist;
            n1 = int(d);
            n2 = int(n / d);
        }
    }
    return n1 > 0;
}

}  // namespace ws
}  // namespace rang
