// Batched mixed-radix FFT held in LDS, for the grid transforms of the
// gridder. Lengths are ducc0's good sizes 2^a 3^b 5^c 7^d 11^e
// (core/src/wgridder.rs:185-222); the transform is the unnormalised DFT
// X[k] = sum_j x[j] exp(sgn 2 pi i j k / n), sgn -1 forward, +1 inverse, as
// rustfft's (the reference's FFT, rustfft 6.4.1).
//
// Scheme: Stockham autosort, one pass per radix (3, 5, 7, 11, then 8, 4, 2).
// A workgroup holds `ncol` transforms of length n back to back in LDS; each
// pass reads every value into registers, synchronises and writes the pass's
// output in place, so LDS holds one copy of the data. Twiddles are a per-length
// table in global memory computed on the host in double precision.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "wg_fft_consts.h"

namespace rang {

constexpr int kFftMaxPass = 24;

template <class T>
struct cx {
    T re, im;
};

template <class T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <class T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
// (sgn i) * a
template <class T>
__device__ __forceinline__ cx<T> cmul_si(cx<T> a, int sgn) {
    return sgn > 0 ? cx<T>{-a.im, a.re} : cx<T>{a.im, -a.re};
}

// Passes of one length: radix, stride Ns before the pass and offset of the
// pass's twiddles (Ns * (R - 1) values: [k * (R - 1) + t - 1] =
// exp(-2 pi i k t / (Ns R))).
struct FftDesc {
    int n = 0;
    int npass = 0;
    int radix[kFftMaxPass] = {};
    int ns[kFftMaxPass] = {};
    int twoff[kFftMaxPass] = {};
    // Multipliers of udiv for d = n / radix and d = ns.
    unsigned mnb[kFftMaxPass] = {};
    unsigned mns[kFftMaxPass] = {};
};

// x / d as the high half of x * ceil(2^32 / d), exact for x, d < 2^16 (every
// index here is below the transform length). d = 1 has no 32-bit multiplier
// (it would be 2^32) and is taken apart.
inline unsigned fft_magic(unsigned d) {
    return d <= 1 ? 0u : unsigned(((1ull << 32) + d - 1) / d);
}

__host__ __device__ __forceinline__ unsigned udiv(unsigned x, unsigned d, unsigned m) {
    return d == 1 ? x : unsigned((unsigned long long)x * m >> 32);
}

// Host: the passes of length n and its twiddle table (double precision).
// With fuse21, 3 and 7 together go as one radix-21 pass; such a plan is for
// the transforms run by fft_lds<..., true> only.
inline bool fft_describe(int n, FftDesc& d, std::vector<double>& tw_re_im, bool fuse21 = false) {
    d = FftDesc();
    d.n = n;
    int m = n;
    auto push = [&](int r) {
        if (d.npass >= kFftMaxPass) return false;
        d.radix[d.npass++] = r;
        m /= r;
        return true;
    };
    // Odd radices first: the first pass writes at a stride of its radix, and
    // an odd stride spreads over the LDS banks where 8 would not.
    while (fuse21 && m % 21 == 0)
        if (!push(21)) return false;
    for (int r : {3, 5, 7, 11})
        while (m % r == 0)
            if (!push(r)) return false;
    while (m % 8 == 0)
        if (!push(8)) return false;
    while (m % 4 == 0)
        if (!push(4)) return false;
    while (m % 2 == 0)
        if (!push(2)) return false;
    if (m != 1 || n >= 65536) return false;
    int ns = 1, total = 0;
    for (int s = 0; s < d.npass; ++s) {
        d.ns[s] = ns;
        d.twoff[s] = total;
        total += ns * (d.radix[s] - 1);
        d.mnb[s] = fft_magic(unsigned(n / d.radix[s]));
        d.mns[s] = fft_magic(unsigned(ns));
        ns *= d.radix[s];
    }
    tw_re_im.assign(std::size_t(2) * (total > 0 ? total : 1), 0.0);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int s = 0; s < d.npass; ++s) {
        const int r = d.radix[s];
        for (int k = 0; k < d.ns[s]; ++k)
            for (int t = 1; t < r; ++t) {
                // reduce k t / (Ns R) to [0, 1) exactly before the angle
                const long long num = (long long)k * t, den = (long long)d.ns[s] * r;
                const long double frac = (long double)(num % den) / (long double)den;
                const long double ang = -2.0L * pi * frac;
                const std::size_t at = std::size_t(d.twoff[s] + k * (r - 1) + (t - 1));
                tw_re_im[2 * at] = double(cosl(ang));
                tw_re_im[2 * at + 1] = double(sinl(ang));
            }
    }
    return true;
}

// Small DFTs, out[q] = sum_t v[t] exp(sgn 2 pi i t q / R), in place.
template <class T>
__device__ __forceinline__ void dft2(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <class T>
__device__ __forceinline__ void dft4(cx<T>* v, int sgn) {
    cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cx<T> t2 = cadd(v[1], v[3]), t3 = cmul_si(csub(v[1], v[3]), sgn);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

// out[2q] = DFT4(x_t + x_{t+4}), out[2q+1] = DFT4((x_t - x_{t+4}) W8^t).
template <class T>
__device__ __forceinline__ void dft8(cx<T>* v, int sgn) {
    const T h = T(0.70710678118654752440084436210484903928);
    cx<T> a[4], b[4];
    for (int t = 0; t < 4; ++t) {
        a[t] = cadd(v[t], v[t + 4]);
        b[t] = csub(v[t], v[t + 4]);
    }
    // W8^1 = h (1 + sgn i), W8^2 = sgn i, W8^3 = h (-1 + sgn i)
    b[1] = cmul(b[1], cx<T>{h, T(sgn) * h});
    b[2] = cmul_si(b[2], sgn);
    b[3] = cmul(b[3], cx<T>{-h, T(sgn) * h});
    dft4(a, sgn);
    dft4(b, sgn);
    for (int q = 0; q < 4; ++q) {
        v[2 * q] = a[q];
        v[2 * q + 1] = b[q];
    }
}

// Odd R: pairs (t, R - t) give out[q] = v0 + sum a_t cos + sgn i sum b_t sin.
template <class T, int R>
__device__ __forceinline__ void dft_odd(cx<T>* v, int sgn) {
    constexpr int m = (R - 1) / 2;
    const double* cs = R == 3 ? kCos3 : R == 5 ? kCos5 : R == 7 ? kCos7 : kCos11;
    const double* sn = R == 3 ? kSin3 : R == 5 ? kSin5 : R == 7 ? kSin7 : kSin11;
    cx<T> a[m + 1], b[m + 1], out[R];
    cx<T> s0 = v[0];
#pragma unroll
    for (int t = 1; t <= m; ++t) {
        a[t] = cadd(v[t], v[R - t]);
        b[t] = csub(v[t], v[R - t]);
        s0 = cadd(s0, a[t]);
    }
    out[0] = s0;
#pragma unroll
    for (int q = 1; q <= m; ++q) {
        cx<T> A = v[0], B = {T(0), T(0)};
#pragma unroll
        for (int t = 1; t <= m; ++t) {
            const int at = (t * q) % R;
            const T c = T(cs[at]), s = T(sn[at]);
            A = {A.re + a[t].re * c, A.im + a[t].im * c};
            B = {B.re + b[t].re * s, B.im + b[t].im * s};
        }
        const cx<T> iB = cmul_si(B, sgn);
        out[q] = cadd(A, iB);
        out[R - q] = csub(A, iB);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
}

// Radix 21 as 3 x 7 in registers by the prime-factor (Good-Thomas) mapping,
// so no twiddles between its two stages: a[n1][n2] = v[(7 n1 + 3 n2) mod 21],
// DFT3 along n1, DFT7 along n2, out[(7 k1 + 15 k2) mod 21] = a[k1][k2]. One
// pass (and its two barriers) fewer wherever 3 and 7 both divide a length.
template <class T>
__device__ __forceinline__ void dft21(cx<T>* v, int sgn) {
    cx<T> a[3][7];
#pragma unroll
    for (int n1 = 0; n1 < 3; ++n1)
#pragma unroll
        for (int n2 = 0; n2 < 7; ++n2) a[n1][n2] = v[(7 * n1 + 3 * n2) % 21];
#pragma unroll
    for (int n2 = 0; n2 < 7; ++n2) {
        cx<T> c[3] = {a[0][n2], a[1][n2], a[2][n2]};
        dft_odd<T, 3>(c, sgn);
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) a[k1][n2] = c[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        dft_odd<T, 7>(a[k1], sgn);
#pragma unroll
        for (int k2 = 0; k2 < 7; ++k2) v[(7 * k1 + 15 * k2) % 21] = a[k1][k2];
    }
}

template <class T, int R>
__device__ __forceinline__ void dft_small(cx<T>* v, int sgn) {
    if constexpr (R == 2)
        dft2(v);
    else if constexpr (R == 4)
        dft4(v, sgn);
    else if constexpr (R == 8)
        dft8(v, sgn);
    else if constexpr (R == 21)
        dft21(v, sgn);
    else
        dft_odd<T, R>(v, sgn);
}

// One Stockham pass of radix R over `ncol` transforms of length n in LDS.
// BMAX bounds the butterflies per thread (ceil(ncol n / (R NT))). With ONE
// (a single transform, as every row and long-column kernel) the butterfly's
// line is 0, so no division by the butterfly count is made, and its output
// base (position q within the output stride, k within the twiddles) is formed
// once in the load phase and kept for the store: the divisions are
// multiply-high, quarter-rate on the VALU. Several transforms recompute their
// indices in the store phase instead, which keeps no index registers live
// across the barrier (their kernels hold two transforms' registers).
template <class T, int NT, int R, int BMAX, bool ONE>
__device__ __forceinline__ void fft_pass(cx<T>* buf, int n, int ncol, int ns, unsigned mnb,
                                         unsigned mns, const cx<T>* tw, int sgn) {
    const int nb = n / R;
    const int total = ONE ? nb : nb * ncol;
    cx<T> v[BMAX][R];
    int cb[ONE ? BMAX : 1];
#pragma unroll
    for (int b = 0; b < BMAX; ++b) {
        const int j = int(threadIdx.x) + b * NT;
        if (j < total) {
            int c = 0, jj = j;
            if constexpr (!ONE) {
                c = int(udiv(unsigned(j), unsigned(nb), mnb));
                jj = j - c * nb;
            }
            const cx<T>* col = buf + c * n;
#pragma unroll
            for (int t = 0; t < R; ++t) v[b][t] = col[jj + t * nb];
            int k = 0;
            if constexpr (ONE) {
                const int q = int(udiv(unsigned(jj), unsigned(ns), mns));
                k = jj - q * ns;
                cb[b] = q * ns * R + k;
            } else if (ns > 1) {
                k = jj - int(udiv(unsigned(jj), unsigned(ns), mns)) * ns;
            }
            if (ns > 1) {
                const cx<T>* w = tw + k * (R - 1);
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    cx<T> f = w[t - 1];
                    if (sgn > 0) f.im = -f.im;
                    v[b][t] = cmul(v[b][t], f);
                }
            }
            dft_small<T, R>(v[b], sgn);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BMAX; ++b) {
        const int j = int(threadIdx.x) + b * NT;
        if (j < total) {
            int base;
            cx<T>* col = buf;
            if constexpr (ONE) {
                base = cb[b];
            } else {
                const int c = int(udiv(unsigned(j), unsigned(nb), mnb)), jj = j - c * nb;
                const int q = int(udiv(unsigned(jj), unsigned(ns), mns)), k = jj - q * ns;
                col = buf + c * n;
                base = q * ns * R + k;
            }
#pragma unroll
            for (int t = 0; t < R; ++t) col[base + t * ns] = v[b][t];
        }
    }
    __syncthreads();
}

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

template <class T, int NT, int MAXTOT, bool ONE, bool R21>
__device__ void fft_passes(cx<T>* buf, const FftDesc& d, int ncol, const cx<T>* tw, int sgn) {
    for (int s = 0; s < d.npass; ++s) {
        const cx<T>* tws = tw + d.twoff[s];
        const unsigned a = d.mnb[s], b = d.mns[s];
        switch (d.radix[s]) {
        case 2: fft_pass<T, NT, 2, cdiv(MAXTOT, 2 * NT), ONE>(buf, d.n, ncol, d.ns[s], a, b, tws, sgn); break;
        case 3: fft_pass<T, NT, 3, cdiv(MAXTOT, 3 * NT), ONE>(buf, d.n, ncol, d.ns[s], a, b, tws, sgn); break;
        case 4: fft_pass<T, NT, 4, cdiv(MAXTOT, 4 * NT), ONE>(buf, d.n, ncol, d.ns[s], a, b, tws, sgn); break;
        case 5: fft_pass<T, NT, 5, cdiv(MAXTOT, 5 * NT), ONE>(buf, d.n, ncol, d.ns[s], a, b, tws, sgn); break;
        case 7: fft_pass<T, NT, 7, cdiv(MAXTOT, 7 * NT), ONE>(buf, d.n, ncol, d.ns[s], a, b, tws, sgn); break;
        case 8: fft_pass<T, NT, 8, cdiv(MAXTOT, 8 * NT), ONE>(buf, d.n, ncol, d.ns[s], a, b, tws, sgn); break;
        case 21:
            if constexpr (R21) fft_pass<T, NT, 21, cdiv(MAXTOT, 21 * NT), ONE>(buf, d.n, ncol, d.ns[s], a, b, tws, sgn);
            break;
        default: fft_pass<T, NT, 11, cdiv(MAXTOT, 11 * NT), ONE>(buf, d.n, ncol, d.ns[s], a, b, tws, sgn); break;
        }
    }
}

// All passes of `d` over ncol * d.n <= MAXTOT values in LDS (synchronised
// on entry by the caller, on exit here). PLANE (the exact gridder's
// single-precision plane transforms): the plan may hold radix-21 passes
// (fft_describe with fuse21), and a single line takes the single-line pass.
// Other kernels keep one pass form: the extra code would take registers they
// do not have to spare (the surrogate's column kernels, two transforms each,
// spilled with it).
template <class T, int NT, int MAXTOT, bool PLANE = false>
__device__ void fft_lds(cx<T>* buf, const FftDesc& d, int ncol, const cx<T>* tw, int sgn) {
    if constexpr (PLANE) {
        if (ncol == 1)
            fft_passes<T, NT, MAXTOT, true, true>(buf, d, 1, tw, sgn);
        else
            fft_passes<T, NT, MAXTOT, false, true>(buf, d, ncol, tw, sgn);
    } else {
        fft_passes<T, NT, MAXTOT, false, false>(buf, d, ncol, tw, sgn);
    }
}

// Plane-wise transpose of complex values, [P][R][C] -> [P][C][R], through
// LDS tiles (64 x 64 in single precision, 32 x 32 in double, a column of
// padding against bank conflicts): both sides coalesced.
template <class T>
__global__ __launch_bounds__(256) void k_tile_transpose(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, int R,
                                                        int C) {
    constexpr int TS = sizeof(T) == 4 ? 64 : 32;
    __shared__ cx<T> t[TS][TS + 1];
    const int c0 = blockIdx.x * TS, r0 = blockIdx.y * TS;
    const long long base = (long long)blockIdx.z * R * C;
    for (int k = threadIdx.x; k < TS * TS; k += 256) {
        const int rr = k / TS, cc = k - rr * TS, r = r0 + rr, c = c0 + cc;
        if (r < R && c < C) t[rr][cc] = in[base + (long long)r * C + c];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TS * TS; k += 256) {
        const int cc = k / TS, rr = k - cc * TS, r = r0 + rr, c = c0 + cc;
        if (r < R && c < C) out[base + (long long)c * R + r] = t[rr][cc];
    }
}

// The same for single precision with even R and C: each lane moves two
// neighbouring values (16 bytes) per access on both sides.
template <class T>
__global__ __launch_bounds__(256) void k_tile_transpose_pairs(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                                                              int R, int C) {
    static_assert(sizeof(T) == 4, "pairs of single-precision values");
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int TS = 64;
    __shared__ cx<T> t[TS][TS + 1];
    const int c0 = blockIdx.x * TS, r0 = blockIdx.y * TS;
    const long long base = (long long)blockIdx.z * R * C;
    for (int k = threadIdx.x; k < TS * TS / 2; k += 256) {
        const int rr = k / (TS / 2), cp = 2 * (k - rr * (TS / 2)), r = r0 + rr, c = c0 + cp;
        if (r < R && c < C) {
            const f4 v = *reinterpret_cast<const f4*>(in + base + (long long)r * C + c);
            t[rr][cp] = cx<T>{v.x, v.y};
            t[rr][cp + 1] = cx<T>{v.z, v.w};
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TS * TS / 2; k += 256) {
        const int cc = k / (TS / 2), rp = 2 * (k - cc * (TS / 2)), r = r0 + rp, c = c0 + cc;
        if (r < R && c < C) {
            const cx<T> a = t[rp][cc], b = t[rp + 1][cc];
            *reinterpret_cast<f4*>(out + base + (long long)c * R + r) = f4{a.re, a.im, b.re, b.im};
        }
    }
}

}  // namespace rang
