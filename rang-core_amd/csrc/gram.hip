// The normal operator without visibilities (rang-core core/src/gridder_gram.rs)
// on the GPU.
//
// G(l, l') = Re sum_p W_p exp(2 pi i (u_p . (l - l') + w_p (x(l) - x(l')))),
// x = n - 1. With exp(2 pi i w xi) ~ sum_k psi_k(w) exp(2 pi i w_k xi) over
// fixed nodes w_k (the compact expansion of WExpansion),
//   G m ~ Re sum_k e_k (K_k (*) (conj(e_k) m)),  e_k(l) = exp(2 pi i w_k x(l)),
//   K_k(delta) = sum_p W_p psi_k(w_p) exp(2 pi i u_p . delta)
// (Fessler et al. 2005, IEEE TSP 53, 3393, eqs 24-26). Each K_k is needed at
// |delta| < n only, so it lives on the 2n grid and (being Hermitian) has a
// real transform. Setup: per node, the node's weights are gridded by the
// exact gridder onto a 2n x 2n image with the constant visibility 1 - i
// (real output sum W (cos + sin): its even and odd parts are Re K and Im K),
// and the 2-D transform of K is kept, real, column-major, scaled by
// 1/(2n)^2. Apply, per node: rows of conj(e_k) m zero-padded to 2n and
// transformed (k_ap_rows_fwd), columns transformed, multiplied by the
// spectrum and transformed back (k_ap_cols), rows transformed back and the
// real part of the product with e_k added into the output (k_ap_rows_inv).
//
// The transforms are the in-LDS mixed-radix passes of wg_fft.h; sides of
// 2n with prime factors above 11 are refused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>

#include "gram.h"
#include "wg_fft.h"

namespace rang {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct GramError {
    std::string msg;
    int status = 2;
};

void gchk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw GramError{std::string(what) + ": " + hipGetErrorString(e)};
}
#define GCHK(x) gchk((x), #x)

// ------------------------------------------------------------ w-expansion

// Lower Cholesky factor of a symmetric positive definite n x n matrix, in place.
void cholesky(std::vector<double>& a, int n) {
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < j; ++k) {
            const double f = a[j * n + k];
            for (int i = j; i < n; ++i) a[i * n + j] -= a[i * n + k] * f;
        }
        const double d = std::sqrt(a[j * n + j]);
        for (int i = j; i < n; ++i) a[i * n + j] /= d;
    }
}

void cholesky_solve(const std::vector<double>& l, int n, double* c) {
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < i; ++k) c[i] -= l[i * n + k] * c[k];
        c[i] /= l[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        for (int k = i + 1; k < n; ++k) c[i] -= l[k * n + i] * c[k];
        c[i] /= l[i * n + i];
    }
}

}  // namespace

double WExpansion::coefficient(double s, int j) const {
    const double f = (s + 0.5) * kWexpBins;
    const int b = std::min(std::max(int(f), 0), kWexpBins - 1);
    const double a = f - b;
    const double lo = table[std::size_t(b) * stencil + j], hi = table[std::size_t(b + 1) * stencil + j];
    return lo + a * (hi - lo);
}

// The regularised least-squares fit of gridder_gram.rs:88-141: nodes t_j at
// unit spacing centred on 0; the error measured over Chebyshev-distributed
// xi in [-1, 1] (64 samples per node), so the normal matrix is
// N_ij = mean_xi cos(theta xi (t_i - t_j)), its diagonal scaled by 1 + ridge,
// and the coefficients at s solve N psi = [mean_xi cos(theta xi (s - t_j))]_j.
// sup_error: the largest |exp(i theta xi s) - sum_j psi_j(s) exp(i theta xi
// t_j)| over 257 values of s and 513 of xi in [0, 1] (:143-163);
// lebesgue = max_s sum_j |psi_j(s)|.
bool wexpansion_build(int stencil, double theta, double ridge, WExpansion& out, std::string& err) {
    if (stencil < 2 || stencil % 2) {
        err = "the w stencil must be even and at least 2";
        return false;
    }
    if (!(theta > 0.0 && theta < kPi)) {
        err = "the w phase step must lie in (0, pi)";
        return false;
    }
    const int s = stencil;
    std::vector<double> t(s);
    for (int j = 0; j < s; ++j) t[j] = j - (s - 1.0) / 2.0;
    const int m = 64 * s;
    std::vector<double> xi(m);
    for (int i = 0; i < m; ++i) xi[i] = std::cos(kPi * (i + 0.5) / m);
    auto kernel = [&](double d) {
        double acc = 0.0;
        for (int i = 0; i < m; ++i) acc += std::cos(theta * xi[i] * d);
        return acc / m;
    };
    std::vector<double> normal(std::size_t(s) * s);
    for (int i = 0; i < s; ++i) {
        for (int j = 0; j < s; ++j) normal[i * s + j] = kernel(t[i] - t[j]);
        normal[i * s + i] *= 1.0 + ridge;
    }
    cholesky(normal, s);
    out = WExpansion();
    out.stencil = s;
    out.theta = theta;
    out.table.assign(std::size_t(kWexpBins + 1) * s, 0.0);
    for (int b = 0; b <= kWexpBins; ++b) {
        const double sb = -0.5 + double(b) / kWexpBins;
        double* row = &out.table[std::size_t(b) * s];
        for (int j = 0; j < s; ++j) row[j] = kernel(sb - t[j]);
        cholesky_solve(normal, s, row);
    }
    double leb = 0.0;
    for (int b = 0; b <= kWexpBins; ++b) {
        double a = 0.0;
        for (int j = 0; j < s; ++j) a += std::fabs(out.table[std::size_t(b) * s + j]);
        leb = std::max(leb, a);
    }
    out.lebesgue = leb;
    double sup = 0.0;
    std::vector<double> c(s);
    for (int sb = 0; sb <= 256; ++sb) {
        const double sv = -0.5 + sb / 256.0 + (sb < 256 ? 0.5 / 8192.0 : 0.0);
        for (int j = 0; j < s; ++j) c[j] = out.coefficient(sv, j);
        for (int xb = 0; xb <= 512; ++xb) {
            const double x = xb / 512.0;
            double re = std::cos(theta * x * sv), im = std::sin(theta * x * sv);
            for (int j = 0; j < s; ++j) {
                re -= c[j] * std::cos(theta * x * t[j]);
                im -= c[j] * std::sin(theta * x * t[j]);
            }
            sup = std::max(sup, std::hypot(re, im));
        }
    }
    out.sup_error = sup;
    return true;
}

namespace {

// ------------------------------------------------------------ kernels

// n - 1 of a pixel, -r^2 / (sqrt(1 - r^2) + 1) (gridder_gram.rs:166-170).
__host__ __device__ inline double nm1_of(double l, double m) {
    const double r2 = l * l + m * m;
    return -r2 / (sqrt(fmax(1.0 - r2, 0.0)) + 1.0);
}

// uvw with each row negated where flip * w < 0 (the real part of the Gram
// matrix does not change), and |flip * w| per row.
__global__ __launch_bounds__(256) void k_fold(const double* __restrict__ uvw, long long nrow, double fw,
                                              double* __restrict__ out, double* __restrict__ absw) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    double u = uvw[3 * r], v = uvw[3 * r + 1], w = uvw[3 * r + 2];
    const double wf = fw * w;
    if (wf < 0.0) {
        u = -u;
        v = -v;
        w = -w;
    }
    out[3 * r] = u;
    out[3 * r + 1] = v;
    out[3 * r + 2] = w;
    absw[r] = fabs(wf);
}

// The weights of node k (gridder_gram.rs:314-337): W psi_{k - cell}(frac)
// where the visibility's w position p = (|w| f/c - wmin)/dw falls in cell
// c = min(max(floor p, 0), cells) and c <= k < c + stencil; zero otherwise.
template <class T, class W>
__global__ __launch_bounds__(256) void k_node_wgt(const double* __restrict__ absw, const double* __restrict__ foc,
                                                  const W* __restrict__ wgt, long long n, int nchan, double wmin,
                                                  double dw, long long cells, int stencil, int k,
                                                  const double* __restrict__ table, T* __restrict__ out,
                                                  unsigned long long* __restrict__ active) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long mine = 0;
    if (i < n) {
        const long long row = i / nchan;
        const int ch = int(i - row * nchan);
        const double weight = double(wgt[i]);
        T o = T(0);
        if (weight != 0.0) {
            const double p = (absw[row] * foc[ch] - wmin) / dw;
            const long long c0 = (long long)fmax(floor(p), 0.0);
            const long long cell = c0 < cells ? c0 : cells;
            if (k >= cell && k < cell + stencil) {
                const double frac = fmin(fmax(p - double(cell) - 0.5, -0.5), 0.5);
                const double f = (frac + 0.5) * kWexpBins;
                const int b = min(max(int(f), 0), kWexpBins - 1);
                const double a = f - b;
                const int j = int(k - cell);
                const double lo = table[std::size_t(b) * stencil + j], hi = table[std::size_t(b + 1) * stencil + j];
                o = T(weight * (lo + a * (hi - lo)));
                mine = 1;
            }
        }
        out[i] = o;
    }
    // one atomic per wave
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
    if ((threadIdx.x & 63) == 0 && mine) atomicAdd(active, mine);
}

template <class T, class W>
__global__ __launch_bounds__(256) void k_copy_wgt(const W* __restrict__ in, T* __restrict__ out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = T(in[i]);
}

template <class T>
__global__ __launch_bounds__(256) void k_const_vis(cx<T>* __restrict__ v, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = cx<T>{T(1), T(-1)};
}

// K at delta = (i - n, j - n) stored at delta mod 2n (gridder_gram.rs:350-368).
// With the constant visibility 1 - i the image is sum W (cos + sin); its even
// part is Re K, its odd part Im K. One node without w (`real_only`): the
// image is the PSF, K real. The row and column delta = -n are zero.
template <class T>
__global__ __launch_bounds__(256) void k_kgrid(const T* __restrict__ img, int n, int real_only, cx<T>* __restrict__ kg) {
    const int big = 2 * n;
    const int kj = blockIdx.x * blockDim.x + threadIdx.x;
    const int ki = blockIdx.y;
    if (kj >= big) return;
    const int i = (ki + n) % big, j = (kj + n) % big;
    cx<T> o{T(0), T(0)};
    if (i != 0 && j != 0) {
        const T a = img[(long long)i * big + j];
        if (real_only) {
            o.re = a;
        } else {
            const T b = img[(long long)(big - i) * big + (big - j)];
            o = cx<T>{(a + b) * T(0.5), (a - b) * T(0.5)};
        }
    }
    kg[(long long)ki * big + kj] = o;
}

// Transforms of `rows` contiguous rows of length d.n in place, one row per
// workgroup.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_fft_rows(cx<T>* __restrict__ data, FftDesc d, const cx<T>* __restrict__ tw,
                                                 int sgn) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(lds_raw);
    cx<T>* row = data + (long long)blockIdx.x * d.n;
    for (int j = threadIdx.x; j < d.n; j += NT) buf[j] = row[j];
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, 1, tw, sgn);
    for (int j = threadIdx.x; j < d.n; j += NT) row[j] = buf[j];
}

// Transforms of the columns of a rows x d.n row-major array (rows == d.n),
// C columns per workgroup, in place.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_fft_cols(cx<T>* __restrict__ data, int C, FftDesc d, const cx<T>* __restrict__ tw,
                                                 int sgn) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(lds_raw);
    const int len = d.n;
    const int j0 = blockIdx.x * C;
    const int nc = min(C, len - j0);
    for (int e = threadIdx.x; e < nc * len; e += NT) {
        const int i = e / nc, c = e - i * nc;
        buf[c * len + i] = data[(long long)i * len + j0 + c];
    }
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, nc, tw, sgn);
    for (int e = threadIdx.x; e < nc * len; e += NT) {
        const int i = e / nc, c = e - i * nc;
        data[(long long)i * len + j0 + c] = buf[c * len + i];
    }
}

// spec[j * big + i] = Re(kg[i][j]) * scale; per-block largest |Re| and |Im|.
template <class T>
__global__ __launch_bounds__(256) void k_spectrum(const cx<T>* __restrict__ kg, int big, double scale, T* __restrict__ spec,
                                                  double* __restrict__ bmax) {
    __shared__ double sre[256], sim[256];
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    double are = 0.0, aim = 0.0;
    if (e < (long long)big * big) {
        const int j = int(e / big), i = int(e - (long long)j * big);
        const cx<T> v = kg[(long long)i * big + j];
        spec[e] = T(double(v.re) * scale);
        are = fabs(double(v.re));
        aim = fabs(double(v.im));
    }
    sre[threadIdx.x] = are;
    sim[threadIdx.x] = aim;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sre[threadIdx.x] = fmax(sre[threadIdx.x], sre[threadIdx.x + s]);
            sim[threadIdx.x] = fmax(sim[threadIdx.x], sim[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bmax[2 * blockIdx.x] = sre[0];
        bmax[2 * blockIdx.x + 1] = sim[0];
    }
}

// Cells of a spectrum below `floor` in magnitude.
// A grid-stride loop over a bounded number of blocks, the count reduced in
// the block and added with one atomic per block (an atomic per wave on the
// one counter serialised the launch: 38 ms for 2^28 cells).
constexpr int kCountBlocks = 2048;
template <class T>
__global__ __launch_bounds__(256) void k_count_small(const T* __restrict__ spec, long long n, double floor_,
                                                     unsigned long long* __restrict__ cnt) {
    unsigned long long mine = 0ull;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
        mine += fabs(double(spec[e])) < floor_ ? 1ull : 0ull;
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
    __shared__ unsigned long long wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t) atomicAdd(cnt, t);
    }
}

// The screen factor exp(2 pi i w x(l)) of pixel (i, j): the phase reduced
// mod 1 in FP64, then the sine and cosine in T.
template <class T>
__device__ __forceinline__ cx<T> screen(double wk, int i, int j, int n, double dx, double dy) {
    const double l = double(i - n / 2) * dx, m = double(j - n / 2) * dy;
    double ph = wk * nm1_of(l, m);
    ph -= floor(ph);
    T s, c;
    if constexpr (sizeof(T) == 4)
        sincosf(float(6.283185307179586 * ph), &s, &c);
    else
        sincos(6.283185307179586 * ph, &s, &c);
    return cx<T>{c, s};
}

// Row i of conj(e_k) m, zero-padded to 2n, transformed forward; tmp holds n
// rows of 2n.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_ap_rows_fwd(const T* __restrict__ m, int n, int use_w, double w0, double dw,
                                                    double dx, double dy, cx<T>* __restrict__ tmp_all, FftDesc d,
                                                    const cx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(lds_raw);
    const int i = blockIdx.x;
    const int big = 2 * n;
    // node blockIdx.y: its w and its own n x 2n slice of the half-transformed array
    const double wk = w0 + double(blockIdx.y) * dw;
    cx<T>* tmp = tmp_all + (long long)blockIdx.y * n * big;
    for (int j = threadIdx.x; j < big; j += NT) {
        cx<T> z{T(0), T(0)};
        if (j < n) {
            const T v = m[(long long)i * n + j];
            if (use_w && v != T(0)) {
                const cx<T> e = screen<T>(wk, i, j, n, dx, dy);
                z = cx<T>{v * e.re, -v * e.im};
            } else {
                z.re = v;
            }
        }
        buf[j] = z;
    }
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, 1, tw, -1);
    cx<T>* row = tmp + (long long)i * big;
    for (int j = threadIdx.x; j < big; j += NT) row[j] = buf[j];
}

// Columns j0 .. j0 + C of tmp (n rows held, rows n .. 2n - 1 zero), forward,
// times the spectrum, inverse, the first n rows written back.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_ap_cols(cx<T>* __restrict__ tmp_all, int n, const T* const* __restrict__ specs,
                                                int C, FftDesc d, const cx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(lds_raw);
    const int big = 2 * n;
    cx<T>* tmp = tmp_all + (long long)blockIdx.y * n * big;
    const T* spec = specs[blockIdx.y];
    const int j0 = blockIdx.x * C;
    const int nc = min(C, big - j0);
    for (int e = threadIdx.x; e < nc * big; e += NT) {
        const int i = e / nc, c = e - i * nc;
        buf[c * big + i] = i < n ? tmp[(long long)i * big + j0 + c] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, nc, tw, -1);
    for (int e = threadIdx.x; e < nc * big; e += NT) {
        const int c = e / big, i = e - c * big;
        const T g = spec[(long long)(j0 + c) * big + i];
        buf[e] = cx<T>{buf[e].re * g, buf[e].im * g};
    }
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, nc, tw, +1);
    for (int e = threadIdx.x; e < nc * n; e += NT) {
        const int i = e / nc, c = e - i * nc;
        tmp[(long long)i * big + j0 + c] = buf[c * big + i];
    }
}

// The same on the transposed slice: column j of the half-transformed array
// held as n contiguous values at tmpT[j * n] (rows n .. 2n - 1 zero), so
// both the load and the store of a column are coalesced.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_ap_cols_t(cx<T>* __restrict__ tmpt_all, int n, const T* const* __restrict__ specs,
                                                  int C, FftDesc d, const cx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(lds_raw);
    const int big = 2 * n;
    cx<T>* tt = tmpt_all + (long long)blockIdx.y * n * big;
    const T* spec = specs[blockIdx.y];
    const int j0 = blockIdx.x * C;
    const int nc = min(C, big - j0);
    for (int e = threadIdx.x; e < nc * big; e += NT) {
        const int c = e / big, i = e - c * big;
        buf[e] = i < n ? tt[(long long)(j0 + c) * n + i] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, nc, tw, -1);
    for (int e = threadIdx.x; e < nc * big; e += NT) {
        const int c = e / big, i = e - c * big;
        const T g = spec[(long long)(j0 + c) * big + i];
        buf[e] = cx<T>{buf[e].re * g, buf[e].im * g};
    }
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, nc, tw, +1);
    for (int e = threadIdx.x; e < nc * big; e += NT) {
        const int c = e / big, i = e - c * big;
        if (i < n) tt[(long long)(j0 + c) * n + i] = buf[e];
    }
}

// Row i of every node's slice transformed back; out = sum over nodes of
// Re(y_k e_k) over the first n values, summed in registers and written once.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_ap_rows_inv(const cx<T>* __restrict__ tmp_all, int n, int nodes, int use_w,
                                                    double w0, double dw, double dx, double dy, T* __restrict__ out,
                                                    int first, FftDesc d, const cx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(lds_raw);
    constexpr int PER = (MT / 2 + NT - 1) / NT;  // image values per thread (n <= MT / 2)
    const int i = blockIdx.x;
    const int big = 2 * n;
    T acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = T(0);
    for (int k = 0; k < nodes; ++k) {
        const double wk = w0 + double(k) * dw;
        const cx<T>* row = tmp_all + ((long long)k * n + i) * big;
        for (int j = threadIdx.x; j < big; j += NT) buf[j] = row[j];
        __syncthreads();
        fft_lds<T, NT, MT>(buf, d, 1, tw, +1);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = int(threadIdx.x) + q * NT;
            if (j < n) {
                const cx<T> y = buf[j];
                T val = y.re;
                if (use_w) {
                    const cx<T> e = screen<T>(wk, i, j, n, dx, dy);
                    val = y.re * e.re - y.im * e.im;
                }
                acc[q] += val;
            }
        }
        __syncthreads();  // the buffer is refilled by the next node
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = int(threadIdx.x) + q * NT;
        if (j < n) {
            T& o = out[(long long)i * n + j];
            o = first ? acc[q] : o + acc[q];
        }
    }
}

// 1 / norm / (max(v norm, 0) + rho) (gridder_gram.rs:488-509).
template <class T>
__global__ __launch_bounds__(256) void k_inverse_filter(const T* __restrict__ in, long long n, double norm, double rho,
                                                        T* __restrict__ out) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) out[e] = T(1.0 / norm / (fmax(double(in[e]) * norm, 0.0) + rho));
}

// ------------------------------------------------------------ host side

template <class T>
constexpr int lds_limit() {
    return sizeof(T) == 4 ? 16384 : 8192;
}

template <int V>
using gic = std::integral_constant<int, V>;

// f(NT, MT) with the smallest in-LDS size holding `need` values.
template <class T, class F>
void dispatch(int need, F&& f) {
    if (need <= 1024)
        f(gic<256>{}, gic<1024>{});
    else if (need <= 2048)
        f(gic<256>{}, gic<2048>{});
    else if (need <= 4096)
        f(gic<256>{}, gic<4096>{});
    else if (need <= 8192)
        f(gic<512>{}, gic<8192>{});
    else if constexpr (sizeof(T) == 4) {
        if (need <= 16384)  // a line filling most of a CU's LDS: one workgroup per CU, 16 waves
            f(gic<1024>{}, gic<16384>{});
        else
            throw GramError{"transform too long for LDS", 1};
    } else {
        throw GramError{"transform too long for LDS", 1};
    }
}

template <class K>
void lds_attr(K kernel, std::size_t bytes) {
    GCHK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
}

template <class T>
struct GramImpl final : GramBase {
    int device = 0;
    int n = 0;
    double dx = 0, dy = 0;
    double w0 = 0, dw = 0;
    bool use_w = false;
    std::vector<T*> spectra;
    FftDesc d;
    cx<T>* tw = nullptr;
    cx<T>* tmp = nullptr;      // n x 2n for the spectra's set-up
    cx<T>* tmp_all = nullptr;  // one n x 2n slice per node for applications
    cx<T>* tmpt_all = nullptr;  // the slices transposed (2n x n), for large images
    T** d_spec = nullptr;      // the spectra's device addresses
    double st[kGramStats] = {};
    std::mutex mu;

    ~GramImpl() override {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != device) (void)hipSetDevice(device);
        for (T* s : spectra) (void)hipFree(s);
        if (tw) (void)hipFree(tw);
        if (tmp) (void)hipFree(tmp);
        if (tmp_all) (void)hipFree(tmp_all);
        if (tmpt_all) (void)hipFree(tmpt_all);
        if (d_spec) (void)hipFree(d_spec);
        if (cur != device) (void)hipSetDevice(cur);
    }
    bool single() const override { return sizeof(T) == 4; }
    int side() const override { return n; }
    void stats(double* out) const override { std::copy(st, st + kGramStats, out); }
    int cols_per_block() const { return std::max(1, 8192 / (2 * n)); }

    void init_fft() {
        std::vector<double> t;
        if (!fft_describe(2 * n, d, t)) throw GramError{"twice the image side is not a product of 2, 3, 5, 7 and 11", 1};
        if (2 * n > lds_limit<T>()) throw GramError{"twice the image side exceeds the in-LDS transform limit", 1};
        std::vector<T> tt(t.begin(), t.end());
        GCHK(hipMalloc(&tw, sizeof(T) * tt.size()));
        GCHK(hipMemcpy(tw, tt.data(), sizeof(T) * tt.size(), hipMemcpyHostToDevice));
        GCHK(hipMalloc(&tmp, sizeof(cx<T>) * std::size_t(n) * 2 * n));
    }

    // 2-D forward transform of a 2n x 2n complex array in place.
    void fft2(cx<T>* a, hipStream_t s) {
        const int big = 2 * n;
        dispatch<T>(big, [&](auto nt, auto mt) {
            constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
            const std::size_t lds = sizeof(cx<T>) * big;
            lds_attr(k_fft_rows<T, NT, MT>, lds);
            k_fft_rows<T, NT, MT><<<big, NT, lds, s>>>(a, d, tw, -1);
        });
        GCHK(hipGetLastError());
        const int C = cols_per_block();
        dispatch<T>(big * C, [&](auto nt, auto mt) {
            constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
            const std::size_t lds = sizeof(cx<T>) * big * C;
            lds_attr(k_fft_cols<T, NT, MT>, lds);
            k_fft_cols<T, NT, MT><<<(big + C - 1) / C, NT, lds, s>>>(a, C, d, tw, -1);
        });
        GCHK(hipGetLastError());
    }

    // The nodes in passes of nodes_per_pass(), one launch per stage and pass:
    // the forward rows of the pass's nodes, their columns (forward, times the
    // node's spectrum, inverse), and the inverse rows summed over the nodes
    // into `out` (written by the first pass, added to by the others).
    // Nodes per pass: their half-transformed slices within 2 GiB
    // (RANG_GRAM_PASS_BYTES lowers the budget, a test knob; it is read at an
    // operator's first application, which sizes the slices).
    //
    // A column pass reading the n x 2n slice in place touches one cell per
    // row; once a slice outgrows the caches (256 MiB and more, n >= 4096 in
    // single precision) the slices are transposed through LDS tiles around
    // the column pass instead, which then reads and writes whole columns
    // (RANG_GRAM_TRANSPOSE=0 or 1 forces either form, a test knob read with
    // the pass budget).
    int pass_nodes = 0;
    bool transposed = false;
    static bool transpose_for(std::size_t slice) {
        if (const char* e = std::getenv("RANG_GRAM_TRANSPOSE")) {
            if (e[0] == '0') return false;
            if (e[0] == '1') return true;
        }
        return slice >= (std::size_t(256) << 20);
    }
    int nodes_per_pass() const {
        std::size_t budget = std::size_t(2) << 30;
        if (const char* e = std::getenv("RANG_GRAM_PASS_BYTES")) {
            const long long v = std::atoll(e);
            if (v > 0 && std::size_t(v) < budget) budget = std::size_t(v);
        }
        const std::size_t slice = sizeof(cx<T>) * std::size_t(n) * 2 * n * (transposed ? 2 : 1);
        const std::size_t k = std::max<std::size_t>(1, budget / slice);
        return int(std::min<std::size_t>(k, spectra.size()));
    }

    // [P][R][C] -> [P][C][R] through LDS tiles.
    static void transpose(const cx<T>* in, cx<T>* out, int P, int R, int C, hipStream_t s) {
        constexpr int TS = sizeof(T) == 4 ? 64 : 32;
        const dim3 grid(unsigned((C + TS - 1) / TS), unsigned((R + TS - 1) / TS), unsigned(P));
        if constexpr (sizeof(T) == 4) {
            if (R % 2 == 0 && C % 2 == 0) {
                k_tile_transpose_pairs<T><<<grid, 256, 0, s>>>(in, out, R, C);
                GCHK(hipGetLastError());
                return;
            }
        }
        k_tile_transpose<T><<<grid, 256, 0, s>>>(in, out, R, C);
        GCHK(hipGetLastError());
    }

    void apply(const T* x, T* out, hipStream_t s) {
        const int big = 2 * n;
        const int C = cols_per_block();
        const int K = int(spectra.size());
        if (!tmp_all) {
            transposed = transpose_for(sizeof(cx<T>) * std::size_t(n) * big);
            pass_nodes = nodes_per_pass();
            GCHK(hipMalloc(&tmp_all, sizeof(cx<T>) * std::size_t(pass_nodes) * n * big));
            if (transposed) GCHK(hipMalloc(&tmpt_all, sizeof(cx<T>) * std::size_t(pass_nodes) * n * big));
            GCHK(hipMalloc(&d_spec, sizeof(T*) * std::size_t(K)));
            GCHK(hipMemcpy(d_spec, spectra.data(), sizeof(T*) * std::size_t(K), hipMemcpyHostToDevice));
        }
        const int KP = pass_nodes;
        const int uw = use_w ? 1 : 0;
        for (int k0 = 0; k0 < K; k0 += KP) {
            const int kn = std::min(KP, K - k0);
            const double wk0 = w0 + double(k0) * dw;
            dispatch<T>(big, [&](auto nt, auto mt) {
                constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                const std::size_t lds = sizeof(cx<T>) * big;
                lds_attr(k_ap_rows_fwd<T, NT, MT>, lds);
                k_ap_rows_fwd<T, NT, MT><<<dim3(n, kn), NT, lds, s>>>(x, n, uw, wk0, dw, dx, dy, tmp_all, d, tw);
            });
            GCHK(hipGetLastError());
            if (transposed) {
                transpose(tmp_all, tmpt_all, kn, n, big, s);
                dispatch<T>(big * C, [&](auto nt, auto mt) {
                    constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                    const std::size_t lds = sizeof(cx<T>) * big * C;
                    lds_attr(k_ap_cols_t<T, NT, MT>, lds);
                    k_ap_cols_t<T, NT, MT><<<dim3((big + C - 1) / C, kn), NT, lds, s>>>(tmpt_all, n, d_spec + k0, C,
                                                                                        d, tw);
                });
                GCHK(hipGetLastError());
                transpose(tmpt_all, tmp_all, kn, big, n, s);
            } else {
                dispatch<T>(big * C, [&](auto nt, auto mt) {
                    constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                    const std::size_t lds = sizeof(cx<T>) * big * C;
                    lds_attr(k_ap_cols<T, NT, MT>, lds);
                    k_ap_cols<T, NT, MT><<<dim3((big + C - 1) / C, kn), NT, lds, s>>>(tmp_all, n, d_spec + k0, C, d,
                                                                                      tw);
                });
                GCHK(hipGetLastError());
            }
            dispatch<T>(big, [&](auto nt, auto mt) {
                constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                const std::size_t lds = sizeof(cx<T>) * big;
                lds_attr(k_ap_rows_inv<T, NT, MT>, lds);
                k_ap_rows_inv<T, NT, MT><<<n, NT, lds, s>>>(tmp_all, n, kn, uw, wk0, dw, dx, dy, out, k0 == 0 ? 1 : 0,
                                                            d, tw);
            });
            GCHK(hipGetLastError());
        }
    }
};

unsigned nblocks(long long n) { return unsigned((n + 255) / 256); }

template <class T>
GramBase* gram_build(const Options& o, const DevInputs& in, const double* freq_host, const WExpansion* ex,
                     double epsilon, hipStream_t st) {
    if (o.nx != o.ny || o.nx % 2 || o.nx == 0) throw GramError{"the Gram operator needs a square image of even side", 1};
    if (o.divide_by_n || o.center_x != 0.0 || o.center_y != 0.0)
        throw GramError{"the Gram operator handles divide_by_n = false, no shift", 1};
    if (in.wgt_kind != kWgtF32 && in.wgt_kind != kWgtF64) throw GramError{"the Gram operator needs f32 or f64 weights", 1};
    const std::size_t nrow = in.nrow, nchan = in.nchan;
    if (nrow == 0 || nchan == 0) throw GramError{"no visibilities", 1};
    std::string err;
    if (!check_freq(freq_host, nchan, err)) throw GramError{err, 1};
    std::unique_ptr<GramImpl<T>> g(new GramImpl<T>());
    GCHK(hipGetDevice(&g->device));
    g->n = int(o.nx);
    g->dx = o.pixsize_x;
    g->dy = o.pixsize_y;
    g->init_fft();
    const int n = g->n, big = 2 * n;
    const long long nvis = (long long)(nrow * nchan);
    // folded rows and |w|
    double *fold = nullptr, *absw = nullptr, *foc = nullptr;
    T* pw = nullptr;
    T* img = nullptr;
    cx<T>* kg = nullptr;
    cx<T>* unit = nullptr;
    double* bmax = nullptr;
    unsigned long long* cnt = nullptr;
    double* table = nullptr;
    struct Frees {
        std::vector<void*> p;
        ~Frees() {
            for (void* q : p)
                if (q) (void)hipFree(q);
        }
    } frees;
    auto dalloc = [&](auto** ptr, std::size_t bytes) {
        GCHK(hipMalloc(reinterpret_cast<void**>(ptr), std::max<std::size_t>(bytes, 256)));
        frees.p.push_back(*ptr);
    };
    dalloc(&fold, sizeof(double) * 3 * nrow);
    dalloc(&absw, sizeof(double) * nrow);
    dalloc(&foc, sizeof(double) * nchan);
    dalloc(&pw, sizeof(T) * nvis);
    dalloc(&img, sizeof(T) * std::size_t(big) * big);
    dalloc(&kg, sizeof(cx<T>) * std::size_t(big) * big);
    const long long nspec = (long long)big * big;
    const unsigned sblocks = nblocks(nspec);
    dalloc(&bmax, sizeof(double) * 2 * sblocks);
    dalloc(&cnt, sizeof(unsigned long long) * 2);
    const double fw = o.flip_w ? -1.0 : 1.0;
    k_fold<<<nblocks(nrow), 256, 0, st>>>(in.uvw, (long long)nrow, fw, fold, absw);
    GCHK(hipGetLastError());
    std::vector<double> fh(nchan);
    for (std::size_t c = 0; c < nchan; ++c) fh[c] = freq_host[c] / kCLight;
    GCHK(hipMemcpyAsync(foc, fh.data(), sizeof(double) * nchan, hipMemcpyHostToDevice, st));
    // the w range over every row (gridder_gram.rs:275-282) and x over the image
    std::vector<double> habs(nrow);
    GCHK(hipMemcpyAsync(habs.data(), absw, sizeof(double) * nrow, hipMemcpyDeviceToHost, st));
    GCHK(hipStreamSynchronize(st));
    double wmin = INFINITY, wmax = 0.0;
    for (double a : habs) {
        wmin = std::min(wmin, a * fh[0]);
        wmax = std::max(wmax, a * fh[nchan - 1]);
    }
    // x = n - 1 is largest (0) at the centre pixel and smallest at the
    // corner farthest from it, (0, 0)
    const double xmin = std::min(0.0, nm1_of(-(n / 2) * o.pixsize_x, -(n / 2) * o.pixsize_y));
    const double xmax = 0.0;
    const double x_range = std::max(xmax - xmin, 1e-300);
    Options go = o;
    go.nx = go.ny = std::size_t(big);
    go.do_wgridding = false;
    go.epsilon = epsilon;
    go.kernel_index = -1;
    go.nu = go.nv = 0;
    DevInputs gi;
    gi.nrow = nrow;
    gi.nchan = nchan;
    gi.uvw = fold;
    gi.freq = in.freq;
    gi.wgt = pw;
    gi.wgt_kind = sizeof(T) == 4 ? kWgtF32 : kWgtF64;
    double info[kNInfo];
    const double scale = 1.0 / (double(big) * big);
    int planes = 1;
    long long cells = 0;
    int s = 0;
    if (ex) {
        s = ex->stencil;
        g->use_w = true;
        g->dw = ex->theta / (2.0 * kPi * x_range);
        const double ncell = std::floor((wmax - wmin) / g->dw);
        if (!(ncell < 4096.0)) throw GramError{"more than 4096 w-nodes for the Gram operator", 1};
        cells = (long long)ncell;
        planes = int(cells) + s;
        g->w0 = wmin - double(s / 2 - 1) * g->dw;
        dalloc(&table, sizeof(double) * ex->table.size());
        GCHK(hipMemcpyAsync(table, ex->table.data(), sizeof(double) * ex->table.size(), hipMemcpyHostToDevice, st));
        dalloc(&unit, sizeof(cx<T>) * nvis);
        k_const_vis<T><<<nblocks(nvis), 256, 0, st>>>(unit, nvis);
        GCHK(hipGetLastError());
        gi.vis = unit;
        gi.vis_kind = sizeof(T) == 4 ? kVisC64 : kVisC128;
    } else {
        if (in.wgt_kind == kWgtF32)
            k_copy_wgt<T, float><<<nblocks(nvis), 256, 0, st>>>(static_cast<const float*>(in.wgt), pw, nvis);
        else
            k_copy_wgt<T, double><<<nblocks(nvis), 256, 0, st>>>(static_cast<const double*>(in.wgt), pw, nvis);
        GCHK(hipGetLastError());
        gi.vis = nullptr;
        gi.vis_kind = kVisUnit;
    }
    double herm = 0.0, peak = 0.0;
    unsigned long long active = 0, small = 0;
    for (int k = 0; k < planes; ++k) {
        GCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * 2, st));
        if (ex) {
            if (in.wgt_kind == kWgtF32)
                k_node_wgt<T, float><<<nblocks(nvis), 256, 0, st>>>(absw, foc, static_cast<const float*>(in.wgt), nvis,
                                                                    int(nchan), wmin, g->dw, cells, s, k, table, pw, cnt);
            else
                k_node_wgt<T, double><<<nblocks(nvis), 256, 0, st>>>(absw, foc, static_cast<const double*>(in.wgt),
                                                                     nvis, int(nchan), wmin, g->dw, cells, s, k, table,
                                                                     pw, cnt);
            GCHK(hipGetLastError());
        }
        std::string gerr;
        const int status = ex ? device_vis2dirty<T>(go, gi, freq_host, img, nullptr, info, st, gerr)
                              : device_vis2dirty<T>(go, gi, freq_host, nullptr, img, info, st, gerr);
        if (status) throw GramError{"gridding a Gram node: " + gerr, status};
        unsigned long long hc[2] = {0, 0};
        GCHK(hipMemcpyAsync(hc, cnt, sizeof hc, hipMemcpyDeviceToHost, st));
        k_kgrid<T><<<dim3((big + 255) / 256, big), 256, 0, st>>>(img, n, ex ? 0 : 1, kg);
        GCHK(hipGetLastError());
        g->fft2(kg, st);
        T* spec = nullptr;
        GCHK(hipMalloc(&spec, sizeof(T) * nspec));
        g->spectra.push_back(spec);
        k_spectrum<T><<<sblocks, 256, 0, st>>>(kg, big, scale, spec, bmax);
        GCHK(hipGetLastError());
        std::vector<double> hb(2 * std::size_t(sblocks));
        GCHK(hipMemcpyAsync(hb.data(), bmax, sizeof(double) * hb.size(), hipMemcpyDeviceToHost, st));
        GCHK(hipStreamSynchronize(st));
        double re_max = 0.0, im_max = 0.0;
        for (unsigned b = 0; b < sblocks; ++b) {
            re_max = std::max(re_max, hb[2 * b]);
            im_max = std::max(im_max, hb[2 * b + 1]);
        }
        herm = std::max(herm, im_max / std::max(re_max, 1e-300));
        peak = std::max(peak, re_max);
        k_count_small<T><<<unsigned(std::min<long long>(kCountBlocks, (nspec + 255) / 256)), 256, 0, st>>>(
            spec, nspec, 1e-7 * re_max * scale, cnt + 1);
        GCHK(hipGetLastError());
        unsigned long long hs = 0;
        GCHK(hipMemcpyAsync(&hs, cnt + 1, sizeof hs, hipMemcpyDeviceToHost, st));
        GCHK(hipStreamSynchronize(st));
        active += ex ? hc[0] : 0;
        small += hs;
    }
    double* o_ = g->st;
    o_[0] = planes;
    o_[1] = s;
    o_[2] = ex ? ex->theta : 0.0;
    o_[3] = ex ? wmin : 0.0;
    o_[4] = ex ? wmax : 0.0;
    o_[5] = ex ? x_range : 0.0;
    o_[6] = ex ? ex->sup_error : 0.0;
    o_[7] = ex ? ex->lebesgue : 0.0;
    o_[8] = double(active);
    o_[9] = double(planes) * double(nspec) * sizeof(T);
    o_[10] = ex ? herm : 0.0;
    o_[11] = double(small) / (double(planes) * double(nspec));
    o_[12] = peak;
    return g.release();
}

}  // namespace

template <class T>
GramBase* gram_new(const Options& o, const DevInputs& in, const double* freq_host, const WExpansion* ex,
                   double epsilon, hipStream_t st, int& status, std::string& err) {
    try {
        status = 0;
        return gram_build<T>(o, in, freq_host, ex, epsilon, st);
    } catch (const GramError& e) {
        status = e.status;
        err = e.msg;
        return nullptr;
    }
}

GramBase* gram_inverse_filter(const GramBase* base, double rho, hipStream_t st, int& status, std::string& err) {
    try {
        status = 0;
        if (!(rho > 0.0)) throw GramError{"the filter needs a one-node operator and a positive rho", 1};
        auto make = [&](auto tag) -> GramBase* {
            using T = decltype(tag);
            const auto& src = *static_cast<const GramImpl<T>*>(base);
            if (src.spectra.size() != 1 || src.use_w)
                throw GramError{"the filter needs a one-node operator and a positive rho", 1};
            std::unique_ptr<GramImpl<T>> g(new GramImpl<T>());
            GCHK(hipGetDevice(&g->device));
            g->n = src.n;
            g->dx = src.dx;
            g->dy = src.dy;
            g->init_fft();
            const long long nspec = 4LL * g->n * g->n;
            T* spec = nullptr;
            GCHK(hipMalloc(&spec, sizeof(T) * nspec));
            g->spectra.push_back(spec);
            k_inverse_filter<T><<<nblocks(nspec), 256, 0, st>>>(src.spectra[0], nspec, double(nspec), rho, spec);
            GCHK(hipGetLastError());
            GCHK(hipStreamSynchronize(st));
            std::copy(src.st, src.st + kGramStats, g->st);
            return g.release();
        };
        if (base->single()) return make(float());
        return make(double());
    } catch (const GramError& e) {
        status = e.status;
        err = e.msg;
        return nullptr;
    }
}

template <class T>
int gram_apply_dev(GramBase* base, const T* x, T* out, hipStream_t st, std::string& err) {
    try {
        auto& g = *static_cast<GramImpl<T>*>(base);
        std::lock_guard<std::mutex> lk(g.mu);
        g.apply(x, out, st);
        return 0;
    } catch (const GramError& e) {
        err = e.msg;
        return e.status;
    }
}

template GramBase* gram_new<float>(const Options&, const DevInputs&, const double*, const WExpansion*, double,
                                   hipStream_t, int&, std::string&);
template GramBase* gram_new<double>(const Options&, const DevInputs&, const double*, const WExpansion*, double,
                                    hipStream_t, int&, std::string&);
template int gram_apply_dev<float>(GramBase*, const float*, float*, hipStream_t, std::string&);
template int gram_apply_dev<double>(GramBase*, const double*, double*, hipStream_t, std::string&);

}  // namespace rang
