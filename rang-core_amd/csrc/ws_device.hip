// The w-stacking gridder (core/src/grid.rs, rang-core's port of WSClean 3.6's
// inversion) on the GPU: the w range and the counters, the placement of
// every sample on its w-layer's grid, then per batch of layers the gridding
// with the oversampled Kaiser-Bessel kernel, the backward 2-D transforms in
// LDS and the projection with the w-screen, added into the image in layer
// order; the normalisation, kernel correction, trim and PSF normalisation
// close. Grid and transforms in T (float or double), sums in double.
//
// Arithmetic is not contracted (no a*b+c fused unless written), as in the
// reference's Rust.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "wg_fft.h"
#include "ws_device.h"
#include "ws_fft.h"

namespace rang {
namespace ws {

namespace {

struct Fail {
    int status;
    std::string msg;
};

void chk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Fail{2, std::string(what) + ": " + hipGetErrorString(e)};
}
#define WSCHK(x) chk((x), #x)

constexpr double kC = 299792458.0;

struct Geo {
    int W, H, ks, os, nl;
    double psx, psy, min_w, max_w;  // max_w as the gridder uses it (never equal to min_w)
};

struct In {
    const double* uvw;
    const double* freq;
    const float* data;
    const float* wt;
    const float* iw;
    long long n;
    int nchan;
};

// A sample as WSClean's AddDataSample grids it (grid.rs:393-448).
struct Sample {
    int x, y;
    unsigned short xk, yk;
    int layer;
    float re, im, weight;
    int pad;
};

// WToLayer (grid.rs:334-347): -1 where the rounded value is negative.
__device__ __forceinline__ long long layer_of(const Geo& g, double w) {
    if (g.nl == 1) return 0;
    const double v = round((fabs(w) - g.min_w) * double(g.nl - 1) / (g.max_w - g.min_w));
    return v < 0.0 ? -1 : (long long)v;
}

// sqrt(1 - l^2 - m^2) - 1 at grid index (x, y) (grid.rs:364-391).
__device__ __forceinline__ double sqrt_lm(const Geo& g, int x, int y) {
    int ys = (g.H - y) + g.H / 2;
    if (ys >= g.H) ys -= g.H;
    const double m = (double(ys) - double(g.H / 2)) * g.psy;
    int xs = x + g.W / 2;
    if (xs >= g.W) xs -= g.W;
    const double l = (double(g.W / 2) - double(xs)) * g.psx;
    const double r = l * l + m * m;
    return r < 1.0 ? sqrt(1.0 - l * l - m * m) - 1.0 : 0.0;
}

__device__ __forceinline__ void block_reduce2(double& a, double& b, bool use_min) {
    __shared__ double sa[256], sb[256];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            sa[threadIdx.x] = use_min ? fmin(sa[threadIdx.x], sa[threadIdx.x + h]) : sa[threadIdx.x] + sa[threadIdx.x + h];
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + h]);
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
    __syncthreads();
}

// CalculateMsLimits (grid.rs:652-693): |w| range of the samples with both
// weights nonzero whose cell lies in the image, both bounds from u as in
// WSClean 3.6. Per block: min, max.
__global__ __launch_bounds__(256) void k_limits(In in, Geo g, double* part) {
    double lo = 1e100, hi = 0.0;
    const double hw = 0.5 * double(g.W), hh = 0.5 * double(g.H);
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < in.n; i += (long long)gridDim.x * 256) {
        if (in.wt[i] == 0.0f) continue;
        const long long row = i / in.nchan;
        const int ch = int(i - row * in.nchan);
        const double lam = kC / in.freq[ch];
        const double u = in.uvw[3 * row] / lam, w = in.uvw[3 * row + 2] / lam;
        const double x = u * g.psx * double(g.W), y = u * g.psy * double(g.H);
        if (in.iw[i] != 0.0f && floor(x) > -hw && ceil(x) < hw && floor(y) > -hh && ceil(y) < hh) {
            hi = fmax(hi, fabs(w));
            lo = fmin(lo, fabs(w));
        }
    }
    block_reduce2(lo, hi, true);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

// The placement of sample i (grid.rs:763-810 and 393-448): its layer, or -1
// where it adds nothing or falls outside every layer or the grid.
__device__ __forceinline__ bool place(const In& in, const Geo& g, long long i, Sample& s, float& weight_out,
                                      float& cw_out) {
    const long long row = i / in.nchan;
    const int ch = int(i - row * in.nchan);
    const double lam = kC / in.freq[ch];
    double u = in.uvw[3 * row] / lam, v = in.uvw[3 * row + 1] / lam;
    const double w = in.uvw[3 * row + 2] / lam;
    const long long layer = layer_of(g, w);
    const bool inrange = layer >= 0 && layer < g.nl;
    const float weight = inrange ? in.wt[i] : 0.0f;
    const float cw = weight * in.iw[i];
    weight_out = weight;
    cw_out = cw;
    const float re = in.data[2 * i] * cw;
    float im = in.data[2 * i + 1] * cw;
    if ((cw == 0.0f && re == 0.0f && im == 0.0f) || !inrange) return false;
    if (w < 0.0) {
        u = -u;
        v = -v;
        im = -im;
    }
    const long long W = g.W, H = g.H, OS = g.os;
    const double xe = u * g.psx * double(g.W), ye = v * g.psy * double(g.H);
    long long x = (long long)round(xe), y = (long long)round(ye);
    long long xk = (long long)round((xe - double(x)) * double(g.os));
    long long yk = (long long)round((ye - double(y)) * double(g.os));
    xk = (xk + (OS * 3) / 2) % OS;
    yk = (yk + (OS * 3) / 2) % OS;
    if (!(x > -W / 2 && y > -H / 2 && x <= W / 2 && y <= H / 2)) return false;
    if (x < 0) x += W;
    if (y < 0) y += H;
    s.x = int(x);
    s.y = int(y);
    s.xk = (unsigned short)xk;
    s.yk = (unsigned short)yk;
    s.layer = int(layer);
    s.re = re;
    s.im = im;
    s.weight = cw;
    s.pad = 0;
    return true;
}

// Per sample its layer (or -1) and per block the counters' partials: image
// weight, visibility weight sum, gridded count, largest gridded weight.
__global__ __launch_bounds__(256) void k_keys(In in, Geo g, int* keys, double* part) {
    double iwsum = 0.0, vsum = 0.0, cnt = 0.0, maxw = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < in.n; i += (long long)gridDim.x * 256) {
        Sample s;
        float weight, cw;
        const bool ok = place(in, g, i, s, weight, cw);
        keys[i] = ok ? s.layer : -1;
        if (cw != 0.0f) {
            vsum += double(weight);
            cnt += 1.0;
        }
        maxw = fmax(maxw, double(cw));
        iwsum += double(cw);
    }
    double a = iwsum, b = maxw;
    block_reduce2(a, b, false);
    double c = vsum, d = 0.0;
    block_reduce2(c, d, false);
    double e = cnt, f = 0.0;
    block_reduce2(e, f, false);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = a;
        part[4 * blockIdx.x + 1] = c;
        part[4 * blockIdx.x + 2] = e;
        part[4 * blockIdx.x + 3] = b;
    }
}

// The counting sort by layer in contiguous chunks, one block each: per
// block the histogram of its chunk's layers in LDS (no global atomics on the
// few layer counters), and the scatter through LDS cursors that start at the
// block's offsets within each layer (chunks in order; within a chunk the
// order of the LDS atomics).
__global__ __launch_bounds__(256) void k_hist_chunks(const int* keys, long long n, long long chunk, int L,
                                                     unsigned* blockcounts) {
    extern __shared__ unsigned hist[];
    for (int l = threadIdx.x; l < L; l += 256) hist[l] = 0;
    __syncthreads();
    const long long a = (long long)blockIdx.x * chunk, b = a + chunk < n ? a + chunk : n;
    for (long long i = a + threadIdx.x; i < b; i += 256)
        if (keys[i] >= 0) atomicAdd(&hist[keys[i]], 1u);
    __syncthreads();
    for (int l = threadIdx.x; l < L; l += 256) blockcounts[(long long)blockIdx.x * L + l] = hist[l];
}

__global__ __launch_bounds__(256) void k_scatter_chunks(In in, Geo g, const int* keys, long long chunk, int L,
                                                        const unsigned* blockoff, Sample* out) {
    extern __shared__ unsigned cur[];
    for (int l = threadIdx.x; l < L; l += 256) cur[l] = blockoff[(long long)blockIdx.x * L + l];
    __syncthreads();
    const long long a = (long long)blockIdx.x * chunk, b = a + chunk < in.n ? a + chunk : in.n;
    for (long long i = a + threadIdx.x; i < b; i += 256) {
        if (keys[i] < 0) continue;
        Sample s;
        float weight, cw;
        place(in, g, i, s, weight, cw);
        out[atomicAdd(&cur[s.layer], 1u)] = s;
    }
}

// grid_layer (grid.rs:450-495): one wave per sample, a lane per tap; the
// weighted sample into the dirty grid and its weight into the PSF grid of
// the sample's slot in the batch. Taps in double, added in T.
template <class T>
__global__ __launch_bounds__(256) void k_grid(const Sample* s, long long s0, long long s1, const int* slot,
                                              Geo g, const double* kern, cx<T>* gd, cx<T>* gp) {
    const long long k = s0 + ((long long)blockIdx.x * 256 + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (k >= s1) return;
    const Sample q = s[k];
    const long long plane = (long long)g.W * g.H;
    const long long base = (long long)slot[q.layer] * plane;
    const int mid = g.ks / 2;
    const double re = double(q.re), im = double(q.im), wt = double(q.weight);
    for (int tap = lane; tap < g.ks * g.ks; tap += 64) {
        const int j = tap / g.ks, i = tap - j * g.ks;
        const int cy = (q.y + j + g.H - mid) % g.H;
        const int cx = (q.x + i + g.W - mid) % g.W;
        const double kv = kern[q.yk * g.ks + j] * kern[q.xk * g.ks + i];
        const long long at = base + cx + (long long)cy * g.W;
        atomicAdd(&gd[at].re, T(re * kv));
        atomicAdd(&gd[at].im, T(im * kv));
        atomicAdd(&gp[at].re, T(wt * kv));
    }
}

// Unnormalised backward transform of contiguous rows of length d.n, in place.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_rows(cx<T>* data, FftDesc d, const cx<T>* tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    cx<T>* row = data + (long long)blockIdx.x * d.n;
    for (int e = threadIdx.x; e < d.n; e += NT) buf[e] = row[e];
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, 1, tw, +1);
    for (int e = threadIdx.x; e < d.n; e += NT) row[e] = buf[e];
}

// A linear block index remapped so that the workgroups one XCD holds at a
// time take consecutive indices (consecutive workgroups go to the eight
// XCDs in turn, each with its own L2).
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    constexpr int X = 8;
    const int per = nblocks / X, rem = nblocks - per * X;
    const int x = bid % X, k = bid / X;
    return x * per + min(x, rem) + k;
}

// Backward transform of C neighbouring columns (length d.n = H, stride W) of
// one of the grids, in place.
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_cols(cx<T>* data, FftDesc d, const cx<T>* tw, int W, int C, int ncb) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int gi = lin / ncb, x0 = (lin - gi * ncb) * C, nc = min(C, W - x0);
    const int H = d.n;
    cx<T>* base = data + (long long)gi * W * H;
    for (int e = threadIdx.x; e < H * nc; e += NT) {
        const int y = e / nc, c = e - y * nc;
        buf[c * H + y] = base[(long long)y * W + x0 + c];
    }
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d, nc, tw, +1);
    for (int e = threadIdx.x; e < H * nc; e += NT) {
        const int y = e / nc, c = e - y * nc;
        base[(long long)y * W + x0 + c] = buf[c * H + y];
    }
}

// project (grid.rs:497-529) of the batch's layers, their terms added to the
// running sums in layer order.
template <class T>
__global__ __launch_bounds__(256) void k_project(const cx<T>* gd, const cx<T>* gp, int nb, const double* wl,
                                                 Geo g, double* sd, double* sp) {
    const long long plane = (long long)g.W * g.H;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= plane) return;
    const int y = int(e / g.W), x = int(e - (long long)y * g.W);
    int ys = (g.H - y) + g.H / 2;
    if (ys >= g.H) ys -= g.H;
    int xs = x + g.W / 2;
    if (xs >= g.W) xs -= g.W;
    const long long t = xs + (long long)ys * g.W;
    const double slm = sqrt_lm(g, x, y);
    double ad = sd[t], ap = sp[t];
    for (int b = 0; b < nb; ++b) {
        const double rad = (-2.0 * M_PI * wl[b]) * slm;
        double sn, cs;
        sincos(rad, &sn, &cs);
        const cx<T> vd = gd[b * plane + e], vp = gp[b * plane + e];
        ad += double(vd.re) * cs - double(vd.im) * sn;
        ap += double(vp.re) * cs - double(vp.im) * sn;
    }
    sd[t] = ad;
    sp[t] = ap;
}

// normalise and correct_for_kernel (grid.rs:597-650), then the trim
// (grid.rs:91-107) into the output images.
__global__ __launch_bounds__(256) void k_finish(const double* sd, const double* sp, Geo g, double factor,
                                                const double* ox, const double* oy, int ow, int oh,
                                                double* od, double* op) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)ow * oh) return;
    const int yo = int(e / ow), xo = int(e - (long long)yo * ow);
    const int x = (g.W - ow) / 2 + xo, y = (g.H - oh) / 2 + yo;
    const double yv = y >= g.H / 2 ? oy[y - g.H / 2] : oy[g.H / 2 - y];
    const double xv = x >= g.W / 2 ? ox[x - g.W / 2] : ox[g.W / 2 - x];
    const double div = xv * yv * (1.0 / (double(g.os) * double(g.os)));
    const long long s = x + (long long)y * g.W;
    od[e] = (sd[s] * factor) / div;
    op[e] = (sp[s] * factor) / div;
}

// dct3_prefix (grid.rs:263-279): out[k] = in[0] + 2 sum_j in[j] cos(pi j (k + 1/2) / n),
// one block per k.
__global__ __launch_bounds__(256) void k_dct3(const double* in, int len, double n, double* out) {
    const double kk = double(blockIdx.x) + 0.5;
    double s = 0.0, unused = 0.0;
    for (int j = 1 + int(threadIdx.x); j < len; j += 256) s += 2.0 * in[j] * cos(M_PI * double(j) * kk / n);
    block_reduce2(s, unused, false);
    if (threadIdx.x == 0) out[blockIdx.x] = in[0] + s;
}

template <int V>
using ic = std::integral_constant<int, V>;

template <class T, class F>
void fft_sizes(int need, F&& f) {
    if (need <= 1024)
        f(ic<256>{}, ic<1024>{});
    else if (need <= 2048)
        f(ic<256>{}, ic<2048>{});
    else if (need <= 4096)
        f(ic<256>{}, ic<4096>{});
    else if (need <= 8192)
        f(ic<512>{}, ic<8192>{});
    else if constexpr (sizeof(T) == 4) {
        if (need <= 16384)
            f(ic<512>{}, ic<16384>{});
        else
            throw Fail{1, "padded image side above the GPU transform limit"};
    } else {
        throw Fail{1, "padded image side above the GPU transform limit"};
    }
}

template <class K>
void set_lds(K kernel, std::size_t bytes) {
    WSCHK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
}

// Device buffers of one inversion, freed on every exit.
// Device buffers kept per device between calls (allocating and freeing
// gigabytes per call cost more than the gridding): the k-th request of a
// call reuses the k-th slot, grown when too small. Calls on one device are
// serialised by the pool's mutex.
struct Pool {
    std::mutex mu;
    std::vector<std::pair<void*, std::size_t>> slots;
};

Pool& pool_of_current_device() {
    static std::mutex mu;
    static std::map<int, Pool*> pools;
    int dev = 0;
    WSCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    Pool*& p = pools[dev];
    if (!p) p = new Pool;  // kept until the process ends, as the buffers are
    return *p;
}

struct Bufs {
    Pool& pool;
    std::size_t next = 0;
    std::size_t bytes = 0;
    explicit Bufs(Pool& p) : pool(p) {}
    template <class P>
    P* get(std::size_t n) {
        const std::size_t need = std::max<std::size_t>(n * sizeof(P), 16);
        if (next >= pool.slots.size()) pool.slots.push_back({nullptr, 0});
        auto& s = pool.slots[next++];
        if (s.second < need) {
            if (s.first) WSCHK(hipFree(s.first));
            s.first = nullptr;
            s.second = 0;
            WSCHK(hipMalloc(&s.first, need));
            s.second = need;
        }
        bytes += n * sizeof(P);
        return static_cast<P*>(s.first);
    }
};

// Transforms of contiguous lines of any length N in place (ws_fft.h): kind
// 0 one line in LDS, 1 four steps of N = M1 M2, 2 Bluestein in LDS, 3
// Bluestein over four steps of M = M1 M2.
template <class T>
struct LineFft {
    int N = 0, kind = 0, sgn = 1;
    long long M = 0;
    int M1 = 0, M2 = 0;
    FftDesc d, d1, d2;
    cx<T>*tw = nullptr, *tw1 = nullptr, *tw2 = nullptr, *chirp = nullptr, *bhat = nullptr;
    static constexpr int kBudget = 8192;  // values per workgroup in the four-step passes

    static cx<T>* upload_tw(Bufs& b, const std::vector<double>& t, hipStream_t st, std::vector<std::vector<cx<T>>>& keep) {
        keep.emplace_back(t.size() / 2);
        auto& h = keep.back();
        for (std::size_t q = 0; q < h.size(); ++q) h[q] = {T(t[2 * q]), T(t[2 * q + 1])};
        cx<T>* dptr = b.get<cx<T>>(std::max<std::size_t>(h.size(), 1));
        if (!h.empty()) WSCHK(hipMemcpyAsync(dptr, h.data(), sizeof(cx<T>) * h.size(), hipMemcpyHostToDevice, st));
        return dptr;
    }

    void describe(int n, FftDesc& dd, cx<T>*& t, Bufs& b, hipStream_t st, std::vector<std::vector<cx<T>>>& keep) {
        std::vector<double> tt;
        if (!fft_describe(n, dd, tt)) throw Fail{1, "transform length not a product of 2, 3, 5, 7 and 11"};
        t = upload_tw(b, tt, st, keep);
    }

    void setup(int n, int sign, Bufs& b, hipStream_t st, std::vector<std::vector<cx<T>>>& keep) {
        N = n;
        sgn = sign;
        const int L = lds_values<T>();
        if (is_good(N) && N <= L) {
            kind = 0;
            describe(N, d, tw, b, st, keep);
            return;
        }
        if (is_good(N)) {
            kind = 1;
            M = N;
            if (!split_two(N, L, M1, M2)) throw Fail{1, "no two-pass split of the transform length"};
            describe(M1, d1, tw1, b, st, keep);
            describe(M2, d2, tw2, b, st, keep);
            return;
        }
        M = good_length(2LL * N - 1);
        // the chirp c_k = exp(sgn pi i k^2 / N) and the filter b_m = conj(c_m) at +-m
        std::vector<cx<T>> hc(N);
        std::vector<cx<double>> hb(std::size_t(M), cx<double>{0.0, 0.0});
        const long double pi = 3.141592653589793238462643383279502884L;
        for (long long k = 0; k < N; ++k) {
            const long long r = (k * k) % (2LL * N);
            const long double a = pi * (long double)r / (long double)N;
            const double c = double(cosl(a)), sn = double(sinl(a)) * double(sgn);
            hc[k] = {T(c), T(sn)};
            hb[k] = {c, -sn};
            if (k > 0) hb[M - k] = {c, -sn};
        }
        keep.emplace_back(hc);
        chirp = b.get<cx<T>>(N);
        WSCHK(hipMemcpyAsync(chirp, keep.back().data(), sizeof(cx<T>) * N, hipMemcpyHostToDevice, st));
        keep.emplace_back(std::size_t(M));
        auto& hbt = keep.back();
        for (long long q = 0; q < M; ++q) hbt[q] = {T(hb[q].re), T(hb[q].im)};
        bhat = b.get<cx<T>>(std::size_t(M));
        WSCHK(hipMemcpyAsync(bhat, hbt.data(), sizeof(cx<T>) * M, hipMemcpyHostToDevice, st));
        if (M <= L) {
            kind = 2;
            describe(int(M), d, tw, b, st, keep);
            forward_lds(bhat, 1, st);
        } else {
            kind = 3;
            if (!split_two(M, L, M1, M2)) throw Fail{1, "no two-pass split of the chirp transform length"};
            describe(M1, d1, tw1, b, st, keep);
            describe(M2, d2, tw2, b, st, keep);
            cols(bhat, 1, -1, true, st);
            rows(bhat, 1, -1, false, st);
        }
        k_scale<T><<<unsigned((M + 255) / 256), 256, 0, st>>>(bhat, M, 1.0 / double(M));
        WSCHK(hipGetLastError());
    }

    void forward_lds(cx<T>* data, long long lines, hipStream_t st) {
        fft_sizes<T>(d.n, [&](auto nt, auto mt) {
            constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
            const std::size_t lds = sizeof(cx<T>) * d.n;
            set_lds(k_line_lds<T, NT, MT>, lds);
            for (long long l0 = 0; l0 < lines; l0 += (1LL << 30)) {
                const long long nl = std::min(lines - l0, 1LL << 30);
                k_line_lds<T, NT, MT><<<unsigned(nl), NT, lds, st>>>(data + l0 * d.n, d, tw, -1);
            }
        });
        WSCHK(hipGetLastError());
    }

    void cols(cx<T>* buf, long long lines, int s, bool twid, hipStream_t st) {
        const int C = std::max(1, std::min(M2, kBudget / M1));
        const long long nb = lines * ((M2 + C - 1) / C);
        if (nb >= (1LL << 31)) throw Fail{1, "transform batch too large"};
        fft_sizes<T>(M1 * C, [&](auto nt, auto mt) {
            constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
            const std::size_t lds = sizeof(cx<T>) * M1 * C;
            set_lds(k_fs_cols<T, NT, MT>, lds);
            k_fs_cols<T, NT, MT><<<unsigned(nb), NT, lds, st>>>(buf, M1, M2, C, d1, tw1, s, twid ? 1 : 0);
        });
        WSCHK(hipGetLastError());
    }

    void rows(cx<T>* buf, long long lines, int s, bool twid, hipStream_t st) {
        const int K = std::max(1, std::min(M1, kBudget / M2));
        const long long nb = lines * ((M1 + K - 1) / K);
        if (nb >= (1LL << 31)) throw Fail{1, "transform batch too large"};
        fft_sizes<T>(M2 * K, [&](auto nt, auto mt) {
            constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
            const std::size_t lds = sizeof(cx<T>) * M2 * K;
            set_lds(k_fs_rows<T, NT, MT>, lds);
            k_fs_rows<T, NT, MT><<<unsigned(nb), NT, lds, st>>>(buf, M1, M2, K, d2, tw2, s, twid ? 1 : 0);
        });
        WSCHK(hipGetLastError());
    }

    static void transpose(const cx<T>* in, cx<T>* out, long long lines, int A, int B, hipStream_t st) {
        for (long long l0 = 0; l0 < lines; l0 += 65535) {
            const long long nl = std::min<long long>(65535, lines - l0);
            k_transpose<T><<<dim3((B + 31) / 32, (A + 31) / 32, unsigned(nl)), 256, 0, st>>>(
                in + l0 * (long long)A * B, out + l0 * (long long)A * B, A, B);
        }
        WSCHK(hipGetLastError());
    }

    // Lines of `scratch_values` complex values of work space at least
    // N (kind 1) or M (kind 3) times one line.
    void apply(cx<T>* data, long long lines, cx<T>* scratch, long long scratch_values, hipStream_t st) {
        if (kind == 0) {
            fft_sizes<T>(N, [&](auto nt, auto mt) {
                constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                const std::size_t lds = sizeof(cx<T>) * N;
                set_lds(k_line_lds<T, NT, MT>, lds);
                k_line_lds<T, NT, MT><<<unsigned(lines), NT, lds, st>>>(data, d, tw, sgn);
            });
            WSCHK(hipGetLastError());
            return;
        }
        if (kind == 2) {
            fft_sizes<T>(int(M), [&](auto nt, auto mt) {
                constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                const std::size_t lds = sizeof(cx<T>) * M;
                set_lds(k_blue_lds<T, NT, MT>, lds);
                k_blue_lds<T, NT, MT><<<unsigned(lines), NT, lds, st>>>(data, N, d, tw, chirp, bhat);
            });
            WSCHK(hipGetLastError());
            return;
        }
        const long long per = std::max<long long>(1, scratch_values / M);
        for (long long l0 = 0; l0 < lines; l0 += per) {
            const long long nl = std::min(per, lines - l0);
            cx<T>* x = data + l0 * N;
            if (kind == 1) {
                cols(x, nl, sgn, true, st);
                rows(x, nl, sgn, false, st);
                transpose(x, scratch, nl, M1, M2, st);
                WSCHK(hipMemcpyAsync(x, scratch, sizeof(cx<T>) * nl * N, hipMemcpyDeviceToDevice, st));
            } else {
                const unsigned gb = unsigned(std::min<long long>((nl * M + 255) / 256, 65536));
                k_blue_in<T><<<gb, 256, 0, st>>>(x, N, M, nl, chirp, scratch);
                cols(scratch, nl, -1, true, st);
                rows(scratch, nl, -1, false, st);
                k_mul_spec<T><<<gb, 256, 0, st>>>(scratch, M, nl, bhat);
                rows(scratch, nl, +1, true, st);
                cols(scratch, nl, +1, false, st);
                k_blue_out<T><<<gb, 256, 0, st>>>(scratch, N, M, nl, chirp, x);
                WSCHK(hipGetLastError());
            }
        }
    }
    // work space values apply() needs for one line
    long long scratch_per_line() const { return kind == 1 ? N : kind == 3 ? M : 0; }
};

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

std::size_t pad_side(std::size_t n, double padding) {
    std::size_t p = std::size_t(std::ceil(double(n) * padding));
    return p + (4 - p % 4) % 4;
}

double bessel0(double x, double precision) {
    double d = 0.0, ds = 1.0, sum = 1.0;
    for (;;) {
        d += 2.0;
        ds *= x * x / (d * d);
        sum += ds;
        if (ds <= sum * precision) break;
    }
    return sum;
}

template <class T>
int invert(const GridInputs& i, const GridOutputs& o, double* timings, std::string& err) {
    const auto t_all = Clock::now();
    double t_prepare = 0, t_bucket = 0, t_upload = 0, t_grid = 0, t_fft = 0, t_project = 0, t_download = 0,
           t_finish = 0;
    const std::size_t nrow = i.nrow, nchan = i.nchan;
    const std::size_t n = nrow * nchan;
    const std::size_t W = pad_side(i.width, i.padding), H = pad_side(i.height, i.padding);
    Geo g{int(W), int(H), int(i.kernel_size), int(i.oversampling), int(i.n_w_layers), i.pixel_size_x,
          i.pixel_size_y, 0.0, 0.0};
    hipStream_t st = nullptr;
    WSCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    Pool& pool = pool_of_current_device();
    std::lock_guard<std::mutex> pool_lock(pool.mu);
    Bufs b(pool);
    // upload
    auto t0 = Clock::now();
    double* d_uvw = b.get<double>(3 * nrow);
    double* d_freq = b.get<double>(nchan);
    float* d_data = b.get<float>(2 * n);
    float* d_wt = b.get<float>(n);
    float* d_iw = b.get<float>(n);
    WSCHK(hipMemcpyAsync(d_uvw, i.uvw, sizeof(double) * 3 * nrow, hipMemcpyHostToDevice, st));
    WSCHK(hipMemcpyAsync(d_freq, i.frequencies, sizeof(double) * nchan, hipMemcpyHostToDevice, st));
    WSCHK(hipMemcpyAsync(d_data, i.data, sizeof(float) * 2 * n, hipMemcpyHostToDevice, st));
    WSCHK(hipMemcpyAsync(d_wt, i.weights, sizeof(float) * n, hipMemcpyHostToDevice, st));
    WSCHK(hipMemcpyAsync(d_iw, i.imaging_weights, sizeof(float) * n, hipMemcpyHostToDevice, st));
    WSCHK(hipStreamSynchronize(st));
    t_upload = since(t0);
    In in{d_uvw, d_freq, d_data, d_wt, d_iw, (long long)n, int(nchan)};
    // w limits
    t0 = Clock::now();
    const int nblk = int(std::max<long long>(1, std::min<long long>(1024, ((long long)n + 255) / 256)));
    double* d_part = b.get<double>(4 * nblk);
    std::vector<double> part(4 * nblk);
    k_limits<<<nblk, 256, 0, st>>>(in, g, d_part);
    WSCHK(hipGetLastError());
    WSCHK(hipMemcpyAsync(part.data(), d_part, sizeof(double) * 2 * nblk, hipMemcpyDeviceToHost, st));
    WSCHK(hipStreamSynchronize(st));
    double min_w = 1e100, max_w = 0.0;
    for (int k = 0; k < nblk; ++k) {
        min_w = std::min(min_w, part[2 * k]);
        max_w = std::max(max_w, part[2 * k + 1]);
    }
    if (min_w == 1e100) min_w = max_w = 0.0;
    g.min_w = min_w;
    g.max_w = min_w == max_w ? max_w + 1.0 : max_w;
    // placement, counters, counting sort by layer
    int* d_keys = b.get<int>(n);
    k_keys<<<nblk, 256, 0, st>>>(in, g, d_keys, d_part);
    WSCHK(hipGetLastError());
    WSCHK(hipMemcpyAsync(part.data(), d_part, sizeof(double) * 4 * nblk, hipMemcpyDeviceToHost, st));
    const int L = int(i.n_w_layers);
    const int nch = 1024;
    const long long chunk = ((long long)n + nch - 1) / nch;
    unsigned* d_bc = b.get<unsigned>(std::size_t(nch) * L);
    if (std::size_t(L) * sizeof(unsigned) > 64 * 1024) throw Fail{1, "too many w-layers for the GPU sort"};
    k_hist_chunks<<<nch, 256, sizeof(unsigned) * L, st>>>(d_keys, (long long)n, chunk, L, d_bc);
    WSCHK(hipGetLastError());
    std::vector<unsigned> bc(std::size_t(nch) * L);
    WSCHK(hipMemcpyAsync(bc.data(), d_bc, sizeof(unsigned) * bc.size(), hipMemcpyDeviceToHost, st));
    WSCHK(hipStreamSynchronize(st));
    double image_weight = 0, vsum = 0, gridded = 0, maxw = 0;
    for (int k = 0; k < nblk; ++k) {
        image_weight += part[4 * k];
        vsum += part[4 * k + 1];
        gridded += part[4 * k + 2];
        maxw = std::max(maxw, part[4 * k + 3]);
    }
    if (image_weight == 0.0) throw Fail{1, "no visibility carries weight"};
    std::vector<unsigned> cnt(std::size_t(L) + 1, 0), off(std::size_t(L) + 1, 0);
    for (int c = 0; c < nch; ++c)
        for (int l = 0; l < L; ++l) cnt[l] += bc[std::size_t(c) * L + l];
    for (int l = 0; l < L; ++l) off[l + 1] = off[l] + cnt[l];
    // block c's first slot in layer l: the layer's offset plus the earlier chunks' counts
    {
        std::vector<unsigned> run(off.begin(), off.end() - 1);
        for (int c = 0; c < nch; ++c)
            for (int l = 0; l < L; ++l) {
                const unsigned k = bc[std::size_t(c) * L + l];
                bc[std::size_t(c) * L + l] = run[l];
                run[l] += k;
            }
    }
    const std::size_t ns = off[L];
    Sample* d_s = b.get<Sample>(std::max<std::size_t>(ns, 1));
    WSCHK(hipMemcpyAsync(d_bc, bc.data(), sizeof(unsigned) * bc.size(), hipMemcpyHostToDevice, st));
    k_scatter_chunks<<<nch, 256, sizeof(unsigned) * L, st>>>(in, g, d_keys, chunk, L, d_bc, d_s);
    WSCHK(hipGetLastError());
    WSCHK(hipStreamSynchronize(st));
    t_bucket = since(t0);
    t_prepare = t_bucket;
    // kernels (grid.rs:165-202)
    const std::size_t ks = i.kernel_size, os = i.oversampling, klen = ks * os, mid = klen / 2;
    std::vector<double> k1(klen), kern(klen);
    {
        const double fr = 1.0 / double(os), norm = double(os) / bessel0(8.6, 1e-8);
        for (std::size_t q = 0; q <= mid && mid + q < klen; ++q) {
            const double x = double(q);
            const double sn = q == 0 ? fr : std::sin(M_PI * fr * x) / (M_PI * x);
            const double t = double(q) / double(mid);
            k1[mid + q] = sn * bessel0(8.6 * std::sqrt(1.0 - t * t), 1e-8) * norm;
        }
        for (std::size_t q = 0; q < mid; ++q) k1[q] = k1[klen - 1 - q];
        for (std::size_t s = 0; s < os; ++s)
            for (std::size_t x = 0; x < ks; ++x) kern[(os - s - 1) * ks + x] = k1[x * os + s];
    }
    double* d_kern = b.get<double>(klen);
    WSCHK(hipMemcpyAsync(d_kern, kern.data(), sizeof(double) * klen, hipMemcpyHostToDevice, st));
    // backward transforms along x (rows of W) and y (columns of H): in LDS
    // where the side allows, else in four steps or by Bluestein (ws_fft.h)
    std::vector<std::vector<cx<T>>> keep;
    LineFft<T> fx, fy;
    fx.setup(int(W), +1, b, st, keep);
    fy.setup(int(H), +1, b, st, keep);
    const int Cw = std::max(1, std::min(8, 8192 / int(H)));
    // columns in place in LDS blocks when H is a good size there; otherwise
    // each layer is transposed so that its columns become lines
    const bool cols_direct = fy.kind == 0 && int(H) * Cw <= lds_values<T>();
    WSCHK(hipStreamSynchronize(st));
    // batches of the layers holding samples
    std::vector<int> layers;
    for (std::size_t l = 0; l < i.n_w_layers; ++l)
        if (cnt[l]) layers.push_back(int(l));
    const long long plane = (long long)W * H;
    // a layer's dirty and PSF grids, plus a transposed copy of one grid when
    // the columns are not transformed in place
    const std::size_t per_layer = (cols_direct ? 2 : 3) * sizeof(cx<T>) * std::size_t(plane);
    const std::size_t budget = std::size_t(4) << 30;
    const int batch = int(std::max<std::size_t>(1, std::min<std::size_t>(layers.size(), budget / per_layer)));
    cx<T>* d_gd = b.get<cx<T>>(std::size_t(batch) * plane);
    cx<T>* d_gp = b.get<cx<T>>(std::size_t(batch) * plane);
    cx<T>* d_tr = cols_direct ? nullptr : b.get<cx<T>>(std::size_t(batch) * plane);
    // work space of the four-step and chirp transforms: at most 1 GiB
    const long long per_line = std::max(fx.scratch_per_line(), fy.scratch_per_line());
    const long long scratch_values =
        per_line ? std::max<long long>(per_line, std::min<long long>((1LL << 30) / sizeof(cx<T>),
                                                                     per_line * std::max<long long>(W, H) * batch))
                 : 0;
    cx<T>* d_fs = per_line ? b.get<cx<T>>(std::size_t(scratch_values)) : nullptr;
    double* d_sd = b.get<double>(plane);
    double* d_sp = b.get<double>(plane);
    WSCHK(hipMemsetAsync(d_sd, 0, sizeof(double) * plane, st));
    WSCHK(hipMemsetAsync(d_sp, 0, sizeof(double) * plane, st));
    int* d_slot = b.get<int>(i.n_w_layers);
    double* d_wl = b.get<double>(batch);
    std::vector<int> slot(i.n_w_layers, -1);
    std::vector<double> wl(batch);
    const int ncb = (int(W) + Cw - 1) / Cw;
    int nbatches = 0;
    for (std::size_t l0 = 0; l0 < layers.size(); l0 += batch) {
        const std::size_t l1 = std::min(layers.size(), l0 + batch);
        const int nb = int(l1 - l0);
        std::fill(slot.begin(), slot.end(), -1);
        for (std::size_t q = l0; q < l1; ++q) {
            slot[layers[q]] = int(q - l0);
            const int L = layers[q];
            // LayerToW (grid.rs:354-361)
            wl[q - l0] = i.n_w_layers == 1 ? 0.0 : double(L) * (g.max_w - g.min_w) / double(i.n_w_layers - 1) + g.min_w;
        }
        WSCHK(hipMemcpyAsync(d_slot, slot.data(), sizeof(int) * i.n_w_layers, hipMemcpyHostToDevice, st));
        WSCHK(hipMemcpyAsync(d_wl, wl.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st));
        WSCHK(hipMemsetAsync(d_gd, 0, sizeof(cx<T>) * nb * plane, st));
        WSCHK(hipMemsetAsync(d_gp, 0, sizeof(cx<T>) * nb * plane, st));
        auto tg = Clock::now();
        const long long s0 = off[layers[l0]], s1 = off[layers[l1 - 1] + 1];
        // a wave per sample: launches of at most 2^24 samples keep each
        // dispatch within the 2^32 work-items of one grid dimension
        for (long long c0 = s0; c0 < s1; c0 += (1LL << 24)) {
            const long long c1 = std::min(s1, c0 + (1LL << 24));
            k_grid<T><<<(c1 - c0 + 3) / 4, 256, 0, st>>>(d_s, c0, c1, d_slot, g, d_kern, d_gd, d_gp);
            WSCHK(hipGetLastError());
        }
        WSCHK(hipStreamSynchronize(st));
        t_grid += since(tg);
        tg = Clock::now();
        for (cx<T>* grid : {d_gd, d_gp}) {
            fx.apply(grid, (long long)nb * H, d_fs, scratch_values, st);
            if (cols_direct) {
                fft_sizes<T>(int(H) * Cw, [&](auto nt, auto mt) {
                    constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                    set_lds(k_cols<T, NT, MT>, sizeof(cx<T>) * H * Cw);
                    k_cols<T, NT, MT><<<nb * ncb, NT, sizeof(cx<T>) * H * Cw, st>>>(grid, fy.d, fy.tw, int(W), Cw,
                                                                                   ncb);
                });
                WSCHK(hipGetLastError());
            } else {
                LineFft<T>::transpose(grid, d_tr, nb, int(H), int(W), st);
                fy.apply(d_tr, (long long)nb * W, d_fs, scratch_values, st);
                LineFft<T>::transpose(d_tr, grid, nb, int(W), int(H), st);
            }
        }
        WSCHK(hipStreamSynchronize(st));
        t_fft += since(tg);
        tg = Clock::now();
        k_project<T><<<(plane + 255) / 256, 256, 0, st>>>(d_gd, d_gp, nb, d_wl, g, d_sd, d_sp);
        WSCHK(hipGetLastError());
        WSCHK(hipStreamSynchronize(st));
        t_project += since(tg);
        ++nbatches;
    }
    // normalise, correct, trim (grid.rs:597-650, 837-873)
    t0 = Clock::now();
    const std::size_t half = klen / 2, rlen = klen / 2 + 1;
    double* d_k1 = b.get<double>(rlen);
    WSCHK(hipMemcpyAsync(d_k1, k1.data() + half, sizeof(double) * rlen, hipMemcpyHostToDevice, st));
    double* d_ox = b.get<double>(W / 2 + 1);
    k_dct3<<<int(W / 2 + 1), 256, 0, st>>>(d_k1, int(rlen), double(W * os / 2), d_ox);
    WSCHK(hipGetLastError());
    double* d_oy = d_ox;
    if (H != W) {
        d_oy = b.get<double>(H / 2 + 1);
        k_dct3<<<int(H / 2 + 1), 256, 0, st>>>(d_k1, int(rlen), double(H * os / 2), d_oy);
        WSCHK(hipGetLastError());
    }
    const std::size_t npix = i.width * i.height;
    double* d_od = b.get<double>(npix);
    double* d_op = b.get<double>(npix);
    k_finish<<<(npix + 255) / 256, 256, 0, st>>>(d_sd, d_sp, g, 1.0 / image_weight, d_ox, d_oy, int(i.width),
                                                 int(i.height), d_od, d_op);
    WSCHK(hipGetLastError());
    WSCHK(hipStreamSynchronize(st));
    t_finish = since(t0);
    t0 = Clock::now();
    WSCHK(hipMemcpyAsync(o.dirty, d_od, sizeof(double) * npix, hipMemcpyDeviceToHost, st));
    WSCHK(hipMemcpyAsync(o.psf, d_op, sizeof(double) * npix, hipMemcpyDeviceToHost, st));
    WSCHK(hipStreamSynchronize(st));
    t_download = since(t0);
    t0 = Clock::now();
    const std::size_t centre = i.width / 2 + (i.height / 2) * i.width;
    const double pn = o.psf[centre] != 0.0 ? 1.0 / o.psf[centre] : 0.0;
    for (std::size_t q = 0; q < npix; ++q) {
        o.dirty[q] *= pn;
        o.psf[q] *= pn;
    }
    t_finish += since(t0);
    const double stats[8] = {image_weight, gridded, vsum, maxw, min_w, max_w, pn, double(i.n_w_layers)};
    std::memcpy(o.stats, stats, sizeof(stats));
    if (timings) {
        const double total = since(t_all);
        // the order of Timings::as_array (core/src/gpu/invert.rs:57-74)
        const double tm[14] = {0.0,
                               0.0,
                               t_prepare,
                               t_upload,
                               total - t_upload - t_prepare - t_download - t_finish,
                               t_download,
                               t_finish,
                               double(b.bytes),
                               double(nbatches),
                               double(batch),
                               t_grid,
                               t_fft,
                               t_project,
                               t_bucket};
        std::memcpy(timings, tm, sizeof(tm));
    }
    return 0;
}

}  // namespace

int invert_gpu(const GridInputs& i, const GridOutputs& o, bool single, double* timings, std::string& err) {
    try {
        return single ? invert<float>(i, o, timings, err) : invert<double>(i, o, timings, err);
    } catch (const Fail& f) {
        err = f.msg;
        return f.status;
    }
}

int kernel(std::size_t ks, std::size_t os, double* out) {
    const std::size_t n = ks * os, mid = n / 2;
    const double fr = 1.0 / double(os), norm = double(os) / bessel0(8.6, 1e-8);
    for (std::size_t q = 0; q <= mid && mid + q < n; ++q) {
        const double x = double(q);
        const double sn = q == 0 ? fr : std::sin(M_PI * fr * x) / (M_PI * x);
        const double t = double(q) / double(mid);
        out[mid + q] = sn * bessel0(8.6 * std::sqrt(1.0 - t * t), 1e-8) * norm;
    }
    for (std::size_t q = 0; q < mid; ++q) out[q] = out[n - 1 - q];
    return 0;
}

}  // namespace ws
}  // namespace rang
