// WSClean imaging weights and Stokes I formation on the GPU (rang-core
// core/src/weights.rs, C ABI core/src/ffi_weights.rs).
//
// imaging weights: one pass over the samples puts each selected sample's
// float weight into its cell of the half-plane grid (double atomics) and the
// in-grid total (k_wt_grid); the finish turns each non-zero cell into
// 1 / (1 + v s^2) (Briggs, with the mean of v^2 over the total from a
// reduction, k_wt_sq / k_wt_finish), 1 / v (uniform) or 1 (natural); the
// rank filter clips each non-zero cell to `level` times the mean of the
// non-zero cells in its window (k_wt_rank); the lookup reads every sample's
// cell as a float (k_wt_lookup) and the header sums reduce the products of
// visibility and imaging weights (k_wt_sums). Stokes I: one thread per cell
// (k_stokes_i). The sample-to-cell arithmetic is the reference's, so the
// cells match exactly; a cell's sum is formed in atomic order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rang_core_weights.h"

namespace {

struct WtError {
    int status;
};

void wchk(hipError_t e) {
    if (e != hipSuccess) throw WtError{2};
}

constexpr int kRed = 1024;  // workgroups of the reductions

__device__ __forceinline__ bool finite2(float re, float im) { return isfinite(re) && isfinite(im); }

// core/src/weights.rs:507-550
__global__ __launch_bounds__(256) void k_stokes_i(long long ncell, const float* __restrict__ data,
                                                  const float* __restrict__ w, const std::uint8_t* __restrict__ f,
                                                  float* __restrict__ vis, float* __restrict__ wout) {
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell) return;
    const float xr = data[8 * c], xi = data[8 * c + 1], yr = data[8 * c + 6], yi = data[8 * c + 7];
    const float re = __fmul_rn(__fadd_rn(yr, xr), 0.5f), im = __fmul_rn(__fadd_rn(yi, xi), 0.5f);
    const bool ok = finite2(re, im);
    vis[2 * c] = ok ? re : 0.0f;
    vis[2 * c + 1] = ok ? im : 0.0f;
    float weight = (f[4 * c] == 0 && finite2(xr, xi)) ? __fmul_rn(w[4 * c], 4.0f) : 0.0f;
    if (f[4 * c + 3] == 0 && finite2(yr, yi)) {
        const float yy = __fmul_rn(w[4 * c + 3], 4.0f);
        weight = yy < weight ? yy : weight;  // std::min: the second only when strictly smaller
    } else {
        weight = 0.0f;
    }
    wout[c] = weight;
}

struct WGrid {
    long long width, height;  // after the super weight
    double sx, sy;
};

// ImageWeights' uv -> cell (core/src/weights.rs:131-145): the v < 0 half
// mirrored, x = floor(u width sx + width / 2), y = floor(v height sy);
// -1 outside the half-plane grid.
__device__ __forceinline__ long long cell_of(const WGrid& g, double u, double v) {
    if (v < 0.0) {
        u = -u;
        v = -v;
    }
    // products and sum rounded one by one, as the reference evaluates them
    const long long x = (long long)floor(__dadd_rn(__dmul_rn(__dmul_rn(u, double(g.width)), g.sx), double(g.width / 2)));
    const long long y = (long long)floor(__dmul_rn(__dmul_rn(v, double(g.height)), g.sy));
    if (x >= 0 && x < g.width && y < g.height / 2) return x + y * g.width;
    return -1;
}

// Workgroup sum of a double (256 threads) into part[block].
__device__ __forceinline__ void wg_sum(double v, double* part) {
    __shared__ double s[256];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// grid_rows (core/src/weights.rs:147-184): the row's u, v in metres mirrored
// when v < 0, then per channel divided by the wavelength.
__global__ __launch_bounds__(256) void k_wt_grid(WGrid g, long long nrow, int nchan, const double* __restrict__ uvw,
                                                 const double* __restrict__ wl, const float* __restrict__ wts,
                                                 const std::uint8_t* __restrict__ sel, double* __restrict__ grid,
                                                 double* __restrict__ part) {
    double tot = 0.0;
    const long long n = nrow * nchan;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const long long row = i / nchan;
        const int ch = int(i - row * nchan);
        if (!sel[row]) continue;
        double um = uvw[3 * row], vm = uvw[3 * row + 1];
        if (vm < 0.0) {
            um = -um;
            vm = -vm;
        }
        const long long c = cell_of(g, um / wl[ch], vm / wl[ch]);
        if (c >= 0) {
            const double w = double(wts[i]);
            atomicAdd(&grid[c], w);
            tot += w;
        }
    }
    wg_sum(tot, part);
}

__global__ __launch_bounds__(256) void k_wt_sq(const double* __restrict__ grid, long long n, double* __restrict__ part) {
    double q = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        q += grid[i] * grid[i];
    wg_sum(q, part);
}

// totals[0] = sum of n0 partials p0, totals[1] = sum of n1 partials p1.
__global__ __launch_bounds__(256) void k_wt_totals(const double* __restrict__ p0, int n0, const double* __restrict__ p1,
                                                   int n1, double* __restrict__ totals) {
    __shared__ double s[2][256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n0; i += 256) a += p0[i];
    for (int i = threadIdx.x; i < n1; i += 256) b += p1[i];
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            s[0][threadIdx.x] += s[0][threadIdx.x + h];
            s[1][threadIdx.x] += s[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        totals[0] = s[0][0];
        totals[1] = s[1][0];
    }
}

// FinishGridding (core/src/weights.rs:186-223)
// num2 = (5 exp(-ln(10) robust))^2, formed on the host.
__global__ __launch_bounds__(256) void k_wt_finish(double* __restrict__ grid, long long n, int cls, double num2,
                                                   const double* __restrict__ totals) {
    double s_sq = 0.0;
    if (cls == 2) {
        const double avg = __ddiv_rn(totals[1], totals[0]);
        s_sq = __ddiv_rn(num2, avg);
    }
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double v = grid[i];
        if (v == 0.0) continue;
        grid[i] = cls == 2 ? __ddiv_rn(1.0, __dadd_rn(1.0, __dmul_rn(v, s_sq))) : cls == 1 ? __ddiv_rn(1.0, v) : 1.0;
    }
}

// RankFilter (core/src/weights.rs:225-265): out = min(v, level * window
// mean) over [x - d, x + d) x [y - d, y + d) clipped to the grid, d = window / 2.
__global__ __launch_bounds__(256) void k_wt_rank(const double* __restrict__ grid, WGrid g, double level, int window,
                                                 double* __restrict__ out) {
    const long long half = g.height / 2;
    const long long n = g.width * half;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double w = grid[i];
    double o = w;
    if (w != 0.0) {
        const long long y = i / g.width, x = i - y * g.width;
        const long long d = window / 2;
        const long long x1 = x > d ? x - d : 0, y1 = y > d ? y - d : 0;
        const long long x2 = x + d >= g.width ? g.width : x + d;
        const long long y2 = y + d >= half ? half : y + d;
        long long count = 0;
        double sum = 0.0;
        for (long long yi = y1; yi < y2; ++yi)
            for (long long xi = x1; xi < x2; ++xi) {
                const double v = grid[yi * g.width + xi];
                if (v != 0.0) {
                    ++count;
                    sum += v;
                }
            }
        const double mean = __ddiv_rn(sum, double(count));
        const double lim = __dmul_rn(mean, level);
        if (w > lim) o = lim;
    }
    out[i] = o;
}

// lookup_rows (core/src/weights.rs:282-304), then the header sums of the
// products (:552-572): per workgroup [visibility weight sum, total, max,
// count].
__global__ __launch_bounds__(256) void k_wt_lookup(WGrid g, long long nrow, int nchan, const double* __restrict__ uvw,
                                                   const double* __restrict__ wl, const float* __restrict__ wts,
                                                   const std::uint8_t* __restrict__ sel, const double* __restrict__ grid,
                                                   float* __restrict__ out, double* __restrict__ part) {
    double vsum = 0.0, tot = 0.0, mx = 0.0, cnt = 0.0;
    const long long n = nrow * nchan;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const long long row = i / nchan;
        const int ch = int(i - row * nchan);
        float v = 0.0f;
        if (sel[row]) {
            const long long c = cell_of(g, uvw[3 * row] / wl[ch], uvw[3 * row + 1] / wl[ch]);
            v = c >= 0 ? float(grid[c]) : 0.0f;
        }
        out[i] = v;
        const float w = wts[i];
        const float cum = __fmul_rn(w, v);
        if (cum != 0.0f) {
            vsum += double(w);
            cnt += 1.0;
        }
        mx = fmax(mx, double(cum));
        tot += double(cum);
    }
    __shared__ double s[4][256];
    s[0][threadIdx.x] = vsum;
    s[1][threadIdx.x] = tot;
    s[2][threadIdx.x] = mx;
    s[3][threadIdx.x] = cnt;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            s[0][threadIdx.x] += s[0][threadIdx.x + h];
            s[1][threadIdx.x] += s[1][threadIdx.x + h];
            s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + h]);
            s[3][threadIdx.x] += s[3][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = s[k][0];
}

__global__ __launch_bounds__(256) void k_wt_sums(const double* __restrict__ part, int n, double* __restrict__ sums) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0, m = 0.0, c = 0.0;
    for (int i = 0; i < n; ++i) {
        a += part[4 * i];
        b += part[4 * i + 1];
        m = fmax(m, part[4 * i + 2]);
        c += part[4 * i + 3];
    }
    sums[0] = a;
    sums[1] = b;
    sums[2] = m;
    sums[3] = c;
}

__global__ __launch_bounds__(256) void k_zero_d(double* p, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = 0.0;
}

bool shape_of(std::size_t width, std::size_t height, double sw, std::size_t& w, std::size_t& h) {
    if (!(sw > 0.0)) return false;
    w = std::size_t(std::llround(double(width) / sw));
    h = std::size_t(std::llround(double(height) / sw));
    if (w % 2) ++w;
    if (h % 2) ++h;
    return true;
}

unsigned blocks(long long n) { return unsigned(std::max<long long>(1, std::min<long long>((n + 255) / 256, kRed))); }

// Device buffers of the weights pass, per device, grown on demand.
struct WtWork {
    std::mutex mu;
    void* p[3] = {};
    std::size_t s[3] = {};
    void* get(int k, std::size_t bytes) {
        bytes = std::max<std::size_t>(bytes, 256);
        if (s[k] < bytes) {
            if (p[k]) wchk(hipFree(p[k]));
            p[k] = nullptr;
            s[k] = 0;
            wchk(hipMalloc(&p[k], bytes));
            s[k] = bytes;
        }
        return p[k];
    }
};

WtWork& work() {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<WtWork>> all;
    int dev = 0;
    wchk(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto& w = all[dev];
    if (!w) w.reset(new WtWork());
    return *w;
}

// The whole pass on device pointers (in, out), on stream st.
int imaging_weights_dev(const ImagingWeightInputs& in, const ImagingWeightOutputs& out, hipStream_t st) {
    if (in.width == 0 || in.height == 0 || !(in.super_weight > 0.0) || in.pixel_scale_x <= 0.0 ||
        in.pixel_scale_y <= 0.0 || in.class_ > 2 || !out.weights)
        return 1;
    if (!in.uvw || !in.wavelength || !in.weights || !in.row_selected) return 1;
    std::size_t gw, gh;
    shape_of(in.width, in.height, in.super_weight, gw, gh);
    WGrid g{(long long)gw, (long long)gh, in.pixel_scale_x, in.pixel_scale_y};
    const long long ncells = (long long)gw * (long long)(gh / 2);
    const long long n = (long long)(in.nrow * in.nchan);
    WtWork& w = work();
    std::lock_guard<std::mutex> lk(w.mu);
    double* grid = static_cast<double*>(w.get(0, sizeof(double) * 2 * std::size_t(std::max<long long>(ncells, 1))));
    double* grid2 = grid + std::max<long long>(ncells, 1);
    double* part = static_cast<double*>(w.get(1, sizeof(double) * (4 * kRed + 8)));
    double* totals = part + 4 * kRed;
    double* sums = static_cast<double*>(w.get(2, sizeof(double) * 4));
    k_zero_d<<<blocks(ncells), 256, 0, st>>>(grid, ncells);
    const unsigned nb = blocks(n);
    k_wt_grid<<<nb, 256, 0, st>>>(g, (long long)in.nrow, int(in.nchan), in.uvw, in.wavelength, in.weights,
                                  in.row_selected, grid, part);
    const unsigned cb = blocks(ncells);
    double* sq = part + kRed;
    if (in.class_ == 2) k_wt_sq<<<cb, 256, 0, st>>>(grid, ncells, sq);
    k_wt_totals<<<1, 256, 0, st>>>(part, int(nb), sq, in.class_ == 2 ? int(cb) : 0, totals);
    wchk(hipGetLastError());
    const double num = 5.0 * std::exp(-2.302585092994045684017991454684364208 * in.robust);  // ln 10
    k_wt_finish<<<cb, 256, 0, st>>>(grid, ncells, int(in.class_), num * num, totals);
    double* fin = grid;
    if (in.rank_filter_level >= 1.0) {
        k_wt_rank<<<unsigned((ncells + 255) / 256), 256, 0, st>>>(grid, g, in.rank_filter_level,
                                                                 int(in.rank_filter_size), grid2);
        fin = grid2;
    }
    wchk(hipGetLastError());
    k_wt_lookup<<<nb, 256, 0, st>>>(g, (long long)in.nrow, int(in.nchan), in.uvw, in.wavelength, in.weights,
                                    in.row_selected, fin, out.weights, part);
    k_wt_sums<<<1, 256, 0, st>>>(part, int(nb), out.sums ? out.sums : sums);
    wchk(hipGetLastError());
    if (out.grid) wchk(hipMemcpyAsync(out.grid, fin, sizeof(double) * ncells, hipMemcpyDeviceToDevice, st));
    wchk(hipStreamSynchronize(st));
    return 0;
}

int stokes_dev(const StokesIInputs& in, const StokesIOutputs& out, hipStream_t st) {
    if (!in.data || !in.weights || !in.flags || !out.vis || !out.weight) return 1;
    if (in.ncell) {
        k_stokes_i<<<unsigned((in.ncell + 255) / 256), 256, 0, st>>>((long long)in.ncell, in.data, in.weights, in.flags,
                                                                     out.vis, out.weight);
        wchk(hipGetLastError());
    }
    wchk(hipStreamSynchronize(st));
    return 0;
}

// Host buffers staged to the device and back on a stream of the library's.
struct Staged {
    std::vector<void*> p;
    ~Staged() {
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
    template <class T>
    T* up(const T* h, std::size_t count, hipStream_t st) {
        void* d = nullptr;
        wchk(hipMalloc(&d, std::max<std::size_t>(count * sizeof(T), 16)));
        p.push_back(d);
        if (h && count) wchk(hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, st));
        return static_cast<T*>(d);
    }
};

hipStream_t lib_stream() {
    static std::mutex mu;
    static std::map<int, hipStream_t> all;
    int dev = 0;
    wchk(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = all.find(dev);
    if (it != all.end()) return it->second;
    hipStream_t s;
    wchk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    all[dev] = s;
    return s;
}

}  // namespace

extern "C" {

int32_t rang_core_weight_grid_shape(size_t width, size_t height, double super_weight, size_t* shape) {
    if (!shape) return 1;
    std::size_t w, h;
    if (!shape_of(width, height, super_weight, w, h)) return 1;
    shape[0] = w;
    shape[1] = h;
    return 0;
}

size_t rang_core_padded_size(size_t trimmed, double padding) {
    std::size_t padded = std::size_t(std::ceil(double(trimmed) * padding));
    padded += (4 - (padded % 4)) % 4;
    return padded;
}

int32_t rang_amd_stokes_i_dev(const StokesIInputs* inputs, StokesIOutputs* outputs, void* stream) {
    try {
        if (!inputs || !outputs) return 1;
        return stokes_dev(*inputs, *outputs, static_cast<hipStream_t>(stream));
    } catch (const WtError& e) {
        return e.status;
    } catch (...) {
        return 2;
    }
}

int32_t rang_amd_imaging_weights_dev(const ImagingWeightInputs* inputs, ImagingWeightOutputs* outputs,
                                     void* stream) {
    try {
        if (!inputs || !outputs) return 1;
        return imaging_weights_dev(*inputs, *outputs, static_cast<hipStream_t>(stream));
    } catch (const WtError& e) {
        return e.status;
    } catch (...) {
        return 2;
    }
}

int32_t rang_core_stokes_i(const StokesIInputs* inputs, StokesIOutputs* outputs) {
    try {
        if (!inputs || !outputs || !outputs->vis || !outputs->weight) return 1;
        const StokesIInputs& i = *inputs;
        if (!i.data || !i.weights || !i.flags) return 1;
        hipStream_t st = lib_stream();
        Staged s;
        StokesIInputs d{i.ncell, s.up(i.data, 8 * i.ncell, st), s.up(i.weights, 4 * i.ncell, st),
                        s.up(i.flags, 4 * i.ncell, st)};
        StokesIOutputs o{s.up<float>(nullptr, 2 * i.ncell, st), s.up<float>(nullptr, i.ncell, st)};
        const int r = stokes_dev(d, o, st);
        if (r) return r;
        if (i.ncell) {
            wchk(hipMemcpyAsync(outputs->vis, o.vis, sizeof(float) * 2 * i.ncell, hipMemcpyDeviceToHost, st));
            wchk(hipMemcpyAsync(outputs->weight, o.weight, sizeof(float) * i.ncell, hipMemcpyDeviceToHost, st));
        }
        wchk(hipStreamSynchronize(st));
        return 0;
    } catch (const WtError& e) {
        return e.status;
    } catch (...) {
        return 2;
    }
}

int32_t rang_core_imaging_weights(const ImagingWeightInputs* inputs, ImagingWeightOutputs* outputs) {
    try {
        if (!inputs || !outputs || !outputs->weights) return 1;
        const ImagingWeightInputs& i = *inputs;
        if (!i.uvw || !i.wavelength || !i.weights || !i.row_selected) return 1;
        std::size_t gw, gh;
        if (!shape_of(i.width, i.height, i.super_weight, gw, gh)) return 1;
        hipStream_t st = lib_stream();
        Staged s;
        ImagingWeightInputs d = i;
        const std::size_t n = i.nrow * i.nchan;
        d.uvw = s.up(i.uvw, 3 * i.nrow, st);
        d.wavelength = s.up(i.wavelength, i.nchan, st);
        d.weights = s.up(i.weights, n, st);
        d.row_selected = s.up(i.row_selected, i.nrow, st);
        const std::size_t ncells = gw * (gh / 2);
        ImagingWeightOutputs o{s.up<float>(nullptr, n, st), outputs->grid ? s.up<double>(nullptr, ncells, st) : nullptr,
                               outputs->sums ? s.up<double>(nullptr, 4, st) : nullptr};
        const int r = imaging_weights_dev(d, o, st);
        if (r) return r;
        if (n) wchk(hipMemcpyAsync(outputs->weights, o.weights, sizeof(float) * n, hipMemcpyDeviceToHost, st));
        if (outputs->grid && ncells)
            wchk(hipMemcpyAsync(outputs->grid, o.grid, sizeof(double) * ncells, hipMemcpyDeviceToHost, st));
        if (outputs->sums) wchk(hipMemcpyAsync(outputs->sums, o.sums, sizeof(double) * 4, hipMemcpyDeviceToHost, st));
        wchk(hipStreamSynchronize(st));
        return 0;
    } catch (const WtError& e) {
        return e.status;
    } catch (...) {
        return 2;
    }
}

}  // extern "C"
