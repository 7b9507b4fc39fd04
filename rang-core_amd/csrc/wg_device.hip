// Device pipeline of the exact w-gridder for CDNA4 (gfx950).
//
// vis2dirty (rang-core core/src/wgridder.rs:2436-2683):
//   scan      active count and |w| range                    (k_scan)
//   plan      kernel, grid and w-planes on the host, FP64    (wg_plan.cpp)
//   sort      bucket key per visibility, counts, prefix sum, (k_keys, scan,
//             one record per active visibility in bucket order  k_scatter)
//   grid      one workgroup per (uv tile, group of w-planes, chunk of up to
//             4096 visibilities) gathers the visibilities of its tile and of
//             the three tiles whose footprints reach into it, accumulates in
//             LDS and writes its tile of each plane once; the chunks of a
//             split tile add the cells they reached with global atomics, and
//             cells no visibility reaches are cleared by a fill kernel
//             (k_grid_items, k_zero_split, k_grid)
//   v FFT     inverse along v of every row of every plane, keeping the image
//             columns                                     (k_rows_inv)
//   u FFT     inverse along u of the image columns, projected with the
//             plane's w-screen                             (k_cols_inv)
//   sum       plane terms added in plane order, global correction (k_sum)
// dirty2vis (wgridder.rs:2998-3163) runs the transposed chain: corrected
// image, u FFT with the w-screen, v FFT, then degridding from the planes with
// the weight and shift phase applied per visibility (k_degrid).
//
// FP64 is used for the uvw x f/c coordinates, the fractional grid position,
// the w coordinate of each plane, the w-screen phase (reduced mod 1 in single
// precision as the reference does) and the correction factors; gridding,
// degridding and the FFTs run in the working precision T.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "gram.h"
#include "wg_device.h"
#include "wg_fft.h"

namespace rang {

namespace {

struct DevError {
    std::string msg;
    int status = 2;  // 2: failure on the device; 1: a request beyond a fixed limit
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DevError{std::string(what) + ": " + hipGetErrorString(e)};
}

#define HIPCHK(x) check((x), #x)

// Blocks of a one-dimensional launch: a dispatch holds fewer than 2^32
// work-items per grid dimension, so larger launches are refused here rather
// than wrapped by the runtime.
unsigned blocks_of(long long blocks, int threads) {
    if (blocks <= 0 || blocks * threads >= (1LL << 32)) throw DevError{"launch above 2^32 work-items"};
    return unsigned(blocks);
}

// Zero fill by a kernel rather than hipMemsetAsync, so a stream that is
// captured into a graph records a kernel node: 16-byte stores over the body
// when the buffer is 16-byte aligned, 4-byte stores otherwise and for the tail.
__global__ __launch_bounds__(256) void k_fill0(std::uint32_t* p, long long words, int vec) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long quads = vec ? words / 4 : 0;
    uint4* q = reinterpret_cast<uint4*>(p);
    for (long long k = i; k < quads; k += stride) q[k] = make_uint4(0u, 0u, 0u, 0u);
    for (long long k = quads * 4 + i; k < words; k += stride) p[k] = 0u;
}

// Raises a kernel's dynamic LDS limit once per size: later launches (and
// launches under graph capture) make no attribute call.
void lds_limit(const void* kernel, std::size_t bytes) {
    static std::mutex mu;
    static std::map<const void*, std::size_t> done;
    std::lock_guard<std::mutex> lk(mu);
    auto it = done.find(kernel);
    if (it != done.end() && it->second >= bytes) return;
    check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)),
          "hipFuncSetAttribute");
    done[kernel] = bytes;
}

// total = sum of n partial sums in a fixed order (one workgroup of 256).
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, long long n,
                                                   double* __restrict__ total) {
    __shared__ double s[256];
    double v = 0.0;
    for (long long i = threadIdx.x; i < n; i += 256) v += part[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s[0];
}

void fill_zero(void* p, std::size_t bytes, hipStream_t st) {
    if (bytes == 0) return;
    if (bytes % 4 || reinterpret_cast<std::uintptr_t>(p) % 4) throw DevError{"zero fill of an unaligned buffer"};
    const long long words = (long long)(bytes / 4);
    const int vec = reinterpret_cast<std::uintptr_t>(p) % 16 == 0 ? 1 : 0;
    const long long work = vec ? (words + 3) / 4 : words;
    const long long blocks = std::max<long long>(1, std::min<long long>((work + 255) / 256, 8192));
    k_fill0<<<unsigned(blocks), 256, 0, st>>>(static_cast<std::uint32_t*>(p), words, vec);
    check(hipGetLastError(), "k_fill0");
}

// ------------------------------------------------------------ workspace

enum Slot : int {
    kSlotFoc,
    kSlotCounters,
    kSlotKeys,
    kSlotCounts,
    kSlotOffsets,
    kSlotCursor,
    kSlotScan,
    kSlotRecs,
    kSlotCoeff,
    kSlotTwU,
    kSlotTwV,
    kSlotGL,
    kSlotCfu,
    kSlotCfv,
    kSlotScreen,
    kSlotCorr,
    kSlotGrid,
    kSlotTmp,
    kSlotProj,
    kSlotImg,
    kSlotOrder,
    kSlotScanPart,
    kSlotItems,
    kSlotItemOff,
    // a second sorted set (dirty2vis tiling) and the order of its tiles
    kSlotKeys2,
    kSlotCounts2,
    kSlotOffsets2,
    kSlotCursor2,
    kSlotRecs2,
    kSlotOrder2,
    // prepared calls: visibility buffer, reductions, solver vectors
    kSlotVbuf,
    kSlotRed,
    kSlotSolve,
    kSlotLong,  // twiddles of the transforms longer than one LDS
    kSlotFista,  // the image vectors of the FISTA solve
    kSlotScalars,  // device scalars of a call (sums)
    kSlotItems2,   // degridding items of a prepared plan (gridding keeps kSlotItems)
    kSlotItemOff2,
    kSlotOcc,  // per (plane group, tile): whether any gridding item reaches it
    kSlotRowNeed,  // per grid row: whether a degridding item reads it
    kSlotDefect,  // the image vectors of the major cycles (x, G x, G~ x, corrected data)
    kSlotCount
};

// Device buffers of one GPU kept between calls and grown on demand; calls on
// one device are serialised by its mutex. In arena mode (`arena` set) every
// slot is carved from one fixed device block instead, host staging is a
// fixed host buffer, and nothing is allocated: a request beyond the block
// fails with status 1.
struct Workspace {
    void* ptr[kSlotCount] = {};
    std::size_t size[kSlotCount] = {};
    std::mutex mu;
    // Pinned host staging for the small transfers of a call.
    void* pinned = nullptr;
    std::size_t pinned_size = 0;
    char* arena = nullptr;
    std::size_t arena_cap = 0, arena_used = 0;
    unsigned gen = 0;  // counts reallocations (a graph holds the old addresses)
    alignas(64) unsigned char arena_host[8192];
    // Signature of the plan whose tables (twiddles, coefficients, screen and
    // correction tables, tile order) the table slots hold, and their layout.
    std::vector<double> table_key;
    FftDesc du, dv;
    std::vector<double> foc_key;
    std::vector<int> order_key[2];
    std::vector<double> long_key;
    void* host(std::size_t bytes) {
        if (arena) {
            if (bytes > sizeof(arena_host)) throw DevError{"host staging beyond the fixed buffer", 1};
            return arena_host;
        }
        if (pinned_size < bytes) {
            if (pinned) HIPCHK(hipHostFree(pinned));
            pinned = nullptr;
            pinned_size = 0;
            HIPCHK(hipHostMalloc(&pinned, bytes, hipHostMallocDefault));
            pinned_size = bytes;
            ++gen;
        }
        return pinned;
    }
    void* get(int slot, std::size_t bytes) {
        bytes = std::max<std::size_t>(bytes, 256);
        if (size[slot] < bytes && arena) {
            bytes = (bytes + 255) & ~std::size_t(255);
            if (bytes > arena_cap - arena_used)
                throw DevError{"the call needs more than the " + std::to_string(arena_cap >> 20) +
                                   " MiB of fixed scratch memory",
                               1};
            ptr[slot] = arena + arena_used;
            size[slot] = bytes;
            arena_used += bytes;
        }
        if (size[slot] < bytes) {
            if (ptr[slot]) HIPCHK(hipFree(ptr[slot]));
            ptr[slot] = nullptr;
            size[slot] = 0;
            HIPCHK(hipMalloc(&ptr[slot], bytes));
            size[slot] = bytes;
            ++gen;
        }
        return ptr[slot];
    }
};

// The fixed device block of the allocation-free entry point: every GPU the
// library runs on holds one copy from module load until the process ends.
constexpr std::size_t kScratchBytes = std::size_t(256) << 20;

}  // namespace
}  // namespace rang

static __device__ __align__(256) unsigned char scratch[rang::kScratchBytes];

namespace rang {
namespace {

// Arena mode: forget every slot and cached table (contents of the block are
// rewritten by the call that follows).
void arena_reset(Workspace& ws) {
    for (int k = 0; k < kSlotCount; ++k) {
        ws.ptr[k] = nullptr;
        ws.size[k] = 0;
    }
    ws.arena_used = 0;
    ws.table_key.clear();
    ws.foc_key.clear();
    ws.order_key[0].clear();
    ws.order_key[1].clear();
    ws.long_key.clear();
}

// The workspace over the device's copy of `scratch`. Its device address is
// read once per device, with that device current; host code passes that
// address, never the array's name, to kernels and copies.
Workspace& scratch_workspace(int device) {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<Workspace>> all;
    std::lock_guard<std::mutex> lk(mu);
    auto& w = all[device];
    if (!w) {
        std::unique_ptr<Workspace> n(new Workspace());
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw DevError{"no device address for the fixed scratch block"};
        n->arena = static_cast<char*>(ptr);
        n->arena_cap = kScratchBytes;
        w = std::move(n);
    }
    return *w;
}

std::mutex g_ws_mu;
std::map<int, std::unique_ptr<Workspace>> g_ws_all;

Workspace& workspace(int device) {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto& w = g_ws_all[device];
    if (!w) w.reset(new Workspace());
    return *w;
}

// Frees a (non-arena) workspace's device and pinned buffers and forgets its
// cached tables; its generation moves on, so graphs holding the old
// addresses are dropped before their next replay. The caller holds ws.mu.
void free_workspace(Workspace& ws) {
    for (int k = 0; k < kSlotCount; ++k) {
        if (ws.ptr[k] && !ws.arena) (void)hipFree(ws.ptr[k]);
        ws.ptr[k] = nullptr;
        ws.size[k] = 0;
    }
    if (ws.pinned) (void)hipHostFree(ws.pinned);
    ws.pinned = nullptr;
    ws.pinned_size = 0;
    ws.table_key.clear();
    ws.foc_key.clear();
    ws.order_key[0].clear();
    ws.order_key[1].clear();
    ws.long_key.clear();
    ++ws.gen;
}

// ------------------------------------------------------------ stage timing

thread_local bool g_timing = false;
thread_local std::vector<std::pair<std::string, double>> g_times;

struct StageTimer {
    hipStream_t st;
    std::vector<std::pair<std::string, hipEvent_t>> marks;
    explicit StageTimer(hipStream_t s) : st(s) {
        if (g_timing) mark("start");
    }
    void mark(const char* name) {
        if (!g_timing) return;
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(e, st));
        marks.emplace_back(name, e);
    }
    void finish() {
        if (!g_timing) return;
        g_times.clear();
        HIPCHK(hipEventSynchronize(marks.back().second));
        for (std::size_t i = 1; i < marks.size(); ++i) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, marks[i - 1].second, marks[i].second));
            g_times.emplace_back(marks[i].first, double(ms));
        }
    }
    ~StageTimer() {
        for (auto& m : marks) (void)hipEventDestroy(m.second);
    }
};

// ------------------------------------------------------------ geometry

// What the kernels read of the call's geometry (wg_plan.h Geometry plus the
// tiling), by value.
struct Geo {
    long long n;
    int nrow, nchan;
    int nx, ny, nu, nv, supp;
    double pixsize_x, pixsize_y, ushift, vshift;
    long long maxiu0, maxiv0;
    double wmin, dw, xdw, wshift;
    int nplanes, nb, wgrid;
    double r_l, r_m, lshift, mshift, nshift;
    int shifting, lmshift;
    double fu, fv, fw;
    int Tu, Tv, ntu, ntv;
    int tnx, tny;
    int divide_by_n;
    // Image-side arrays column-major (index j nx + i): the w-screen and
    // correction tables, the per-plane projections and the corrected image.
    // Set with one-pass u transforms, whose workgroups each take whole
    // image columns.
    int cm;
    // Transforms longer than one LDS along u or v: n = n1 n2 (0: in LDS).
    // Their forward results sit at (k mod n1) n2 + k / n1.
    int lu1, lu2, lv1, lv2;
    // Gridding on the matrix cores (16 x 16 tiles, k_grid_mfma).
    int mfma;
};

__device__ __forceinline__ int upos(const Geo& g, int u) { return g.lu1 ? (u % g.lu1) * g.lu2 + u / g.lu1 : u; }
__device__ __forceinline__ int vpos(const Geo& g, int v) { return g.lv1 ? (v % g.lv1) * g.lv2 + v / g.lv1 : v; }

struct DevIn {
    const double* uvw;
    const double* foc;
    const void* vis;
    const void* wgt;
    const void* wgt2;
    const std::uint8_t* mask;
    int vis_kind, wgt_kind;
    // Gridding from values held in the records' order (complex T, one per
    // record of the sorted set): each record's data value is formed from it
    // when the record is staged.
    const void* stored = nullptr;
};

// One active visibility in bucket order. v2d: (re, im) its stored value
// (wgridder.rs:1439-1475); dirty2vis: the weight times the conjugate shift
// phase. idx: row * nchan + channel, bit 31 set where FixW flipped the row.
template <class T>
struct Rec {
    double w;
    T x0, y0;
    T re, im;
    std::uint32_t pos;  // ou | ov << 8 | first plane << 16
    std::uint32_t idx;
};

constexpr std::uint32_t kFlip = 0x80000000u;
constexpr std::uint32_t kNoKey = 0xffffffffu;

__device__ __forceinline__ void coord_of(const DevIn& in, const Geo& g, int row, double c[3]) {
    const double tu = g.fu * in.uvw[3 * row], tv = g.fv * in.uvw[3 * row + 1];
    const double tw = g.fw * in.uvw[3 * row + 2];
    c[0] = tu + g.r_l * tw;
    c[1] = tv + g.r_m * tw;
    c[2] = tw;
}

__device__ __forceinline__ void sincos_t(float a, float& s, float& c) { sincosf(a, &s, &c); }
__device__ __forceinline__ void sincos_t(double a, double& s, double& c) { sincos(a, &s, &c); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }

template <class T>
__device__ __forceinline__ cx<T> vis_get(const DevIn& in, long long i) {
    if (in.vis_kind == kVisC64) {
        const float2 f = static_cast<const float2*>(in.vis)[i];
        return {T(double(f.x)), T(double(f.y))};
    }
    if (in.vis_kind == kVisC128) {
        const double2 f = static_cast<const double2*>(in.vis)[i];
        return {T(f.x), T(f.y)};
    }
    return {T(1), T(0)};
}

template <class T>
__device__ __forceinline__ bool wgt_get(const DevIn& in, long long i, T& w) {
    switch (in.wgt_kind) {
    case kWgtF32:
        w = T(double(static_cast<const float*>(in.wgt)[i]));
        return true;
    case kWgtF64:
        w = T(static_cast<const double*>(in.wgt)[i]);
        return true;
    case kWgtF32Product: {
        const float p = __fmul_rn(static_cast<const float*>(in.wgt)[i],
                                  static_cast<const float*>(in.wgt2)[i]);
        w = T(double(p));
        return true;
    }
    default:
        return false;
    }
}

// Activity (wgridder.rs:1033-1041): |vis|^2 wgt mask != 0 for a dirty
// image alone, wgt mask != 0 otherwise.
template <class T, bool USE_VIS>
__device__ __forceinline__ bool active(const DevIn& in, long long i) {
    double tnorm = 1.0;
    if (USE_VIS) {
        // |v|^2 in T without contraction, as num_complex's norm_sqr
        const cx<T> v = vis_get<T>(in, i);
        tnorm = double(add_rn(mul_rn(v.re, v.re), mul_rn(v.im, v.im)));
    }
    T wv;
    const double twgt = wgt_get<T>(in, i, wv) ? double(wv) : 1.0;
    const double tmask = in.mask ? double(in.mask[i]) : 1.0;
    return tnorm * twgt * tmask != 0.0;
}

__device__ __forceinline__ void getpix(const Geo& g, double uin, double vin, double& uf,
                                       double& vf, long long& iu0, long long& iv0) {
    double u = uin * g.pixsize_x;
    u = (u - floor(u)) * double(g.nu);
    iu0 = min((long long)(u + g.ushift) - g.nu, g.maxiu0);
    uf = u - double(iu0);
    double v = vin * g.pixsize_y;
    v = (v - floor(v)) * double(g.nv);
    iv0 = min((long long)(v + g.vshift) - g.nv, g.maxiv0);
    vf = v - double(iv0);
}

__device__ __forceinline__ int shifted(long long i0, int n) {
    long long v = (i0 + n / 2) % n;
    return int(v < 0 ? v + n : v);
}

__device__ __forceinline__ int minplane(const Geo& g, double w) {
    if (!g.wgrid) return 0;
    const double p = (w + g.wshift) * g.xdw;
    const int m = p > 0.0 ? int(min(p, 2147483647.0)) : 0;
    return min(m, g.nb - 1);
}

// Shift phase of a visibility (wgridder.rs:1395-1408).
template <class T>
__device__ __forceinline__ cx<T> shift_phase(const Geo& g, const double b[3], double imflip,
                                             double foc, double sign) {
    const double fct = sign * imflip * (b[0] * g.lshift + b[1] * g.mshift + b[2] * g.nshift);
    const double tmp = fct * foc;
    T ang;
    if constexpr (sizeof(T) == 4)
        ang = T(2.0 * M_PI * (tmp - floor(tmp)));
    else
        ang = T(2.0 * M_PI * tmp);
    T s, c;
    sincos_t(ang, s, c);
    return {c, s};
}


// ------------------------------------------------------------ scan

constexpr int kScanBlocks = 1024;

// Active count and |w| range per block (wgridder.rs:1051-1077), reduced by
// k_scan_final; the index arithmetic stays in 32 bits (n < 2^31).
template <class T, bool USE_VIS>
__global__ __launch_bounds__(256) void k_scan(DevIn in, Geo g, unsigned long long* bcnt,
                                              double* blo, double* bhi) {
    unsigned long long cnt = 0;
    double lo = 1e300, hi = -1e300;
    const unsigned n = unsigned(g.n), nchan = unsigned(g.nchan);
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (!active<T, USE_VIS>(in, i)) continue;
        const unsigned row = i / nchan, ch = i - row * nchan;
        const double w = fabs(g.fw * in.uvw[3 * row + 2] * in.foc[ch]);
        ++cnt;
        lo = fmin(lo, w);
        hi = fmax(hi, w);
    }
    __shared__ unsigned long long s_cnt[256];
    __shared__ double s_lo[256], s_hi[256];
    s_cnt[threadIdx.x] = cnt;
    s_lo[threadIdx.x] = lo;
    s_hi[threadIdx.x] = hi;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
            s_lo[threadIdx.x] = fmin(s_lo[threadIdx.x], s_lo[threadIdx.x + h]);
            s_hi[threadIdx.x] = fmax(s_hi[threadIdx.x], s_hi[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bcnt[blockIdx.x] = s_cnt[0];
        blo[blockIdx.x] = s_lo[0];
        bhi[blockIdx.x] = s_hi[0];
    }
}

__global__ __launch_bounds__(256) void k_scan_final(const unsigned long long* bcnt, const double* blo,
                                                    const double* bhi, int nblk,
                                                    unsigned long long* out) {
    __shared__ unsigned long long s_cnt[256];
    __shared__ double s_lo[256], s_hi[256];
    unsigned long long c = 0;
    double lo = 1e300, hi = -1e300;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        c += bcnt[b];
        lo = fmin(lo, blo[b]);
        hi = fmax(hi, bhi[b]);
    }
    s_cnt[threadIdx.x] = c;
    s_lo[threadIdx.x] = lo;
    s_hi[threadIdx.x] = hi;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
            s_lo[threadIdx.x] = fmin(s_lo[threadIdx.x], s_lo[threadIdx.x + h]);
            s_hi[threadIdx.x] = fmax(s_hi[threadIdx.x], s_hi[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_cnt[0];
        out[1] = (unsigned long long)__double_as_longlong(s_lo[0]);
        out[2] = (unsigned long long)__double_as_longlong(s_hi[0]);
    }
}

// ------------------------------------------------------------ sort

// Bucket of an active visibility: (tile, spill class, first plane). The
// spill class says whether its footprint reaches the next tile in u (2) and
// in v (1).
template <class T, bool USE_VIS>
__device__ __forceinline__ std::uint32_t key_of(const DevIn& in, const Geo& g, unsigned i,
                                                unsigned row, unsigned ch) {
    if (!active<T, USE_VIS>(in, i)) return kNoKey;
    double c[3];
    coord_of(in, g, int(row), c);
    const double sgn = c[2] < 0.0 ? -1.0 : 1.0;
    const double f = in.foc[ch];
    const double u = c[0] * sgn * f, v = c[1] * sgn * f, w = c[2] * sgn * f;
    double uf, vf;
    long long iu0, iv0;
    getpix(g, u, v, uf, vf, iu0, iv0);
    const int su = shifted(iu0, g.nu), sv = shifted(iv0, g.nv);
    const int tu = su / g.Tu, tv = sv / g.Tv;
    const int ou = su - tu * g.Tu, ov = sv - tv * g.Tv;
    const int cat = ((ou + g.supp > g.Tu) ? 2 : 0) | ((ov + g.supp > g.Tv) ? 1 : 0);
    return std::uint32_t((((long long)tu * g.ntv + tv) * 4 + cat) * g.nb + minplane(g, w));
}

// Runs of equal keys among the consecutive lanes of a wave (channels of one
// row usually share a bucket): whether this lane starts a run, the run's
// length (for the first lane) and this lane's position in its run.
__device__ __forceinline__ void wave_runs(std::uint32_t key, bool& head, int& len, int& pos,
                                          int& headlane) {
    const int lane = threadIdx.x & 63;
    const std::uint32_t prev = __shfl_up(key, 1, 64);
    head = lane == 0 || prev != key;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    headlane = 63 - __clzll(heads & upto);
    pos = lane - headlane;
    const unsigned long long after = heads & ~upto;
    len = (after ? __ffsll((long long)after) - 1 : 64) - lane;
}

// One thread per visibility (coalesced); the first lane of each run adds the
// run's length to its bucket's count.
template <class T, bool USE_VIS>
__global__ __launch_bounds__(256) void k_keys(DevIn in, Geo g, std::uint32_t* keys,
                                              std::uint32_t* counts) {
    const unsigned n = unsigned(g.n), nchan = unsigned(g.nchan);
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool inside = i < n;
    std::uint32_t key = kNoKey;
    if (inside) {
        const unsigned row = i / nchan;
        key = key_of<T, USE_VIS>(in, g, i, row, i - row * nchan);
        keys[i] = key;
    }
    bool head;
    int len, pos, headlane;
    wave_runs(key, head, len, pos, headlane);
    if (head && key != kNoKey) atomicAdd(&counts[key], std::uint32_t(len));
}

constexpr int kScanItems = 4096;  // per block of 1024 threads

// Exclusive prefix sum of n values in blocks of kScanItems; block totals to
// sums.
__global__ __launch_bounds__(1024) void k_scan_block(const std::uint32_t* in, std::uint32_t* out,
                                                     std::uint32_t* sums, long long n) {
    __shared__ std::uint32_t s[1024];
    const long long base = (long long)blockIdx.x * kScanItems + threadIdx.x * 4;
    std::uint32_t v[4], tot = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        tot += v[k];
    }
    s[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        std::uint32_t add = threadIdx.x >= unsigned(off) ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    std::uint32_t run = s[threadIdx.x] - tot;
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 1023) sums[blockIdx.x] = s[1023];
}

__global__ __launch_bounds__(256) void k_scan_add(std::uint32_t* out, const std::uint32_t* offs,
                                                  long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += offs[i / kScanItems];
}

// One record per active visibility in bucket order. MODE 0: a dirty image
// alone (the data value), 1: with a PSF (the unit value; the data value is
// formed while gridding), 2: dirty2vis (weight times conjugate shift phase).
// One thread per row; each run of channels in one bucket takes a contiguous
// block of slots with one atomic, in channel order.
template <class T, int MODE>
__device__ __forceinline__ Rec<T> make_rec(const DevIn& in, const Geo& g, unsigned i, int ch,
                                           const double c[3], std::uint32_t key) {
    const double imflip = c[2] < 0.0 ? -1.0 : 1.0;  // FixW (wgridder.rs:583-586)
    const double b[3] = {c[0] * imflip, c[1] * imflip, c[2] * imflip};
    const double f = in.foc[ch];
    const double u = b[0] * f, v = b[1] * f, w = b[2] * f;
    double uf, vf;
    long long iu0, iv0;
    getpix(g, u, v, uf, vf, iu0, iv0);
    const int su = shifted(iu0, g.nu), sv = shifted(iv0, g.nv);
    const int ou = su % g.Tu, ov = sv % g.Tv;
    Rec<T> r;
    r.w = w;
    r.x0 = T(-uf * 2.0 + (double(g.supp) - 1.0));
    r.y0 = T(-vf * 2.0 + (double(g.supp) - 1.0));
    r.pos = std::uint32_t(ou) | (std::uint32_t(ov) << 8) | (std::uint32_t(key % std::uint32_t(g.nb)) << 16);
    r.idx = i | (imflip < 0.0 ? kFlip : 0u);
    T wv = T(1);
    const bool haswgt = wgt_get<T>(in, i, wv);
    if constexpr (MODE == 2) {
        // wgridder.rs:3146-3159
        if (g.shifting) {
            const cx<T> ph = shift_phase<T>(g, b, imflip, f, -1.0);
            r.re = ph.re * wv;
            r.im = ph.im * wv;
        } else {
            r.re = wv;
            r.im = T(0);
        }
    } else {
        // unit value (wgridder.rs:1414-1424)
        cx<T> p = {T(1), T(0)};
        if (g.shifting) p = cmul(p, shift_phase<T>(g, b, imflip, f, 1.0));
        if (haswgt) p = {p.re * wv, p.im * wv};
        p.im = p.im * T(imflip);
        if constexpr (MODE == 0) {
            // data value (wgridder.rs:1432-1435)
            const cx<T> d = cmul(vis_get<T>(in, i), cx<T>{p.re, p.im * T(imflip)});
            p = {d.re, d.im * T(imflip)};
        }
        r.re = p.re;
        r.im = p.im;
    }
    return r;
}

template <class T, int MODE>
__global__ __launch_bounds__(256) void k_scatter(DevIn in, Geo g, const std::uint32_t* keys,
                                                 const std::uint32_t* offsets,
                                                 std::uint32_t* cursor, Rec<T>* recs) {
    const unsigned n = unsigned(g.n), nchan = unsigned(g.nchan);
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const std::uint32_t key = i < n ? keys[i] : kNoKey;
    bool head;
    int len, pos, headlane;
    wave_runs(key, head, len, pos, headlane);
    std::uint32_t base = 0;
    if (head && key != kNoKey) base = offsets[key] + atomicAdd(&cursor[key], std::uint32_t(len));
    base = __shfl(base, headlane, 64);
    if (key == kNoKey) return;
    const unsigned row = i / nchan, ch = i - row * nchan;
    double c[3];
    coord_of(in, g, int(row), c);
    recs[base + pos] = make_rec<T, MODE>(in, g, i, int(ch), c, key);
}

// ------------------------------------------------------------ kernel taps

template <int S>
constexpr int ne_of() {
    return (((S + 3) | 1) + 1) / 2;
}

// u and v taps of one visibility by Horner's rule in x^2 on the even and odd
// parts, the second half mirrored (wgridder.rs:1514-1543).
template <class T, int S>
__device__ __forceinline__ void eval_taps2(const T* c, T x, T y, T* ru, T* rv) {
    constexpr int NE = ne_of<S>(), H = (S + 1) / 2;
    const T x2 = x * x, y2 = y * y;
    T ex[H], ey[H], ox[H], oy[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
        ex[i] = ey[i] = c[i];
        ox[i] = oy[i] = c[NE * S + i];
    }
#pragma unroll
    for (int k = 1; k < NE; ++k) {
#pragma unroll
        for (int i = 0; i < H; ++i) {
            ex[i] = fma(ex[i], x2, c[k * S + i]);
            ey[i] = fma(ey[i], y2, c[k * S + i]);
            ox[i] = fma(ox[i], x2, c[(NE + k) * S + i]);
            oy[i] = fma(oy[i], y2, c[(NE + k) * S + i]);
        }
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        ru[i] = fma(ox[i], x, ex[i]);
        rv[i] = fma(oy[i], y, ey[i]);
    }
#pragma unroll
    for (int i = 0; i < S - H; ++i) {
        ru[S - 1 - i] = fma(ox[i], -x, ex[i]);
        rv[S - 1 - i] = fma(oy[i], -y, ey[i]);
    }
}

// Tap `nth` at x (the w kernel of one plane), mirrored (wgridder.rs:1548-1560).
template <class T, int S>
__device__ __forceinline__ T eval_tap(const T* c, int nth, T x) {
    constexpr int NE = ne_of<S>(), H = (S + 1) / 2;
    if (nth >= H) {
        nth = S - 1 - nth;
        x = -x;
    }
    const T x2 = x * x;
    T e = c[nth], o = c[NE * S + nth];
#pragma unroll
    for (int k = 1; k < NE; ++k) {
        e = fma(e, x2, c[k * S + nth]);
        o = fma(o, x2, c[(NE + k) * S + nth]);
    }
    return fma(o, x, e);
}

// w factor of a visibility at w on plane p, tap p - minplane (wgridder.rs:1580-1591).
template <class T, int S>
__device__ __forceinline__ T w_factor(const Geo& g, const T* c, int nth, int p, double w) {
    const double w0 = g.wmin + double(p) * g.dw;
    const T z = T(g.xdw * (w0 - w));
    const T locz = (z - T(nth)) * T(2) + T(double(S) - 1.0);
    return eval_tap<T, S>(c, nth, locz);
}

// Wave-uniform copies of values every lane holds alike (scalar registers),
// and a value read from one lane.
__device__ __forceinline__ std::uint32_t uni(std::uint32_t x) {
    return std::uint32_t(__builtin_amdgcn_readfirstlane(int(x)));
}
__device__ __forceinline__ float uni(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uni(double x) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = unsigned(__builtin_amdgcn_readfirstlane(int(b)));
    const unsigned hi = unsigned(__builtin_amdgcn_readfirstlane(int(b >> 32)));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_value(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lane_value(double x, int l) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(b), l));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(b >> 32), l));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One tap's polynomial held in registers: a lane that evaluates the same tap
// for every visibility loads its even and odd coefficients once, mirrored as
// eval_tap does, and evaluates with the same operations.
template <class T, int S>
struct TapReg {
    static constexpr int NE = ne_of<S>();
    T e[NE], o[NE];
    bool neg;
    __device__ __forceinline__ void load(const T* c, int nth) {
        constexpr int H = (S + 1) / 2;
        neg = nth >= H;
        if (neg) nth = S - 1 - nth;
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            e[k] = c[k * S + nth];
            o[k] = c[(NE + k) * S + nth];
        }
    }
    __device__ __forceinline__ T eval(T x) const {
        if (neg) x = -x;
        const T x2 = x * x;
        T a = e[0], b = o[0];
#pragma unroll
        for (int k = 1; k < NE; ++k) {
            a = fma(a, x2, e[k]);
            b = fma(b, x2, o[k]);
        }
        return fma(b, x, a);
    }
};

// ------------------------------------------------------------ gridding

// Threads per gridding workgroup: one workgroup per CU, 16 waves in single
// precision (4 per SIMD) and 8 in double, whose kernels need more registers.
template <class T>
constexpr int grid_threads() {
    return sizeof(T) == 4 ? 1024 : 512;
}

// Work items of the gridding: per (tile rank, plane group) the visibilities
// of the nine source ranges that reach the tile's cells on the group's
// planes (see k_grid), cut into chunks of at most `rc`. Ranks follow
// `order` (densest tiles first), so the longest items start first.
__device__ __forceinline__ void source_of(int src, int& du, int& dv, int& cat) {
    // own tile with every spill class, the tile before in u with the classes
    // spilling in u, the one before in v with those spilling in v, and the
    // diagonal one spilling in both
    if (src < 4) {
        du = 0, dv = 0, cat = src;
    } else if (src < 6) {
        du = 1, dv = 0, cat = 2 + (src - 4);
    } else if (src < 8) {
        du = 0, dv = 1, cat = 1 + 2 * (src - 6);
    } else {
        du = 1, dv = 1, cat = 3;
    }
}

__device__ __forceinline__ void source_range(const Geo& g, const std::uint32_t* off, int tu, int tv,
                                             int src, int mlo, int mhi, std::uint32_t& a,
                                             std::uint32_t& b) {
    int du, dv, cat;
    source_of(src, du, dv, cat);
    const int stile = ((tu - du + g.ntu) % g.ntu) * g.ntv + (tv - dv + g.ntv) % g.ntv;
    const long long k0 = ((long long)stile * 4 + cat) * g.nb;
    a = off[k0 + mlo];
    b = off[k0 + mhi + 1];
}

__global__ __launch_bounds__(256) void k_grid_items(Geo g, const std::uint32_t* __restrict__ off,
                                                    const std::uint32_t* __restrict__ order, int G,
                                                    int ngroups, int S, std::uint32_t rc,
                                                    std::uint32_t* __restrict__ nitems) {
    const int ntg = g.ntu * g.ntv * ngroups;
    const int tg = blockIdx.x * blockDim.x + threadIdx.x;
    if (tg > ntg) return;
    if (tg == ntg) {
        nitems[ntg] = 0;
        return;
    }
    const int rank = tg / ngroups, grp = tg - rank * ngroups;
    const int tile = int(order[rank]);
    const int tu = tile / g.ntv, tv = tile - tu * g.ntv;
    const int P = g.wgrid ? g.nplanes : 1;
    const int p0 = grp * G, p1 = min(P, p0 + G);
    const int mlo = g.wgrid ? max(0, p0 - S + 1) : 0, mhi = g.wgrid ? min(g.nb - 1, p1 - 1) : 0;
    std::uint32_t total = 0;
    for (int src = 0; src < 9; ++src) {
        std::uint32_t a, b;
        source_range(g, off, tu, tv, src, mlo, mhi, a, b);
        total += b - a;
    }
    // a unit without visibilities takes no item: k_zero_split clears its cells
    nitems[tg] = (total + rc - 1) / rc;
}

// Cells of the (tile, plane group)s gridded by more than one item, zeroed
// before those items add into them, and (unless `skip_empty`: the row
// transform then takes them as zero) of those gridded by none.
template <class T, int NOUT>
__global__ __launch_bounds__(256) void k_zero_split(Geo g, const std::uint32_t* __restrict__ item_off,
                                                    const std::uint32_t* __restrict__ order, int G,
                                                    int ngroups, cx<T>* __restrict__ grid, int skip_empty) {
    const int tg = blockIdx.x;
    const std::uint32_t n = item_off[tg + 1] - item_off[tg];
    if (n == 1 || (n == 0 && skip_empty)) return;
    const int rank = tg / ngroups, grp = tg - rank * ngroups;
    const int tile = int(order[rank]);
    const int tu = tile / g.ntv, tv = tile - tu * g.ntv;
    const int P = g.wgrid ? g.nplanes : 1;
    const int p0 = grp * G, np = min(P, p0 + G) - p0;
    const long long plane = (long long)g.nu * g.nv;
    // a fixed column per thread, rows in steps (no division per cell)
    const int istep = int(blockDim.x) / g.Tv;
    const int i0 = int(threadIdx.x) / g.Tv, j = int(threadIdx.x) - i0 * g.Tv;
    if (i0 >= istep) return;
    int gcol = tv * g.Tv + j + g.nv / 2;
    if (gcol >= g.nv) gcol -= g.nv;
    for (int o = 0; o < NOUT; ++o)
        for (int pl = 0; pl < np; ++pl) {
            cx<T>* base = grid + (long long)o * P * plane + (long long)(p0 + pl) * plane + gcol;
            for (int i = i0; i < g.Tu; i += istep) {
                int grow = tu * g.Tu + i + g.nu / 2;
                if (grow >= g.nu) grow -= g.nu;
                base[(long long)grow * g.nv] = cx<T>{T(0), T(0)};
            }
        }
}

// occ[group][tile] = 1 when a gridding item reaches the (tile, plane group).
__global__ __launch_bounds__(256) void k_occupancy(const std::uint32_t* __restrict__ item_off,
                                                   const std::uint32_t* __restrict__ order, int ntg, int ngroups,
                                                   int ntiles, std::uint8_t* __restrict__ occ) {
    const int tg = blockIdx.x * 256 + threadIdx.x;
    if (tg >= ntg) return;
    const int rank = tg / ngroups, grp = tg - rank * ngroups;
    occ[(long long)grp * ntiles + order[rank]] = item_off[tg + 1] > item_off[tg] ? 1 : 0;
}

// The item `item` of a prefix array of per-unit item counts: the unit and
// the chunk within it, or false past the last item.
__device__ __forceinline__ bool find_item(const std::uint32_t* item_off, int nunits, std::uint32_t item,
                                          int& unit, std::uint32_t& chunk, std::uint32_t& nchunks) {
    if (item >= item_off[nunits]) return false;
    int lo = 0, hi = nunits;  // item_off[lo] <= item < item_off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (item_off[mid] <= item)
            lo = mid;
        else
            hi = mid;
    }
    unit = lo;
    chunk = item - item_off[lo];
    nchunks = item_off[lo + 1] - item_off[lo];
    return true;
}

constexpr int kGridBatch = 512;  // records staged in LDS at a time

// Gridding of one item: a chunk of the visibilities reaching one tile on
// one group of planes (see k_grid_items), accumulated in LDS and added to the
// planes (written when the tile has one item, added with global atomics into
// zeroed cells otherwise). The workgroup stages the chunk's records in LDS a
// batch at a time, one coalesced load per thread, the next batch prefetched
// into registers while the current one is gridded. One wave takes one
// visibility at a time (RPW of them for supports up to 5), one lane per tap:
// the lanes of an LDS atomic always address distinct cells, however many
// visibilities share a position (the channels and times of a short baseline
// do). NOUT = 2 grids a dirty image and a PSF together
// (wgridder.rs:2128-2152). Row stride TvP is padded so that the S x S
// footprint spreads over the banks. The accumulators are double in both
// precisions: LDS adds of doubles run some twenty times faster than those of
// floats on CDNA4, and the sums only gain accuracy.
using GridAcc = double;

// Record j of an item's concatenated source ranges, with its in-tile offsets
// made relative to the item's tile (biased by 64: tiles are at most 64 wide,
// so the offsets lie in [-64, 64)). With w-gridding its w becomes the
// argument of the w kernel's taps, formed once per record here (one lane per
// record) rather than per tap in the gridding loop (one wave per record).
// w_factor's argument (z - nth) 2 + S - 1 with z = xdw (wmin + (mp + nth) dw
// - w) is, as xdw = 1 / dw, the same for every tap nth:
// 2 xdw (wmin + mp dw - w) + S - 1 (wgridder.rs:1580-1591, 696-698).
// Its first plane mp becomes, for the item's planes p0 .. p1 - 1, the range
// [tlo, thi) of its taps on them and the item plane q0 of tap tlo (5 bits
// each above the offsets; mp = p0 + q0 - tlo).
// The w taps' argument of a staged record, in the T just before x0 (the
// upper half of w in single precision, w itself in double), so that the
// arguments of the u, v and w taps lie at x0, x0 + 1 and x0 - 1.
template <class T>
__device__ __forceinline__ T warg_of(const Rec<T>& r) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(&r) + offsetof(Rec<T>, x0) - sizeof(T));
}

__device__ __forceinline__ void taps_on_item(std::uint32_t pos, int p0, int& tlo, int& thi, int& q0, int& mp) {
    tlo = int((pos >> 16) & 31u);
    thi = int((pos >> 21) & 31u);
    q0 = int((pos >> 26) & 31u);
    mp = p0 + q0 - tlo;
}

template <class T>
__device__ __forceinline__ Rec<T> grid_fetch(const Geo& g, const Rec<T>* __restrict__ recs,
                                             const std::uint32_t* s_start,
                                             const std::uint32_t* s_pre, std::uint32_t j, int p0,
                                             int p1, const cx<T>* __restrict__ stored) {
    int src = 0;
    while (j >= s_pre[src + 1]) ++src;
    const std::uint32_t at = s_start[src] + (j - s_pre[src]);
    Rec<T> r = recs[at];
    if (stored) {
        // the data value of a stored value v and the record's unit value p
        // (wgridder.rs:1432-1435): v p with p's imaginary part FixW-signed
        const T fl = (r.idx & kFlip) ? T(-1) : T(1);
        const cx<T> d = cmul(stored[at], cx<T>{r.re, r.im * fl});
        r.re = d.re;
        r.im = d.im * fl;
    }
    int du, dv, cat;
    source_of(src, du, dv, cat);
    const int ou = int(r.pos & 255u) - (du ? g.Tu : 0) + 64;
    const int ov = int((r.pos >> 8) & 255u) - (dv ? g.Tv : 0) + 64;
    const int mp = int(r.pos >> 16);
    r.pos = std::uint32_t(ou) | (std::uint32_t(ov) << 8);
    if (g.wgrid) {
        const T wa = T(2.0 * g.xdw * (g.wmin + double(mp) * g.dw - r.w) + (double(g.supp) - 1.0));
        if constexpr (sizeof(T) == 4)  // the upper half of w (little-endian)
            r.w = __longlong_as_double((long long)((unsigned long long)__float_as_uint(float(wa)) << 32));
        else
            r.w = double(wa);
        const int tlo = max(0, p0 - mp), thi = min(g.supp, p1 - mp);
        r.pos |= (std::uint32_t(tlo) << 16) | (std::uint32_t(thi) << 21) | (std::uint32_t(mp + tlo - p0) << 26);
    }
    return r;
}

// The gridding workgroup's LDS besides its accumulators: the kernel's
// coefficients for supports up to S, the item's nine source ranges and the
// staged records.
template <class T, int S>
struct GridShared {
    T coef[2 * ne_of<S>() * S];
    std::uint32_t start[9], pre[10];
    Rec<T> rec[kGridBatch];
};

template <class T, int S, int NOUT, bool FORM>
__device__ __forceinline__ void grid_item(Geo g, const Rec<T>* __restrict__ recs,
                                          const std::uint32_t* __restrict__ off,
                                          const T* __restrict__ coeff, DevIn in,
                                          int G, int ngroups, int TvP, std::uint32_t rc,
                                          const std::uint32_t* __restrict__ item_off,
                                          const std::uint32_t* __restrict__ order,
                                          cx<T>* __restrict__ grid, std::uint32_t item, T* coef,
                                          std::uint32_t* s_start, std::uint32_t* s_pre, Rec<T>* s_rec) {
    constexpr int NE = ne_of<S>();
    constexpr int SS = S * S;
    constexpr int RPW = SS <= 16 ? 4 : SS <= 32 ? 2 : 1;
    constexpr int LPG = 64 / RPW;
    constexpr int CPL = (SS + LPG - 1) / LPG;
    constexpr int NT = grid_threads<T>(), NW = NT / 64;
    // 16-byte aligned: the dynamic segment follows the static arrays above,
    // and LDS double atomics need naturally aligned addresses.
    extern __shared__ __align__(16) unsigned char smem[];
    GridAcc* acc = reinterpret_cast<GridAcc*>(smem);
    const int ntg = g.ntu * g.ntv * ngroups;
    int tg;
    std::uint32_t chunk, nchunks;
    if (!find_item(item_off, ntg, item, tg, chunk, nchunks)) return;
    const int rank = tg / ngroups, grp = tg - rank * ngroups;
    const int tile = int(order[rank]);
    const int tu = tile / g.ntv, tv = tile - tu * g.ntv;
    const int P = g.wgrid ? g.nplanes : 1;
    const int p0 = grp * G, p1 = min(P, p0 + G), np = p1 - p0;
    const int pstride = g.Tu * TvP;  // one plane of one output, one of re / im
    for (int e = threadIdx.x; e < NOUT * np * 2 * pstride; e += NT) acc[e] = GridAcc(0);
    for (int e = threadIdx.x; e < 2 * NE * S; e += NT) coef[e] = coeff[e];
    const int mlo = g.wgrid ? max(0, p0 - S + 1) : 0;
    const int mhi = g.wgrid ? min(g.nb - 1, p1 - 1) : 0;
    if (threadIdx.x < 9) {
        std::uint32_t a, b;
        source_range(g, off, tu, tv, int(threadIdx.x), mlo, mhi, a, b);
        s_start[threadIdx.x] = a;
        s_pre[threadIdx.x + 1] = b - a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_pre[0] = 0;
        for (int k = 1; k <= 9; ++k) s_pre[k] += s_pre[k - 1];
    }
    __syncthreads();
    const std::uint32_t jbeg = chunk * rc, jend = min(s_pre[9], jbeg + rc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int group = lane / LPG, local = lane - group * LPG;
    int cu[CPL], cv[CPL];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
        const int q = min(local + k * LPG, SS - 1);
        cu[k] = q / S;
        cv[k] = q - (q / S) * S;
    }
    const int srclane = group * LPG;  // lane of the group holding w tap `local`
    // With one tap per lane its three polynomials live in registers.
    constexpr bool REG = CPL == 1;
    TapReg<T, S> tru, trv, trw;
    if constexpr (REG) {
        tru.load(coef, cu[0]);
        trv.load(coef, cv[0]);
        trw.load(coef, min(local, S - 1));
    }
    // One visibility per wave (the w-gridding fast path): lane L evaluates
    // tap L mod S of kind L / S, the u taps at x0, the v taps at y0 and the w
    // taps of planes mp .. mp + S - 1; the u-v lanes fetch theirs across lanes.
    const int tk = lane % S, kind = lane / S;
    TapReg<T, S> trl;
    if constexpr (REG && RPW == 1) trl.load(coef, tk);
    // the lane's tap argument: x0 (u), x0 + 1 (y0, v) or x0 - 1 (w), in T
    // units from x0; the lane's cell offset from the visibility's first cell
    const int argoff = kind == 0 ? 0 : kind == 1 ? 1 : kind == 2 ? -1 : 0;
    unsigned lane8 = unsigned(cu[0] * TvP + cv[0]) * unsigned(sizeof(GridAcc));
    // opaque to the optimiser, which would otherwise fold it into a per-visit
    // (cu + ou) TvP multiply
    asm volatile("" : "+v"(lane8));
    const unsigned pstride8 = unsigned(pstride) * unsigned(sizeof(GridAcc));
    const bool stager = threadIdx.x < kGridBatch;
    Rec<T> nxt{};
    const cx<T>* stored = static_cast<const cx<T>*>(in.stored);
    if (stager && jbeg + threadIdx.x < jend)
        nxt = grid_fetch(g, recs, s_start, s_pre, jbeg + threadIdx.x, p0, p1, stored);
    for (std::uint32_t base = jbeg; base < jend; base += kGridBatch) {
        const int nbat = int(min(jend - base, std::uint32_t(kGridBatch)));
        __syncthreads();  // the previous batch is consumed
        if (int(threadIdx.x) < nbat) s_rec[threadIdx.x] = nxt;
        __syncthreads();
        const std::uint32_t jn = base + kGridBatch + threadIdx.x;
        if (stager && jn < jend) nxt = grid_fetch(g, recs, s_start, s_pre, jn, p0, p1, stored);
        if constexpr (REG && RPW == 1) {
            if (g.wgrid) {
                // One visibility per wave: its position is wave-uniform, and
                // only its taps on the item's planes are visited, each w tap
                // read from its lane. The loops are scalar: the scalar unit
                // (one instruction per SIMD every four clocks) bounds this
                // loop before the vector unit or the LDS does.
                const int wv = __builtin_amdgcn_readfirstlane(wave);
                for (int r = wv; r < nbat; r += NW) {
                    const Rec<T>& rc_ = s_rec[r];
                    // the position wave-uniform (scalar control of the plane
                    // loop); the other fields are broadcast reads in every lane
                    const std::uint32_t pos = uni(rc_.pos);
                    const int ou = int(pos & 255u) - 64, ov = int((pos >> 8) & 255u) - 64;
                    int tlo, thi, q0, mp;
                    taps_on_item(pos, p0, tlo, thi, q0, mp);
                    const T arg = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(&rc_) +
                                                              offsetof(Rec<T>, x0) + argoff * int(sizeof(T)));
                    cx<T> val[NOUT];
                    if constexpr (NOUT == 2 || FORM) {
                        const std::uint32_t idx = uni(rc_.idx);
                        const T re = rc_.re, im = rc_.im;
                        const T fl = (idx & kFlip) ? T(-1) : T(1);
                        const cx<T> d = cmul(vis_get<T>(in, idx & ~kFlip), cx<T>{re, im * fl});
                        val[0] = {d.re, d.im * fl};
                        if constexpr (NOUT == 2) val[1] = {re, im};
                    } else {
                        val[0] = {rc_.re, rc_.im};
                    }
                    const T tval = trl.eval(arg);
                    const T ku = __shfl(tval, cu[0], 64), kv = __shfl(tval, S + cv[0], 64);
                    const GridAcc kd = GridAcc(ku * kv);
                    const int iu = ou + cu[0], iv = ov + cv[0];
                    const bool inside = local < SS && unsigned(iu) < unsigned(g.Tu) &&
                                        unsigned(iv) < unsigned(g.Tv);
                    GridAcc a[NOUT], b[NOUT];
#pragma unroll
                    for (int o = 0; o < NOUT; ++o) {
                        a[o] = GridAcc(val[o].re) * kd;
                        b[o] = GridAcc(val[o].im) * kd;
                    }
                    // taps t in [tlo, thi) fall on the item's planes q0, q0 + 1, ...
                    if (inside) {
                        // byte offsets: the lane's cell, plus the wave-uniform
                        // first cell and plane
                        unsigned base8 = unsigned((q0 * 2 * pstride + ou * TvP + ov)) * unsigned(sizeof(GridAcc));
                        for (int t = 2 * S + tlo; t < 2 * S + thi; ++t) {
                            const GridAcc kw = GridAcc(lane_value(tval, t));
                            const unsigned re8 = lane8 + base8;
#pragma unroll
                            for (int o = 0; o < NOUT; ++o) {
                                const unsigned oo = unsigned(o * np * 2) * pstride8;
                                atomicAdd(reinterpret_cast<GridAcc*>(smem + re8 + oo), a[o] * kw);
                                atomicAdd(reinterpret_cast<GridAcc*>(smem + re8 + oo + pstride8), b[o] * kw);
                            }
                            base8 += 2 * pstride8;
                        }
                    }
                }
                continue;
            }
        }
        for (int rw = wave * RPW; rw < nbat; rw += NW * RPW) {
            const int r = rw + group;
            const bool valid = r < nbat;
            const Rec<T>& rc_ = s_rec[valid ? r : rw];
            const std::uint32_t pos = rc_.pos;
            const int ou = int(pos & 255u) - 64, ov = int((pos >> 8) & 255u) - 64;
            int tlo, thi, q0, mp;
            taps_on_item(pos, p0, tlo, thi, q0, mp);
            const T x0 = rc_.x0, y0 = rc_.y0;
            cx<T> val[NOUT];
            if constexpr (NOUT == 2 || FORM) {
                const T fl = (rc_.idx & kFlip) ? T(-1) : T(1);
                const cx<T> d = cmul(vis_get<T>(in, rc_.idx & ~kFlip), cx<T>{rc_.re, rc_.im * fl});
                val[0] = {d.re, d.im * fl};
                if constexpr (NOUT == 2) val[1] = {rc_.re, rc_.im};
            } else {
                val[0] = {rc_.re, rc_.im};
            }
            T kuv[CPL];
            int at[CPL];
            bool inside[CPL];
#pragma unroll
            for (int k = 0; k < CPL; ++k) {
                if constexpr (REG)
                    kuv[k] = tru.eval(x0) * trv.eval(y0);
                else
                    kuv[k] = eval_tap<T, S>(coef, cu[k], x0) * eval_tap<T, S>(coef, cv[k], y0);
                const int iu = ou + cu[k], iv = ov + cv[k];
                inside[k] = valid && local + k * LPG < SS && unsigned(iu) < unsigned(g.Tu) &&
                            unsigned(iv) < unsigned(g.Tv);
                at[k] = iu * TvP + iv;
            }
            // The products with the taps are formed in the accumulators'
            // precision (as fast as single precision on CDNA4, and no
            // conversion per plane).
            GridAcc vre[NOUT], vim[NOUT], kd[CPL];
#pragma unroll
            for (int o = 0; o < NOUT; ++o) {
                vre[o] = GridAcc(val[o].re);
                vim[o] = GridAcc(val[o].im);
            }
#pragma unroll
            for (int k = 0; k < CPL; ++k) kd[k] = GridAcc(kuv[k]);
            if (g.wgrid) {
                // lane `local` < S evaluates the w tap `local` (plane mp + local)
                const int nth = min(local, S - 1);
                T kwv;
                // the record holds the taps' argument (grid_fetch)
                if constexpr (REG)
                    kwv = trw.eval(warg_of(rc_));
                else
                    kwv = eval_tap<T, S>(coef, nth, warg_of(rc_));
#pragma unroll
                for (int nth2 = 0; nth2 < S; ++nth2) {
                    const GridAcc kw = GridAcc(__shfl(kwv, srclane + nth2, 64));
                    const int p = mp + nth2;
                    if (p < p0 || p >= p1) continue;
#pragma unroll
                    for (int k = 0; k < CPL; ++k) {
                        if (!inside[k]) continue;
                        const GridAcc f = kd[k] * kw;
#pragma unroll
                        for (int o = 0; o < NOUT; ++o) {
                            GridAcc* are = acc + ((o * np + (p - p0)) * 2) * pstride + at[k];
                            atomicAdd(are, vre[o] * f);
                            atomicAdd(are + pstride, vim[o] * f);
                        }
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < CPL; ++k) {
                    if (!inside[k]) continue;
#pragma unroll
                    for (int o = 0; o < NOUT; ++o) {
                        GridAcc* are = acc + (o * 2) * pstride + at[k];
                        atomicAdd(are, vre[o] * kd[k]);
                        atomicAdd(are + pstride, vim[o] * kd[k]);
                    }
                }
            }
        }
    }
    __syncthreads();
    const long long plane = (long long)g.nu * g.nv;
    // a fixed column per thread, rows in steps (no division per cell)
    const int istep = NT / g.Tv;
    const int i0 = int(threadIdx.x) / g.Tv, j = int(threadIdx.x) - i0 * g.Tv;
    if (i0 >= istep) return;
    int gcol = tv * g.Tv + j + g.nv / 2;
    if (gcol >= g.nv) gcol -= g.nv;
    for (int pl = 0; pl < np; ++pl)
        for (int i = i0; i < g.Tu; i += istep) {
            int grow = tu * g.Tu + i + g.nu / 2;
            if (grow >= g.nu) grow -= g.nu;
#pragma unroll
            for (int o = 0; o < NOUT; ++o) {
                const GridAcc* are = acc + ((o * np + pl) * 2) * pstride + i * TvP + j;
                cx<T>* dst = grid + (long long)o * P * plane + (long long)(p0 + pl) * plane +
                             (long long)grow * g.nv + gcol;
                if (nchunks == 1) {
                    *dst = cx<T>{T(are[0]), T(are[pstride])};
                } else if (are[0] != GridAcc(0) || are[pstride] != GridAcc(0)) {
                    // a chunk adds only the cells its visibilities reached
                    atomicAdd(&dst->re, T(are[0]));
                    atomicAdd(&dst->im, T(are[pstride]));
                }
            }
        }
}

template <class T, int S, int NOUT, bool FORM>
__global__ __launch_bounds__(grid_threads<T>()) void k_grid(Geo g, const Rec<T>* __restrict__ recs,
                                                       const std::uint32_t* __restrict__ off,
                                                       const T* __restrict__ coeff, DevIn in,
                                                       int G, int ngroups, int TvP, std::uint32_t rc,
                                                       const std::uint32_t* __restrict__ item_off,
                                                       const std::uint32_t* __restrict__ order,
                                                       cx<T>* __restrict__ grid) {
    __shared__ GridShared<T, S> sh;
    grid_item<T, S, NOUT, FORM>(g, recs, off, coeff, in, G, ngroups, TvP, rc, item_off, order, grid, blockIdx.x,
                                sh.coef, sh.start, sh.pre, sh.rec);
}

// One image's gridding in a launch that grids several (a sub-band batch):
// its geometry (with its support), sorted records, tables and items, and its
// first workgroup.
template <class T>
struct GridJob {
    Geo g;
    const Rec<T>* recs;
    const std::uint32_t* off;
    const T* coeff;
    DevIn in;
    int G, ngroups, TvP;
    std::uint32_t rc;
    const std::uint32_t* item_off;
    const std::uint32_t* order;
    cx<T>* grid;
    std::uint32_t first;
};

template <class T>
constexpr int max_support() {
    return sizeof(T) == 4 ? 8 : 16;
}

// Every job's items in one launch: workgroup b takes item b - first of the
// last job whose first workgroup is at most b (jobs in order of first), and
// runs the gridding of that job's support. The LDS besides the accumulators
// is laid out once for the largest support.
template <class T, int NOUT, bool FORM>
__global__ __launch_bounds__(grid_threads<T>()) void k_grid_jobs(const GridJob<T>* __restrict__ jobs, int njobs) {
    __shared__ GridShared<T, max_support<T>()> sh;
    int j = 0;
    for (int k = 1; k < njobs; ++k)
        if (jobs[k].first <= blockIdx.x) j = k;
    const GridJob<T>& J = jobs[j];
    const std::uint32_t item = blockIdx.x - J.first;
#define JOB_CASE(SS)                                                                                   \
    case SS:                                                                                           \
        if constexpr (SS <= max_support<T>())                                                          \
            grid_item<T, SS, NOUT, FORM>(J.g, J.recs, J.off, J.coeff, J.in, J.G, J.ngroups, J.TvP, J.rc, \
                                         J.item_off, J.order, J.grid, item, sh.coef, sh.start, sh.pre,  \
                                         sh.rec);                                                      \
        break;
    switch (J.g.supp) {
        JOB_CASE(4) JOB_CASE(5) JOB_CASE(6) JOB_CASE(7) JOB_CASE(8) JOB_CASE(9) JOB_CASE(10) JOB_CASE(11)
        JOB_CASE(12) JOB_CASE(13) JOB_CASE(14) JOB_CASE(15) JOB_CASE(16)
    default:
        break;
    }
#undef JOB_CASE
}

// ------------------------------------------------------------ gridding on the matrix cores

// Single-precision w-gridding on 16 x 16 tiles with the matrix cores: one
// wave per item (tile, group of up to kMfmaPlanes planes, chunk), whose
// accumulators are the item's planes of the tile held in the MFMA result
// registers, so that no LDS atomics are needed. Four visibilities at a
// time: with A[i][k] = ku_k(i - ou_k) kw_k(plane) (lane 16 k + i) and
// B[k][j] = kv_k(j - ov_k) v_k (lane 16 k + j), one
// v_mfma_f32_16x16x4_f32 per plane and part adds sum_k A[i][k] B[k][j] to
// cell (i, j) of the plane, which is the gridding sum (wgridder.rs:2098-2156)
// for those four visibilities; taps outside a visibility's footprint or
// planes are zero. Lane 16 k + t evaluates tap t of visibility k at its u,
// v and w arguments once; the operands fetch theirs across lanes. The
// result register r of lane l holds cell (4 (l / 16) + r, l mod 16).
// Planes per item and the occupancy target, measured on the 464M workload
// (gridding stage): 8 planes at 4 waves per SIMD 103.5 ms (117.4 without the
// target), 6 at 5 waves 105.9, 4 at 6 waves 111.4.
constexpr int kMfmaPlanes = 8;

// Lane q of each row of 16 lanes, broadcast to its row (DPP row_newbcast;
// q is a constant once the plane loop is unrolled).
__device__ __forceinline__ int row_bcast(int v, int q) {
    switch (q) {
    case 0: return __builtin_amdgcn_mov_dpp(v, 0x150, 0xf, 0xf, false);
    case 1: return __builtin_amdgcn_mov_dpp(v, 0x151, 0xf, 0xf, false);
    case 2: return __builtin_amdgcn_mov_dpp(v, 0x152, 0xf, 0xf, false);
    case 3: return __builtin_amdgcn_mov_dpp(v, 0x153, 0xf, 0xf, false);
    case 4: return __builtin_amdgcn_mov_dpp(v, 0x154, 0xf, 0xf, false);
    case 5: return __builtin_amdgcn_mov_dpp(v, 0x155, 0xf, 0xf, false);
    case 6: return __builtin_amdgcn_mov_dpp(v, 0x156, 0xf, 0xf, false);
    default: return __builtin_amdgcn_mov_dpp(v, 0x157, 0xf, 0xf, false);
    }
}
static_assert(kMfmaPlanes <= 8, "row_bcast covers planes 0-7");
constexpr int kMfmaTile = 16;
typedef float mfma_v4 __attribute__((ext_vector_type(4)));

template <int S>
struct MfmaShared {
    float coef[2 * ne_of<S>() * S];
    std::uint32_t start[9], pre[10];
    Rec<float> rec[64];
    int qr[32];  // per group of 4 staged records: its planes [lo, hi) in the item
};

template <int S, bool FORM>
__device__ __forceinline__ void grid_item_mfma(Geo g, const Rec<float>* __restrict__ recs,
                                               const std::uint32_t* __restrict__ off,
                                               const float* __restrict__ coeff, DevIn in, int G, int ngroups,
                                               std::uint32_t rc, const std::uint32_t* __restrict__ item_off,
                                               const std::uint32_t* __restrict__ order,
                                               cx<float>* __restrict__ grid, std::uint32_t item, float* coef,
                                               std::uint32_t* s_start, std::uint32_t* s_pre, Rec<float>* s_rec,
                                               int* s_qr) {
    constexpr int NE = ne_of<S>();
    const int lane = threadIdx.x;
    const int ntg = g.ntu * g.ntv * ngroups;
    int tg;
    std::uint32_t chunk, nchunks;
    if (!find_item(item_off, ntg, item, tg, chunk, nchunks)) return;
    const int rank = tg / ngroups, grp = tg - rank * ngroups;
    const int tile = int(order[rank]);
    const int tu = tile / g.ntv, tv = tile - tu * g.ntv;
    const int P = g.nplanes;
    const int p0 = grp * G, p1 = min(P, p0 + G), np = p1 - p0;
    for (int e = lane; e < 2 * NE * S; e += 64) coef[e] = coeff[e];
    const int mlo = max(0, p0 - S + 1), mhi = min(g.nb - 1, p1 - 1);
    if (lane < 9) {
        std::uint32_t a, b;
        source_range(g, off, tu, tv, lane, mlo, mhi, a, b);
        s_start[lane] = a;
        s_pre[lane + 1] = b - a;
    }
    __syncthreads();
    if (lane == 0) {
        s_pre[0] = 0;
        for (int k = 1; k <= 9; ++k) s_pre[k] += s_pre[k - 1];
    }
    __syncthreads();
    const std::uint32_t jbeg = chunk * rc, jend = min(s_pre[9], jbeg + rc);
    const int k = lane >> 4, i = lane & 15;
    // lane 16 k + i holds tap i (u and w) for i < S and tap i - S (v) for
    // S <= i < 2 S of visibility k: the u and v taps in one evaluation
    static_assert(2 * S <= 16, "u and v taps share a row of 16 lanes");
    const bool lane_u = i < S;
    TapReg<float, S> tr;
    tr.load(coef, lane_u ? i : min(i - S, S - 1));
    const cx<float>* stored = static_cast<const cx<float>*>(in.stored);
    mfma_v4 are[kMfmaPlanes], aim[kMfmaPlanes];
#pragma unroll
    for (int q = 0; q < kMfmaPlanes; ++q) {
        are[q] = mfma_v4{0.f, 0.f, 0.f, 0.f};
        aim[q] = mfma_v4{0.f, 0.f, 0.f, 0.f};
    }
    for (std::uint32_t base = jbeg; base < jend; base += 64) {
        const int nrec = int(min(jend - base, 64u));
        Rec<float> r{};
        if (lane < nrec) {
            r = grid_fetch(g, recs, s_start, s_pre, base + lane, p0, p1, stored);
            if constexpr (FORM) {
                const float fl = (r.idx & kFlip) ? -1.f : 1.f;
                const cx<float> d = cmul(vis_get<float>(in, r.idx & ~kFlip), cx<float>{r.re, r.im * fl});
                r.re = d.re;
                r.im = d.im * fl;
            }
        } else {
            // no taps anywhere: offsets far outside the tile, an empty plane span
            r.pos = 0u;
            r.re = r.im = 0.f;
        }
        // the planes of each group of four records: the union of their
        // item planes, reduced over the four lanes once per batch
        {
            int tlo, thi, q0, mp;
            taps_on_item(r.pos, p0, tlo, thi, q0, mp);
            int qa = thi > tlo ? q0 : kMfmaPlanes, qb = thi > tlo ? q0 + thi - tlo : 0;
            qa = min(qa, __shfl_xor(qa, 1, 64));
            qb = max(qb, __shfl_xor(qb, 1, 64));
            qa = min(qa, __shfl_xor(qa, 2, 64));
            qb = max(qb, __shfl_xor(qb, 2, 64));
            __syncthreads();  // the previous batch is consumed
            s_rec[lane] = r;
            if ((lane & 3) == 0) {
                s_qr[lane >> 1] = qa;
                s_qr[(lane >> 1) + 1] = min(qb, np);
            }
        }
        __syncthreads();
        for (int g4 = 0; g4 < nrec; g4 += 4) {
            const Rec<float>& rk = s_rec[g4 + k];
            const std::uint32_t pos = rk.pos;
            const int ou = int(pos & 255u) - 64, ov = int((pos >> 8) & 255u) - 64;
            int tlo, thi, q0, mp;
            taps_on_item(pos, p0, tlo, thi, q0, mp);
            const float tUV = tr.eval(lane_u ? rk.x0 : rk.y0);
            // w taps zero on the lanes past the support: a fetch of tap t
            // of visibility k at lane 16 k + t for t in [-8, 16) then reads
            // zero outside [0, S) (lane indices wrap modulo 64)
            const float tW = lane_u ? tr.eval(warg_of(rk)) : 0.f;
            const int au = i - ou, av = i - ov;
            const bool inu = au >= 0 && au < S, inv = av >= 0 && av < S;
            const float ku = __shfl(tUV, 16 * k + (inu ? au : 0), 64);
            const float kv = __shfl(tUV, 16 * k + S + (inv ? av : 0), 64);
            const float A0 = inu ? ku : 0.f;
            const float Bk = inv ? kv : 0.f;
            const float bre = Bk * rk.re, bim = Bk * rk.im;
            // the group's planes: the union of its visibilities' item planes
            const int qlo = __builtin_amdgcn_readfirstlane(s_qr[g4 >> 1]);
            const int qhi = __builtin_amdgcn_readfirstlane(s_qr[(g4 >> 1) + 1]);
            // plane q of the item takes tap t = q - q0 + tlo of visibility
            // k: in [tlo, thi) on its planes, outside [0, S) (a zero lane)
            // on the item's other planes (q < q0 means tlo = 0; q past its
            // planes means thi = S or q >= np)
            // One permute per group puts the tap of plane j at lane 16 k + j
            // (j < kMfmaPlanes); plane q then broadcasts lane q of each row of
            // 16 lanes (DPP row_newbcast, a register move on the VALU) instead
            // of waiting on a permute through the LDS per plane.
            const int wbase = 4 * (16 * k + tlo - q0);
            const int wq = __builtin_amdgcn_ds_bpermute(wbase + 4 * i, __float_as_int(tW));
#pragma unroll
            for (int q = 0; q < kMfmaPlanes; ++q) {
                if (q < qlo || q >= qhi) continue;
                const float kw = __int_as_float(row_bcast(wq, q));
                const float a = A0 * kw;
                are[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bre, are[q], 0, 0, 0);
                aim[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bim, aim[q], 0, 0, 0);
            }
        }
    }
    // flush: register r of lane l is cell (4 (l / 16) + r, l mod 16)
    const long long plane = (long long)g.nu * g.nv;
    int gcol = tv * kMfmaTile + i + g.nv / 2;
    if (gcol >= g.nv) gcol -= g.nv;
#pragma unroll
    for (int q = 0; q < kMfmaPlanes; ++q) {
        if (q >= np) break;
        cx<float>* dst = grid + (long long)(p0 + q) * plane + gcol;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            int grow = tu * kMfmaTile + 4 * k + rr + g.nu / 2;
            if (grow >= g.nu) grow -= g.nu;
            const float vre = are[q][rr], vim = aim[q][rr];
            cx<float>* c = dst + (long long)grow * g.nv;
            if (nchunks == 1) {
                *c = cx<float>{vre, vim};
            } else if (vre != 0.f || vim != 0.f) {
                atomicAdd(&c->re, vre);
                atomicAdd(&c->im, vim);
            }
        }
    }
}

template <int S, bool FORM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_grid_mfma(Geo g, const Rec<float>* __restrict__ recs,
                                                  const std::uint32_t* __restrict__ off,
                                                  const float* __restrict__ coeff, DevIn in, int G,
                                                  int ngroups, std::uint32_t rc,
                                                  const std::uint32_t* __restrict__ item_off,
                                                  const std::uint32_t* __restrict__ order,
                                                  cx<float>* __restrict__ grid) {
    __shared__ MfmaShared<S> sh;
    grid_item_mfma<S, FORM>(g, recs, off, coeff, in, G, ngroups, rc, item_off, order, grid, blockIdx.x, sh.coef,
                            sh.start, sh.pre, sh.rec, sh.qr);
}

// Every job's items in one launch on the matrix cores (jobs as k_grid_jobs,
// each gridded by k_grid_mfma's item body at its support).
template <bool FORM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_grid_mfma_jobs(const GridJob<float>* __restrict__ jobs, int njobs) {
    __shared__ MfmaShared<8> sh;
    int j = 0;
    for (int k = 1; k < njobs; ++k)
        if (jobs[k].first <= blockIdx.x) j = k;
    const GridJob<float>& J = jobs[j];
    const std::uint32_t item = blockIdx.x - J.first;
    switch (J.g.supp) {
#define MJOB_CASE(SS)                                                                                    \
    case SS:                                                                                             \
        grid_item_mfma<SS, FORM>(J.g, J.recs, J.off, J.coeff, J.in, J.G, J.ngroups, J.rc, J.item_off,    \
                                 J.order, J.grid, item, sh.coef, sh.start, sh.pre, sh.rec, sh.qr);       \
        break;
        MJOB_CASE(4) MJOB_CASE(5) MJOB_CASE(6) MJOB_CASE(7) MJOB_CASE(8)
#undef MJOB_CASE
    default:
        break;
    }
}

// ------------------------------------------------------------ transforms

// Largest transform held in LDS: 128 KiB of complex values.
template <class T>
constexpr int fft_maxtot() {
    return sizeof(T) == 4 ? 16384 : 8192;
}

// Per-pixel tables (wgridder.rs:806-861): n-1+nshift for the w-screens and
// the global correction factor, on the quarter (or full with an l/m shift)
// table of the reference.
__device__ __forceinline__ double nm1_of(double x, double y, double r_l, double r_m) {
    const double xsq = x * x, ysq = y * y, tmp = 1.0 - xsq - ysq;
    const double nm1 = tmp >= 0.0 ? (-xsq - ysq) / (sqrt(tmp) + 1.0) : -sqrt(-tmp) - 1.0;
    return nm1 + r_l * x + r_m * y;
}

__device__ __forceinline__ double corfunc(const double* glx, const double* glw, int ngl, double v) {
    double t = 0.0;
    for (int k = 0; k < ngl; ++k) t += glw[k] * cos(glx[k] * v);
    return 1.0 / t;
}

__global__ __launch_bounds__(256) void k_tables(Geo g, double* screen, double* corr,
                                                const double* cfu, const double* cfv,
                                                const double* glx, const double* glw, int ngl) {
    const long long total = (long long)g.tnx * g.tny;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int i, j;
    if (g.cm) {
        j = int(t / g.tnx);
        i = int(t - (long long)j * g.tnx);
    } else {
        i = int(t / g.tny);
        j = int(t - (long long)i * g.tny);
    }
    const double x0 = g.lshift - 0.5 * double(g.nx) * g.pixsize_x;
    const double y0 = g.mshift - 0.5 * double(g.ny) * g.pixsize_y;
    const double x = x0 + double(i) * g.pixsize_x, y = y0 + double(j) * g.pixsize_y;
    screen[t] = nm1_of(x, y, g.r_l, g.r_m) + g.nshift;
    const double xsq = x * x, ysq = y * y, tmp = 1.0 - xsq - ysq;
    double fct;
    if (tmp >= 0.0) {
        const double nm1 = (-xsq - ysq) / (sqrt(tmp) + 1.0);
        fct = corfunc(glx, glw, ngl, (nm1 + g.r_l * x + g.r_m * y + g.nshift) * g.dw);
        if (g.divide_by_n) fct /= nm1 + 1.0;
    } else if (g.divide_by_n) {
        fct = 0.0;
    } else {
        fct = corfunc(glx, glw, ngl, (-sqrt(-tmp) - 1.0 + g.nshift) * g.dw);
    }
    const int iu = g.nx / 2 - min(i, g.nx - i), iv = g.ny / 2 - min(j, g.ny - j);
    corr[t] = fct * (cfu[iu] * cfv[iv]);
}

__device__ __forceinline__ int table_index(const Geo& g, int i, int j) {
    if (g.cm) {
        if (g.lmshift) return j * g.tnx + i;
        return min(j, g.ny - j) * g.tnx + min(i, g.nx - i);
    }
    if (g.lmshift) return i * g.tny + j;
    return min(i, g.nx - i) * g.tny + min(j, g.ny - j);
}

// The w-screen of plane w at table value s (wgridder.rs:2296-2314).
template <class T>
__device__ __forceinline__ cx<T> screen_at(double w, double s, bool adjoint) {
    double phs = w * s;
    if (adjoint) phs = -phs;
    T ang;
    if constexpr (sizeof(T) == 4)
        ang = T(2.0 * M_PI * (phs - floor(phs)));
    else
        ang = T(2.0 * M_PI * phs);
    T sn, cs;
    sincos_t(ang, sn, cs);
    return {cs, sn};
}

// Fills LDS from memory with every load of the block in flight before the
// first LDS write (a transform's loads are otherwise serialised on their
// latency): element e < count (count <= MT) is src[at(e).first], stored at
// buf[at(e).second].
template <class T, int NT, int MT, class At>
__device__ __forceinline__ void staged_fill(cx<T>* buf, const cx<T>* src, int count, At at) {
    constexpr int PER = (MT + NT - 1) / NT;
    if constexpr (sizeof(T) == 4) {
        cx<T> r[PER];
        int dst[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = threadIdx.x + k * NT;
            if (e < count) {
                const auto sd = at(e);
                r[k] = src[sd.first];
                dst[k] = sd.second;
            }
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = threadIdx.x + k * NT;
            if (e < count) buf[dst[k]] = r[k];
        }
    } else {
        // double: the values staged as two-lane vectors (staging arrays of the
        // complex struct are placed in scratch memory by the compiler)
        using V = __attribute__((ext_vector_type(2))) double;
        V r[PER];
        int dst[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = threadIdx.x + k * NT;
            if (e < count) {
                const auto sd = at(e);
                r[k] = *reinterpret_cast<const V*>(src + sd.first);
                dst[k] = sd.second;
            }
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = threadIdx.x + k * NT;
            if (e < count) *reinterpret_cast<V*>(buf + dst[k]) = r[k];
        }
    }
}

// A linear block index remapped so that the workgroups one XCD holds at a time
// take consecutive indices: the dispatcher deals consecutive workgroups to the
// eight XCDs in turn, and each XCD has its own L2, so blocks that read
// neighbouring columns should share an XCD.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    constexpr int X = 8;
    const int per = nblocks / X, rem = nblocks - per * X;
    const int x = bid % X, k = bid / X;
    return x * per + min(x, rem) + k;
}

// Which tiles of a grid row were gridded: with `occ` (per plane group and
// tile, nonzero when a gridding item reached it) the row transform reads
// only those tiles' cells and takes the others as zero, so cells no
// visibility reaches are neither cleared before gridding nor read after.
struct RowOcc {
    const std::uint8_t* occ;  // null: every cell is read
    int G;                    // planes per group
    unsigned mtv;             // udiv multiplier for the tile side Tv
};

// v2d: inverse transform along v of one grid row, keeping the image columns
// (wgridder.rs:2331-2338, 1745-1749).
template <class T, int NT, int MAXTOT>
__global__ __launch_bounds__(NT) void k_rows_inv(const cx<T>* __restrict__ grid,
                                                          cx<T>* __restrict__ tmp, FftDesc d,
                                                          const cx<T>* __restrict__ tw, int ny, Geo g, RowOcc ro) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nv = d.n;
    const long long row = blockIdx.x;
    const cx<T>* src = grid + row * nv;
    int nz = 0;
    if (ro.occ) {
        // the row's tiles: grid row u is tile row (u - nu/2) mod nu / Tu of
        // plane p, column v tile column (v - nv/2) mod nv / Tv
        const int p = int(row / g.nu), u = int(row - (long long)p * g.nu);
        int us = u - g.nu / 2;
        if (us < 0) us += g.nu;
        const std::uint8_t* orow = ro.occ + ((long long)(p / ro.G) * g.ntu + us / g.Tu) * g.ntv;
        // thread t takes the shifted columns vs = t + k NT (physical column
        // vs + nv/2 mod nv): its tile column is t / Tv + k NT / Tv when Tv
        // divides NT, one division per thread
        constexpr int PER = (MAXTOT + NT - 1) / NT;
        const int half = nv / 2;
        const bool even = NT % g.Tv == 0;
        const int t0 = int(udiv(threadIdx.x, unsigned(g.Tv), ro.mtv)), tstep = NT / g.Tv;
        // every occupancy byte first, then the loads of the occupied cells
        // (no wait between them)
        std::uint8_t oc[PER];
        if (even) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int vs = int(threadIdx.x) + k * NT;
                const std::uint8_t o = orow[min(t0 + k * tstep, g.ntv - 1)];
                oc[k] = vs < nv ? o : std::uint8_t(0);
            }
        } else {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int vs = int(threadIdx.x) + k * NT;
                const std::uint8_t o = orow[min(int(udiv(unsigned(vs), unsigned(g.Tv), ro.mtv)), g.ntv - 1)];
                oc[k] = vs < nv ? o : std::uint8_t(0);
            }
        }
        cx<T> r[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int vs = int(threadIdx.x) + k * NT;
            const int e = vs + half < nv ? vs + half : vs + half - nv;
            r[k] = oc[k] ? src[e] : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int vs = int(threadIdx.x) + k * NT;
            if (vs < nv) {
                buf[vs + half < nv ? vs + half : vs + half - nv] = r[k];
                nz |= (r[k].re != T(0) || r[k].im != T(0)) ? 1 : 0;
            }
        }
    } else {
        staged_fill<T, NT, MAXTOT>(buf, src, nv, [](int e) { return std::pair<long long, int>(e, e); });
        __syncthreads();
        for (int e = threadIdx.x; e < nv; e += NT) nz |= (buf[e].re != T(0) || buf[e].im != T(0)) ? 1 : 0;
    }
    // a row no visibility reached (beyond the coverage on this plane)
    // transforms to zero: write the zeros, skip the passes
    if (!__syncthreads_or(nz)) {
        for (int j = threadIdx.x; j < ny; j += NT) tmp[row * ny + j] = cx<T>{T(0), T(0)};
        return;
    }
    fft_lds<T, NT, MAXTOT, sizeof(T) == 4>(buf, d, 1, tw, +1);
    const int sh = nv - ny / 2;
    for (int j = threadIdx.x; j < ny; j += NT) {
        const int e = sh + j;
        tmp[row * ny + j] = buf[e >= nv ? e - nv : e];
    }
}

// v2d: inverse transform along u of C image columns of one plane, then the
// projection Re(g screen) of the image rows (wgridder.rs:2342-2397), one term
// per plane and pixel. SCM: tmp column-major ([P][ny][nu], each column
// contiguous), else row-major ([P][nu][ny]).
template <class T, int NT, int MAXTOT, bool SCM = false>
__global__ __launch_bounds__(NT) void k_cols_inv(const cx<T>* __restrict__ tmp,
                                                          T* __restrict__ proj, FftDesc d,
                                                          const cx<T>* __restrict__ tw, Geo g,
                                                          const double* __restrict__ screen, int C) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nu = d.n, ny = g.ny, nx = g.nx;
    const int ncb = (ny + C - 1) / C;
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int p = lin / ncb;
    const int j0 = (lin - p * ncb) * C, nc = min(C, ny - j0);
    const cx<T>* src = tmp + (long long)p * nu * ny;
    const unsigned mnc = nc > 1 ? unsigned(((1ull << 32) + nc - 1) / nc) : 0u;
    // nu * nc <= MAXTOT
    if constexpr (SCM) {
        staged_fill<T, NT, MAXTOT>(buf, src, nu * nc, [&](int e) {
            return std::pair<long long, int>((long long)j0 * nu + e, e);
        });
    } else {
        staged_fill<T, NT, MAXTOT>(buf, src, nu * nc, [&](int e) {
            const int u = int(udiv(unsigned(e), unsigned(nc), mnc)), c = e - u * nc;
            return std::pair<long long, int>((long long)u * ny + j0 + c, c * nu + u);
        });
    }
    __syncthreads();
    fft_lds<T, NT, MAXTOT, sizeof(T) == 4>(buf, d, nc, tw, +1);
    const double w0 = g.wmin + double(p) * g.dw;
    T* dst = proj + (long long)p * nx * ny;
    const int sh = nu - nx / 2;
    const unsigned mnx = unsigned(((1ull << 32) + nx - 1) / nx);
    for (int e = threadIdx.x; e < nx * nc; e += NT) {
        // column-major: i fastest, so the block writes whole columns
        int i, c;
        if (g.cm) {
            c = int(udiv(unsigned(e), unsigned(nx), mnx));
            i = e - c * nx;
        } else {
            i = int(udiv(unsigned(e), unsigned(nc), mnc));
            c = e - i * nc;
        }
        const int j = j0 + c;
        const int ui = sh + i;
        const cx<T> gg = buf[c * nu + (ui >= nu ? ui - nu : ui)];
        T term;
        if (g.wgrid) {
            const cx<T> ph = screen_at<T>(w0, screen[table_index(g, i, j)], true);
            term = gg.re * ph.re - gg.im * ph.im;
        } else {
            term = gg.re;
        }
        dst[g.cm ? (long long)j * nx + i : (long long)i * ny + j] = term;
    }
}

// v2d: the plane terms added in plane order, then the global correction
// (wgridder.rs:2655-2659, 2403-2428).
template <class T>
__global__ __launch_bounds__(256) void k_sum(const T* __restrict__ proj, T* __restrict__ out, Geo g,
                                             int P, const double* __restrict__ corr,
                                             const double* __restrict__ cfu,
                                             const double* __restrict__ cfv) {
    const long long npix = (long long)g.nx * g.ny;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npix) return;
    T acc = proj[e];
    for (int p = 1; p < P; ++p) acc = acc + proj[(long long)p * npix + e];
    const int i = int(e / g.ny), j = int(e - (long long)i * g.ny);
    if (g.wgrid) {
        out[e] = acc * T(corr[table_index(g, i, j)]);
    } else {
        const int iu = g.nx / 2 - min(i, g.nx - i), iv = g.ny / 2 - min(j, g.ny - j);
        out[e] = T(double(acc) * cfu[iu] * cfv[iv]);
    }
}

// k_sum for column-major projections: a 32 x 32 tile of pixels per
// workgroup, the planes summed along i and the tile written back along j.
template <class T>
__global__ __launch_bounds__(256) void k_sum_cm(const T* __restrict__ proj, T* __restrict__ out, Geo g,
                                                int P, const double* __restrict__ corr,
                                                const double* __restrict__ cfu,
                                                const double* __restrict__ cfv) {
    __shared__ T tile[32][33];
    const long long npix = (long long)g.nx * g.ny;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int i = i0 + tx;
    for (int jj = ty; jj < 32; jj += 8) {
        const int j = j0 + jj;
        if (i < g.nx && j < g.ny) {
            const long long e = (long long)j * g.nx + i;
            T acc = proj[e];
            for (int p = 1; p < P; ++p) acc = acc + proj[(long long)p * npix + e];
            T v;
            if (g.wgrid) {
                v = acc * T(corr[table_index(g, i, j)]);
            } else {
                const int iu = g.nx / 2 - min(i, g.nx - i), iv = g.ny / 2 - min(j, g.ny - j);
                v = T(double(acc) * cfu[iu] * cfv[iv]);
            }
            tile[jj][tx] = v;
        }
    }
    __syncthreads();
    for (int ii = ty; ii < 32; ii += 8) {
        const int i2 = i0 + ii, j2 = j0 + tx;
        if (i2 < g.nx && j2 < g.ny) out[(long long)i2 * g.ny + j2] = tile[tx][ii];
    }
}

// d2v: the corrected copy of the image (wgridder.rs:3064-3091).
template <class T>
__global__ __launch_bounds__(256) void k_correct(const T* __restrict__ dirty, T* __restrict__ img,
                                                 Geo g, const double* __restrict__ corr,
                                                 const double* __restrict__ cfu,
                                                 const double* __restrict__ cfv) {
    const long long npix = (long long)g.nx * g.ny;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npix) return;
    const int i = int(e / g.ny), j = int(e - (long long)i * g.ny);
    if (g.wgrid) {
        img[e] = dirty[e] * T(corr[table_index(g, i, j)]);
    } else {
        const int iu = g.nx / 2 - min(i, g.nx - i), iv = g.ny / 2 - min(j, g.ny - j);
        img[e] = dirty[e] * T(cfu[iu] * cfv[iv]);
    }
}

// k_correct writing the column-major copy: a 32 x 32 tile per workgroup,
// read along j and written along i.
template <class T>
__global__ __launch_bounds__(256) void k_correct_cm(const T* __restrict__ dirty, T* __restrict__ img,
                                                    Geo g, const double* __restrict__ corr,
                                                    const double* __restrict__ cfu,
                                                    const double* __restrict__ cfv) {
    __shared__ T tile[32][33];
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int ii = ty; ii < 32; ii += 8) {
        const int i = i0 + ii, j = j0 + tx;
        if (i < g.nx && j < g.ny) tile[ii][tx] = dirty[(long long)i * g.ny + j];
    }
    __syncthreads();
    for (int jj = ty; jj < 32; jj += 8) {
        const int i = i0 + tx, j = j0 + jj;
        if (i < g.nx && j < g.ny) {
            const T d = tile[tx][jj];
            T v;
            if (g.wgrid) {
                v = d * T(corr[table_index(g, i, j)]);
            } else {
                const int iu = g.nx / 2 - min(i, g.nx - i), iv = g.ny / 2 - min(j, g.ny - j);
                v = d * T(cfu[iu] * cfv[iv]);
            }
            img[(long long)j * g.nx + i] = v;
        }
    }
}

template <class T>
void transpose_planes(const cx<T>* in, cx<T>* out, int P, int R, int C, hipStream_t st) {
    constexpr int TS = sizeof(T) == 4 ? 64 : 32;
    if (P > 65535) throw DevError{"too many planes for the transpose grid"};
    const dim3 grid(unsigned((C + TS - 1) / TS), unsigned((R + TS - 1) / TS), unsigned(P));
    if constexpr (sizeof(T) == 4) {
        if (R % 2 == 0 && C % 2 == 0) {
            k_tile_transpose_pairs<float><<<grid, 256, 0, st>>>(in, out, R, C);
            HIPCHK(hipGetLastError());
            return;
        }
    }
    k_tile_transpose<T><<<grid, 256, 0, st>>>(in, out, R, C);
    HIPCHK(hipGetLastError());
}

// d2v: C image columns times the plane's w-screen placed on the grid rows,
// forward transform along u (wgridder.rs:2931-2967). TCM: the output
// column-major ([P][ny][nu], each column's transform written contiguously,
// transposed afterwards), else row-major ([P][nu][ny]).
template <class T, int NT, int MAXTOT, bool TCM = false>
__global__ __launch_bounds__(NT) void k_cols_fwd(const T* __restrict__ img,
                                                          cx<T>* __restrict__ tmp, FftDesc d,
                                                          const cx<T>* __restrict__ tw, Geo g,
                                                          const double* __restrict__ screen, int C) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nu = d.n, ny = g.ny, nx = g.nx;
    const int ncb = (ny + C - 1) / C;
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int p = lin / ncb;
    const int j0 = (lin - p * ncb) * C, nc = min(C, ny - j0);
    const unsigned mnc = nc > 1 ? unsigned(((1ull << 32) + nc - 1) / nc) : 0u;
    // the image values of the block in flight while the buffer is cleared
    constexpr int PER = (MAXTOT + NT - 1) / NT;
    const unsigned mnx = unsigned(((1ull << 32) + nx - 1) / nx);
    T r[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + k * NT;
        if (e < nx * nc) {
            if (g.cm) {
                // column-major image: the block reads whole columns
                const int c = int(udiv(unsigned(e), unsigned(nx), mnx)), i = e - c * nx;
                r[k] = img[(long long)(j0 + c) * nx + i];
            } else {
                const int i = int(udiv(unsigned(e), unsigned(nc), mnc)), c = e - i * nc;
                r[k] = img[(long long)i * ny + j0 + c];
            }
        }
    }
    for (int e = threadIdx.x; e < nu * nc; e += NT) buf[e] = cx<T>{T(0), T(0)};
    __syncthreads();
    const double w0 = g.wmin + double(p) * g.dw;
    const int sh = nu - nx / 2;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + k * NT;
        if (e < nx * nc) {
            int i, c;
            if (g.cm) {
                c = int(udiv(unsigned(e), unsigned(nx), mnx));
                i = e - c * nx;
            } else {
                i = int(udiv(unsigned(e), unsigned(nc), mnc));
                c = e - i * nc;
            }
            const int j = j0 + c;
            const T v = r[k];
            cx<T> val;
            if (g.wgrid) {
                const cx<T> ph = screen_at<T>(w0, screen[table_index(g, i, j)], false);
                val = {v * ph.re, v * ph.im};
            } else {
                val = {v, T(0)};
            }
            const int ui = sh + i;
            buf[c * nu + (ui >= nu ? ui - nu : ui)] = val;
        }
    }
    __syncthreads();
    fft_lds<T, NT, MAXTOT, sizeof(T) == 4>(buf, d, nc, tw, -1);
    cx<T>* dst = tmp + (long long)p * nu * ny;
    if constexpr (TCM) {
        for (int e = threadIdx.x; e < nu * nc; e += NT) dst[(long long)j0 * nu + e] = buf[e];
    } else {
        for (int e = threadIdx.x; e < nu * nc; e += NT) {
            const int u = int(udiv(unsigned(e), unsigned(nc), mnc)), c = e - u * nc;
            dst[(long long)u * ny + j0 + c] = buf[c * nu + u];
        }
    }
}

// d2v: one grid row from its image columns, forward transform along v
// (wgridder.rs:2922-2929, 2968-2991).
template <class T, int NT, int MAXTOT>
__global__ __launch_bounds__(NT) void k_rows_fwd(const cx<T>* __restrict__ tmp,
                                                          cx<T>* __restrict__ grid, FftDesc d,
                                                          const cx<T>* __restrict__ tw, Geo g,
                                                          const std::uint8_t* __restrict__ need) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nv = d.n, ny = g.ny;
    const long long row = blockIdx.x;
    const long long p = row / g.nu;
    // a row no degridding item reads is neither transformed nor written
    if (need && !need[row - p * g.nu]) return;
    const long long src = p * g.nu + upos(g, int(row - p * g.nu));
    // the image row in flight while the buffer is cleared
    constexpr int PER = (MAXTOT + NT - 1) / NT;
    cx<T> r[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int j = threadIdx.x + k * NT;
        if (j < ny) r[k] = tmp[src * ny + j];
    }
    for (int e = threadIdx.x; e < nv; e += NT) buf[e] = cx<T>{T(0), T(0)};
    __syncthreads();
    const int sh = nv - ny / 2;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int j = threadIdx.x + k * NT;
        if (j < ny) {
            const int e = sh + j;
            buf[e >= nv ? e - nv : e] = r[k];
        }
    }
    __syncthreads();
    fft_lds<T, NT, MAXTOT, sizeof(T) == 4>(buf, d, 1, tw, -1);
    cx<T>* dst = grid + row * nv;
    for (int e = threadIdx.x; e < nv; e += NT) dst[e] = buf[e];
}

// ------------------------------------------------------------ long transforms

// Transforms longer than one LDS, n = n1 n2, in two passes over memory:
// X[k1 + n1 k2] = sum_j2 w_n2^(j2 k2) w_n^(j2 k1) sum_j1 x[j1 n2 + j2] w_n1^(j1 k1)
// (w_m = exp(sgn 2 pi i / m)). Pass 1 transforms along j1 and applies the
// twiddle in place; pass 2 transforms along j2 and hands X[k1 + n1 k2] to
// its consumer: in place at k1 n2 + k2, or gathered to the pruned output.
// `twn` holds exp(-2 pi i k / n) for k < n (conjugated for sgn +1).

template <class T>
__device__ __forceinline__ cx<T> twn_at(const cx<T>* twn, int e, int sgn) {
    cx<T> f = twn[e];
    if (sgn > 0) f.im = -f.im;
    return f;
}

// Rows of length n = n1 n2 at `data + row * n`; block: (row, C2 lines j2).
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_lrow_p1(cx<T>* data, int n1, int n2, int C2, FftDesc d1,
                                                const cx<T>* tw1, const cx<T>* twn, int sgn) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nb = (n2 + C2 - 1) / C2;
    const long long row = blockIdx.x / nb;
    const int j20 = int(blockIdx.x - row * nb) * C2, nc = min(C2, n2 - j20);
    cx<T>* x = data + row * (long long)n1 * n2;
    staged_fill<T, NT, MT>(buf, x, n1 * nc, [&](int e) {
        const int j1 = e / nc, c = e - j1 * nc;
        return std::pair<long long, int>((long long)j1 * n2 + j20 + c, c * n1 + j1);
    });
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d1, nc, tw1, sgn);
    for (int e = threadIdx.x; e < n1 * nc; e += NT) {
        const int k1 = e / nc, c = e - k1 * nc;
        x[(long long)k1 * n2 + j20 + c] = cmul(buf[c * n1 + k1], twn_at(twn, (j20 + c) * k1, sgn));
    }
}

// Pass 2 over rows; block: (row, K1 lines k1). MODE 0: in place; MODE 1: the
// image columns j < ny of X[(n - ny/2 + j) mod n] into out (rows of ny).
template <class T, int NT, int MT, int MODE>
__global__ __launch_bounds__(NT) void k_lrow_p2(cx<T>* data, int n1, int n2, int K1, FftDesc d2,
                                                const cx<T>* tw2, int sgn, cx<T>* out, int ny) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nb = (n1 + K1 - 1) / K1;
    const long long row = blockIdx.x / nb;
    const int k10 = int(blockIdx.x - row * nb) * K1, nc = min(K1, n1 - k10);
    const int n = n1 * n2;
    cx<T>* x = data + row * (long long)n;
    staged_fill<T, NT, MT>(buf, x, n2 * nc,
                           [&](int e) { return std::pair<long long, int>((long long)k10 * n2 + e, e); });
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d2, nc, tw2, sgn);
    for (int e = threadIdx.x; e < n2 * nc; e += NT) {
        const int c = e / n2, k2 = e - c * n2;
        if constexpr (MODE == 0) {
            x[(long long)(k10 + c) * n2 + k2] = buf[e];
        } else {
            const int k = k10 + c + n1 * k2;
            const int j = (k + ny / 2) % n;
            if (j < ny) out[row * ny + j] = buf[e];
        }
    }
}

// Columns of [P][n][M] (lines (p, m), element j at (p n + j) M + m); block:
// (p, Cm columns, C2 lines j2).
template <class T, int NT, int MT>
__global__ __launch_bounds__(NT) void k_lcol_p1(cx<T>* data, int n1, int n2, int M, int Cm, int C2,
                                                FftDesc d1, const cx<T>* tw1, const cx<T>* twn, int sgn) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nmb = (M + Cm - 1) / Cm, njb = (n2 + C2 - 1) / C2;
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int p = lin / (nmb * njb), rem = lin - p * nmb * njb;
    const int mb = rem / njb, jb = rem - mb * njb;
    const int m0 = mb * Cm, ncm = min(Cm, M - m0), j20 = jb * C2, nc2 = min(C2, n2 - j20);
    const int lines = ncm * nc2;
    cx<T>* x = data + (long long)p * n1 * n2 * M;
    staged_fill<T, NT, MT>(buf, x, n1 * lines, [&](int e) {
        const int j1 = e / lines, r = e - j1 * lines, c2 = r / ncm, cm = r - c2 * ncm;
        return std::pair<long long, int>(((long long)j1 * n2 + j20 + c2) * M + m0 + cm, r * n1 + j1);
    });
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d1, lines, tw1, sgn);
    for (int e = threadIdx.x; e < n1 * lines; e += NT) {
        const int k1 = e / lines, r = e - k1 * lines, c2 = r / ncm, cm = r - c2 * ncm;
        x[((long long)k1 * n2 + j20 + c2) * M + m0 + cm] = cmul(buf[r * n1 + k1], twn_at(twn, (j20 + c2) * k1, sgn));
    }
}

// Pass 2 over columns; block: (p, Cm columns, K1 lines k1). MODE 0: in place;
// MODE 2: the image rows i < nx of X[(n - nx/2 + i) mod n] projected with the
// plane's w-screen into proj (as k_cols_inv).
template <class T, int NT, int MT, int MODE>
__global__ __launch_bounds__(NT) void k_lcol_p2(cx<T>* data, int n1, int n2, int M, int Cm, int K1,
                                                FftDesc d2, const cx<T>* tw2, int sgn, T* proj, Geo g,
                                                const double* screen) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nmb = (M + Cm - 1) / Cm, nkb = (n1 + K1 - 1) / K1;
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int p = lin / (nmb * nkb), rem = lin - p * nmb * nkb;
    const int mb = rem / nkb, kb = rem - mb * nkb;
    const int m0 = mb * Cm, ncm = min(Cm, M - m0), k10 = kb * K1, nc1 = min(K1, n1 - k10);
    const int lines = ncm * nc1, n = n1 * n2;
    cx<T>* x = data + (long long)p * n * M;
    staged_fill<T, NT, MT>(buf, x, n2 * lines, [&](int e) {
        const int j2 = e / lines, r = e - j2 * lines, c1 = r / ncm, cm = r - c1 * ncm;
        return std::pair<long long, int>(((long long)(k10 + c1) * n2 + j2) * M + m0 + cm, r * n2 + j2);
    });
    __syncthreads();
    fft_lds<T, NT, MT>(buf, d2, lines, tw2, sgn);
    const double w0 = g.wmin + double(p) * g.dw;
    for (int e = threadIdx.x; e < n2 * lines; e += NT) {
        const int k2 = e / lines, r = e - k2 * lines, c1 = r / ncm, cm = r - c1 * ncm;
        const cx<T> v = buf[r * n2 + k2];
        if constexpr (MODE == 0) {
            x[((long long)(k10 + c1) * n2 + k2) * M + m0 + cm] = v;
        } else {
            const int k = k10 + c1 + n1 * k2;
            const int i = (k + g.nx / 2) % n;
            if (i < g.nx) {
                const int j = m0 + cm;
                T term;
                if (g.wgrid) {
                    const cx<T> ph = screen_at<T>(w0, screen[table_index(g, i, j)], true);
                    term = v.re * ph.re - v.im * ph.im;
                } else {
                    term = v.re;
                }
                proj[(long long)p * g.nx * g.ny + (long long)i * g.ny + j] = term;
            }
        }
    }
}

// d2v with long u: the image rows times the plane's w-screen placed on the
// grid rows of tmp ([P][nu][ny], zeros elsewhere), as k_cols_fwd places them.
// One workgroup per grid row (p, u): a flat launch over P nu ny values
// would pass the 2^32 work-items one dispatch dimension holds.
template <class T>
__global__ __launch_bounds__(256) void k_place_cols(const T* __restrict__ img, cx<T>* tmp, Geo g,
                                                    const double* __restrict__ screen) {
    const long long row = blockIdx.x;
    const int p = int(row / g.nu), u = int(row - (long long)p * g.nu);
    const int i = (u + g.nx / 2) % g.nu;
    cx<T>* dst = tmp + row * g.ny;
    if (i >= g.nx) {
        for (int j = threadIdx.x; j < g.ny; j += 256) dst[j] = cx<T>{T(0), T(0)};
        return;
    }
    const double w0 = g.wmin + double(p) * g.dw;
    const T* src = img + (long long)i * g.ny;
    for (int j = threadIdx.x; j < g.ny; j += 256) {
        const T v = src[j];
        cx<T> val;
        if (g.wgrid) {
            const cx<T> ph = screen_at<T>(w0, screen[table_index(g, i, j)], false);
            val = {v * ph.re, v * ph.im};
        } else {
            val = {v, T(0)};
        }
        dst[j] = val;
    }
}

// d2v with long v: grid rows from the image columns of tmp (its rows at their
// physical place), zeros elsewhere, as k_rows_fwd places them. One workgroup
// per grid row.
template <class T>
__global__ __launch_bounds__(256) void k_place_rows(const cx<T>* __restrict__ tmp, cx<T>* grid, Geo g) {
    const long long row = blockIdx.x;
    const long long p = row / g.nu;
    const cx<T>* src = tmp + (p * g.nu + upos(g, int(row - p * g.nu))) * g.ny;
    cx<T>* dst = grid + row * g.nv;
    for (int v = threadIdx.x; v < g.nv; v += 256) {
        const int j = (v + g.ny / 2) % g.nv;
        dst[v] = j < g.ny ? src[j] : cx<T>{T(0), T(0)};
    }
}

// Work items of the degridding: per tile rank (`order`, densest first) its
// visibilities in chunks of at most `rc` (none for an empty tile).
__global__ __launch_bounds__(256) void k_degrid_items(Geo g, const std::uint32_t* __restrict__ off,
                                                      const std::uint32_t* __restrict__ order,
                                                      std::uint32_t rc, std::uint32_t* __restrict__ nitems) {
    const int nt = g.ntu * g.ntv;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nt) return;
    if (r == nt) {
        nitems[nt] = 0;
        return;
    }
    const int tile = int(order[r]);
    const std::uint32_t cnt = off[(long long)(tile + 1) * 4 * g.nb] - off[(long long)tile * 4 * g.nb];
    nitems[r] = (cnt + rc - 1) / rc;
}

// need[u] = 1 for the grid rows a degridding item reads: the rows of its
// tile and halo (taken S rows wider on both sides), for every tile with
// items; rows no visibility reaches are left 0 and the forward row
// transform skips them.
__global__ __launch_bounds__(256) void k_row_need(Geo g, const std::uint32_t* __restrict__ item_off,
                                                  const std::uint32_t* __restrict__ order,
                                                  std::uint8_t* __restrict__ need) {
    const int nt = g.ntu * g.ntv;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt || item_off[r + 1] == item_off[r]) return;
    const int tu = int(order[r]) / g.ntv;
    for (int us = tu * g.Tu - g.supp; us < (tu + 1) * g.Tu + g.supp; ++us) {
        int u = us + g.nu / 2;
        u %= g.nu;
        if (u < 0) u += g.nu;
        need[u] = 1;
    }
}

// Items of rc records per tile in tile order (an identity tile order):
// items[t] = ceil(count_t / rc), the last entry 0 for the exclusive scan.
__global__ __launch_bounds__(256) void k_chunk_counts(Geo g, const std::uint32_t* __restrict__ off, std::uint32_t rc,
                                                      std::uint32_t* __restrict__ items) {
    const int nt = g.ntu * g.ntv;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > nt) return;
    if (t == nt) {
        items[nt] = 0;
        return;
    }
    const std::uint32_t cnt = off[(long long)(t + 1) * 4 * g.nb] - off[(long long)t * 4 * g.nb];
    items[t] = (cnt + rc - 1) / rc;
}

// The degridding records of each item (chunk c of rc records of a tile,
// [t0 + c rc, min(t1, t0 + (c + 1) rc))) regrouped by their cell in the tile
// (row, then column). The degridder's lanes take neighbouring records, so
// its staged reads then hit equal or neighbouring cells instead of cells
// scattered over the tile (LDS bank conflicts were 57 % of its LDS cycles;
// 61.5 -> 41.4 ms on the 464M workload). Records stay within their chunk: the
// item's plane range (the bucket bins overlapping the chunk) is unchanged.
// One 1024-thread workgroup per item, rc <= kSortChunk records held in
// registers: a counting sort over the Tu Tv cells in LDS (slots within a
// cell in atomic order).
constexpr int kSortChunk = 4096;
template <class T>
__global__ __launch_bounds__(1024) void k_sort_chunks(Geo g, Rec<T>* __restrict__ recs,
                                                      const std::uint32_t* __restrict__ off,
                                                      const std::uint32_t* __restrict__ item_off, std::uint32_t rc) {
    __shared__ unsigned cnt[kSortChunk];
    __shared__ unsigned wsum[16];
    constexpr int PER = kSortChunk / 1024;
    int tile;
    std::uint32_t chunk, nchunks;
    if (!find_item(item_off, g.ntu * g.ntv, blockIdx.x, tile, chunk, nchunks)) return;
    const std::uint32_t t0 = off[(long long)tile * 4 * g.nb], t1 = off[(long long)(tile + 1) * 4 * g.nb];
    const std::uint32_t a = t0 + chunk * rc, b = min(t1, a + rc);
    const int n = int(b - a);
    const int ncell = g.Tu * g.Tv;  // <= kSortChunk, checked by the caller
    for (int c = threadIdx.x; c < ncell; c += 1024) cnt[c] = 0u;
    __syncthreads();
    Rec<T> held[PER];
    int cell[PER];
    unsigned slot[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int sl = int(threadIdx.x) + k * 1024;
        cell[k] = -1;
        if (sl < n) {
            held[k] = recs[a + sl];
            // pos: row | column << 8 | first plane << 16, row < Tu and column < Tv
            // (clamped: the count stays inside the table whatever the record)
            const std::uint32_t pos = held[k].pos;
            cell[k] = min(int(pos & 255u), g.Tu - 1) * g.Tv + min(int((pos >> 8) & 255u), g.Tv - 1);
            slot[k] = atomicAdd(&cnt[cell[k]], 1u);
        }
    }
    __syncthreads();
    // exclusive scan of the cell counts: four per thread, then over the
    // threads (waves, then lanes)
    constexpr int CPT = kSortChunk / 1024;
    unsigned v[CPT], run = 0u;
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = int(threadIdx.x) * CPT + q;
        v[q] = c < ncell ? cnt[c] : 0u;
        run += v[q];
    }
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    unsigned incl = run;
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    unsigned base = incl - run;
    for (int w = 0; w < wave; ++w) base += wsum[w];
    __syncthreads();  // every count read before the bases overwrite them
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = int(threadIdx.x) * CPT + q;
        if (c < ncell) cnt[c] = base;
        base += v[q];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (cell[k] >= 0) recs[a + cnt[cell[k]] + slot[k]] = held[k];
}

// d2v: degridding (wgridder.rs:2793-2901) of one item: per visibility of the
// chunk, the S x S taps of each of its planes against the plane grid, the
// planes added in order, then the weight and shift phase
// (wgridder.rs:3146-3159). With STAGE the tile and its halo of every plane
// are first copied to LDS (P (Tu+S-1) (Tv+S-1) values); otherwise the taps
// read the planes in global memory. Each thread takes a contiguous part of
// the chunk.
// Threads per degridding workgroup: staged single precision takes a whole
// CU (its taps fit 128 registers); the rest run 256-thread workgroups.
template <class T, bool STAGE>
constexpr int degrid_threads() {
    return STAGE && sizeof(T) == 4 ? 1024 : 256;
}

// Where the degridder puts each record's value c (the prediction times the
// record's value): mode 0, vis[index] = c (dense); mode 1, stored[pos[r]] = c
// (the values in another sorted set's order); mode 2, the residual
// stored[pos[r]] = d - c of the data d (dense at index, or `sorted` in that
// order), with sum w |d - c|^2 (the data's weights, dense) per workgroup into
// part[workgroup].
template <class T>
struct DegridOut {
    cx<T>* vis = nullptr;
    int mode = 0;
    const std::uint32_t* pos = nullptr;
    cx<T>* stored = nullptr;
    DevIn data{};
    int sorted = 0;
    double* part = nullptr;
};

template <class T, int S, bool STAGE>
__global__ __launch_bounds__((degrid_threads<T, STAGE>())) void k_degrid(Geo g, const Rec<T>* __restrict__ recs,
                                                const std::uint32_t* __restrict__ off,
                                                const T* __restrict__ coeff,
                                                const cx<T>* __restrict__ grid, std::uint32_t rc,
                                                const std::uint32_t* __restrict__ item_off,
                                                const std::uint32_t* __restrict__ order,
                                                DegridOut<T> o) {
    constexpr int NE = ne_of<S>();
    __shared__ T coef[2 * NE * S];
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    int rank;
    std::uint32_t chunk, nchunks;
    if (!find_item(item_off, g.ntu * g.ntv, blockIdx.x, rank, chunk, nchunks)) return;
    const int tile = int(order[rank]);
    const std::uint32_t t0 = off[(long long)tile * 4 * g.nb], t1 = off[(long long)(tile + 1) * 4 * g.nb];
    const std::uint32_t a = t0 + chunk * rc, b = min(t1, a + rc);
    for (int e = threadIdx.x; e < 2 * NE * S; e += blockDim.x) coef[e] = coeff[e];
    const int tu = tile / g.ntv, tv = tile - (tile / g.ntv) * g.ntv;
    const int P = g.wgrid ? g.nplanes : 1;
    const long long plane = (long long)g.nu * g.nv;
    const int SU = g.Tu + S - 1, SV = g.Tv + S - 1;
    // With STAGE, the planes [plo, phi) the chunk's visibilities reach: their
    // first planes are the bucket bins (tile, spill class, first plane) whose
    // ranges of records overlap the chunk's [a, b).
    int plo = 0, phi = P;
    if constexpr (STAGE) {
        if (g.wgrid) {
            __shared__ int s_lo, s_hi;
            if (threadIdx.x == 0) {
                s_lo = P;
                s_hi = -1;
            }
            __syncthreads();
            const long long k0 = (long long)tile * 4 * g.nb;
            for (int k = threadIdx.x; k < 4 * g.nb; k += blockDim.x) {
                const std::uint32_t ks = off[k0 + k], ke = off[k0 + k + 1];
                if (ke > ks && ks < b && ke > a) {
                    const int m = k - (k / g.nb) * g.nb;
                    atomicMin(&s_lo, m);
                    atomicMax(&s_hi, m);
                }
            }
            __syncthreads();
            plo = s_lo;
            phi = min(P, s_hi + S);
        }
        // a fixed column per thread, rows (plane, i) in steps: no division per cell
        const int rstep = int(blockDim.x) / SV;
        const int r0 = int(threadIdx.x) / SV, j = int(threadIdx.x) - r0 * SV;
        if (r0 < rstep) {
            int gcol = tv * g.Tv + j + g.nv / 2;
            if (gcol >= g.nv) gcol -= g.nv;
            const int col = vpos(g, gcol);
            const int nrows = (phi - plo) * SU;
            int pl = r0 / SU, i = r0 - pl * SU;
            for (int rr = r0; rr < nrows; rr += rstep) {
                int grow = tu * g.Tu + i + g.nu / 2;
                if (grow >= g.nu) grow -= g.nu;
                buf[rr * SV + j] = grid[(long long)(plo + pl) * plane + (long long)grow * g.nv + col];
                i += rstep;
                while (i >= SU) {
                    i -= SU;
                    ++pl;
                }
            }
        }
    }
    __syncthreads();
    // neighbouring lanes take neighbouring records (a row's channels, whose
    // footprints overlap): their staged reads spread over the banks and their
    // record loads coalesce
    double mis = 0.0;
    for (std::uint32_t r = a + threadIdx.x; r < b; r += blockDim.x) {
        const Rec<T> rc_ = recs[r];
        const int ou = int(rc_.pos & 255u), ov = int((rc_.pos >> 8) & 255u), mp = int(rc_.pos >> 16);
        T ku[S], kv[S];
        eval_taps2<T, S>(coef, rc_.x0, rc_.y0, ku, kv);
        int rows[S], cols[S];
        if constexpr (!STAGE) {
#pragma unroll
            for (int k = 0; k < S; ++k) {
                rows[k] = (tu * g.Tu + ou + k + g.nu / 2) % g.nu;
                cols[k] = vpos(g, (tv * g.Tv + ov + k + g.nv / 2) % g.nv);
            }
        }
        const T fl = (rc_.idx & kFlip) ? T(-1) : T(1);
        cx<T> out = {T(0), T(0)};
        const int pa = g.wgrid ? mp : 0, pb = g.wgrid ? min(P, mp + S) : 1;
        for (int p = pa; p < pb; ++p) {
            T kw = T(1);
            if (g.wgrid) kw = w_factor<T, S>(g, coef, p - mp, p, rc_.w);
            T rr = T(0), ri = T(0);
#pragma unroll
            for (int cu = 0; cu < S; ++cu) {
                const T kuc = g.wgrid ? ku[cu] * kw : ku[cu];
                T tr = T(0), ti = T(0);
                if constexpr (STAGE) {
                    const cx<T>* line = buf + ((long long)(p - plo) * SU + ou + cu) * SV + ov;
#pragma unroll
                    for (int cv = 0; cv < S; ++cv) {
                        const cx<T> v = line[cv];
                        tr = fma(kv[cv], v.re, tr);
                        ti = fma(kv[cv], v.im, ti);
                    }
                } else {
                    const cx<T>* line = grid + (long long)p * plane + (long long)rows[cu] * g.nv;
#pragma unroll
                    for (int cv = 0; cv < S; ++cv) {
                        const cx<T> v = line[cols[cv]];
                        tr = fma(kv[cv], v.re, tr);
                        ti = fma(kv[cv], v.im, ti);
                    }
                }
                rr = fma(kuc, tr, rr);
                ri = fma(kuc, ti, ri);
            }
            out = {rr + out.re, ri * fl + out.im};
        }
        cx<T> res;
        if (g.shifting)
            res = cmul(out, cx<T>{rc_.re, rc_.im});
        else
            res = {out.re * rc_.re, out.im * rc_.re};
        const std::uint32_t idx = rc_.idx & ~kFlip;
        if (o.mode == 0) {
            o.vis[idx] = res;
        } else {
            const std::uint32_t k0 = o.pos[r];
            if (o.mode == 1) {
                o.stored[k0] = res;
            } else {
                const cx<T> d = vis_get<T>(o.data, o.sorted ? (long long)k0 : (long long)idx);
                const cx<T> q = {d.re - res.re, d.im - res.im};
                o.stored[k0] = q;
                T wv = T(1);
                wgt_get<T>(o.data, idx, wv);
                mis += double(wv) * (double(q.re) * double(q.re) + double(q.im) * double(q.im));
            }
        }
    }
    if (o.mode == 2) {
        // the workgroup's misfit: lanes, then waves
        for (int h = 32; h > 0; h >>= 1) mis += __shfl_down(mis, h, 64);
        __shared__ double wsum[16];
        const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
        if (lane == 0) wsum[wave] = mis;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int k = 0; k < int(blockDim.x) / 64; ++k) t += wsum[k];
            o.part[blockIdx.x] = t;
        }
    }
}

// ------------------------------------------------------------ dispatch

template <class T, int NOUT, bool FORM>
void launch_grid(int S, int nblocks, std::size_t lds, hipStream_t st, const Geo& g,
                 const Rec<T>* recs, const std::uint32_t* off, const T* coeff, const DevIn& in,
                 int G, int ngroups, int TvP, std::uint32_t rc, const std::uint32_t* item_off,
                 const std::uint32_t* order, cx<T>* out) {
#define GRID_CASE(SS)                                                                           \
    case SS:                                                                                    \
        if constexpr (SS <= max_support<T>()) {                                                 \
            lds_limit((const void*)k_grid<T, SS, NOUT, FORM>, lds);                             \
            k_grid<T, SS, NOUT, FORM><<<nblocks, grid_threads<T>(), lds, st>>>(g, recs, off, coeff, in, G,  \
                                                                     ngroups, TvP, rc, item_off, \
                                                                     order, out);               \
        } else {                                                                                \
            throw DevError{"kernel support too large for this precision"};                      \
        }                                                                                       \
        break;
    switch (S) {
        GRID_CASE(4) GRID_CASE(5) GRID_CASE(6) GRID_CASE(7) GRID_CASE(8) GRID_CASE(9)
        GRID_CASE(10) GRID_CASE(11) GRID_CASE(12) GRID_CASE(13) GRID_CASE(14) GRID_CASE(15)
        GRID_CASE(16)
    default:
        throw DevError{"unsupported kernel support"};
    }
#undef GRID_CASE
    HIPCHK(hipGetLastError());
}

void launch_grid_mfma(int S, int nblocks, bool form, hipStream_t st, const Geo& g, const Rec<float>* recs,
                      const std::uint32_t* off, const float* coeff, const DevIn& in, int G, int ngroups,
                      std::uint32_t rc, const std::uint32_t* item_off, const std::uint32_t* order,
                      cx<float>* out) {
#define MFMA_CASE(SS)                                                                                   \
    case SS:                                                                                            \
        if (form)                                                                                       \
            k_grid_mfma<SS, true><<<nblocks, 64, 0, st>>>(g, recs, off, coeff, in, G, ngroups, rc, item_off, \
                                                          order, out);                                  \
        else                                                                                            \
            k_grid_mfma<SS, false><<<nblocks, 64, 0, st>>>(g, recs, off, coeff, in, G, ngroups, rc, item_off, \
                                                           order, out);                                 \
        break;
    switch (S) {
        MFMA_CASE(4) MFMA_CASE(5) MFMA_CASE(6) MFMA_CASE(7) MFMA_CASE(8)
    default:
        throw DevError{"unsupported kernel support"};
    }
#undef MFMA_CASE
    HIPCHK(hipGetLastError());
}

template <class T, bool STAGE>
void launch_degrid(int S, int nblocks, std::size_t lds, hipStream_t st, const Geo& g,
                   const Rec<T>* recs, const std::uint32_t* off, const T* coeff,
                   const cx<T>* grid, std::uint32_t rc, const std::uint32_t* item_off,
                   const std::uint32_t* order, const DegridOut<T>& vis) {
#define DEGRID_CASE(SS)                                                                        \
    case SS:                                                                                   \
        if constexpr (SS <= max_support<T>()) {                                                \
            if (STAGE) lds_limit((const void*)k_degrid<T, SS, STAGE>, lds);                   \
            k_degrid<T, SS, STAGE><<<nblocks, degrid_threads<T, STAGE>(), lds, st>>>(g, recs, off, coeff, grid, rc, \
                                                              item_off, order, vis);           \
        } else {                                                                               \
            throw DevError{"kernel support too large for this precision"};                     \
        }                                                                                      \
        break;
    switch (S) {
        DEGRID_CASE(4) DEGRID_CASE(5) DEGRID_CASE(6) DEGRID_CASE(7) DEGRID_CASE(8) DEGRID_CASE(9)
        DEGRID_CASE(10) DEGRID_CASE(11) DEGRID_CASE(12) DEGRID_CASE(13) DEGRID_CASE(14)
        DEGRID_CASE(15) DEGRID_CASE(16)
    default:
        throw DevError{"unsupported kernel support"};
    }
#undef DEGRID_CASE
    HIPCHK(hipGetLastError());
}

// Tile side along a grid axis of n cells: a divisor of n (so that no tile
// is partial) of at least the support and at most `hi`, nearest `target`.
// Tile side: the divisor of n in [max(supp, 8), hi] nearest `target`, else
// the smallest one up to 64 (the kernels keep in-tile offsets below 64).
int choose_tile(int n, int supp, int hi, int target) {
    int best = -1;
    for (int t = std::max(supp, 8); t <= std::min(hi, n); ++t)
        if (n % t == 0 && (best < 0 || std::abs(t - target) < std::abs(best - target))) best = t;
    for (int t = std::max(supp, 8); best < 0 && t <= std::min(64, n); ++t)
        if (n % t == 0) best = t;
    return best;
}

void exclusive_scan(Workspace& ws, const std::uint32_t* in, std::uint32_t* out, long long n,
                    hipStream_t st) {
    std::vector<long long> sizes{n};
    while (sizes.back() > kScanItems) sizes.push_back((sizes.back() + kScanItems - 1) / kScanItems);
    long long total = 0;
    for (std::size_t i = 1; i < sizes.size(); ++i) total += 2 * sizes[i];
    std::uint32_t* tmp = static_cast<std::uint32_t*>(ws.get(kSlotScan, sizeof(std::uint32_t) * (total + 2)));
    std::vector<std::uint32_t*> sums, sums_scanned;
    std::uint32_t* p = tmp;
    for (std::size_t i = 1; i < sizes.size(); ++i) {
        sums.push_back(p);
        p += sizes[i];
        sums_scanned.push_back(p);
        p += sizes[i];
    }
    std::vector<const std::uint32_t*> ins{in};
    std::vector<std::uint32_t*> outs{out};
    for (std::size_t i = 1; i < sizes.size(); ++i) {
        ins.push_back(sums[i - 1]);
        outs.push_back(sums_scanned[i - 1]);
    }
    std::uint32_t* dummy = tmp + total;  // the top level's total
    for (std::size_t lv = 0; lv < sizes.size(); ++lv) {
        const long long m = sizes[lv];
        const long long nblk = (m + kScanItems - 1) / kScanItems;
        std::uint32_t* sm = lv + 1 < sizes.size() ? sums[lv] : dummy;
        k_scan_block<<<nblk, 1024, 0, st>>>(ins[lv], outs[lv], sm, m);
        HIPCHK(hipGetLastError());
    }
    for (std::size_t lv = sizes.size() - 1; lv-- > 0;) {
        const long long m = sizes[lv];
        k_scan_add<<<(m + 255) / 256, 256, 0, st>>>(outs[lv], outs[lv + 1], m);
        HIPCHK(hipGetLastError());
    }
}

template <class T>
struct Tables {
    FftDesc du, dv;
    cx<T>* twu = nullptr;
    cx<T>* twv = nullptr;
    T* coeff = nullptr;
    double* screen = nullptr;
    double* corr = nullptr;
    double* cfu = nullptr;
    double* cfv = nullptr;
    std::uint32_t* order = nullptr;
};

// Host-side values kept alive until the call's stream is synchronised.
struct HostKeep {
    std::vector<std::vector<double>> d;
    std::vector<std::vector<float>> f;
    std::vector<std::vector<std::uint32_t>> u;
    std::vector<std::vector<char>> b;
};

template <class P>
P* upload_vec(Workspace& ws, int slot, std::vector<P> h, hipStream_t st, std::vector<std::vector<P>>& keep) {
    keep.push_back(std::move(h));
    auto& v = keep.back();
    P* d = static_cast<P*>(ws.get(slot, sizeof(P) * std::max<std::size_t>(v.size(), 1)));
    if (!v.empty()) HIPCHK(hipMemcpyAsync(d, v.data(), sizeof(P) * v.size(), hipMemcpyHostToDevice, st));
    return d;
}

template <class T>
T* upload_real(Workspace& ws, int slot, const std::vector<double>& h, hipStream_t st, HostKeep& keep) {
    if constexpr (sizeof(T) == 8)
        return reinterpret_cast<T*>(upload_vec<double>(ws, slot, h, st, keep.d));
    else
        return reinterpret_cast<T*>(upload_vec<float>(ws, slot, std::vector<float>(h.begin(), h.end()), st, keep.f));
}

// The signature of everything the tables depend on.
template <class T>
std::vector<double> table_key(const Geometry& G, const Geo& g) {
    return {double(sizeof(T)), double(G.nu), double(G.nv), double(G.nx), double(G.ny),
            double(G.kernel.index), G.pixsize_x, G.pixsize_y, G.lshift, G.mshift, G.r_l, G.r_m,
            G.nshift, G.dw, double(G.divide_by_n), double(g.wgrid), double(g.lmshift)};
}

// Tiles ordered by distance of their centre from the grid centre (u = v = 0,
// where the visibility density peaks), densest first.
std::vector<std::uint32_t> tile_order(const Geo& g) {
    std::vector<std::pair<double, std::uint32_t>> d;
    d.reserve(std::size_t(g.ntu) * g.ntv);
    for (int tu = 0; tu < g.ntu; ++tu)
        for (int tv = 0; tv < g.ntv; ++tv) {
            const double cu = (tu + 0.5) * g.Tu - 0.5 * g.nu, cv = (tv + 0.5) * g.Tv - 0.5 * g.nv;
            d.emplace_back(cu * cu + cv * cv, std::uint32_t(tu * g.ntv + tv));
        }
    std::stable_sort(d.begin(), d.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<std::uint32_t> out(d.size());
    for (std::size_t i = 0; i < d.size(); ++i) out[i] = d[i].second;
    return out;
}

template <class T>
void make_tables(Workspace& ws, const Geometry& G, const Geo& g, hipStream_t st, HostKeep& keep,
                 Tables<T>& tb, int oset) {
    const std::vector<double> key = table_key<T>(G, g);
    const bool cached = ws.table_key == key;
    tb.twu = static_cast<cx<T>*>(ws.ptr[kSlotTwU]);
    tb.twv = static_cast<cx<T>*>(ws.ptr[kSlotTwV]);
    tb.coeff = static_cast<T*>(ws.ptr[kSlotCoeff]);
    tb.cfu = static_cast<double*>(ws.ptr[kSlotCfu]);
    tb.cfv = static_cast<double*>(ws.ptr[kSlotCfv]);
    tb.screen = static_cast<double*>(ws.ptr[kSlotScreen]);
    tb.corr = static_cast<double*>(ws.ptr[kSlotCorr]);
    const std::vector<int> okey{g.Tu, g.Tv, g.ntu, g.ntv};
    const int oslot = oset ? kSlotOrder2 : kSlotOrder;
    if (ws.order_key[oset] != okey) {
        upload_vec<std::uint32_t>(ws, oslot, tile_order(g), st, keep.u);
        ws.order_key[oset] = okey;
    }
    tb.order = static_cast<std::uint32_t*>(ws.ptr[oslot]);
    if (cached) {
        tb.du = ws.du;
        tb.dv = ws.dv;
        return;
    }
    ws.table_key.clear();
    std::vector<double> twu, twv;
    // single-precision plane transforms take the radix-21 pass (their kernels
    // instantiate it; double precision keeps its registers)
    if (!fft_describe(int(G.nu), tb.du, twu, sizeof(T) == 4) || !fft_describe(int(G.nv), tb.dv, twv, sizeof(T) == 4))
        throw DevError{"grid size is not a product of 2, 3, 5, 7 and 11"};
    tb.twu = reinterpret_cast<cx<T>*>(upload_real<T>(ws, kSlotTwU, twu, st, keep));
    tb.twv = reinterpret_cast<cx<T>*>(upload_real<T>(ws, kSlotTwV, twv, st, keep));
    int ne = 0;
    tb.coeff = upload_real<T>(ws, kSlotCoeff, G.kernel.even_odd(ne), st, keep);
    // correction functions along u and v (wgridder.rs:824-829)
    std::vector<double> cfu(G.nx / 2 + 1), cfv(G.ny / 2 + 1);
    for (std::size_t i = 0; i < cfu.size(); ++i) cfu[i] = G.kernel.corfunc(double(i) * (1.0 / double(G.nu)));
    for (std::size_t i = 0; i < cfv.size(); ++i) cfv[i] = G.kernel.corfunc(double(i) * (1.0 / double(G.nv)));
    tb.cfu = upload_vec<double>(ws, kSlotCfu, cfu, st, keep.d);
    tb.cfv = upload_vec<double>(ws, kSlotCfv, cfv, st, keep.d);
    if (g.wgrid) {
        std::vector<double> gl(G.kernel.corr_x);
        gl.insert(gl.end(), G.kernel.corr_w.begin(), G.kernel.corr_w.end());
        double* dgl = upload_vec<double>(ws, kSlotGL, gl, st, keep.d);
        const long long tn = (long long)g.tnx * g.tny;
        tb.screen = static_cast<double*>(ws.get(kSlotScreen, sizeof(double) * tn));
        tb.corr = static_cast<double*>(ws.get(kSlotCorr, sizeof(double) * tn));
        const int ngl = int(G.kernel.corr_x.size());
        k_tables<<<(tn + 255) / 256, 256, 0, st>>>(g, tb.screen, tb.corr, tb.cfu, tb.cfv, dgl, dgl + ngl, ngl);
        HIPCHK(hipGetLastError());
    }
    ws.du = tb.du;
    ws.dv = tb.dv;
    ws.table_key = key;
}

// Compute units of the current device.
int device_cus() {
    int dev = 0, n = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    return std::max(n, 1);
}

// One count from the device (the stream is synchronised).
long long read_u32(Workspace& ws, const std::uint32_t* d, hipStream_t st) {
    std::uint32_t* h = static_cast<std::uint32_t*>(ws.host(sizeof(std::uint32_t)));
    HIPCHK(hipMemcpyAsync(h, d, sizeof(std::uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return (long long)*h;
}

template <class T, bool USE_VIS>
unsigned long long run_scan(Workspace& ws, const DevIn& din, const Geo& g, hipStream_t st,
                            double& wlo, double& whi) {
    if (g.n == 0) return 0;
    const int nblk = int(std::min<long long>((g.n + 255) / 256, kScanBlocks));
    unsigned long long* cnt = static_cast<unsigned long long*>(ws.get(kSlotCounters, 3 * sizeof(unsigned long long)));
    char* part = static_cast<char*>(ws.get(kSlotScanPart, 24 * std::size_t(kScanBlocks)));
    unsigned long long* bcnt = reinterpret_cast<unsigned long long*>(part);
    double* blo = reinterpret_cast<double*>(part + 8 * kScanBlocks);
    double* bhi = reinterpret_cast<double*>(part + 16 * kScanBlocks);
    k_scan<T, USE_VIS><<<nblk, 256, 0, st>>>(din, g, bcnt, blo, bhi);
    HIPCHK(hipGetLastError());
    k_scan_final<<<1, 256, 0, st>>>(bcnt, blo, bhi, nblk, cnt);
    HIPCHK(hipGetLastError());
    unsigned long long* res = static_cast<unsigned long long*>(ws.host(3 * sizeof(unsigned long long)));
    HIPCHK(hipMemcpyAsync(res, cnt, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::memcpy(&wlo, &res[1], sizeof(double));
    std::memcpy(&whi, &res[2], sizeof(double));
    return res[0];
}

// LDS bytes the staged degridding needs for tiles of tu x tv cells.
template <class T>
std::size_t degrid_lds(int tu, int tv, int supp, int P) {
    return std::size_t(P) * (tu + supp - 1) * (tv + supp - 1) * sizeof(cx<T>);
}

// Gridding accumulators per workgroup: one workgroup per CU, the 160 KiB of
// a gfx950 CU's LDS less the staged records and the largest coefficient
// table.
template <class T>
constexpr std::size_t grid_lds() {
    return 160 * 1024 - kGridBatch * sizeof(Rec<T>) - 2 * 10 * 16 * sizeof(T) - 256;
}
// Visibilities per gridding item, measured on the 464M-visibility workload
// once a split tile's chunks add only the cells they reach (gridding stage:
// 512 -> 280.5 ms, 1024 -> 251.6, 2048 -> 238.0, 4096 -> 233.4, 8192 -> 234.5).
constexpr std::uint32_t kGridChunk = 4096;
constexpr std::uint32_t kDegridChunk = 1024;     // visibilities per degridding item
// Staged degridding holds a tile and its halo on every plane in LDS: in
// single precision one 1024-thread workgroup per CU with up to 150 KiB and
// 4096 visibilities per item, in double precision 256-thread workgroups with
// up to 64 KiB. Measured on the 464M-visibility workload (26 planes, 16 x 16
// tiles): 1024 visibilities per item 86.8 ms, 2048 76.3, 4096 73.9, 8192
// 73.9, 16384 75.6; unstaged (64 KiB limit) 174.7.
template <class T>
constexpr std::uint32_t degrid_chunk_staged() {
    return sizeof(T) == 4 ? 4096 : kDegridChunk;
}
template <class T>
constexpr std::size_t degrid_lds_max() {
    return sizeof(T) == 4 ? 150 * 1024 : 64 * 1024;
}

// Items per degridding launch and their records reordered within each item
// (k_sort_chunks), before the degridder first reads them: the chunk size the
// degrid pipeline will use for this geometry. RANG_DEGRID_SORT=0 leaves the
// bucket order (a measurement knob).
template <class T>
void sort_degrid_chunks(Workspace& ws, const Geo& g, Rec<T>* recs, const std::uint32_t* off, hipStream_t st) {
    static const bool on = [] {
        const char* e = std::getenv("RANG_DEGRID_SORT");
        return !(e && e[0] == '0');
    }();
    if (!on) return;
    const int P = g.wgrid ? g.nplanes : 1;
    const bool staged = degrid_lds<T>(g.Tu, g.Tv, g.supp, P) <= degrid_lds_max<T>();
    const std::uint32_t rc = staged ? degrid_chunk_staged<T>() : kDegridChunk;
    if (rc > std::uint32_t(kSortChunk) || g.Tu * g.Tv > kSortChunk) return;
    const int nt = g.ntu * g.ntv;
    std::uint32_t* items = static_cast<std::uint32_t*>(ws.get(kSlotItems, sizeof(std::uint32_t) * (nt + 1)));
    std::uint32_t* item_off = static_cast<std::uint32_t*>(ws.get(kSlotItemOff, sizeof(std::uint32_t) * (nt + 1)));
    k_chunk_counts<<<(nt + 1 + 255) / 256, 256, 0, st>>>(g, off, rc, items);
    HIPCHK(hipGetLastError());
    exclusive_scan(ws, items, item_off, nt + 1, st);
    const long long nitems = read_u32(ws, item_off + nt, st);
    if (nitems >= (1LL << 31)) throw DevError{"too many degridding items"};
    if (nitems == 0) return;
    k_sort_chunks<T><<<unsigned(nitems), 1024, 0, st>>>(g, recs, off, item_off, rc);
    HIPCHK(hipGetLastError());
}

// Longest transform done in one LDS pass (fft_maxtot); longer ones take two
// passes over memory. RANG_FFT_LDS_MAX lowers the limit (a tuning and test
// knob; it never raises it).
template <class T>
int lds_fft_limit() {
    int lim = fft_maxtot<T>();
    if (const char* e = std::getenv("RANG_FFT_LDS_MAX")) {
        const long v = std::atol(e);
        if (v >= 16 && v < lim) lim = int(v);
    }
    return lim;
}

// n = n1 n2 with both factors at most lim, n1 nearest sqrt(n).
bool long_split(int n, int lim, int& n1, int& n2) {
    n1 = n2 = 0;
    double best = 1e300;
    for (int d = 2; d <= lim && d < n; ++d) {
        if (n % d || n / d > lim) continue;
        const double dist = std::fabs(double(d) - std::sqrt(double(n)));
        if (dist < best) {
            best = dist;
            n1 = d;
            n2 = n / d;
        }
    }
    return n1 > 0;
}

// Geometry and tiling into the kernels' Geo. Gridding keeps a plane of the
// tile's accumulators within 40 KiB; degridding prefers tiles whose halo of
// every plane fits the staged LDS limit.
// The row transforms read only the gridded tiles (vis2dirty) and write only
// the rows degridding reads (dirty2vis) when the planes take `bytes` >=
// 512 MiB (measured: 8.6 + 5 ms less per step at 45 GB of planes; at
// 316 MB the per-cell lookups cost more than they save), unless
// RANG_GRID_OCC is 0 (never) or 1 (always).
bool grid_occ(std::size_t bytes) {
    const char* e = std::getenv("RANG_GRID_OCC");
    if (e && e[0]) return std::atoi(e) != 0;
    return bytes >= (std::size_t(1) << 29);
}

// Column transforms on whole columns with a tiled transpose beside them
// when the planes' image columns take `bytes` >= 512 MiB (measured: at
// 25 GB 41 -> 20 + 10 ms per direction; at 136 MB the strided form is
// 0.07 ms faster per step), unless RANG_FFT_TCM is 0 (never) or 1 (always).
bool fft_tcm(std::size_t bytes) {
    const char* e = std::getenv("RANG_FFT_TCM");  // read per call: tests switch it
    if (e && e[0]) return std::atoi(e) != 0;
    return bytes >= (std::size_t(1) << 29);
}

// Gridding on the matrix cores unless RANG_GRID_MFMA is 0.
bool mfma_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("RANG_GRID_MFMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <class T>
bool fill_geo(const Geometry& G, Geo& g, int nout, bool degrid, std::string& err, int tile = 0) {
    g.nx = int(G.nx);
    g.ny = int(G.ny);
    g.nu = int(G.nu);
    g.nv = int(G.nv);
    g.supp = G.supp;
    g.pixsize_x = G.pixsize_x;
    g.pixsize_y = G.pixsize_y;
    g.ushift = G.ushift;
    g.vshift = G.vshift;
    g.maxiu0 = G.maxiu0;
    g.maxiv0 = G.maxiv0;
    g.wmin = G.wmin;
    g.dw = G.dw;
    g.xdw = G.xdw;
    g.wshift = G.wshift;
    g.wgrid = G.do_wgridding ? 1 : 0;
    g.nplanes = int(G.nplanes);
    g.nb = int(std::max<std::size_t>(G.nplanes, 1));
    g.r_l = G.r_l;
    g.r_m = G.r_m;
    g.lshift = G.lshift;
    g.mshift = G.mshift;
    g.nshift = G.nshift;
    g.shifting = G.shifting ? 1 : 0;
    g.lmshift = G.lmshift ? 1 : 0;
    g.divide_by_n = G.divide_by_n ? 1 : 0;
    g.tnx = G.lmshift ? g.nx : g.nx / 2 + 1;
    g.tny = G.lmshift ? g.ny : g.ny / 2 + 1;
    const int P = g.wgrid ? g.nplanes : 1;
    int hi = int(std::sqrt(40960.0 / (double(nout) * 2.0 * sizeof(GridAcc))));
    int target = 32;
    if (degrid) {
        // largest side whose staged planes fit, within 48
        const double side = std::sqrt(double(degrid_lds_max<T>()) / (double(P) * sizeof(cx<T>))) - (g.supp - 1);
        hi = std::min(48, std::max(int(side), sizeof(T) == 4 ? 8 : 16));
        target = hi;
    } else if (tile > 0) {
        target = std::min(tile, hi);  // the tuner's tile side
    }
    // single-precision w-gridding of one image on 16 x 16 tiles on the
    // matrix cores, unless a tile side is asked for
    g.mfma = (sizeof(T) == 4 && !degrid && nout == 1 && g.wgrid && g.supp <= 8 && tile == 0 &&
              g.nu % kMfmaTile == 0 && g.nv % kMfmaTile == 0 && mfma_enabled())
                 ? 1
                 : 0;
    g.Tu = g.mfma ? kMfmaTile : choose_tile(g.nu, g.supp, std::min(hi, 64), target);
    g.Tv = g.mfma ? kMfmaTile : choose_tile(g.nv, g.supp, std::min(hi, 64), target);
    if (g.Tu < 0 || g.Tv < 0) {
        err = "no tile size divides the grid";
        return false;
    }
    g.ntu = g.nu / g.Tu;
    g.ntv = g.nv / g.Tv;
    const long long nkeys = (long long)g.ntu * g.ntv * 4 * g.nb;
    if (nkeys >= 0xffffffffll) {
        err = "grid too large for one call";
        return false;
    }
    g.lu1 = g.lu2 = g.lv1 = g.lv2 = 0;
    const int lim = lds_fft_limit<T>();
    if (g.nu > lim && !long_split(g.nu, lim, g.lu1, g.lu2)) {
        err = "no two-pass split of the grid size for the transforms";
        return false;
    }
    if (g.nv > lim && !long_split(g.nv, lim, g.lv1, g.lv2)) {
        err = "no two-pass split of the grid size for the transforms";
        return false;
    }
    g.cm = g.lu1 == 0 ? 1 : 0;
    return true;
}

// The active visibilities (activity of `din`) as records in bucket order, in
// the slots of record set `set` (0 or 1); the records' values are formed
// from `vals` (din when null).
template <class T, int MODE, bool USE_VIS>
void sort_records(Workspace& ws, const DevIn& din, const Geo& g, long long nvis, hipStream_t st,
                  Rec<T>** recs_out, std::uint32_t** off_out, int set = 0,
                  const DevIn* vals = nullptr) {
    const long long nkeys = (long long)g.ntu * g.ntv * 4 * g.nb;
    std::uint32_t* keys = static_cast<std::uint32_t*>(ws.get(set ? kSlotKeys2 : kSlotKeys, sizeof(std::uint32_t) * g.n));
    std::uint32_t* counts = static_cast<std::uint32_t*>(ws.get(set ? kSlotCounts2 : kSlotCounts, sizeof(std::uint32_t) * (nkeys + 1)));
    std::uint32_t* off = static_cast<std::uint32_t*>(ws.get(set ? kSlotOffsets2 : kSlotOffsets, sizeof(std::uint32_t) * (nkeys + 1)));
    std::uint32_t* cursor = static_cast<std::uint32_t*>(ws.get(set ? kSlotCursor2 : kSlotCursor, sizeof(std::uint32_t) * (nkeys + 1)));
    Rec<T>* recs = static_cast<Rec<T>*>(ws.get(set ? kSlotRecs2 : kSlotRecs, sizeof(Rec<T>) * std::max<long long>(nvis, 1)));
    fill_zero(counts, sizeof(std::uint32_t) * (nkeys + 1), st);
    fill_zero(cursor, sizeof(std::uint32_t) * (nkeys + 1), st);
    const int eblocks = int((g.n + 255) / 256);
    k_keys<T, USE_VIS><<<eblocks, 256, 0, st>>>(din, g, keys, counts);
    HIPCHK(hipGetLastError());
    exclusive_scan(ws, counts, off, nkeys + 1, st);
    k_scatter<T, MODE><<<eblocks, 256, 0, st>>>(vals ? *vals : din, g, keys, off, cursor, recs);
    HIPCHK(hipGetLastError());
    *recs_out = recs;
    *off_out = off;
}

Geo base_geo(std::size_t nrow, std::size_t nchan, const Options& o) {
    Geo g{};
    g.n = (long long)(nrow * nchan);
    g.nrow = int(nrow);
    g.nchan = int(nchan);
    double r_l, r_m;
    recentring(o, r_l, r_m);
    g.r_l = r_l;
    g.r_m = r_m;
    g.fu = o.flip_u ? -1.0 : 1.0;
    g.fv = o.flip_v ? -1.0 : 1.0;
    g.fw = !o.flip_w ? -1.0 : 1.0;  // the core's w flip (wgridder.rs:2461)
    return g;
}

int device_of() {
    int d = 0;
    HIPCHK(hipGetDevice(&d));
    return d;
}

bool common_checks(const Options& o, const DevInputs& in, const double* freq_host, std::string& err) {
    if (!check_common(o, in.nrow, in.nchan, err)) return false;
    if (in.nrow * in.nchan >= (std::size_t(1) << 31)) {
        err = "too many visibilities for one call (at most 2^31 - 1)";
        return false;
    }
    return check_freq(freq_host, in.nchan, err);
}

double* upload_foc(Workspace& ws, const double* freq_host, std::size_t nchan, hipStream_t st,
                   HostKeep& keep) {
    std::vector<double> foc(nchan);
    for (std::size_t c = 0; c < nchan; ++c) foc[c] = freq_host[c] / kCLight;
    double* d = static_cast<double*>(ws.get(kSlotFoc, sizeof(double) * nchan));
    if (ws.foc_key == foc) return d;
    ws.foc_key = foc;
    keep.d.push_back(std::move(foc));
    HIPCHK(hipMemcpyAsync(d, keep.d.back().data(), sizeof(double) * nchan, hipMemcpyHostToDevice, st));
    return d;
}

template <int V>
using ic = std::integral_constant<int, V>;

// Calls f(NT, MAXTOT) with the smallest in-LDS transform size holding `need`
// values and its thread count.
template <class T, class F>
void fft_dispatch(int need, F&& f) {
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
            f(ic<1024>{}, ic<16384>{});
        else
            throw DevError{"transform too long for LDS"};
    } else {
        throw DevError{"transform too long for LDS"};
    }
}

template <class K>
void set_lds(K kernel, std::size_t bytes) {
    lds_limit((const void*)kernel, bytes);
}

// Column block of the u transforms: as many columns as fit 8192 values.
int col_block(int nu) { return std::max(1, std::min(8, 8192 / nu)); }

// Twiddles of the two-pass transforms of one call.
template <class T>
struct LongTabs {
    FftDesc u1, u2, v1, v2;
    cx<T>*tu1 = nullptr, *tu2 = nullptr, *tun = nullptr, *tv1 = nullptr, *tv2 = nullptr, *tvn = nullptr;
};

template <class T>
void make_long(Workspace& ws, const Geo& g, hipStream_t st, HostKeep& keep, LongTabs<T>& lt) {
    std::vector<double> all;
    long long ofs[6] = {0, 0, 0, 0, 0, 0};
    auto add = [&](int k, const std::vector<double>& v) {
        ofs[k] = (long long)all.size() / 2;
        all.insert(all.end(), v.begin(), v.end());
    };
    auto roots = [](int n) {
        std::vector<double> r(2 * std::size_t(n));
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int k = 0; k < n; ++k) {
            const long double a = -2.0L * pi * (long double)k / (long double)n;
            r[2 * k] = double(cosl(a));
            r[2 * k + 1] = double(sinl(a));
        }
        return r;
    };
    std::vector<double> t;
    if (g.lu1) {
        if (!fft_describe(g.lu1, lt.u1, t)) throw DevError{"two-pass transform factor"};
        add(0, t);
        if (!fft_describe(g.lu2, lt.u2, t)) throw DevError{"two-pass transform factor"};
        add(1, t);
        add(2, roots(g.nu));
    }
    if (g.lv1) {
        if (!fft_describe(g.lv1, lt.v1, t)) throw DevError{"two-pass transform factor"};
        add(3, t);
        if (!fft_describe(g.lv2, lt.v2, t)) throw DevError{"two-pass transform factor"};
        add(4, t);
        add(5, roots(g.nv));
    }
    if (all.empty()) return;
    // uploaded once per geometry: a call that repeats it reads no host memory
    const std::vector<double> key{double(sizeof(T)), double(g.nu), double(g.nv), double(g.lu1), double(g.lv1)};
    cx<T>* d;
    if (ws.long_key == key) {
        d = static_cast<cx<T>*>(ws.ptr[kSlotLong]);
    } else {
        d = reinterpret_cast<cx<T>*>(upload_real<T>(ws, kSlotLong, all, st, keep));
        ws.long_key = key;
    }
    if (g.lu1) {
        lt.tu1 = d + ofs[0];
        lt.tu2 = d + ofs[1];
        lt.tun = d + ofs[2];
    }
    if (g.lv1) {
        lt.tv1 = d + ofs[3];
        lt.tv2 = d + ofs[4];
        lt.tvn = d + ofs[5];
    }
}

// Two-pass transforms of `rows` contiguous rows of length n1 n2 in place
// (MODE 0) or, for MODE 1, with pass 2 gathering the ny image columns into out.
template <class T, int MODE>
void long_rows(cx<T>* data, long long rows, int n1, int n2, const FftDesc& d1, const FftDesc& d2,
               const cx<T>* t1, const cx<T>* t2, const cx<T>* tn, int sgn, cx<T>* out, int ny,
               hipStream_t st) {
    const int budget = fft_maxtot<T>() / 2;
    const int C2 = std::max(1, std::min(n2, budget / n1));
    fft_dispatch<T>(n1 * C2, [&](auto nt, auto mt) {
        constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
        const std::size_t lds = sizeof(cx<T>) * n1 * C2;
        set_lds(k_lrow_p1<T, NT, MT>, lds);
        k_lrow_p1<T, NT, MT><<<blocks_of(rows * ((n2 + C2 - 1) / C2), NT), NT, lds, st>>>(data, n1, n2, C2, d1, t1, tn, sgn);
    });
    HIPCHK(hipGetLastError());
    const int K1 = std::max(1, std::min(n1, budget / n2));
    fft_dispatch<T>(n2 * K1, [&](auto nt, auto mt) {
        constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
        const std::size_t lds = sizeof(cx<T>) * n2 * K1;
        set_lds(k_lrow_p2<T, NT, MT, MODE>, lds);
        k_lrow_p2<T, NT, MT, MODE><<<blocks_of(rows * ((n1 + K1 - 1) / K1), NT), NT, lds, st>>>(data, n1, n2, K1, d2, t2, sgn,
                                                                               out, ny);
    });
    HIPCHK(hipGetLastError());
}

// Two-pass transforms along the middle axis of [P][n1 n2][M] in place (MODE
// 0) or, for MODE 2, with pass 2 projecting the image rows into proj.
template <class T, int MODE>
void long_cols(cx<T>* data, int P, int M, int n1, int n2, const FftDesc& d1, const FftDesc& d2,
               const cx<T>* t1, const cx<T>* t2, const cx<T>* tn, int sgn, T* proj, const Geo& g,
               const double* screen, hipStream_t st) {
    const int budget = fft_maxtot<T>() / 2;
    int Cm = std::max(1, std::min(std::min(16, M), budget / n1));
    const int C2 = std::max(1, std::min(n2, budget / (n1 * Cm)));
    long long nblk = (long long)P * ((M + Cm - 1) / Cm) * ((n2 + C2 - 1) / C2);
    fft_dispatch<T>(n1 * Cm * C2, [&](auto nt, auto mt) {
        constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
        const std::size_t lds = sizeof(cx<T>) * n1 * Cm * C2;
        set_lds(k_lcol_p1<T, NT, MT>, lds);
        k_lcol_p1<T, NT, MT><<<blocks_of(nblk, NT), NT, lds, st>>>(data, n1, n2, M, Cm, C2, d1, t1, tn, sgn);
    });
    HIPCHK(hipGetLastError());
    Cm = std::max(1, std::min(std::min(16, M), budget / n2));
    const int K1 = std::max(1, std::min(n1, budget / (n2 * Cm)));
    nblk = (long long)P * ((M + Cm - 1) / Cm) * ((n1 + K1 - 1) / K1);
    fft_dispatch<T>(n2 * Cm * K1, [&](auto nt, auto mt) {
        constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
        const std::size_t lds = sizeof(cx<T>) * n2 * Cm * K1;
        set_lds(k_lcol_p2<T, NT, MT, MODE>, lds);
        k_lcol_p2<T, NT, MT, MODE><<<blocks_of(nblk, NT), NT, lds, st>>>(data, n1, n2, M, Cm, K1, d2, t2, sgn, proj, g,
                                                              screen);
    });
    HIPCHK(hipGetLastError());
}



// ------------------------------------------------------------ prepared calls

// Partial sums of a block of 256 threads into part[blockIdx.x].
__device__ __forceinline__ void block_sum(double v, double* part) {
    __shared__ double s[256];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

constexpr int kRedBlocks = 512;

// The residual visibilities r = data - pred (pred null: the data) and the
// misfit partials sum w |r|^2, formed in double from the values in T
// (wgridder.rs:3318-3335, 3337-3353).
template <class T>
__global__ __launch_bounds__(256) void k_residual(DevIn in, const cx<T>* pred, cx<T>* r, long long n,
                                                  double* part) {
    double q = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        cx<T> d = vis_get<T>(in, i);
        if (pred) d = csub(d, pred[i]);
        r[i] = d;
        T w = T(1);
        wgt_get<T>(in, i, w);
        const double re = double(d.re), im = double(d.im);
        q += double(w) * (re * re + im * im);
    }
    block_sum(q, part);
}

// map[index] = k for the sorted records k.
template <class T>
__global__ __launch_bounds__(256) void k_index_map(const Rec<T>* __restrict__ recs, long long n,
                                                   std::uint32_t* __restrict__ map) {
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k < n) map[recs[k].idx & ~kFlip] = std::uint32_t(k);
}

// pos[r] = map[index of record r].
template <class T>
__global__ __launch_bounds__(256) void k_pos_of(const Rec<T>* __restrict__ recs, long long n,
                                                const std::uint32_t* __restrict__ map,
                                                std::uint32_t* __restrict__ pos) {
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r < n) pos[r] = map[recs[r].idx & ~kFlip];
}

// stored[k] = d, the data of sorted record k (dense at its index, or at k
// when `sorted`), and per block sum w |d|^2 with the data's weights (dense).
template <class T>
__global__ __launch_bounds__(256) void k_gather_resid(const Rec<T>* __restrict__ recs, long long n, DevIn data,
                                                      int sorted, cx<T>* __restrict__ stored, double* part) {
    double q = 0.0;
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const long long idx = recs[k].idx & ~kFlip;
        const cx<T> d = vis_get<T>(data, sorted ? k : idx);
        stored[k] = d;
        T w = T(1);
        wgt_get<T>(data, idx, w);
        q += double(w) * (double(d.re) * double(d.re) + double(d.im) * double(d.im));
    }
    block_sum(q, part);
}

// out[k] = dense[index of sorted record k], elements of `bytes` (4, 8 or 16).
template <class T>
__global__ __launch_bounds__(256) void k_gather_sorted(const Rec<T>* __restrict__ recs, long long n,
                                                       const unsigned char* __restrict__ dense, int bytes,
                                                       unsigned char* __restrict__ out) {
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const long long idx = recs[k].idx & ~kFlip;
    if (bytes == 4)
        reinterpret_cast<std::uint32_t*>(out)[k] = reinterpret_cast<const std::uint32_t*>(dense)[idx];
    else if (bytes == 8)
        reinterpret_cast<uint2*>(out)[k] = reinterpret_cast<const uint2*>(dense)[idx];
    else
        reinterpret_cast<uint4*>(out)[k] = reinterpret_cast<const uint4*>(dense)[idx];
}

// vis[i] *= w[i].
template <class T>
__global__ __launch_bounds__(256) void k_weigh(DevIn in, cx<T>* vis, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        T w = T(1);
        if (wgt_get<T>(in, i, w)) vis[i] = {vis[i].re * w, vis[i].im * w};
    }
}

// Partials of sum x y in double.
template <class T>
__global__ __launch_bounds__(256) void k_dot(const T* x, const T* y, long long n, double* part) {
    double q = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        q += double(x[i]) * double(y[i]);
    block_sum(q, part);
}

// x += alpha p, r -= alpha ap, and the partials of |r|^2 (clean.rs:291-296).
template <class T>
__global__ __launch_bounds__(256) void k_cg_step(T* x, T* r, const T* p, const T* ap, double alpha,
                                                 long long n, double* part) {
    double q = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        x[i] = T(double(x[i]) + alpha * double(p[i]));
        const T ri = T(double(r[i]) - alpha * double(ap[i]));
        r[i] = ri;
        q += double(ri) * double(ri);
    }
    block_sum(q, part);
}

// p = r + beta p (clean.rs:298-300).
template <class T>
__global__ __launch_bounds__(256) void k_cg_dir(T* p, const T* r, double beta, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        p[i] = T(double(r[i]) + beta * double(p[i]));
}

// The same steps with the scalars on the device, s = {|r|^2, <p, A p>,
// next |r|^2}: alpha = s0 / s1 (no step where s1 <= 0, where the solve
// stops), beta = s2 / s0.
template <class T>
__global__ __launch_bounds__(256) void k_cg_step_s(T* x, T* r, const T* p, const T* ap, const double* s, long long n,
                                                   double* part) {
    const double alpha = s[1] > 0.0 ? s[0] / s[1] : 0.0;
    double q = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        x[i] = T(double(x[i]) + alpha * double(p[i]));
        const T ri = T(double(r[i]) - alpha * double(ap[i]));
        r[i] = ri;
        q += double(ri) * double(ri);
    }
    block_sum(q, part);
}

template <class T>
__global__ __launch_bounds__(256) void k_cg_dir_s(T* p, const T* r, const double* s, long long n) {
    const double beta = s[1] > 0.0 ? s[2] / s[0] : 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        p[i] = T(double(r[i]) + beta * double(p[i]));
}

// s0 <- s2 after a step; s0 <- v at the start.
__global__ void k_cg_roll(double* s) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && s[1] > 0.0) s[0] = s[2];
}
__global__ void k_set_scalar(double* s, double v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *s = v;
}

// FISTA (core/src/prox_solver.rs:283-445). Per iteration the sums
// <d, x>, <x, G x>, the penalty, |x - x_prev|^2, |x|^2, the flux, the
// nonzero count and the curvature's numerator and denominator, then the
// largest optimality violation and the smallest dual scale bound.
constexpr int kFistaSums = 9;
constexpr int kFistaSlots = kFistaSums + 2;

__device__ __forceinline__ void block_reduce_fista(double (&v)[kFistaSlots], double* part) {
    __shared__ double s[256];
    for (int q = 0; q < kFistaSlots; ++q) {
        s[threadIdx.x] = v[q];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (int(threadIdx.x) < h) {
                const double a = s[threadIdx.x], b = s[threadIdx.x + h];
                s[threadIdx.x] = q < kFistaSums ? a + b : q == kFistaSums ? fmax(a, b) : fmin(a, b);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(long long)blockIdx.x * kFistaSlots + q] = s[0];
        __syncthreads();
    }
}

// G x by linearity from G y (prox_solver.rs:332-341), the curvature terms
// (:342-353) and the certificate of x (:236-281) with the per-pixel
// thresholds lam * factor.
template <class T>
__global__ __launch_bounds__(256) void k_fista_eval(const T* y, const T* y_prev, const T* gy,
                                                    const T* gy_prev, const T* gx_prev, T* gx, const T* x,
                                                    const T* x_prev, const T* d, const T* factors,
                                                    double lam, int sgn, double beta, long long n,
                                                    double* part) {
    double v[kFistaSlots];
    for (int q = 0; q < kFistaSums; ++q) v[q] = 0.0;
    v[kFistaSums] = 0.0;
    v[kFistaSums + 1] = INFINITY;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double gyi = double(gy[i]);
        const double gxi = beta == 0.0 ? gyi : (gyi + beta * double(gx_prev[i])) / (1.0 + beta);
        gx[i] = T(gxi);
        const double dy = double(y[i]) - double(y_prev[i]);
        v[7] += dy * (gyi - double(gy_prev[i]));
        v[8] += dy * dy;
        const double xi = double(x[i]), di = double(d[i]);
        v[0] += di * xi;
        v[1] += xi * gxi;
        const double r = di - gxi;
        const double lami = factors ? lam * double(factors[i]) : lam;
        double viol;
        if (xi > 0.0)
            viol = fabs(r - lami);
        else if (xi < 0.0)
            viol = fabs(r + lami);
        else if (sgn)
            viol = fmax(fabs(r) - lami, 0.0);
        else
            viol = fmax(r - lami, 0.0);
        const double excess = sgn ? fabs(r) : r;
        const double bound = excess > 0.0 ? lami / excess : INFINITY;
        v[2] += lami * fabs(xi);
        v[kFistaSums] = fmax(v[kFistaSums], viol);
        v[kFistaSums + 1] = fmin(v[kFistaSums + 1], bound);
        const double ch = xi - double(x_prev[i]);
        v[3] += ch * ch;
        v[4] += xi * xi;
        v[5] += xi;
        v[6] += xi != 0.0 ? 1.0 : 0.0;
    }
    block_reduce_fista(v, part);
}

// The forward-backward step from y and the prox of the L1 prior
// (prox_solver.rs:407-412, 80-101), and the partials of the restart test
// <y - x_next, x_next - x> (:416-419).
template <class T>
__global__ __launch_bounds__(256) void k_fista_step(const T* y, const T* d, const T* gy, const T* x, T* x_next,
                                                    const T* factors, double lam, int sgn, double step,
                                                    long long n, double* part) {
    double q = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double yi = double(y[i]);
        const double z = yi + step * (double(d[i]) - double(gy[i]));
        const double tau = factors ? step * lam * double(factors[i]) : step * lam;
        // f64::signum: +-1 by the sign bit
        const double xn = sgn ? copysign(1.0, z) * fmax(fabs(z) - tau, 0.0) : fmax(z - tau, 0.0);
        x_next[i] = T(xn);
        q += (yi - xn) * (xn - double(x[i]));
    }
    block_sum(q, part);
}

// y = x + beta (x - x_prev) (prox_solver.rs:438-440).
template <class T>
__global__ __launch_bounds__(256) void k_fista_extrap(T* y, const T* x, const T* x_prev, double beta,
                                                      long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        y[i] = T(double(x[i]) + beta * (double(x[i]) - double(x_prev[i])));
}

// v *= s.
template <class T>
__global__ __launch_bounds__(256) void k_scale(T* v, double s, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        v[i] = T(double(v[i]) * s);
}

}  // namespace

void set_stage_timing(bool on) { g_timing = on; }
int stage_timing_count() { return int(g_times.size()); }
const char* stage_timing_name(int i) {
    return i >= 0 && i < int(g_times.size()) ? g_times[i].first.c_str() : "";
}
double stage_timing_ms(int i) { return i >= 0 && i < int(g_times.size()) ? g_times[i].second : 0.0; }

// The work items of one record set and tiling, kept by a prepared plan: its
// records never change, so the item offsets and their count are formed once
// and later calls read no count back from the device (none of them waits on
// the host, and a stream of them can be captured into a graph).
struct ItemCache {
    bool valid = false;
    long long ub = 0;
};

// A gridding made ready to launch: its tables, grid, plane groups and items.
template <class T>
struct GridPrep {
    Tables<T> tb;
    LongTabs<T> lt;
    int P = 0, G_planes = 0, ngroups = 0, ntg = 0, TvP = 0;
    std::size_t lds = 0;
    std::uint32_t* item_off = nullptr;
    long long ub = 0;  // workgroups of the gridding launch
    cx<T>* grid = nullptr;
    const std::uint8_t* occ = nullptr;  // tiles gridded, for the row transform (null: all cells cleared)
};

// Tables, the grid (its cells reached by several items cleared) and the
// work items of one vis2dirty of `nout` outputs.
template <class T>
void grid_prepare(Workspace& ws, const Geometry& G, const Geo& g, const std::uint32_t* off, int nout,
                  hipStream_t st, HostKeep& keep, StageTimer& timer, ItemCache* cache, GridPrep<T>& gp) {
    make_tables<T>(ws, G, g, st, keep, gp.tb, 0);
    make_long<T>(ws, g, st, keep, gp.lt);
    timer.mark("tables");
    const int P = g.wgrid ? g.nplanes : 1;
    gp.P = P;
    const long long plane = (long long)g.nu * g.nv;
    gp.grid = static_cast<cx<T>*>(ws.get(kSlotGrid, sizeof(cx<T>) * plane * P * nout));
    // padded row stride: TvP = S (mod 32) spreads a footprint's rows over the banks
    gp.TvP = g.Tv + ((g.supp - g.Tv) % 32 + 32) % 32;
    if (g.mfma) {
        // the matrix cores hold kMfmaPlanes planes of a tile in registers
        // (groups of 8 measured faster than balanced ones: 133 vs 139 ms)
        gp.G_planes = std::max(1, std::min(P, kMfmaPlanes));
        gp.lds = 0;
    } else {
        const std::size_t per_plane = std::size_t(nout) * g.Tu * gp.TvP * 2 * sizeof(GridAcc);
        if (per_plane > grid_lds<T>()) throw DevError{"tile too large for LDS"};
        // at most 32 planes per item: a staged record keeps its item plane in 5 bits
        gp.G_planes = std::max(1, std::min({P, int(grid_lds<T>() / per_plane), 32}));
        gp.lds = per_plane * std::min(P, gp.G_planes);
    }
    gp.ngroups = (P + gp.G_planes - 1) / gp.G_planes;
    gp.ntg = g.ntu * g.ntv * gp.ngroups;
    const int ntg = gp.ntg;
    std::uint32_t* items = static_cast<std::uint32_t*>(ws.get(kSlotItems, sizeof(std::uint32_t) * (ntg + 1)));
    gp.item_off = static_cast<std::uint32_t*>(ws.get(kSlotItemOff, sizeof(std::uint32_t) * (ntg + 1)));
    if (!cache || !cache->valid) {
        k_grid_items<<<(ntg + 1 + 255) / 256, 256, 0, st>>>(g, off, gp.tb.order, gp.G_planes, gp.ngroups, g.supp,
                                                             kGridChunk, items);
        HIPCHK(hipGetLastError());
        exclusive_scan(ws, items, gp.item_off, ntg + 1, st);
    }
    // the in-LDS row transform reads only the tiles an item reached
    // (RANG_GRID_OCC=0: every cell cleared and read)
    gp.occ = nullptr;
    if (!g.lv1 && grid_occ(sizeof(cx<T>) * std::size_t(plane) * P)) {
        std::uint8_t* occ = static_cast<std::uint8_t*>(ws.get(kSlotOcc, std::size_t(ntg)));
        k_occupancy<<<(ntg + 255) / 256, 256, 0, st>>>(gp.item_off, gp.tb.order, ntg, gp.ngroups, g.ntu * g.ntv, occ);
        HIPCHK(hipGetLastError());
        gp.occ = occ;
    }
    const int skip = gp.occ ? 1 : 0;
    if (nout == 2)
        k_zero_split<T, 2><<<ntg, 256, 0, st>>>(g, gp.item_off, gp.tb.order, gp.G_planes, gp.ngroups, gp.grid, skip);
    else
        k_zero_split<T, 1><<<ntg, 256, 0, st>>>(g, gp.item_off, gp.tb.order, gp.G_planes, gp.ngroups, gp.grid, skip);
    HIPCHK(hipGetLastError());
    if (cache && cache->valid) {
        gp.ub = cache->ub;
    } else {
        gp.ub = read_u32(ws, gp.item_off + ntg, st);
        if (cache) {
            cache->ub = gp.ub;
            cache->valid = true;
        }
    }
    if (gp.ub >= (1LL << 31)) throw DevError{"too many gridding items"};
}

// The transforms of the gridded planes and their corrected sum into `dirty`
// and/or `psf`.
template <class T>
void grid_finish(Workspace& ws, const Geo& g, const GridPrep<T>& gp, int nout, T* dirty, T* psf, hipStream_t st,
                 StageTimer& timer) {
    const long long npix = (long long)g.nx * g.ny;
    const int P = gp.P;
    const long long plane = (long long)g.nu * g.nv;
    const Tables<T>& tb = gp.tb;
    const LongTabs<T>& lt = gp.lt;
    cx<T>* tmp = static_cast<cx<T>*>(ws.get(kSlotTmp, sizeof(cx<T>) * (long long)P * g.nu * g.ny));
    T* proj = static_cast<T*>(ws.get(kSlotProj, sizeof(T) * (long long)P * npix));
    const int C = col_block(g.nu);
    T* outs[2] = {dirty ? dirty : psf, psf};
    for (int oi = 0; oi < nout; ++oi) {
        cx<T>* gsrc = gp.grid + (long long)oi * P * plane;
        if (g.lv1) {
            long_rows<T, 1>(gsrc, (long long)P * g.nu, g.lv1, g.lv2, lt.v1, lt.v2, lt.tv1, lt.tv2, lt.tvn, +1,
                            tmp, g.ny, st);
        } else {
            fft_dispatch<T>(g.nv, [&](auto nt, auto mt) {
                constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                const std::size_t lds_v = sizeof(cx<T>) * g.nv;
                set_lds(k_rows_inv<T, NT, MT>, lds_v);
                const RowOcc ro{gp.occ, gp.G_planes, fft_magic(unsigned(g.Tv))};
                k_rows_inv<T, NT, MT><<<blocks_of((long long)P * g.nu, NT), NT, lds_v, st>>>(gsrc, tmp, tb.dv, tb.twv,
                                                                                             g.ny, g, ro);
            });
            HIPCHK(hipGetLastError());
        }
        timer.mark("fft_v");
        if (g.lu1) {
            long_cols<T, 2>(tmp, P, g.ny, g.lu1, g.lu2, lt.u1, lt.u2, lt.tu1, lt.tu2, lt.tun, +1, proj, g,
                            tb.screen, st);
        } else {
            // the grid of this output is consumed: it takes the columns of
            // tmp, transposed, so that the u transform reads them whole
            const bool scm = fft_tcm(sizeof(cx<T>) * std::size_t(P) * g.nu * g.ny);
            if (scm) {
                transpose_planes<T>(tmp, gsrc, P, g.nu, g.ny, st);
                timer.mark("transpose");
            }
            fft_dispatch<T>(g.nu * C, [&](auto nt, auto mt) {
                constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
                const std::size_t lds_u = sizeof(cx<T>) * g.nu * C;
                const unsigned nb = unsigned(blocks_of((long long)(g.ny + C - 1) / C * P, NT));
                if (scm) {
                    set_lds(k_cols_inv<T, NT, MT, true>, lds_u);
                    k_cols_inv<T, NT, MT, true><<<nb, NT, lds_u, st>>>(gsrc, proj, tb.du, tb.twu, g, tb.screen, C);
                } else {
                    set_lds(k_cols_inv<T, NT, MT>, lds_u);
                    k_cols_inv<T, NT, MT><<<nb, NT, lds_u, st>>>(tmp, proj, tb.du, tb.twu, g, tb.screen, C);
                }
            });
            HIPCHK(hipGetLastError());
        }
        timer.mark("fft_u");
        if (g.cm)
            k_sum_cm<T><<<dim3((g.nx + 31) / 32, (g.ny + 31) / 32), 256, 0, st>>>(proj, outs[oi], g, P, tb.corr,
                                                                                 tb.cfu, tb.cfv);
        else
            k_sum<T><<<(npix + 255) / 256, 256, 0, st>>>(proj, outs[oi], g, P, tb.corr, tb.cfu, tb.cfv);
        HIPCHK(hipGetLastError());
        timer.mark("sum");
    }
}

// vis2dirty from sorted records: tables, gridding, the transforms and the
// sum over planes into `dirty` and/or `psf` (nx * ny each). With `form` the
// records hold unit values and the data values are formed from din.vis while
// gridding (one output).
template <class T>
void grid_pipeline(Workspace& ws, const Geometry& G, const Geo& g, const DevIn& din,
                   const Rec<T>* recs, const std::uint32_t* off, int nout, bool form, T* dirty,
                   T* psf, hipStream_t st, HostKeep& keep, StageTimer& timer, ItemCache* cache = nullptr) {
    GridPrep<T> gp;
    grid_prepare<T>(ws, G, g, off, nout, st, keep, timer, cache, gp);
    const std::uint32_t rc = kGridChunk;
    if (gp.ub == 0) {
        // nothing to grid: every cell was cleared above
    } else if (g.mfma) {
        if constexpr (sizeof(T) == 4)
            launch_grid_mfma(g.supp, int(gp.ub), form, st, g, recs, off, gp.tb.coeff, din, gp.G_planes, gp.ngroups,
                             rc, gp.item_off, gp.tb.order, gp.grid);
    } else if (nout == 2)
        launch_grid<T, 2, false>(g.supp, int(gp.ub), gp.lds, st, g, recs, off, gp.tb.coeff, din, gp.G_planes,
                                 gp.ngroups, gp.TvP, rc, gp.item_off, gp.tb.order, gp.grid);
    else if (form)
        launch_grid<T, 1, true>(g.supp, int(gp.ub), gp.lds, st, g, recs, off, gp.tb.coeff, din, gp.G_planes,
                                gp.ngroups, gp.TvP, rc, gp.item_off, gp.tb.order, gp.grid);
    else
        launch_grid<T, 1, false>(g.supp, int(gp.ub), gp.lds, st, g, recs, off, gp.tb.coeff, din, gp.G_planes,
                                 gp.ngroups, gp.TvP, rc, gp.item_off, gp.tb.order, gp.grid);
    timer.mark("grid");
    grid_finish<T>(ws, g, gp, nout, dirty, psf, st, timer);
}

// vis2dirty on the device's workspace, or (fixed) on its fixed scratch block.
template <class T>
int vis2dirty_on(bool fixed, const Options& o, const DevInputs& in, const double* freq_host,
                 T* dirty, T* psf, double* info, hipStream_t st, std::string& err, int tile = 0) {
    try {
        if (!dirty && !psf) {
            err = "null array";
            return 1;
        }
        if (!common_checks(o, in, freq_host, err)) return 1;
        Workspace& ws = fixed ? scratch_workspace(device_of()) : workspace(device_of());
        std::lock_guard<std::mutex> lk(ws.mu);
        if (fixed) arena_reset(ws);
        HostKeep keep;
        StageTimer timer(st);
        DevIn din{in.uvw, nullptr, in.vis, in.wgt, in.wgt2, in.mask, in.vis_kind, in.wgt_kind};
        Geo g = base_geo(in.nrow, in.nchan, o);
        din.foc = upload_foc(ws, freq_host, in.nchan, st, keep);
        const bool use_vis = psf == nullptr;
        double wlo = 0, whi = 0;
        const unsigned long long nvis = use_vis ? run_scan<T, true>(ws, din, g, st, wlo, whi)
                                                : run_scan<T, false>(ws, din, g, st, wlo, whi);
        timer.mark("scan");
        const std::size_t npix = o.nx * o.ny;
        std::fill(info, info + kNInfo, 0.0);
        if (nvis == 0) {
            if (dirty) fill_zero(dirty, sizeof(T) * npix, st);
            if (psf) fill_zero(psf, sizeof(T) * npix, st);
            HIPCHK(hipStreamSynchronize(st));
            return 0;
        }
        Geometry G;
        if (!make_geometry(o, sizeof(T) == 4, std::size_t(nvis), wlo, whi, G, err)) return 1;
        const int nout = (dirty ? 1 : 0) + (psf ? 1 : 0);
        if (!fill_geo<T>(G, g, nout, false, err, tile)) return 1;
        Rec<T>* recs;
        std::uint32_t* off;
        if (use_vis)
            sort_records<T, 0, true>(ws, din, g, (long long)nvis, st, &recs, &off);
        else
            sort_records<T, 1, false>(ws, din, g, (long long)nvis, st, &recs, &off);
        timer.mark("sort");
        grid_pipeline<T>(ws, G, g, din, recs, off, nout, false, dirty, psf, st, keep, timer);
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        std::copy(G.info, G.info + kNInfo, info);
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        if (e.status != 2) (void)hipStreamSynchronize(st);
        return e.status;
    }
}

template <class T>
int device_vis2dirty(const Options& o, const DevInputs& in, const double* freq_host, T* dirty,
                     T* psf, double* info, hipStream_t st, std::string& err) {
    return vis2dirty_on<T>(false, o, in, freq_host, dirty, psf, info, st, err);
}

template <class T>
int device_vis2dirty_fixed(const Options& o, const DevInputs& in, const double* freq_host,
                           T* dirty, T* psf, double* info, hipStream_t st, std::string& err) {
    return vis2dirty_on<T>(true, o, in, freq_host, dirty, psf, info, st, err);
}

std::size_t fixed_scratch_bytes() { return kScratchBytes; }


// ------------------------------------------------------------ tuner

namespace {

struct TuneEntry {
    long long kidx = -1;
    std::size_t nu = 0, nv = 0;
    int tile = 0;
    double ms = 0, model_ms = 0;
};

struct TuneCache {
    std::mutex mu;
    std::map<std::string, std::map<std::string, TuneEntry>> files;  // path -> signature -> choice
    std::map<std::string, TuneEntry>& of(const std::string& path) {
        auto it = files.find(path);
        if (it != files.end()) return it->second;
        auto& m = files[path];
        if (!path.empty()) {
            std::ifstream f(path);
            std::string line;
            while (std::getline(f, line)) {
                std::istringstream is(line);
                std::string key;
                TuneEntry e;
                if (is >> key >> e.kidx >> e.nu >> e.nv >> e.tile >> e.ms >> e.model_ms) m[key] = e;
            }
        }
        return m;
    }
};

TuneCache& tune_cache() {
    static TuneCache c;
    return c;
}

}  // namespace

// Times vis2dirty of one image over the cost model's cheapest kernel rows
// (`kernels` of them, one per support) at the default tile, then the best
// row over gridding tile sides 16 to 48, and keeps the fastest choice per
// signature (precision, image, cell, centre, epsilon, w-gridding, sigma
// range, visibility count to half a power of two, planes of the default
// plan) in memory and, when `path` is not empty, in that file (one line per
// signature). Sets the image's kernel_index, nu, nv and tile. A caller's
// explicit kernel_index is kept and only the tile tuned.
template <class T>
int device_tune(BatchImage& im, const char* path, int kernels, double* ms, double* default_ms, int* cached,
                hipStream_t st, std::string& err) {
    try {
        if (!im.dirty) {
            err = "null array";
            return 1;
        }
        if (!common_checks(im.o, im.in, im.freq, err)) return 1;
        if (cached) *cached = 0;
        long long nvis;
        double wlo = 0, whi = 0;
        {
            Workspace& ws = workspace(device_of());
            std::lock_guard<std::mutex> lk(ws.mu);
            HostKeep keep;
            DevIn din{im.in.uvw, nullptr, im.in.vis, im.in.wgt, im.in.wgt2, im.in.mask, im.in.vis_kind,
                      im.in.wgt_kind};
            Geo g = base_geo(im.in.nrow, im.in.nchan, im.o);
            din.foc = upload_foc(ws, im.freq, im.in.nchan, st, keep);
            nvis = (long long)run_scan<T, true>(ws, din, g, st, wlo, whi);
        }
        if (nvis == 0) return 0;
        const bool single = sizeof(T) == 4;
        Geometry G0;
        if (!make_geometry(im.o, single, std::size_t(nvis), wlo, whi, G0, err)) return 1;
        char key[512];
        std::snprintf(key, sizeof(key), "%s:%zux%zu:%.9gx%.9g:c%.9g,%.9g:e%.3g:w%d:s%.3g-%.3g:k%lld:n%.1f:p%zu",
                      single ? "f32" : "f64", im.o.nx, im.o.ny, im.o.pixsize_x, im.o.pixsize_y, im.o.center_x,
                      im.o.center_y, im.o.epsilon, im.o.do_wgridding ? 1 : 0, im.o.sigma_min, im.o.sigma_max,
                      im.o.kernel_index, std::round(2.0 * std::log2(double(nvis))) / 2.0, G0.nplanes);
        const std::string file = path ? path : "";
        TuneCache& tc = tune_cache();
        std::unique_lock<std::mutex> tl(tc.mu);
        auto& table = tc.of(file);
        auto hit = table.find(key);
        TuneEntry best;
        if (hit != table.end()) {
            best = hit->second;
            if (cached) *cached = 1;
        } else {
            tl.unlock();
            auto time_of = [&](const Options& o2, int tile, double& out) {
                double info[kNInfo];
                std::string e2;
                T* d = static_cast<T*>(im.dirty);
                double t = 1e300;
                for (int rep = 0; rep < 3; ++rep) {
                    const auto t0 = std::chrono::steady_clock::now();
                    const int rc = vis2dirty_on<T>(false, o2, im.in, im.freq, d, nullptr, info, st, e2, tile);
                    const auto t1 = std::chrono::steady_clock::now();
                    if (rc != 0) return false;
                    if (rep > 0) t = std::min(t, std::chrono::duration<double, std::milli>(t1 - t0).count());
                }
                out = t;
                return true;
            };
            std::vector<KernelChoice> choices;
            Options base = im.o;
            if (im.o.kernel_index >= 0) {
                choices.push_back(KernelChoice{int(im.o.kernel_index), G0.supp, G0.nu, G0.nv, 0.0});
            } else if (!kernel_choices(im.o, single, std::size_t(nvis), wlo, whi, choices, err)) {
                return 1;
            }
            const int nk = std::max(1, std::min(kernels, int(choices.size())));
            best.ms = 1e300;
            for (int c = 0; c < nk; ++c) {
                Options o2 = base;
                o2.kernel_index = choices[c].index;
                o2.nu = choices[c].nu;
                o2.nv = choices[c].nv;
                double t;
                if (!time_of(o2, 0, t)) continue;
                if (c == 0) best.model_ms = t;
                if (t < best.ms) {
                    best.ms = t;
                    best.kidx = o2.kernel_index;
                    best.nu = o2.nu;
                    best.nv = o2.nv;
                    best.tile = 0;
                }
            }
            if (best.kidx < 0) {
                err = "no kernel row could be timed";
                return 1;
            }
            // tile sides for the chosen row, each distinct tiling once
            Options ob = base;
            ob.kernel_index = best.kidx;
            ob.nu = best.nu;
            ob.nv = best.nv;
            Geometry Gb;
            if (!make_geometry(ob, single, std::size_t(nvis), wlo, whi, Gb, err)) return 1;
            std::vector<std::pair<int, int>> seen;
            {
                Geo gd{};
                if (fill_geo<T>(Gb, gd, 1, false, err, 0)) seen.emplace_back(gd.Tu, gd.Tv);
            }
            for (int tile : {16, 24, 40, 48}) {
                Geo gt{};
                std::string e3;
                if (!fill_geo<T>(Gb, gt, 1, false, e3, tile)) continue;
                if (std::find(seen.begin(), seen.end(), std::make_pair(gt.Tu, gt.Tv)) != seen.end()) continue;
                seen.emplace_back(gt.Tu, gt.Tv);
                double t;
                if (time_of(ob, tile, t) && t < best.ms) {
                    best.ms = t;
                    best.tile = tile;
                }
            }
            tl.lock();
            table[key] = best;
            if (!file.empty()) {
                std::ofstream f(file, std::ios::app);
                char line[256];
                std::snprintf(line, sizeof(line), " %lld %zu %zu %d %.6g %.6g\n", best.kidx, best.nu, best.nv,
                              best.tile, best.ms, best.model_ms);
                f << key << line;
            }
        }
        im.o.kernel_index = best.kidx;
        im.o.nu = best.nu;
        im.o.nv = best.nv;
        im.tile = best.tile;
        if (ms) *ms = best.ms;
        if (default_ms) *default_ms = best.model_ms;
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        if (e.status != 2) (void)hipStreamSynchronize(st);
        return e.status;
    }
}

// ------------------------------------------------------------ sub-band batches

namespace {

// Workspaces of a batch: one per image of a launch group, and one for the
// scans and the job table; kept per device between calls.
struct BatchPool {
    std::mutex mu;
    Workspace control;
    std::vector<std::unique_ptr<Workspace>> images;
    Workspace& image(std::size_t k) {
        while (images.size() <= k) images.emplace_back(new Workspace());
        return *images[k];
    }
};

std::mutex g_pool_mu;
std::map<int, std::unique_ptr<BatchPool>> g_pool_all;

BatchPool& batch_pool(int device) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto& p = g_pool_all[device];
    if (!p) p.reset(new BatchPool());
    return *p;
}

// Device bytes one image's vis2dirty holds: records and their bucket
// tables, and the grid, the half-transformed planes and the projections.
template <class T>
std::size_t image_bytes(const Geo& g, long long nvis) {
    const long long P = g.wgrid ? g.nplanes : 1;
    const long long nkeys = (long long)g.ntu * g.ntv * 4 * g.nb;
    return std::size_t(nvis) * (sizeof(Rec<T>) + 8) + std::size_t(g.n) * 4 + std::size_t(nkeys) * 16 +
           std::size_t(P) * ((std::size_t)g.nu * g.nv * sizeof(cx<T>) + (std::size_t)g.nu * g.ny * sizeof(cx<T>) +
                             (std::size_t)g.nx * g.ny * sizeof(T));
}

}  // namespace

template <class T>
int device_batch_vis2dirty(BatchImage* imgs, int n, std::size_t budget, hipStream_t st, std::string& err) {
    try {
        if (n < 0 || (n > 0 && !imgs)) {
            err = "null array";
            return 1;
        }
        BatchPool& pool = batch_pool(device_of());
        std::lock_guard<std::mutex> lk(pool.mu);
        HostKeep keep;
        StageTimer timer(st);
        struct Plan {
            Geometry G;
            Geo g;
            DevIn din;
            long long nvis = 0;
            std::size_t bytes = 0;
        };
        std::vector<Plan> plans(n);
        // scans and plans of every image
        for (int i = 0; i < n; ++i) {
            BatchImage& im = imgs[i];
            Plan& pl = plans[i];
            if (!im.dirty) {
                err = "image " + std::to_string(i) + ": null array";
                return 1;
            }
            if (!common_checks(im.o, im.in, im.freq, err)) {
                err = "image " + std::to_string(i) + ": " + err;
                return 1;
            }
            std::fill(im.info, im.info + kNInfo, 0.0);
            pl.din = DevIn{im.in.uvw, nullptr, im.in.vis, im.in.wgt, im.in.wgt2, im.in.mask, im.in.vis_kind,
                           im.in.wgt_kind};
            pl.g = base_geo(im.in.nrow, im.in.nchan, im.o);
            pl.din.foc = upload_foc(pool.control, im.freq, im.in.nchan, st, keep);
            double wlo = 0, whi = 0;
            pl.nvis = (long long)run_scan<T, true>(pool.control, pl.din, pl.g, st, wlo, whi);
            if (pl.nvis == 0) continue;
            if (!make_geometry(im.o, sizeof(T) == 4, std::size_t(pl.nvis), wlo, whi, pl.G, err) ||
                !fill_geo<T>(pl.G, pl.g, 1, false, err, im.tile)) {
                err = "image " + std::to_string(i) + ": " + err;
                return 1;
            }
            std::copy(pl.G.info, pl.G.info + kNInfo, im.info);
            pl.bytes = image_bytes<T>(pl.g, pl.nvis);
        }
        timer.mark("scan");
        // images without visibilities: zero images
        std::vector<int> order;
        for (int i = 0; i < n; ++i) {
            if (plans[i].nvis == 0)
                fill_zero(imgs[i].dirty, sizeof(T) * imgs[i].o.nx * imgs[i].o.ny, st);
            else
                order.push_back(i);
        }
        // largest first, so that a later group's image never needs more of a
        // pooled workspace than the first group's image in its place
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return plans[a].bytes > plans[b].bytes; });
        std::size_t at = 0;
        while (at < order.size()) {
            std::size_t end = at, used = 0;
            while (end < order.size() && (end == at || used + plans[order[end]].bytes <= budget))
                used += plans[order[end++]].bytes;
            std::vector<GridPrep<T>> preps(end - at);
            std::vector<GridJob<T>> jobs, mjobs;  // LDS gridding, matrix-core gridding
            std::size_t lds = 0;
            long long total = 0, mtotal = 0;
            for (std::size_t k = at; k < end; ++k) {
                const int i = order[k];
                Plan& pl = plans[i];
                Workspace& ws = pool.image(k - at);
                pl.din.foc = upload_foc(ws, imgs[i].freq, imgs[i].in.nchan, st, keep);
                Rec<T>* recs;
                std::uint32_t* off;
                sort_records<T, 0, true>(ws, pl.din, pl.g, pl.nvis, st, &recs, &off);
                GridPrep<T>& gp = preps[k - at];
                grid_prepare<T>(ws, pl.G, pl.g, off, 1, st, keep, timer, nullptr, gp);
                if (gp.ub == 0) continue;
                long long& tot = pl.g.mfma ? mtotal : total;
                if (tot + gp.ub >= (1LL << 31)) throw DevError{"too many gridding items in one batch"};
                (pl.g.mfma ? mjobs : jobs)
                    .push_back(GridJob<T>{pl.g, recs, off, gp.tb.coeff, pl.din, gp.G_planes, gp.ngroups, gp.TvP,
                                          kGridChunk, gp.item_off, gp.tb.order, gp.grid, std::uint32_t(tot)});
                tot += gp.ub;
                if (!pl.g.mfma) lds = std::max(lds, gp.lds);
            }
            timer.mark("sort");
            // the job tables to the device: one launch per gridding kernel
            auto upload = [&](const std::vector<GridJob<T>>& js, int slot) {
                keep.b.emplace_back(reinterpret_cast<const char*>(js.data()),
                                    reinterpret_cast<const char*>(js.data() + js.size()));
                const std::size_t jb = sizeof(GridJob<T>) * js.size();
                void* dj = pool.control.get(slot, jb);
                HIPCHK(hipMemcpyAsync(dj, keep.b.back().data(), jb, hipMemcpyHostToDevice, st));
                return static_cast<const GridJob<T>*>(dj);
            };
            if (!jobs.empty()) {
                const GridJob<T>* dj = upload(jobs, kSlotItems);
                lds_limit((const void*)k_grid_jobs<T, 1, false>, lds);
                k_grid_jobs<T, 1, false><<<int(total), grid_threads<T>(), lds, st>>>(dj, int(jobs.size()));
                HIPCHK(hipGetLastError());
            }
            if constexpr (sizeof(T) == 4) {
                if (!mjobs.empty()) {
                    const GridJob<T>* dj = upload(mjobs, kSlotItemOff);
                    k_grid_mfma_jobs<false><<<int(mtotal), 64, 0, st>>>(dj, int(mjobs.size()));
                    HIPCHK(hipGetLastError());
                }
            }
            timer.mark("grid");
            for (std::size_t k = at; k < end; ++k) {
                const int i = order[k];
                grid_finish<T>(pool.image(k - at), plans[i].g, preps[k - at], 1, static_cast<T*>(imgs[i].dirty),
                               nullptr, st, timer);
            }
            HIPCHK(hipStreamSynchronize(st));
            at = end;
        }
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        if (e.status != 2) (void)hipStreamSynchronize(st);
        return e.status;
    }
}

// dirty2vis from sorted records: the corrected image, the transforms and the
// degridding into `cvis` (zeroed by the caller; records' values multiply).
template <class T>
void degrid_pipeline(Workspace& ws, const Geometry& G, const Geo& g, const Rec<T>* recs,
                     const std::uint32_t* off, const T* dirty, const DegridOut<T>& outs, hipStream_t st,
                     HostKeep& keep, StageTimer& timer, double* total = nullptr, ItemCache* cache = nullptr) {
    Tables<T> tb;
    make_tables<T>(ws, G, g, st, keep, tb, 1);
    LongTabs<T> lt;
    make_long<T>(ws, g, st, keep, lt);
    timer.mark("tables");
    const int P = g.wgrid ? g.nplanes : 1;
    const long long plane = (long long)g.nu * g.nv;
    const long long npix = (long long)g.nx * g.ny;
    T* img = static_cast<T*>(ws.get(kSlotImg, sizeof(T) * npix));
    if (g.cm)
        k_correct_cm<T><<<dim3((g.nx + 31) / 32, (g.ny + 31) / 32), 256, 0, st>>>(dirty, img, g, tb.corr,
                                                                                  tb.cfu, tb.cfv);
    else
        k_correct<T><<<(npix + 255) / 256, 256, 0, st>>>(dirty, img, g, tb.corr, tb.cfu, tb.cfv);
    HIPCHK(hipGetLastError());
    cx<T>* tmp = static_cast<cx<T>*>(ws.get(kSlotTmp, sizeof(cx<T>) * (long long)P * g.nu * g.ny));
    cx<T>* grid = static_cast<cx<T>*>(ws.get(kSlotGrid, sizeof(cx<T>) * plane * P));
    // the degridding items (before the transforms: the forward row
    // transform writes only the rows they read)
    const int nt = g.ntu * g.ntv;
    // a prepared plan's degridding items stay in slots of their own
    std::uint32_t* items = static_cast<std::uint32_t*>(
        ws.get(cache ? kSlotItems2 : kSlotItems, sizeof(std::uint32_t) * (nt + 1)));
    std::uint32_t* item_off = static_cast<std::uint32_t*>(
        ws.get(cache ? kSlotItemOff2 : kSlotItemOff, sizeof(std::uint32_t) * (nt + 1)));
    const std::size_t dl = degrid_lds<T>(g.Tu, g.Tv, g.supp, P);
    const bool staged = dl <= degrid_lds_max<T>();
    const std::uint32_t rc = staged ? degrid_chunk_staged<T>() : kDegridChunk;
    long long ub;
    if (cache && cache->valid) {
        ub = cache->ub;
    } else {
        k_degrid_items<<<(nt + 1 + 255) / 256, 256, 0, st>>>(g, off, tb.order, rc, items);
        HIPCHK(hipGetLastError());
        exclusive_scan(ws, items, item_off, nt + 1, st);
        ub = read_u32(ws, item_off + nt, st);
        if (cache) {
            cache->ub = ub;
            cache->valid = true;
        }
    }
    if (ub >= (1LL << 31)) throw DevError{"too many degridding items"};
    std::uint8_t* need = nullptr;
    if (!g.lv1 && grid_occ(sizeof(cx<T>) * std::size_t(plane) * P)) {
        const std::size_t nb4 = (std::size_t(g.nu) + 3) / 4 * 4;
        need = static_cast<std::uint8_t*>(ws.get(kSlotRowNeed, nb4));
        fill_zero(need, nb4, st);
        k_row_need<<<(nt + 255) / 256, 256, 0, st>>>(g, item_off, tb.order, need);
        HIPCHK(hipGetLastError());
    }
    const int C = col_block(g.nu);
    // when both transforms stay in LDS, the u transform writes whole
    // columns into the grid (free until the v transform fills it) and a
    // tiled transpose lays them out as rows (RANG_FFT_TCM=0: strided writes
    // straight into tmp)
    const bool tcm = !g.lu1 && !g.lv1 && fft_tcm(sizeof(cx<T>) * std::size_t(P) * g.nu * g.ny);
    if (g.lu1) {
        k_place_cols<T><<<blocks_of((long long)P * g.nu, 256), 256, 0, st>>>(img, tmp, g, tb.screen);
        HIPCHK(hipGetLastError());
        long_cols<T, 0>(tmp, P, g.ny, g.lu1, g.lu2, lt.u1, lt.u2, lt.tu1, lt.tu2, lt.tun, -1, nullptr, g,
                        tb.screen, st);
    } else {
        fft_dispatch<T>(g.nu * C, [&](auto nt, auto mt) {
            constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
            const std::size_t lds_u = sizeof(cx<T>) * g.nu * C;
            const unsigned nb = unsigned(blocks_of((long long)(g.ny + C - 1) / C * P, NT));
            if (tcm) {
                // the columns into the (not yet used) grid, then transposed into tmp
                set_lds(k_cols_fwd<T, NT, MT, true>, lds_u);
                k_cols_fwd<T, NT, MT, true><<<nb, NT, lds_u, st>>>(img, grid, tb.du, tb.twu, g, tb.screen, C);
            } else {
                set_lds(k_cols_fwd<T, NT, MT>, lds_u);
                k_cols_fwd<T, NT, MT><<<nb, NT, lds_u, st>>>(img, tmp, tb.du, tb.twu, g, tb.screen, C);
            }
        });
        HIPCHK(hipGetLastError());
        if (tcm) {
            timer.mark("fft_u");
            transpose_planes<T>(grid, tmp, P, g.ny, g.nu, st);
        }
    }
    timer.mark(tcm ? "transpose" : "fft_u");
    if (g.lv1) {
        const long long rows = (long long)P * g.nu;
        k_place_rows<T><<<blocks_of(rows, 256), 256, 0, st>>>(tmp, grid, g);
        HIPCHK(hipGetLastError());
        long_rows<T, 0>(grid, rows, g.lv1, g.lv2, lt.v1, lt.v2, lt.tv1, lt.tv2, lt.tvn, -1, nullptr, g.ny, st);
    } else {
        fft_dispatch<T>(g.nv, [&](auto nt, auto mt) {
            constexpr int NT = decltype(nt)::value, MT = decltype(mt)::value;
            const std::size_t lds_v = sizeof(cx<T>) * g.nv;
            const unsigned nb = unsigned(blocks_of((long long)P * g.nu, NT));
            set_lds(k_rows_fwd<T, NT, MT>, lds_v);
            k_rows_fwd<T, NT, MT><<<nb, NT, lds_v, st>>>(tmp, grid, tb.dv, tb.twv, g, need);
        });
        HIPCHK(hipGetLastError());
    }
    timer.mark("fft_v");
    DegridOut<T> o = outs;
    if (o.mode == 2) o.part = static_cast<double*>(ws.get(kSlotRed, sizeof(double) * std::max<long long>(ub, 1)));
    if (ub == 0) {
        // nothing to degrid
    } else if (staged)
        launch_degrid<T, true>(g.supp, int(ub), dl, st, g, recs, off, tb.coeff, grid, rc,
                               item_off, tb.order, o);
    else
        launch_degrid<T, false>(g.supp, int(ub), 0, st, g, recs, off, tb.coeff, grid, kDegridChunk,
                                item_off, tb.order, o);
    if (o.mode == 2 && total) {
        k_sum_parts<<<1, 256, 0, st>>>(o.part, ub, total);
        HIPCHK(hipGetLastError());
    }
    timer.mark("degrid");
}

template <class T>
int device_dirty2vis(const Options& o, const DevInputs& in, const double* freq_host,
                     const T* dirty, T* vis, double* info, hipStream_t st, std::string& err) {
    try {
        if (!dirty || !vis) {
            err = "null array";
            return 1;
        }
        if (!common_checks(o, in, freq_host, err)) return 1;
        Workspace& ws = workspace(device_of());
        std::lock_guard<std::mutex> lk(ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        DevIn din{in.uvw, nullptr, nullptr, in.wgt, nullptr, in.mask, kVisUnit, in.wgt_kind};
        Geo g = base_geo(in.nrow, in.nchan, o);
        din.foc = upload_foc(ws, freq_host, in.nchan, st, keep);
        double wlo = 0, whi = 0;
        const unsigned long long nvis = run_scan<T, false>(ws, din, g, st, wlo, whi);
        timer.mark("scan");
        fill_zero(vis, sizeof(T) * 2 * in.nrow * in.nchan, st);
        std::fill(info, info + kNInfo, 0.0);
        if (nvis == 0) {
            HIPCHK(hipStreamSynchronize(st));
            return 0;
        }
        Geometry G;
        if (!make_geometry(o, sizeof(T) == 4, std::size_t(nvis), wlo, whi, G, err)) return 1;
        if (!fill_geo<T>(G, g, 1, true, err)) return 1;
        Rec<T>* recs;
        std::uint32_t* off;
        sort_records<T, 2, false>(ws, din, g, (long long)nvis, st, &recs, &off);
        sort_degrid_chunks<T>(ws, g, recs, off, st);
        timer.mark("sort");
        DegridOut<T> o;
        o.vis = reinterpret_cast<cx<T>*>(vis);
        degrid_pipeline<T>(ws, G, g, recs, off, dirty, o, st, keep, timer);
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        std::copy(G.info, G.info + kNInfo, info);
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}


// ------------------------------------------------------------ plans

// The plan a call would choose, without gridding (core/src/wgridder.rs:
// 3525-3549: plan; 3551-3586: gridding_footprint): the scan on the device,
// then the kernel and grid choice. `footprint` (nout > 0) receives the
// device bytes the gridding pass of this implementation would hold (grids,
// column transforms and plane terms, sorted records).
template <class T>
int device_plan(const Options& o, const DevInputs& in, const double* freq_host, bool use_vis, int nout,
                double* info, std::size_t* footprint, hipStream_t st, std::string& err) {
    try {
        if (!common_checks(o, in, freq_host, err)) return 1;
        Workspace& ws = workspace(device_of());
        std::lock_guard<std::mutex> lk(ws.mu);
        HostKeep keep;
        DevIn din{in.uvw, nullptr, in.vis, in.wgt, in.wgt2, in.mask, in.vis_kind, in.wgt_kind};
        Geo g = base_geo(in.nrow, in.nchan, o);
        din.foc = upload_foc(ws, freq_host, in.nchan, st, keep);
        double wlo = 0, whi = 0;
        const unsigned long long nvis = use_vis ? run_scan<T, true>(ws, din, g, st, wlo, whi)
                                                : run_scan<T, false>(ws, din, g, st, wlo, whi);
        std::fill(info, info + kNInfo, 0.0);
        if (footprint) *footprint = 0;
        if (nvis == 0) return 0;
        Geometry G;
        if (!make_geometry(o, sizeof(T) == 4, std::size_t(nvis), wlo, whi, G, err)) return 1;
        std::copy(G.info, G.info + kNInfo, info);
        if (footprint && nout > 0) {
            const std::size_t P = G.do_wgridding ? std::max<std::size_t>(G.nplanes, 1) : 1;
            const std::size_t plane = G.nu * G.nv * sizeof(cx<T>);
            *footprint = P * nout * plane + P * G.nu * o.ny * sizeof(cx<T>) + P * o.nx * o.ny * sizeof(T) +
                         std::size_t(nvis) * sizeof(Rec<T>);
        }
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

template int device_plan<float>(const Options&, const DevInputs&, const double*, bool, int, double*,
                                std::size_t*, hipStream_t, std::string&);
template int device_plan<double>(const Options&, const DevInputs&, const double*, bool, int, double*,
                                 std::size_t*, hipStream_t, std::string&);

// ------------------------------------------------------------ prepared plans

// The plan and sorted layouts of a fixed set of visibilities kept between
// calls (core/src/wgridder.rs:3165-3499): the scan, the kernel choice and
// both sorts are done once; the active set is the nonzero weights. Owns a
// work space of its own and device copies of uvw and the weights.
template <class T>
struct PreparedImpl final : PreparedBase {
    int device = 0;
    Workspace ws;
    Options o;
    Geometry G;
    Geo gg{}, gd{};
    long long n = 0, nvis = 0;
    std::size_t nrow = 0, nchan = 0;
    std::vector<double> freq;
    double* uvw = nullptr;
    void* wgt = nullptr;
    std::size_t wbytes = 0;
    int wkind = kWgtNone;
    Rec<T>* rg = nullptr;
    std::uint32_t* og = nullptr;
    Rec<T>* rd = nullptr;
    std::uint32_t* od = nullptr;
    double* foc = nullptr;
    // for each degridding record, the position of its visibility among the
    // gridding records: the degridder writes predictions and residuals
    // straight into values held in the gridding order
    std::uint32_t* pos01 = nullptr;
    ItemCache ic_grid, ic_degrid;
    // Graphs of the plan's repeated calls, keyed by the call and its buffers
    // (a graph holds their addresses): the first call of a key runs eagerly
    // and fills every cache, the second is captured on the plan's own stream,
    // later ones replay it. The caller's stream and the plan's are ordered
    // by events.
    hipStream_t gstream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::map<std::vector<std::uintptr_t>, hipGraphExec_t> graphs;
    std::map<std::vector<std::uintptr_t>, int> seen;
    unsigned graph_gen = 0;

    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
        seen.clear();
    }

    ~PreparedImpl() override {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != device) (void)hipSetDevice(device);
        drop_graphs();
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (gstream) (void)hipStreamDestroy(gstream);
        for (int k = 0; k < kSlotCount; ++k)
            if (ws.ptr[k]) (void)hipFree(ws.ptr[k]);
        if (ws.pinned) (void)hipHostFree(ws.pinned);
        if (uvw) (void)hipFree(uvw);
        if (wgt) (void)hipFree(wgt);
        if (pos01) (void)hipFree(pos01);
        if (cur != device) (void)hipSetDevice(cur);
    }
    bool single() const override { return sizeof(T) == 4; }
    std::size_t active() const override { return std::size_t(nvis); }
    std::size_t image_nx() const override { return std::size_t(o.nx); }
    std::size_t image_ny() const override { return std::size_t(o.ny); }
    std::size_t memory() const override {
        std::size_t b = sizeof(double) * 3 * nrow + wbytes + sizeof(std::uint32_t) * std::size_t(nvis);
        for (int k = 0; k < kSlotCount; ++k) b += ws.size[k];
        return b;
    }
    void release_grids() override {
        std::lock_guard<std::mutex> lk(ws.mu);
        drop_graphs();  // they hold the addresses of the buffers freed here
        for (int k : {int(kSlotGrid), int(kSlotTmp), int(kSlotProj), int(kSlotImg), int(kSlotVbuf), int(kSlotSolve)}) {
            if (ws.ptr[k]) (void)hipFree(ws.ptr[k]);
            ws.ptr[k] = nullptr;
            ws.size[k] = 0;
        }
    }
    DevIn din(const void* vis, int vis_kind) const {
        return DevIn{uvw, foc, vis, wgt, nullptr, nullptr, vis_kind, wkind};
    }
};

namespace {

template <class T>
double sum_partials(Workspace& ws, const double* part, int nblk, hipStream_t st) {
    double* h = static_cast<double*>(ws.host(sizeof(double) * nblk));
    HIPCHK(hipMemcpyAsync(h, part, sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double q = 0.0;
    for (int b = 0; b < nblk; ++b) q += h[b];
    return q;
}

int red_blocks(long long n) {
    return int(std::max<long long>(1, std::min<long long>(kRedBlocks, (n + 255) / 256)));
}

// Graph replays of prepared calls; RANG_GRAPHS=0 turns them off (every call
// then runs eagerly), and stage timing does too (its events are not graph
// work).
bool graphs_on() {
    static const bool on = [] {
        const char* e = std::getenv("RANG_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on && !g_timing;
}

// Runs body(stream) for the call `key`: eagerly on `st` the first time, then
// captured once into a graph on the plan's stream and replayed, ordered
// after the work already queued on `st` and before what follows on it. The
// body must make no host round trip and allocate nothing once warm (the
// first, eager, call sizes every buffer and fills every cache).
template <class T, class F>
void run_graph(PreparedImpl<T>& p, const std::vector<std::uintptr_t>& key, hipStream_t st, F&& body) {
    if (!graphs_on()) {
        body(st);
        return;
    }
    if (p.ws.gen != p.graph_gen) {
        // a buffer a graph uses was reallocated since it was captured
        p.drop_graphs();
        p.graph_gen = p.ws.gen;
    }
    auto it = p.graphs.find(key);
    if (it == p.graphs.end() && p.seen[key]++ == 0) {
        body(st);
        return;
    }
    if (!p.gstream) {
        HIPCHK(hipStreamCreateWithFlags(&p.gstream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&p.ev_in, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&p.ev_out, hipEventDisableTiming));
    }
    HIPCHK(hipEventRecord(p.ev_in, st));
    HIPCHK(hipStreamWaitEvent(p.gstream, p.ev_in, 0));
    if (it == p.graphs.end()) {
        HIPCHK(hipStreamBeginCapture(p.gstream, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            body(p.gstream);
        } catch (...) {
            (void)hipStreamEndCapture(p.gstream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(p.gstream, &g));
        if (p.ws.gen != p.graph_gen) {
            // the capture allocated (not warm after all): run it eagerly
            (void)hipGraphDestroy(g);
            p.drop_graphs();
            p.graph_gen = p.ws.gen;
            body(p.gstream);
            HIPCHK(hipEventRecord(p.ev_out, p.gstream));
            HIPCHK(hipStreamWaitEvent(st, p.ev_out, 0));
            return;
        }
        hipGraphExec_t e = nullptr;
        const hipError_t ie = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ie);
        it = p.graphs.emplace(key, e).first;
    }
    HIPCHK(hipGraphLaunch(it->second, p.gstream));
    HIPCHK(hipEventRecord(p.ev_out, p.gstream));
    HIPCHK(hipStreamWaitEvent(st, p.ev_out, 0));
}

std::uintptr_t up(const void* q) { return reinterpret_cast<std::uintptr_t>(q); }

// vis = A dirty (no weights) at the active visibilities, zero elsewhere.
template <class T>
void prepared_predict(PreparedImpl<T>& p, const T* image, cx<T>* vis, hipStream_t st, HostKeep& keep,
                      StageTimer& timer) {
    fill_zero(vis, sizeof(cx<T>) * p.n, st);
    if (p.nvis == 0) return;
    DegridOut<T> o;
    o.vis = vis;
    degrid_pipeline<T>(p.ws, p.G, p.gd, p.rd, p.od, image, o, st, keep, timer, nullptr, &p.ic_degrid);
}

// image = A^H W vis over the active visibilities.
template <class T>
void prepared_image(PreparedImpl<T>& p, const void* vis, int vis_kind, T* image, hipStream_t st,
                    HostKeep& keep, StageTimer& timer) {
    if (p.nvis == 0) {
        fill_zero(image, sizeof(T) * p.o.nx * p.o.ny, st);
        return;
    }
    grid_pipeline<T>(p.ws, p.G, p.gg, p.din(vis, vis_kind), p.rg, p.og, 1, true, image, nullptr, st,
                     keep, timer, &p.ic_grid);
}

constexpr int vis_kind_of(float) { return kVisC64; }
constexpr int vis_kind_of(double) { return kVisC128; }

}  // namespace

template <class T>
PreparedBase* prepared_new(const Options& o, const DevInputs& in, const double* freq_host,
                           hipStream_t st, int& status, std::string& err) {
    status = 1;
    if (!common_checks(o, in, freq_host, err)) return nullptr;
    if (in.wgt_kind != kWgtNone && in.wgt_kind != kWgtF32 && in.wgt_kind != kWgtF64) {
        err = "prepared weights must be float32 or float64";
        return nullptr;
    }
    status = 2;
    std::unique_ptr<PreparedImpl<T>> p(new PreparedImpl<T>());
    try {
        p->device = device_of();
        p->o = o;
        p->nrow = in.nrow;
        p->nchan = in.nchan;
        p->n = (long long)(in.nrow * in.nchan);
        p->freq.assign(freq_host, freq_host + in.nchan);
        HIPCHK(hipMalloc(&p->uvw, sizeof(double) * 3 * std::max<std::size_t>(in.nrow, 1)));
        HIPCHK(hipMemcpyAsync(p->uvw, in.uvw, sizeof(double) * 3 * in.nrow, hipMemcpyDeviceToDevice, st));
        p->wkind = in.wgt_kind;
        if (in.wgt_kind != kWgtNone) {
            p->wbytes = std::size_t(p->n) * (in.wgt_kind == kWgtF32 ? 4 : 8);
            HIPCHK(hipMalloc(&p->wgt, std::max<std::size_t>(p->wbytes, 8)));
            HIPCHK(hipMemcpyAsync(p->wgt, in.wgt, p->wbytes, hipMemcpyDeviceToDevice, st));
        }
        HostKeep keep;
        p->foc = upload_foc(p->ws, freq_host, in.nchan, st, keep);
        DevIn d = p->din(nullptr, kVisUnit);
        Geo g = base_geo(in.nrow, in.nchan, o);
        double wlo = 0, whi = 0;
        p->nvis = (long long)run_scan<T, false>(p->ws, d, g, st, wlo, whi);
        if (p->nvis > 0) {
            if (!make_geometry(o, sizeof(T) == 4, std::size_t(p->nvis), wlo, whi, p->G, err)) {
                status = 1;
                return nullptr;
            }
            p->gg = g;
            p->gd = g;
            if (!fill_geo<T>(p->G, p->gg, 1, false, err) || !fill_geo<T>(p->G, p->gd, 1, true, err)) {
                status = 1;
                return nullptr;
            }
            // gridding: unit values with the weights; degridding: the shift
            // phase alone (the prediction carries no weights)
            sort_records<T, 1, false>(p->ws, d, p->gg, p->nvis, st, &p->rg, &p->og, 0);
            DevIn unweighted = d;
            unweighted.wgt = nullptr;
            unweighted.wgt_kind = kWgtNone;
            sort_records<T, 2, false>(p->ws, d, p->gd, p->nvis, st, &p->rd, &p->od, 1, &unweighted);
            sort_degrid_chunks<T>(p->ws, p->gd, p->rd, p->od, st);
            // position of each visibility among the gridding records, through
            // a visibility-indexed map held only while the plan is made
            std::uint32_t* map = static_cast<std::uint32_t*>(p->ws.get(kSlotKeys, sizeof(std::uint32_t) * p->n));
            HIPCHK(hipMalloc(&p->pos01, sizeof(std::uint32_t) * std::size_t(p->nvis)));
            const unsigned nb = blocks_of((p->nvis + 255) / 256, 256);
            k_index_map<T><<<nb, 256, 0, st>>>(p->rg, p->nvis, map);
            HIPCHK(hipGetLastError());
            k_pos_of<T><<<nb, 256, 0, st>>>(p->rd, p->nvis, map, p->pos01);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(st));
        // the sort's per-visibility keys, counts and cursors are not needed again
        for (int k : {int(kSlotKeys), int(kSlotKeys2), int(kSlotCounts), int(kSlotCounts2), int(kSlotCursor),
                      int(kSlotCursor2)}) {
            if (p->ws.ptr[k]) HIPCHK(hipFree(p->ws.ptr[k]));
            p->ws.ptr[k] = nullptr;
            p->ws.size[k] = 0;
        }
        HIPCHK(hipStreamSynchronize(st));
        status = 0;
        return p.release();
    } catch (const DevError& e) {
        err = e.msg;
        return nullptr;
    }
}

template <class T>
int prepared_vis2dirty(PreparedBase* base, const void* vis, int vis_kind, T* dirty, double* info,
                       hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        run_graph(p, {1, up(vis), std::uintptr_t(vis_kind), up(dirty)}, st,
                  [&](hipStream_t s) { prepared_image<T>(p, vis, vis_kind, dirty, s, keep, timer); });
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        std::fill(info, info + kNInfo, 0.0);
        if (p.nvis) std::copy(p.G.info, p.G.info + kNInfo, info);
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

template <class T>
int prepared_dirty2vis(PreparedBase* base, const T* dirty, bool weighted, T* vis, hipStream_t st,
                       std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        cx<T>* cv = reinterpret_cast<cx<T>*>(vis);
        run_graph(p, {2, up(dirty), std::uintptr_t(weighted), up(vis)}, st, [&](hipStream_t s) {
            prepared_predict<T>(p, dirty, cv, s, keep, timer);
            if (weighted && p.wkind != kWgtNone && p.n > 0)
                k_weigh<T><<<red_blocks(p.n), 256, 0, s>>>(p.din(nullptr, kVisUnit), cv, p.n);
            HIPCHK(hipGetLastError());
        });
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

namespace {

// The values one per active visibility in the gridding records' order, which
// the fused residual and the Gram application pass from the degridder to the
// gridder instead of a visibility-sized buffer (wgridder.rs:3300-3383).
template <class T>
cx<T>* stored_of(PreparedImpl<T>& p) {
    return static_cast<cx<T>*>(p.ws.get(kSlotVbuf, sizeof(cx<T>) * std::max<long long>(p.nvis, 1)));
}

// image = A^H W v for v held in the gridding order.
template <class T>
void image_of_stored(PreparedImpl<T>& p, const cx<T>* stored, T* image, hipStream_t st, HostKeep& keep,
                     StageTimer& timer) {
    DevIn d = p.din(nullptr, kVisUnit);
    d.stored = stored;
    grid_pipeline<T>(p.ws, p.G, p.gg, d, p.rg, p.og, 1, false, image, nullptr, st, keep, timer, &p.ic_grid);
}

// out = A^H W A x: the degridder writes A x into the gridding order, the
// gridder reads it from there.
template <class T>
void gram_body(PreparedImpl<T>& p, const T* x, T* out, hipStream_t st, HostKeep& keep, StageTimer& timer) {
    if (p.nvis == 0) {
        fill_zero(out, sizeof(T) * p.o.nx * p.o.ny, st);
        return;
    }
    cx<T>* stored = stored_of(p);
    DegridOut<T> o;
    o.mode = 1;
    o.pos = p.pos01;
    o.stored = stored;
    degrid_pipeline<T>(p.ws, p.G, p.gd, p.rd, p.od, x, o, st, keep, timer, nullptr, &p.ic_degrid);
    image_of_stored<T>(p, stored, out, st, keep, timer);
}

template <class T>
void gram_apply(PreparedImpl<T>& p, const T* x, T* out, hipStream_t st, HostKeep& keep, StageTimer& timer) {
    run_graph(p, {3, up(x), up(out)}, st, [&](hipStream_t s) { gram_body<T>(p, x, out, s, keep, timer); });
}

// Prepared::residual / residual_sorted (wgridder.rs:3230-3255, 3300-3383):
// the residual d - A model (d alone without a model) of the data `vis`, dense
// or (sorted) in the gridding order, formed in the degridder's epilogue into
// the gridding order; its misfit sum w |.|^2 with the prepared weights left
// on the device; then its image.
template <class T>
int residual_impl(PreparedBase* base, const void* vis, int vis_kind, bool sorted, const T* model, T* dirty,
                  double* misfit, hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        if (p.nvis == 0) {
            fill_zero(dirty, sizeof(T) * p.o.nx * p.o.ny, st);
            HIPCHK(hipStreamSynchronize(st));
            *misfit = 0.0;
            return 0;
        }
        double* total = static_cast<double*>(p.ws.get(kSlotScalars, sizeof(double) * 4));
        const std::vector<std::uintptr_t> key{4, up(vis), std::uintptr_t(vis_kind), std::uintptr_t(sorted), up(model),
                                              up(dirty)};
        run_graph(p, key, st, [&](hipStream_t s) {
            cx<T>* stored = stored_of(p);
            const DevIn data = p.din(vis, vis_kind);
            if (model) {
                DegridOut<T> o;
                o.mode = 2;
                o.pos = p.pos01;
                o.stored = stored;
                o.data = data;
                o.sorted = sorted ? 1 : 0;
                degrid_pipeline<T>(p.ws, p.G, p.gd, p.rd, p.od, model, o, s, keep, timer, total, &p.ic_degrid);
            } else {
                const int nblk = red_blocks(p.nvis);
                double* part = static_cast<double*>(p.ws.get(kSlotRed, sizeof(double) * kRedBlocks));
                k_gather_resid<T><<<nblk, 256, 0, s>>>(p.rg, p.nvis, data, sorted ? 1 : 0, stored, part);
                HIPCHK(hipGetLastError());
                k_sum_parts<<<1, 256, 0, s>>>(part, nblk, total);
                HIPCHK(hipGetLastError());
            }
            timer.mark("residual");
            image_of_stored<T>(p, stored, dirty, s, keep, timer);
        });
        double* h = static_cast<double*>(p.ws.host(sizeof(double)));
        HIPCHK(hipMemcpyAsync(h, total, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        *misfit = *h;
        timer.finish();
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

}  // namespace

template <class T>
int prepared_residual(PreparedBase* base, const void* vis, int vis_kind, const T* model, T* dirty,
                      double* misfit, hipStream_t st, std::string& err) {
    return residual_impl<T>(base, vis, vis_kind, false, model, dirty, misfit, st, err);
}

template <class T>
int prepared_residual_sorted(PreparedBase* base, const void* vis, int vis_kind, const T* model, T* dirty,
                             double* misfit, hipStream_t st, std::string& err) {
    return residual_impl<T>(base, vis, vis_kind, true, model, dirty, misfit, st, err);
}

// Prepared::sorted (wgridder.rs:3269-3286): a dense nrow x nchan array at the
// active visibilities, in the gridding order of the residual's values.
template <class T>
int prepared_sorted(PreparedBase* base, const void* dense, int bytes, void* out, hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        if (p.nvis > 0) {
            k_gather_sorted<T><<<blocks_of((p.nvis + 255) / 256, 256), 256, 0, st>>>(
                p.rg, p.nvis, static_cast<const unsigned char*>(dense), bytes, static_cast<unsigned char*>(out));
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

namespace {

template <class T>
double dot(PreparedImpl<T>& p, const T* x, const T* y, long long n, hipStream_t st) {
    double* part = static_cast<double*>(p.ws.get(kSlotRed, sizeof(double) * kRedBlocks));
    const int nblk = red_blocks(n);
    k_dot<T><<<nblk, 256, 0, st>>>(x, y, n, part);
    HIPCHK(hipGetLastError());
    return sum_partials<T>(p.ws, part, nblk, st);
}

}  // namespace

template <class T>
int prepared_gram(PreparedBase* base, const T* x, T* out, hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        gram_apply<T>(p, x, out, st, keep, timer);
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

// Conjugate gradients for A^H W A x = rhs from zero (core/src/clean.rs:263-306):
// stops after `iterations` Gram applications, when |r| <= tolerance |rhs|,
// or when <p, A p> <= 0. Vectors in T, inner products in double.
template <class T>
int prepared_cg(PreparedBase* base, const T* rhs, T* x, int iterations, double tolerance, int* used,
                double* rel, hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        const long long m = (long long)p.o.nx * p.o.ny;
        T* work = static_cast<T*>(p.ws.get(kSlotSolve, sizeof(T) * 3 * std::max<long long>(m, 1)));
        T *r = work, *dir = work + m, *ap = work + 2 * m;
        double* sc = static_cast<double*>(p.ws.get(kSlotScalars, sizeof(double) * 4));
        double* part = static_cast<double*>(p.ws.get(kSlotRed, sizeof(double) * kRedBlocks));
        double* hb = static_cast<double*>(p.ws.host(sizeof(double) * kRedBlocks));
        const double norm_d = std::sqrt(dot<T>(p, rhs, rhs, m, st));
        fill_zero(x, sizeof(T) * m, st);
        HIPCHK(hipMemcpyAsync(r, rhs, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(dir, rhs, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
        double rr = norm_d * norm_d;
        k_set_scalar<<<1, 64, 0, st>>>(sc, rr);
        HIPCHK(hipGetLastError());
        int k = 0;
        const int nblk = red_blocks(m);
        // one iteration: A p, <p, A p>, the step and |r|^2, the direction, on
        // the device; the two scalars the host tests come back in one copy
        auto iteration = [&](hipStream_t s) {
            gram_body<T>(p, dir, ap, s, keep, timer);
            k_dot<T><<<nblk, 256, 0, s>>>(dir, ap, m, part);
            k_sum_parts<<<1, 256, 0, s>>>(part, nblk, sc + 1);
            k_cg_step_s<T><<<nblk, 256, 0, s>>>(x, r, dir, ap, sc, m, part);
            k_sum_parts<<<1, 256, 0, s>>>(part, nblk, sc + 2);
            k_cg_dir_s<T><<<nblk, 256, 0, s>>>(dir, r, sc, m);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hb, sc + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
            k_cg_roll<<<1, 64, 0, s>>>(sc);
            HIPCHK(hipGetLastError());
        };
        while (k < iterations && std::sqrt(rr) > tolerance * norm_d) {
            run_graph(p, {5, up(x)}, st, iteration);
            HIPCHK(hipStreamSynchronize(st));
            const double pap = hb[0], rr_next = hb[1];
            if (!(pap > 0.0)) break;
            rr = rr_next;
            ++k;
        }
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        *used = k;
        *rel = std::sqrt(rr) / norm_d;
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

// Preconditioned conjugate gradients for A^H W A x = rhs from zero, with a
// symmetric positive definite z = M r: the CG of clean.rs:263-306 with the
// search direction built from z (alpha = <r, z> / <p, A p>, beta = <r', z'> /
// <r, z>) and the same stopping rule on the unpreconditioned residual,
// |r| <= tolerance |rhs|, so the counts of Gram applications compare.
template <class T>
int prepared_pcg(PreparedBase* base, const std::function<void(const T*, T*, hipStream_t)>& precond,
                 const T* rhs, T* x, int iterations, double tolerance, int* used, double* rel,
                 hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        const long long m = (long long)p.o.nx * p.o.ny;
        T* work = static_cast<T*>(p.ws.get(kSlotSolve, sizeof(T) * 4 * std::max<long long>(m, 1)));
        T *r = work, *dir = work + m, *ap = work + 2 * m, *z = work + 3 * m;
        const double norm_d = std::sqrt(dot<T>(p, rhs, rhs, m, st));
        fill_zero(x, sizeof(T) * m, st);
        HIPCHK(hipMemcpyAsync(r, rhs, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
        precond(r, z, st);
        HIPCHK(hipMemcpyAsync(dir, z, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
        double rz = dot<T>(p, r, z, m, st);
        double rr = norm_d * norm_d;
        int k = 0;
        double* part = static_cast<double*>(p.ws.get(kSlotRed, sizeof(double) * kRedBlocks));
        const int nblk = red_blocks(m);
        while (k < iterations && std::sqrt(rr) > tolerance * norm_d) {
            if (!(rz > 0.0)) break;
            gram_apply<T>(p, dir, ap, st, keep, timer);
            const double pap = dot<T>(p, dir, ap, m, st);
            ++k;
            if (!(pap > 0.0)) break;
            const double alpha = rz / pap;
            k_cg_step<T><<<nblk, 256, 0, st>>>(x, r, dir, ap, alpha, m, part);
            HIPCHK(hipGetLastError());
            rr = sum_partials<T>(p.ws, part, nblk, st);
            if (std::sqrt(rr) <= tolerance * norm_d) break;
            precond(r, z, st);
            const double rz_next = dot<T>(p, r, z, m, st);
            const double beta = rz_next / rz;
            k_cg_dir<T><<<nblk, 256, 0, st>>>(dir, z, beta, m);
            HIPCHK(hipGetLastError());
            rz = rz_next;
        }
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        *used = k;
        *rel = norm_d > 0.0 ? std::sqrt(rr) / norm_d : 0.0;
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

template int prepared_pcg<float>(PreparedBase*, const std::function<void(const float*, float*, hipStream_t)>&,
                                 const float*, float*, int, double, int*, double*, hipStream_t, std::string&);
template int prepared_pcg<double>(PreparedBase*, const std::function<void(const double*, double*, hipStream_t)>&,
                                  const double*, double*, int, double, int*, double*, hipStream_t, std::string&);

namespace {

// The reference's seeded start vectors (core/src/prox_solver.rs:871-882):
// xorshift64 from max(seed, 1), each value (state >> 11) / 2^53 - 1/2.
std::vector<double> seeded_vector(std::size_t n, std::uint64_t seed) {
    std::vector<double> v(n);
    std::uint64_t state = std::max<std::uint64_t>(seed, 1);
    for (std::size_t i = 0; i < n; ++i) {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        v[i] = double(state >> 11) / double(1ull << 53) - 0.5;
    }
    return v;
}

// Sums, the largest violation and the smallest bound of k_fista_eval's
// partials, combined on the host in block order.
template <class T>
void fista_partials(Workspace& ws, const double* part, int nblk, hipStream_t st, double (&v)[kFistaSlots]) {
    double* h = static_cast<double*>(ws.host(sizeof(double) * kFistaSlots * nblk));
    HIPCHK(hipMemcpyAsync(h, part, sizeof(double) * kFistaSlots * nblk, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int q = 0; q < kFistaSums; ++q) v[q] = 0.0;
    v[kFistaSums] = 0.0;
    v[kFistaSums + 1] = INFINITY;
    for (int b = 0; b < nblk; ++b) {
        const double* r = h + (long long)b * kFistaSlots;
        for (int q = 0; q < kFistaSums; ++q) v[q] += r[q];
        v[kFistaSums] = std::fmax(v[kFistaSums], r[kFistaSums]);
        v[kFistaSums + 1] = std::fmin(v[kFistaSums + 1], r[kFistaSums + 1]);
    }
}

// The certificate of x (prox_solver.rs:236-281) from k_fista_eval's sums:
// misfit, penalty, duality gap and optimality residual.
struct Certificate {
    double misfit, pen, gap, kkt;
};
Certificate certificate_of(const double (&v)[kFistaSlots], double q0) {
    const double d_x = v[0];
    const double misfit = q0 - d_x + 0.5 * v[1];
    const double pen = v[2], kkt = v[kFistaSums], scale_max = v[kFistaSums + 1];
    // the dual point s e, e = y - Phi x: <y, e> = 2 Q0 - <d, x>, |e|^2 = 2 Q
    const double y_e = 2.0 * q0 - d_x, e_e = 2.0 * misfit;
    const double sc = std::min(std::max(y_e / std::max(e_e, 1e-300), 0.0), scale_max);
    const double dual = sc * y_e - 0.5 * sc * sc * e_e;
    return Certificate{misfit, pen, misfit + pen - dual, kkt};
}

// The FISTA / forward-backward loop of prox_solver.rs:283-445 on any normal
// operator `apply(in, out)` (out = G in, device vectors of m values on `st`),
// from x_io, which receives the last iterate (gx_out, when given, its G x).
// Nine image vectors in `work`; per-block partial sums in `part`.
template <class T, class Apply>
void fista_core(Workspace& ws, Apply&& apply, T* work, double* part, long long m, const T* dirty, double q0,
                double lam, const T* factors, int sgn, const RangFistaSettings& set, T* x_io, T* gx_out,
                RangFistaRecord* records, std::size_t nrecords, RangFistaOutcome& out, hipStream_t st) {
    T* x = work;
    T* x_prev = work + m;
    T* x_next = work + 2 * m;
    T* y = work + 3 * m;
    T* y_prev = work + 4 * m;
    T* gy = work + 5 * m;
    T* gy_prev = work + 6 * m;
    T* gx = work + 7 * m;
    T* gx_prev = work + 8 * m;
    HIPCHK(hipMemcpyAsync(x, x_io, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(x_prev, x_io, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(y, x_io, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
    fill_zero(y_prev, sizeof(T) * m, st);
    fill_zero(gy_prev, sizeof(T) * m, st);
    fill_zero(gx_prev, sizeof(T) * m, st);
    const int nblk = red_blocks(m);
    out = RangFistaOutcome{};
    out.min_curvature = INFINITY;
    out.max_curvature = -INFINITY;
    double t = 1.0, beta = 0.0, last_objective = INFINITY;
    bool restarted = false;
    const bool have_gap = !std::isnan(set.gap_tolerance), have_kkt = !std::isnan(set.kkt_tolerance);
    const bool have_step = !std::isnan(set.step_tolerance);
    for (std::uint64_t k = 0;; ++k) {
        apply(static_cast<const T*>(y), gy);
        k_fista_eval<T><<<nblk, 256, 0, st>>>(y, y_prev, gy, gy_prev, gx_prev, gx, x, x_prev, dirty, factors,
                                               lam, sgn, beta, m, part);
        HIPCHK(hipGetLastError());
        double v[kFistaSlots];
        fista_partials<T>(ws, part, nblk, st, v);
        double curvature = NAN;
        if (k > 0 && v[8] > 0.0) curvature = v[7] / v[8];
        if (std::isfinite(curvature)) {
            out.min_curvature = std::fmin(out.min_curvature, curvature);
            out.max_curvature = std::fmax(out.max_curvature, curvature);
            if (curvature < 0.0) ++out.negative_curvatures;
        }
        const Certificate c = certificate_of(v, q0);
        RangFistaRecord rec{};
        rec.iteration = k;
        rec.objective = c.misfit + c.pen;
        rec.misfit = c.misfit;
        rec.penalty = c.pen;
        rec.gap = c.gap;
        rec.kkt = c.kkt;
        rec.nonzero = std::uint64_t(v[6]);
        rec.flux = v[5];
        rec.curvature = curvature;
        rec.relative_step = std::sqrt(v[3]) / std::max(std::sqrt(v[4]), 1e-300);
        rec.restarted = restarted ? 1 : 0;
        if (records && k < nrecords) records[k] = rec;
        if (std::isfinite(rec.objective)) {
            if (rec.objective > last_objective) {
                ++out.objective_rises;
                out.largest_rise = std::fmax(out.largest_rise, rec.objective - last_objective);
            }
            last_objective = rec.objective;
        }
        out.last = rec;
        out.iterations = k;
        const bool converged = (have_gap && c.gap <= set.gap_tolerance) || (have_kkt && c.kkt <= set.kkt_tolerance) ||
                               (k > 0 && have_step && rec.relative_step <= set.step_tolerance);
        if (converged || k == set.max_iterations) {
            out.converged = converged ? 1 : 0;
            HIPCHK(hipMemcpyAsync(x_io, x, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
            if (gx_out) HIPCHK(hipMemcpyAsync(gx_out, gx, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
            return;
        }
        k_fista_step<T><<<nblk, 256, 0, st>>>(y, dirty, gy, x, x_next, factors, lam, sgn, set.step, m, part);
        HIPCHK(hipGetLastError());
        const double inner = sum_partials<T>(ws, part, nblk, st);
        double next_beta = 0.0;
        restarted = false;
        if (set.momentum) {
            if (set.restart && inner > 0.0) {
                t = 1.0;
                restarted = true;
                ++out.restarts;
            } else {
                const double t_next = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * t * t));
                next_beta = (t - 1.0) / t_next;
                t = t_next;
            }
        }
        std::swap(y, y_prev);
        std::swap(gy, gy_prev);
        std::swap(gx, gx_prev);
        std::swap(x_prev, x);
        std::swap(x, x_next);
        k_fista_extrap<T><<<nblk, 256, 0, st>>>(y, x, x_prev, next_beta, m);
        HIPCHK(hipGetLastError());
        beta = next_beta;
    }
}

// The defect of a cheap operator at x (prox_solver.rs:797-812, 838-850):
// corrected = d - (G x - G~ x); per block the largest |G x - G~ x| and the
// largest residual d - G x (|d - G x| when signed), both floored at 0.
template <class T>
__global__ __launch_bounds__(256) void k_defect(const T* d, const T* g, const T* gc, T* corrected, int sgn,
                                                long long n, double* part) {
    double sup = 0.0, peak = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double gi = double(g[i]), di = double(d[i]);
        const double c = gi - double(gc[i]);
        corrected[i] = T(di - c);
        sup = fmax(sup, fabs(c));
        const double r = di - gi;
        peak = fmax(peak, sgn ? fabs(r) : r);
    }
    __shared__ double s[2][256];
    s[0][threadIdx.x] = sup;
    s[1][threadIdx.x] = peak;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            s[0][threadIdx.x] = fmax(s[0][threadIdx.x], s[0][threadIdx.x + h]);
            s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * (long long)blockIdx.x] = s[0][0];
        part[2 * (long long)blockIdx.x + 1] = s[1][0];
    }
}

}  // namespace

// FISTA or forward-backward on the prepared plan's normal operator
// (core/src/prox_solver.rs:283-445) with the L1 prior lam * factor_i (x >= 0
// unless signed): one Gram application per iteration, G x by linearity, and
// per iterate the record the reference reports (objective, misfit, penalty,
// duality gap, optimality residual, nonzero count, flux, curvature, relative
// step, restart). Vectors in T on the device, sums in double.
template <class T>
int prepared_fista(PreparedBase* base, const T* dirty, double q0, double lam, const T* factors, int sgn,
                   const RangFistaSettings& set, T* x_io, T* gx_out, RangFistaRecord* records,
                   std::size_t nrecords, RangFistaOutcome* outc, hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        const long long m = (long long)p.o.nx * p.o.ny;
        const std::size_t vb = sizeof(T) * std::max<long long>(m, 1);
        T* work = static_cast<T*>(p.ws.get(kSlotFista, vb * 9));
        double* part = static_cast<double*>(p.ws.get(kSlotRed, sizeof(double) * kFistaSlots * kRedBlocks));
        RangFistaOutcome out{};
        fista_core<T>(p.ws, [&](const T* in, T* o) { gram_apply<T>(p, in, o, st, keep, timer); }, work, part, m,
                      dirty, q0, lam, factors, sgn, set, x_io, gx_out, records, nrecords, out, st);
        timer.finish();
        *outc = out;
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

// Major cycles on a cheap normal operator with its defect moved into the data
// (core/src/prox_solver.rs:760-875, defect_correction): each cycle applies the
// exact operator (the prepared plan: vis2dirty of dirty2vis) and the cheap one
// at x_j, judges x_j by the exact certificate at the target lambda, and solves
// min 1/2 <x, G~ x> - <d - (G - G~) x_j, x> + R(x) by FISTA on the cheap
// operator from x_j, at max(lambda, gamma * peak residual) and without a gap
// test. At a fixed point the cheap operator's error cancels. One record per
// cycle; x_io receives the last judged iterate and gx_out its exact G x.
template <class T>
int prepared_defect(PreparedBase* base, const std::function<void(const T*, T*, hipStream_t)>& cheap,
                    const T* dirty, double q0, const T* factors, int sgn, const RangFistaSettings& inner,
                    const RangOuterSchedule& sched, T* x_io, T* gx_out, RangOuterRecord* records,
                    std::size_t nrecords, RangDefectOutcome* outc, hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        const long long m = (long long)p.o.nx * p.o.ny;
        const std::size_t vb = sizeof(T) * std::max<long long>(m, 1);
        T* work = static_cast<T*>(p.ws.get(kSlotFista, vb * 9));
        T* outer = static_cast<T*>(p.ws.get(kSlotDefect, vb * 4));
        T *x = outer, *g = outer + m, *gc = outer + 2 * m, *corrected = outer + 3 * m;
        double* part = static_cast<double*>(p.ws.get(kSlotRed, sizeof(double) * kFistaSlots * kRedBlocks));
        const int nblk = red_blocks(m);
        HIPCHK(hipMemcpyAsync(x, x_io, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
        RangFistaSettings set = inner;
        set.gap_tolerance = NAN;  // the inner objective is not of least-squares form
        RangDefectOutcome out{};
        std::uint64_t inner_iterations = 0;
        for (std::uint64_t j = 0;; ++j) {
            gram_apply<T>(p, x, g, st, keep, timer);
            ++out.exact_applies;
            // the certificate at the target lambda: k_fista_eval with y = x
            // and beta = 0 (its G x written over g, unchanged)
            k_fista_eval<T><<<nblk, 256, 0, st>>>(x, x, g, g, g, g, x, x, dirty, factors, sched.lambda, sgn, 0.0,
                                                   m, part);
            HIPCHK(hipGetLastError());
            double v[kFistaSlots];
            fista_partials<T>(p.ws, part, nblk, st, v);
            const Certificate c = certificate_of(v, q0);
            cheap(x, gc, st);
            ++out.cheap_applies;
            k_defect<T><<<nblk, 256, 0, st>>>(dirty, g, gc, corrected, sgn, m, part);
            HIPCHK(hipGetLastError());
            double* h = static_cast<double*>(p.ws.host(sizeof(double) * 2 * nblk));
            HIPCHK(hipMemcpyAsync(h, part, sizeof(double) * 2 * nblk, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            double sup = 0.0, peak = 0.0;
            for (int b = 0; b < nblk; ++b) {
                sup = std::fmax(sup, h[2 * b]);
                peak = std::fmax(peak, h[2 * b + 1]);
            }
            RangOuterRecord rec{};
            rec.outer = j;
            rec.objective = c.misfit + c.pen;
            rec.misfit = c.misfit;
            rec.gap = c.gap;
            rec.kkt = c.kkt;
            rec.inner_iterations = inner_iterations;
            rec.defect_sup = sup;
            if (records && j < nrecords) records[j] = rec;
            out.last = rec;
            out.cycles = j;
            if (c.kkt <= sched.kkt || j == sched.cycles) {
                out.converged = c.kkt <= sched.kkt ? 1 : 0;
                out.iterations = inner_iterations;
                HIPCHK(hipMemcpyAsync(x_io, x, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
                if (gx_out) HIPCHK(hipMemcpyAsync(gx_out, g, sizeof(T) * m, hipMemcpyDeviceToDevice, st));
                HIPCHK(hipStreamSynchronize(st));
                break;
            }
            const double lam = std::fmax(sched.lambda, sched.gamma * peak);
            RangFistaOutcome io{};
            fista_core<T>(
                p.ws,
                [&](const T* in, T* o) {
                    cheap(in, o, st);
                    ++out.cheap_applies;
                },
                work, part, m, corrected, q0, lam, factors, sgn, set, x, nullptr, nullptr, 0, io, st);
            inner_iterations += io.iterations;
            out.restarts += io.restarts;
            out.negative_curvatures += io.negative_curvatures;
        }
        timer.finish();
        *outc = out;
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

template int prepared_defect<float>(PreparedBase*, const std::function<void(const float*, float*, hipStream_t)>&,
                                    const float*, double, const float*, int, const RangFistaSettings&,
                                    const RangOuterSchedule&, float*, float*, RangOuterRecord*, std::size_t,
                                    RangDefectOutcome*, hipStream_t, std::string&);
template int prepared_defect<double>(PreparedBase*, const std::function<void(const double*, double*, hipStream_t)>&,
                                     const double*, double, const double*, int, const RangFistaSettings&,
                                     const RangOuterSchedule&, double*, double*, RangOuterRecord*, std::size_t,
                                     RangDefectOutcome*, hipStream_t, std::string&);

// The largest eigenvalue of the normal operator by power iteration from the
// reference's seeded start (core/src/prox_solver.rs:886-908): the estimate
// and the applications used.
template <class T>
int prepared_power(PreparedBase* base, std::uint64_t seed, double tolerance, std::uint64_t max_iterations,
                   double* value_out, std::uint64_t* used, hipStream_t st, std::string& err) {
    auto& p = *static_cast<PreparedImpl<T>*>(base);
    try {
        std::lock_guard<std::mutex> lk(p.ws.mu);
        HostKeep keep;
        StageTimer timer(st);
        const long long m = (long long)p.o.nx * p.o.ny;
        T* work = static_cast<T*>(p.ws.get(kSlotSolve, sizeof(T) * 2 * std::max<long long>(m, 1)));
        T *v = work, *g = work + m;
        const std::vector<double> s0 = seeded_vector(std::size_t(m), seed);
        std::vector<T> h(s0.begin(), s0.end());
        HIPCHK(hipMemcpyAsync(v, h.data(), sizeof(T) * m, hipMemcpyHostToDevice, st));
        const int nblk = red_blocks(m);
        const double norm = std::sqrt(dot<T>(p, v, v, m, st));
        k_scale<T><<<nblk, 256, 0, st>>>(v, 1.0 / norm, m);
        HIPCHK(hipGetLastError());
        double value = 0.0;
        std::uint64_t k = 1;
        for (; k <= max_iterations; ++k) {
            gram_apply<T>(p, v, g, st, keep, timer);
            const double next = std::sqrt(dot<T>(p, g, g, m, st));
            k_scale<T><<<nblk, 256, 0, st>>>(g, 1.0 / std::max(next, 1e-300), m);
            HIPCHK(hipGetLastError());
            std::swap(v, g);
            if (k > 1 && std::fabs(next - value) <= tolerance * next) {
                value = next;
                break;
            }
            value = next;
        }
        HIPCHK(hipStreamSynchronize(st));
        timer.finish();
        *value_out = value;
        *used = std::min(k, max_iterations);
        return 0;
    } catch (const DevError& e) {
        err = e.msg;
        return 2;
    }
}

#define RANG_PREPARED_INST(T)                                                                     \
    template PreparedBase* prepared_new<T>(const Options&, const DevInputs&, const double*,         \
                                           hipStream_t, int&, std::string&);                        \
    template int prepared_vis2dirty<T>(PreparedBase*, const void*, int, T*, double*, hipStream_t,  \
                                       std::string&);                                               \
    template int prepared_dirty2vis<T>(PreparedBase*, const T*, bool, T*, hipStream_t, std::string&); \
    template int prepared_residual<T>(PreparedBase*, const void*, int, const T*, T*, double*,       \
                                      hipStream_t, std::string&);                                   \
    template int prepared_residual_sorted<T>(PreparedBase*, const void*, int, const T*, T*, double*, \
                                             hipStream_t, std::string&);                            \
    template int prepared_sorted<T>(PreparedBase*, const void*, int, void*, hipStream_t, std::string&); \
    template int prepared_gram<T>(PreparedBase*, const T*, T*, hipStream_t, std::string&);         \
    template int prepared_cg<T>(PreparedBase*, const T*, T*, int, double, int*, double*, hipStream_t, \
                                std::string&);                                                      \
    template int prepared_fista<T>(PreparedBase*, const T*, double, double, const T*, int,          \
                                   const RangFistaSettings&, T*, T*, RangFistaRecord*, std::size_t, \
                                   RangFistaOutcome*, hipStream_t, std::string&);                   \
    template int prepared_power<T>(PreparedBase*, std::uint64_t, double, std::uint64_t, double*,    \
                                   std::uint64_t*, hipStream_t, std::string&);
RANG_PREPARED_INST(float)
RANG_PREPARED_INST(double)
#undef RANG_PREPARED_INST

// Frees the device buffers every call keeps between calls (the per-device
// workspaces and the sub-band batch pools; not the fixed scratch block,
// nor the buffers of live prepared plans) after the device's work is done.
void release_workspaces() {
    std::vector<Workspace*> all;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        for (auto& kv : g_ws_all)
            if (kv.second) all.push_back(kv.second.get());
    }
    (void)hipDeviceSynchronize();
    for (Workspace* w : all) {
        std::lock_guard<std::mutex> lk(w->mu);
        free_workspace(*w);
    }
    std::vector<BatchPool*> pools;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        for (auto& kv : g_pool_all)
            if (kv.second) pools.push_back(kv.second.get());
    }
    for (BatchPool* p : pools) {
        std::lock_guard<std::mutex> lk(p->mu);
        free_workspace(p->control);
        for (auto& w : p->images) free_workspace(*w);
        p->images.clear();
    }
}

template int device_batch_vis2dirty<float>(BatchImage*, int, std::size_t, hipStream_t, std::string&);
template int device_batch_vis2dirty<double>(BatchImage*, int, std::size_t, hipStream_t, std::string&);
template int device_tune<float>(BatchImage&, const char*, int, double*, double*, int*, hipStream_t, std::string&);
template int device_tune<double>(BatchImage&, const char*, int, double*, double*, int*, hipStream_t, std::string&);
template int device_vis2dirty_fixed<float>(const Options&, const DevInputs&, const double*, float*,
                                           float*, double*, hipStream_t, std::string&);
template int device_vis2dirty<float>(const Options&, const DevInputs&, const double*, float*,
                                     float*, double*, hipStream_t, std::string&);
template int device_vis2dirty<double>(const Options&, const DevInputs&, const double*, double*,
                                      double*, double*, hipStream_t, std::string&);
template int device_dirty2vis<float>(const Options&, const DevInputs&, const double*,
                                     const float*, float*, double*, hipStream_t, std::string&);
template int device_dirty2vis<double>(const Options&, const DevInputs&, const double*,
                                      const double*, double*, double*, hipStream_t, std::string&);

}  // namespace rang
