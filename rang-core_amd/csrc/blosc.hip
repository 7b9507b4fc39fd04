// Blosc-1 chunk decoding on the GPU: zarr chunks (the visibility, flag,
// weight and uvw variables of an MSv4 store) expanded in HBM.
//
// The host reads each chunk's 16-byte header, block starts and stream
// lengths (core/src/blosc.rs:53-79, 276-351: the refusals, the block span,
// the split rule) and uploads the compressed bytes once; the device expands
// every stream at once, one wave per stream: the sequences of a stream are
// parsed by the whole wave in step (the stream's bytes staged in LDS a
// window at a time), each literal run and match copied by the 64 lanes
// together, its recent output kept in an LDS ring. A match (offset o,
// length m) is periodic with period o, so lane i copies byte
// d - o + (i mod o): every byte it reads lies before the match, written by
// earlier sequences. Blocks stored shuffled are
// expanded into scratch and then transposed (stream b holds byte b of every
// element; core/src/blosc.rs:180-196).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rang_core_blosc.h"

namespace rang {
namespace {

constexpr int kHead = 16;
constexpr unsigned kShuffle = 0x01, kMemcpyed = 0x02, kBitShuffle = 0x04, kDontSplit = 0x10;
constexpr int kMinBuffer = 128, kMaxSplits = 16;
constexpr int kWaves = 4;         // streams per workgroup
constexpr int kSlack = 24;        // bytes a sequence head may need past its token

thread_local std::string t_blosc_error;

struct Fail {
    int status;
    std::string msg;
};

[[noreturn]] void refuse(const std::string& m) { throw Fail{1, m}; }

void hipchk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Fail{2, std::string(what) + ": " + hipGetErrorString(e)};
}

std::size_t le32(const std::uint8_t* p) {
    return std::size_t(p[0]) | std::size_t(p[1]) << 8 | std::size_t(p[2]) << 16 | std::size_t(p[3]) << 24;
}

// One stream: `clen` compressed bytes at `src` expand to `dlen` bytes at
// `dst` (LZ4, or a raw copy when the lengths are equal).
struct Stream {
    std::uint64_t src, dst;  // device addresses
    std::uint32_t clen, dlen;
};

// A shuffled block: `len` bytes of `typesize` elements, scratch -> dst.
struct Shuffled {
    std::uint64_t scratch, dst;
    std::uint32_t len, typesize;
};

// The header fields (core/src/blosc.rs:53-79).
struct Header {
    unsigned flags;
    std::size_t typesize, nbytes, blocksize, cbytes, nblocks;
};

Header parse_header(const std::uint8_t* src, std::size_t len) {
    if (len < kHead) refuse("truncated blosc header");
    Header h;
    h.flags = src[2];
    h.typesize = src[3];
    h.nbytes = le32(src + 4);
    h.blocksize = le32(src + 8);
    h.cbytes = le32(src + 12);
    if (src[0] != 2 && src[0] != 1) refuse("unsupported blosc format version");
    if (h.typesize == 0 || (h.nbytes > 0 && h.blocksize == 0)) refuse("invalid blosc typesize or blocksize");
    if ((h.flags & 0x08) || (h.flags & kBitShuffle)) refuse("bit shuffle or unknown blosc flags are not supported");
    if (!(h.flags & kMemcpyed) && ((h.flags & 0xe0) >> 5) != 1)
        refuse("only LZ4-compressed blosc chunks are supported");
    h.nblocks = h.blocksize ? (h.nbytes + h.blocksize - 1) / h.blocksize : 0;
    return h;
}

// ------------------------------------------------------------ device

__device__ int g_blosc_status;  // first failure code of a launch (0 none)

__device__ __forceinline__ void flag_error(int code) { atomicCAS(&g_blosc_status, 0, code); }

// One wave expands one stream. The stream's bytes [wbase, wend) sit in the
// wave's LDS window (refilled with coalesced loads) and its last kRing
// output bytes in the wave's LDS ring. A sequence's token, lengths and
// offset are read from the window and made wave-uniform
// (readfirstlane), so the parse runs on the scalar unit; its literals are
// copied from the window, its match from the ring (a match of offset o
// reads bytes [d - o, d), copied in segments of at most kRing - o bytes so
// that no write lands on a byte still to be read), each byte stored to
// global memory and to the ring. A match reaching further back than the
// ring reads the wave's own earlier output in global memory: every store
// older than the wave's last 63 vector memory instructions has completed
// after `s_waitcnt vmcnt(63)`, and those instructions wrote at most
// 63 x 64 < kRing bytes, so the bytes it reads are complete; the loads are
// work-group-scope atomic loads, coherent with the wave's completed stores.
constexpr int kRing = 4096;    // output bytes kept per wave (a power of two, > 63 x 64)
constexpr int kWin = 1024;     // input bytes staged per wave
constexpr int kWinPad = 0;
constexpr int kLdsBytes = kWaves * (kRing + kWin + kWinPad);
static_assert(kRing > 63 * 64, "the ring must cover the last 63 stores");

__device__ __forceinline__ std::uint32_t uniform(std::uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Global-memory bytes: global (not flat) loads and stores, which count in
// vmcnt only, so waits for the LDS reads of the parse do not wait for them.
typedef __attribute__((address_space(1))) std::uint8_t gbyte;

__global__ __launch_bounds__(64 * kWaves) void k_lz4_streams(const Stream* __restrict__ streams, int nstreams) {
    extern __shared__ __align__(16) std::uint8_t smem[];
    // the wave's index and its stream are wave-uniform: the stream's
    // description is read with scalar loads into scalar registers
    const int lane = threadIdx.x & 63, w = int(uniform(threadIdx.x >> 6));
    const int sid = blockIdx.x * kWaves + w;
    if (sid >= nstreams) return;
    std::uint8_t* ring = smem + w * (kRing + kWin + kWinPad);
    std::uint8_t* win = ring + kRing;
    const Stream S = streams[sid];
    const gbyte* in = reinterpret_cast<const gbyte*>(S.src);
    gbyte* out = reinterpret_cast<gbyte*>(S.dst);
    const std::uint32_t clen = uniform(S.clen), dlen = uniform(S.dlen);
    if (clen == dlen) {
        // stored raw: 16 loads in flight per lane (indices clamped into the
        // stream, so every load is in bounds), then the stores
        constexpr int U = 16;
        for (std::uint32_t base = 0; base < dlen; base += 64 * U) {
            std::uint8_t v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) v[k] = in[min(base + lane + 64 * k, dlen - 1)];
#pragma unroll
            for (int k = 0; k < U; ++k)
                if (base + lane + 64 * k < dlen) out[base + lane + 64 * k] = v[k];
        }
        return;
    }
    constexpr std::uint32_t mask = kRing - 1;
    std::uint32_t wbase = 0, wend = 0;
    // the window at `at` (< clen): every lane's loads issued before the
    // first LDS write, indices clamped into the stream
    auto refill = [&](std::uint32_t at) {
        wbase = at;
        wend = min(clen, at + std::uint32_t(kWin));
        const std::uint32_t n = wend - wbase;
        std::uint8_t v[kWin / 64];
#pragma unroll
        for (int k = 0; k < kWin / 64; ++k) v[k] = in[wbase + min(std::uint32_t(lane + 64 * k), n - 1)];
#pragma unroll
        for (int k = 0; k < kWin / 64; ++k)
            if (std::uint32_t(lane + 64 * k) < n) win[lane + 64 * k] = v[k];
    };
    // byte `at` of the stream (at < clen), from the window, wave-uniform
    auto byte = [&](std::uint32_t at) -> std::uint32_t {
        if (at >= wend || at < wbase) refill(at);
        return uniform(win[at - wbase]);
    };
    refill(0);
    std::uint32_t s = 0, d = 0;
    for (;;) {
        if (s >= clen) return flag_error(3);  // a stream ends with literals
        if (s + kSlack > wend && wend < clen) refill(s);
        const std::uint32_t token = byte(s++);
        std::uint32_t lit = token >> 4;
        if (lit == 15) {
            std::uint32_t b;
            do {
                if (s >= clen) return flag_error(3);
                b = byte(s++);
                lit += b;
            } while (b == 255);
        }
        if (lit > clen - s || lit > dlen - d) return flag_error(3);
        if (lit <= 64 && s + lit <= wend) {
            // a short run inside the window: one predicated copy
            const std::uint32_t o = s - wbase;
            if (std::uint32_t(lane) < lit) {
                const std::uint8_t v = win[o + lane];
                out[d + lane] = v;
                ring[(d + lane) & mask] = v;
            }
            s += lit;
            d += lit;
            lit = 0;
        }
        while (lit > 0) {
            if (s >= wend || s < wbase) refill(s);
            const std::uint32_t n = min(lit, wend - s), o = s - wbase;
            for (std::uint32_t i = lane; i < n; i += 64) {
                const std::uint8_t v = win[o + i];
                out[d + i] = v;
                ring[(d + i) & mask] = v;
            }
            s += n;
            d += n;
            lit -= n;
        }
        if (s == clen) break;
        if (clen - s < 2) return flag_error(3);
        const std::uint32_t off = byte(s) | byte(s + 1) << 8;
        s += 2;
        if (off == 0 || off > d) return flag_error(3);
        std::uint32_t m = token & 15;
        if (m == 15) {
            std::uint32_t b;
            do {
                if (s >= clen) return flag_error(3);
                b = byte(s++);
                m += b;
            } while (b == 255);
        }
        m += 4;
        if (m > dlen - d) return flag_error(3);
        if (m <= 64 && off >= m && off < std::uint32_t(kRing)) {
            // a short match not overlapping itself: one predicated copy
            if (std::uint32_t(lane) < m) {
                const std::uint8_t v = ring[(d - off + lane) & mask];
                out[d + lane] = v;
                ring[(d + lane) & mask] = v;
            }
            d += m;
        } else if (off < std::uint32_t(kRing)) {
            while (m > 0) {
                const std::uint32_t seg = min(m, std::uint32_t(kRing) - off), from = d - off;
                // lane i reads byte from + (i mod off); r steps by 64 mod off
                // (divisions only for offsets below 64)
                std::uint32_t r = lane, step = 64;
                if (off <= 64) {
                    r = lane % off;
                    step = 64 % off;
                }
                for (std::uint32_t i = lane; i < seg; i += 64) {
                    const std::uint8_t v = ring[(from + r) & mask];
                    out[d + i] = v;
                    ring[(d + i) & mask] = v;
                    r += step;
                    if (r >= off) r -= off;
                }
                d += seg;
                m -= seg;
            }
        } else {
            __builtin_amdgcn_s_waitcnt(0xCF7F);  // vmcnt(63), expcnt and lgkmcnt unbounded
            const std::uint32_t from = d - off;
            for (std::uint32_t i = lane; i < m; i += 64) {
                const std::uint8_t v =
                    __hip_atomic_load(out + from + (i < off ? i : i % off), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                out[d + i] = v;
                ring[(d + i) & mask] = v;
            }
            d += m;
        }
    }
    if (d != dlen) flag_error(3);
}

// Byte unshuffle of whole blocks: output byte x = i * ts + b takes scratch
// byte b * q + i (q = len / ts); the tail past q * ts stays in place.
__global__ __launch_bounds__(256) void k_unshuffle(const Shuffled* __restrict__ blocks) {
    const Shuffled B = blocks[blockIdx.x];
    const std::uint8_t* src = reinterpret_cast<const std::uint8_t*>(B.scratch);
    std::uint8_t* dst = reinterpret_cast<std::uint8_t*>(B.dst);
    const std::uint32_t ts = B.typesize, q = B.len / ts, body = q * ts;
    for (std::uint32_t x = threadIdx.x; x < B.len; x += 256) {
        if (x < body) {
            const std::uint32_t i = x / ts, b = x - i * ts;
            dst[x] = src[b * q + i];
        } else {
            dst[x] = src[x];
        }
    }
}

// ------------------------------------------------------------ host

// Device buffers kept between calls (the staged compressed bytes, scratch
// for shuffled blocks, the tables), grown on demand; calls are serialised.
struct Pool {
    std::mutex mu;
    void* buf[4] = {nullptr, nullptr, nullptr, nullptr};
    std::size_t cap[4] = {0, 0, 0, 0};
    void* get(int k, std::size_t n) {
        if (n == 0) n = 1;
        if (cap[k] < n) {
            if (buf[k]) hipchk(hipFree(buf[k]), "hipFree");
            buf[k] = nullptr;
            cap[k] = 0;
            hipchk(hipMalloc(&buf[k], n), "hipMalloc");
            cap[k] = n;
        }
        return buf[k];
    }
};
Pool g_pool;

// The streams of one chunk whose compressed bytes are at host `src` and
// will sit at device `dsrc`; shuffled blocks expand into `scratch`.
void plan_chunk(const std::uint8_t* src, std::size_t len, std::uint64_t dsrc, std::uint64_t dst,
                std::size_t dst_len, std::uint64_t scratch, std::vector<Stream>& streams,
                std::vector<Shuffled>& shuffled, std::size_t& scratch_used) {
    const Header h = parse_header(src, len);
    if (h.cbytes != len) refuse("blosc compressed size disagrees with chunk length");
    if (h.nbytes != dst_len) refuse("output buffer has the wrong size");
    if (h.nbytes >= (std::size_t(1) << 32)) refuse("blosc chunk of 4 GiB or more");
    if (h.flags & kMemcpyed) {
        if (kHead + h.nbytes > len) refuse("truncated memcpyed block");
        for (std::size_t j = 0; j < h.nblocks; ++j) {
            const std::size_t lo = j * h.blocksize, n = std::min(h.blocksize, h.nbytes - lo);
            streams.push_back(Stream{dsrc + kHead + lo, dst + lo, std::uint32_t(n), std::uint32_t(n)});
        }
        return;
    }
    if (h.nblocks > (len - kHead) / 4) refuse("blosc block table out of range");
    std::vector<std::size_t> starts(h.nblocks);
    for (std::size_t k = 0; k < h.nblocks; ++k) starts[k] = le32(src + kHead + 4 * k);
    std::vector<std::size_t> sorted = starts;
    std::sort(sorted.begin(), sorted.end());
    const std::size_t left = h.nbytes % h.blocksize;
    for (std::size_t j = 0; j < h.nblocks; ++j) {
        const bool last = j + 1 == h.nblocks && left > 0;
        const std::size_t bsize = last ? left : h.blocksize, lo = j * h.blocksize;
        // the block runs to the nearest larger block start, or the chunk end
        const std::size_t start = starts[j];
        const auto nx = std::upper_bound(sorted.begin(), sorted.end(), start);
        const std::size_t end = nx == sorted.end() ? h.cbytes : *nx;
        if (end < start || end > len) refuse("blosc block offset beyond the chunk");
        const bool shuf = (h.flags & kShuffle) && h.typesize > 1;
        const std::size_t nsplit = (!(h.flags & kDontSplit) && h.typesize <= std::size_t(kMaxSplits) &&
                                    bsize / h.typesize >= std::size_t(kMinBuffer) && !last)
                                       ? h.typesize
                                       : 1;
        const std::size_t ne = bsize / nsplit;
        const std::uint64_t target = shuf ? scratch + scratch_used : dst + lo;
        std::size_t at = start;
        for (std::size_t k = 0; k < nsplit; ++k) {
            if (at + 4 > end) refuse("truncated blosc stream");
            const std::size_t cb = le32(src + at);
            at += 4;
            if (cb > end - at) refuse("truncated blosc stream");
            streams.push_back(Stream{dsrc + at, target + k * ne, std::uint32_t(cb), std::uint32_t(ne)});
            at += cb;
        }
        if (shuf) {
            shuffled.push_back(Shuffled{scratch + scratch_used, dst + lo, std::uint32_t(bsize), std::uint32_t(h.typesize)});
            scratch_used += bsize;
        }
    }
}

// Bytes of scratch the shuffled blocks of a chunk need.
std::size_t scratch_of(const std::uint8_t* src, std::size_t len) {
    const Header h = parse_header(src, len);
    if ((h.flags & kMemcpyed) || !((h.flags & kShuffle) && h.typesize > 1)) return 0;
    return h.nbytes;
}

// `resident` (optional): each chunk's bytes already in device memory (the
// host copy in `chunks` is read for the plan only).
int decode_batch(const RangBloscChunk* chunks, std::size_t n, const void* const* resident, hipStream_t st) {
    std::lock_guard<std::mutex> lock(g_pool.mu);
    std::size_t total = 0, scratch = 0;
    std::vector<std::size_t> at(n);
    for (std::size_t i = 0; i < n; ++i) {
        if (!chunks[i].src || (!chunks[i].dst && chunks[i].dst_len) || (resident && !resident[i]))
            refuse("null chunk pointer");
        at[i] = total;
        total += (chunks[i].len + 15) / 16 * 16;
        scratch += scratch_of(chunks[i].src, chunks[i].len);
    }
    std::uint8_t* dsrc = resident ? nullptr : static_cast<std::uint8_t*>(g_pool.get(0, total));
    std::uint8_t* dscr = static_cast<std::uint8_t*>(g_pool.get(1, scratch));
    std::vector<Stream> streams;
    std::vector<Shuffled> shuffled;
    std::size_t used = 0;
    for (std::size_t i = 0; i < n; ++i) {
        const std::uint64_t base = resident ? std::uint64_t(resident[i]) : std::uint64_t(dsrc + at[i]);
        plan_chunk(chunks[i].src, chunks[i].len, base, std::uint64_t(chunks[i].dst), chunks[i].dst_len,
                   std::uint64_t(dscr), streams, shuffled, used);
    }
    Stream* dstreams = static_cast<Stream*>(g_pool.get(2, sizeof(Stream) * streams.size()));
    Shuffled* dshuf = static_cast<Shuffled*>(g_pool.get(3, sizeof(Shuffled) * shuffled.size()));
    // one staged upload of every chunk's bytes
    std::vector<std::uint8_t> host;
    if (!resident) {
        host.resize(total);
        for (std::size_t i = 0; i < n; ++i) std::memcpy(host.data() + at[i], chunks[i].src, chunks[i].len);
        hipchk(hipMemcpyAsync(dsrc, host.data(), total, hipMemcpyHostToDevice, st), "upload");
    }
    if (!streams.empty())
        hipchk(hipMemcpyAsync(dstreams, streams.data(), sizeof(Stream) * streams.size(), hipMemcpyHostToDevice, st),
               "upload");
    if (!shuffled.empty())
        hipchk(hipMemcpyAsync(dshuf, shuffled.data(), sizeof(Shuffled) * shuffled.size(), hipMemcpyHostToDevice, st),
               "upload");
    int zero = 0;
    hipchk(hipMemcpyToSymbolAsync(HIP_SYMBOL(g_blosc_status), &zero, sizeof zero, 0, hipMemcpyHostToDevice, st),
           "status reset");
    if (!streams.empty()) {
        const int blocks = int((streams.size() + kWaves - 1) / kWaves);
        static const bool lds_set = [] {
            hipchk(hipFuncSetAttribute(reinterpret_cast<const void*>(k_lz4_streams),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes),
                   "hipFuncSetAttribute");
            return true;
        }();
        (void)lds_set;
        k_lz4_streams<<<blocks, 64 * kWaves, kLdsBytes, st>>>(dstreams, int(streams.size()));
        hipchk(hipGetLastError(), "k_lz4_streams");
    }
    if (!shuffled.empty()) {
        k_unshuffle<<<unsigned(shuffled.size()), 256, 0, st>>>(dshuf);
        hipchk(hipGetLastError(), "k_unshuffle");
    }
    int status = 0;
    hipchk(hipMemcpyFromSymbolAsync(&status, HIP_SYMBOL(g_blosc_status), sizeof status, 0, hipMemcpyDeviceToHost, st),
           "status read");
    hipchk(hipStreamSynchronize(st), "synchronise");
    if (status) refuse("corrupt LZ4 stream");
    return 0;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const Fail& e) {
        t_blosc_error = e.msg;
        return e.status;
    } catch (const std::exception& e) {
        t_blosc_error = e.what();
        return 2;
    }
}

}  // namespace
}  // namespace rang

using namespace rang;

extern "C" {

size_t rang_amd_blosc_last_error(uint8_t* buf, size_t len) {
    if (buf && len) std::memcpy(buf, t_blosc_error.data(), std::min(len, t_blosc_error.size()));
    return t_blosc_error.size();
}

int32_t rang_core_blosc_size(const uint8_t* src, size_t len, size_t* nbytes) {
    if (!src || !nbytes) return 1;
    return guarded([&] {
        const Header h = parse_header(src, len);
        if (h.cbytes != len) refuse("blosc compressed size disagrees with chunk length");
        *nbytes = h.nbytes;
        return 0;
    });
}

int32_t rang_core_blosc_header(const uint8_t* head, size_t len, size_t* fields) {
    if (!head || !fields) return 1;
    return guarded([&] {
        const Header h = parse_header(head, len);
        fields[0] = h.nbytes;
        fields[1] = h.blocksize;
        fields[2] = h.nblocks;
        fields[3] = h.typesize;
        return 0;
    });
}

int32_t rang_core_blosc_decode(const uint8_t* src, size_t len, uint8_t* out, size_t out_len) {
    if (!src || !out) return 1;
    return guarded([&] {
        void* d = nullptr;
        hipchk(hipMalloc(&d, out_len ? out_len : 1), "hipMalloc");
        struct Free {
            void* p;
            ~Free() { (void)hipFree(p); }
        } guard{d};
        RangBloscChunk c{src, len, d, out_len};
        decode_batch(&c, 1, nullptr, nullptr);
        hipchk(hipMemcpy(out, d, out_len, hipMemcpyDeviceToHost), "download");
        return 0;
    });
}

int32_t rang_amd_blosc_decode_batch(const RangBloscChunk* chunks, size_t n, void* stream) {
    if (!chunks && n) return 1;
    return guarded([&] { return decode_batch(chunks, n, nullptr, static_cast<hipStream_t>(stream)); });
}

int32_t rang_amd_blosc_decode_resident(const RangBloscChunk* chunks, const void* const* dev_src, size_t n,
                                       void* stream) {
    if ((!chunks || !dev_src) && n) return 1;
    return guarded([&] { return decode_batch(chunks, n, dev_src, static_cast<hipStream_t>(stream)); });
}

}  // extern "C"
