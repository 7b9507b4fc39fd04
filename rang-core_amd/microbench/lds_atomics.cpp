// LDS atomic-add throughput on the GPU, single against double precision:
// the measurement behind the gridding kernel's double accumulators.
//
// Each 1024-thread workgroup (one per CU, 1024 of them) clears 64 KiB of
// LDS, then every wave adds into it with 49 of its 64 lanes active, laid
// out as the gridding's 7 x 7 footprint at a row stride of `stride` cells
// (39 is the kernel's padded stride: S = 7 mod 32), two adds per step (the
// real and imaginary accumulators, 1024 cells apart), 4096 steps, the
// footprint moved by 131 cells per step within the array. Prints lane-adds
// per clock per CU from the event-timed duration of one launch and the
// device's clock rate. Every address stays inside the array (the bound is
// the footprint's last cell plus 1024 plus the start offset, checked below).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

template <class A>
constexpr int cells() {
    return 65536 / int(sizeof(A));
}

template <class A>
__global__ __launch_bounds__(1024) void k_adds(A* out, int steps, int stride) {
    constexpr int M = cells<A>();
    __shared__ A s[M];
    for (int i = threadIdx.x; i < M; i += 1024) s[i] = A(0);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool on = lane < 49;
    const int cell = (lane / 7) * stride + lane % 7;
    const unsigned span = unsigned(M - 8 * stride - 1100);
    unsigned base = (unsigned(wave) * 1237u) % span;
    const A v = A(1);
    for (int it = 0; it < steps; ++it) {
        if (on) {
            atomicAdd(&s[base + cell], v);
            atomicAdd(&s[base + cell + 1024], v);
        }
        base = (base + 131u) % span;
    }
    __syncthreads();
    A t = A(0);
    for (int i = threadIdx.x; i < M; i += 1024) t += s[i];
    out[blockIdx.x * 1024 + threadIdx.x] = t;
}

template <class A>
void run(const char* name, int stride) {
    constexpr int M = cells<A>();
    // last address: span - 1 + 6 stride + 6 + 1024 < M
    if (!(M - 8 * stride - 1100 > 0 && (M - 8 * stride - 1101) + 6 * stride + 6 + 1024 < M)) {
        std::fprintf(stderr, "stride %d does not fit\n", stride);
        std::exit(1);
    }
    const int blocks = 1024, steps = 4096;
    A* out;
    CHECK(hipMalloc(&out, sizeof(A) * blocks * 1024));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    k_adds<A><<<blocks, 1024>>>(out, 16, stride);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(a));
    k_adds<A><<<blocks, 1024>>>(out, steps, stride);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    int khz = 0, cus = 0;
    CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, 0));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const double adds = double(blocks) * 16 * 49 * 2 * steps;
    const double clocks = ms * 1e-3 * khz * 1e3 * cus;
    std::printf("%s stride %d: %.3f ms, %.2f lane-adds/clk/CU (%d CUs, %d kHz)\n", name, stride, ms,
                adds / clocks, cus, khz);
    CHECK(hipFree(out));
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
}

int main() {
    for (int st : {39, 40, 33}) {
        run<float>("ds_add_f32", st);
        run<double>("ds_add_f64", st);
    }
    return 0;
}
