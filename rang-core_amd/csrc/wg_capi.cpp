// The C ABI of include/rang_core_wgridder.h: rang-core's ffi_wgridder.rs
// entry points on host buffers (staged to the GPU and back), and the
// device-resident rang_amd_* variants.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <vector>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/rang_core_batch.h"
#include "../../include/rang_core_gram.h"
#include "../../include/rang_core_wgridder.h"
#include "gram.h"
#include "wg_device.h"
#include "wg_plan.h"

using namespace rang;

namespace {

thread_local std::string t_last_error;

int32_t fail(int32_t status, const std::string& msg) {
    t_last_error = msg;
    return status;
}

Options to_options(const WgridderOptions& o) {
    Options r;
    r.nx = o.nx;
    r.ny = o.ny;
    r.pixsize_x = o.pixsize_x;
    r.pixsize_y = o.pixsize_y;
    r.epsilon = o.epsilon;
    r.center_x = o.center_x;
    r.center_y = o.center_y;
    r.sigma_min = o.sigma_min;
    r.sigma_max = o.sigma_max;
    r.do_wgridding = o.do_wgridding != 0;
    r.divide_by_n = o.divide_by_n != 0;
    r.flip_u = o.flip_u != 0;
    r.flip_v = o.flip_v != 0;
    r.flip_w = o.flip_w != 0;
    r.allow_nshift = o.allow_nshift != 0;
    r.kernel_index = o.kernel_index;
    r.nu = o.kernel_index >= 0 ? o.nu : 0;
    r.nv = o.kernel_index >= 0 ? o.nv : 0;
    return r;
}

std::size_t vis_bytes(int kind) { return kind == kVisC64 ? 8 : kind == kVisC128 ? 16 : 0; }
std::size_t wgt_bytes(int kind) {
    return kind == kWgtF32 || kind == kWgtF32Product ? 4 : kind == kWgtF64 ? 8 : 0;
}

// Checks of the kinds and pointers (core/src/ffi_wgridder.rs:318-370).
bool kinds_ok(const void* vis, int vis_kind, const void* wgt, const void* wgt2, int wgt_kind,
              std::string& err) {
    if (!(vis_kind == kVisUnit || ((vis_kind == kVisC64 || vis_kind == kVisC128) && vis))) {
        err = "invalid visibility array or kind";
        return false;
    }
    const bool wok = wgt_kind == kWgtNone || ((wgt_kind == kWgtF32 || wgt_kind == kWgtF64) && wgt) ||
                     (wgt_kind == kWgtF32Product && wgt && wgt2);
    if (!wok) {
        err = "invalid weight array or kind";
        return false;
    }
    return true;
}

// The checks that need no GPU, in the reference's order
// (core/src/wgridder.rs:2449-2462): options, then the frequencies.
bool host_checks(const Options& o, std::size_t nrow, std::size_t nchan, const double* freq,
                 std::string& err) {
    if (!check_common(o, nrow, nchan, err)) return false;
    if (nrow * nchan >= (std::size_t(1) << 31)) {
        err = "too many visibilities for one call (at most 2^31 - 1)";
        return false;
    }
    return check_freq(freq, nchan, err);
}

// Host staging buffers of one device and the stream the host-buffer entry
// points run on.
struct Staging {
    std::mutex mu;
    hipStream_t stream = nullptr;
    void* ptr[8] = {};
    std::size_t size[8] = {};
    void* get(int slot, std::size_t bytes) {
        if (bytes == 0) return nullptr;
        if (size[slot] < bytes) {
            if (ptr[slot]) (void)hipFree(ptr[slot]);
            ptr[slot] = nullptr;
            size[slot] = 0;
            if (hipMalloc(&ptr[slot], bytes) != hipSuccess) return nullptr;
            size[slot] = bytes;
        }
        return ptr[slot];
    }
};

Staging* staging() {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<Staging>> all;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    auto& s = all[dev];
    if (!s) {
        s.reset(new Staging());
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
            s.reset();
            return nullptr;
        }
    }
    return s.get();
}

struct Stage {
    Staging* s;
    std::string err;
    bool ok = true;
    template <class P>
    P* up(int slot, const P* host, std::size_t bytes) {
        if (!host || bytes == 0 || !ok) return nullptr;
        void* d = s->get(slot, bytes);
        if (!d || hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
            ok = false;
            err = "staging to the GPU failed";
            return nullptr;
        }
        return static_cast<P*>(d);
    }
    void* out(int slot, const void* host, std::size_t bytes) {
        if (!host || bytes == 0 || !ok) return nullptr;
        void* d = s->get(slot, bytes);
        if (!d) {
            ok = false;
            err = "device allocation failed";
        }
        return d;
    }
    bool down(void* host, const void* dev, std::size_t bytes) {
        if (!host || !dev || bytes == 0) return true;
        if (hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s->stream) != hipSuccess) return false;
        return true;
    }
};

template <class T>
int32_t v2d_host(const WgridderOptions* opts, std::size_t nrow, std::size_t nchan,
                 const double* uvw, const double* freq, const void* vis, int32_t vis_kind,
                 const void* wgt, const void* wgt2, int32_t wgt_kind, const std::uint8_t* mask,
                 T* dirty, T* psf, double* info) {
    try {
        if (!opts) return fail(1, "null options");
        if (!uvw || !freq || !info || (!dirty && !psf)) return fail(1, "null array");
        std::string err;
        if (!kinds_ok(vis, vis_kind, wgt, wgt2, wgt_kind, err)) return fail(1, err);
        const Options o = to_options(*opts);
        if (!host_checks(o, nrow, nchan, freq, err)) return fail(1, err);
        Staging* s = staging();
        if (!s) return fail(2, "no HIP device");
        std::lock_guard<std::mutex> lk(s->mu);
        const std::size_t n = nrow * nchan, npix = o.nx * o.ny;
        Stage st{s};
        DevInputs in;
        in.nrow = nrow;
        in.nchan = nchan;
        in.uvw = st.up(0, uvw, sizeof(double) * 3 * nrow);
        in.freq = st.up(1, freq, sizeof(double) * nchan);
        in.vis = st.up(2, static_cast<const char*>(vis), vis_bytes(vis_kind) * n);
        in.vis_kind = vis_kind;
        in.wgt = st.up(3, static_cast<const char*>(wgt), wgt_bytes(wgt_kind) * n);
        in.wgt2 = wgt_kind == kWgtF32Product ? st.up(4, static_cast<const char*>(wgt2), 4 * n) : nullptr;
        in.wgt_kind = wgt_kind;
        in.mask = st.up(5, mask, n);
        T* ddirty = static_cast<T*>(st.out(6, dirty, sizeof(T) * npix));
        T* dpsf = static_cast<T*>(st.out(7, psf, sizeof(T) * npix));
        if (!st.ok) return fail(2, st.err);
        const int32_t status = device_vis2dirty<T>(o, in, freq, ddirty, dpsf, info, s->stream, err);
        if (status != 0) return fail(status, err);
        if (!st.down(dirty, ddirty, sizeof(T) * npix) || !st.down(psf, dpsf, sizeof(T) * npix) ||
            hipStreamSynchronize(s->stream) != hipSuccess)
            return fail(2, "copy from the GPU failed");
        return 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

template <class T>
int32_t d2v_host(const WgridderOptions* opts, std::size_t nrow, std::size_t nchan,
                 const double* uvw, const double* freq, const T* dirty, const void* wgt,
                 int32_t wgt_kind, const std::uint8_t* mask, T* vis, double* info) {
    try {
        if (!opts) return fail(1, "null options");
        if (!uvw || !freq || !vis || !dirty || !info) return fail(1, "null array");
        std::string err;
        if (wgt_kind == kWgtF32Product || !kinds_ok(nullptr, kVisUnit, wgt, nullptr, wgt_kind, err))
            return fail(1, err.empty() ? "invalid weight array or kind" : err);
        const Options o = to_options(*opts);
        if (!host_checks(o, nrow, nchan, freq, err)) return fail(1, err);
        Staging* s = staging();
        if (!s) return fail(2, "no HIP device");
        std::lock_guard<std::mutex> lk(s->mu);
        const std::size_t n = nrow * nchan, npix = o.nx * o.ny;
        Stage st{s};
        DevInputs in;
        in.nrow = nrow;
        in.nchan = nchan;
        in.uvw = st.up(0, uvw, sizeof(double) * 3 * nrow);
        in.freq = st.up(1, freq, sizeof(double) * nchan);
        in.wgt = st.up(3, static_cast<const char*>(wgt), wgt_bytes(wgt_kind) * n);
        in.wgt_kind = wgt_kind;
        in.mask = st.up(5, mask, n);
        const T* ddirty = st.up(6, dirty, sizeof(T) * npix);
        T* dvis = static_cast<T*>(st.out(7, vis, sizeof(T) * 2 * n));
        if (!st.ok) return fail(2, st.err);
        if (!dvis && n > 0) return fail(2, "device allocation failed");
        if (n == 0) {
            // nothing to degrid; the reference still checks the options
            std::string e2;
            if (!check_common(o, nrow, nchan, e2)) return fail(1, e2);
            if (!check_freq(freq, nchan, e2)) return fail(1, e2);
            std::fill(info, info + kNInfo, 0.0);
            return 0;
        }
        const int32_t status = device_dirty2vis<T>(o, in, freq, ddirty, dvis, info, s->stream, err);
        if (status != 0) return fail(status, err);
        if (!st.down(vis, dvis, sizeof(T) * 2 * n) || hipStreamSynchronize(s->stream) != hipSuccess)
            return fail(2, "copy from the GPU failed");
        return 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

// The channel frequencies from the device, read in order on the call's
// stream (a write to them queued on that stream is seen).
bool freq_to_host(const double* freq_dev, std::size_t nchan, std::vector<double>& out,
                  void* stream, std::string& err) {
    out.assign(nchan, 0.0);
    if (!nchan) return true;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(out.data(), freq_dev, sizeof(double) * nchan, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        err = "cannot read the channel frequencies from the device";
        return false;
    }
    return true;
}

template <class T>
int32_t v2d_dev(const WgridderOptions* opts, std::size_t nrow, std::size_t nchan,
                const double* uvw, const double* freq, const void* vis, int32_t vis_kind,
                const void* wgt, const void* wgt2, int32_t wgt_kind, const std::uint8_t* mask,
                T* dirty, T* psf, double* info, void* stream, bool fixed = false) {
    try {
        if (!opts) return fail(1, "null options");
        if (!uvw || !freq || !info || (!dirty && !psf)) return fail(1, "null array");
        std::string err;
        if (!kinds_ok(vis, vis_kind, wgt, wgt2, wgt_kind, err)) return fail(1, err);
        std::vector<double> fh;
        if (!freq_to_host(freq, nchan, fh, stream, err)) return fail(2, err);
        DevInputs in;
        in.nrow = nrow;
        in.nchan = nchan;
        in.uvw = uvw;
        in.freq = freq;
        in.vis = vis;
        in.vis_kind = vis_kind;
        in.wgt = wgt;
        in.wgt2 = wgt2;
        in.wgt_kind = wgt_kind;
        in.mask = mask;
        int32_t status;
        if constexpr (sizeof(T) == 4) {
            status = fixed ? device_vis2dirty_fixed<T>(to_options(*opts), in, fh.data(), dirty, psf, info,
                                                       static_cast<hipStream_t>(stream), err)
                           : device_vis2dirty<T>(to_options(*opts), in, fh.data(), dirty, psf, info,
                                                 static_cast<hipStream_t>(stream), err);
        } else {
            status = device_vis2dirty<T>(to_options(*opts), in, fh.data(), dirty, psf, info,
                                         static_cast<hipStream_t>(stream), err);
        }
        return status ? fail(status, err) : 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

template <class T>
int32_t d2v_dev(const WgridderOptions* opts, std::size_t nrow, std::size_t nchan,
                const double* uvw, const double* freq, const T* dirty, const void* wgt,
                int32_t wgt_kind, const std::uint8_t* mask, T* vis, double* info, void* stream) {
    try {
        if (!opts) return fail(1, "null options");
        if (!uvw || !freq || !vis || !dirty || !info) return fail(1, "null array");
        std::string err;
        if (wgt_kind == kWgtF32Product || !kinds_ok(nullptr, kVisUnit, wgt, nullptr, wgt_kind, err))
            return fail(1, err.empty() ? "invalid weight array or kind" : err);
        std::vector<double> fh;
        if (!freq_to_host(freq, nchan, fh, stream, err)) return fail(2, err);
        DevInputs in;
        in.nrow = nrow;
        in.nchan = nchan;
        in.uvw = uvw;
        in.freq = freq;
        in.wgt = wgt;
        in.wgt_kind = wgt_kind;
        in.mask = mask;
        const int32_t status = device_dirty2vis<T>(to_options(*opts), in, fh.data(), dirty, vis, info,
                                                   static_cast<hipStream_t>(stream), err);
        return status ? fail(status, err) : 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}


// The plan of a call on host buffers, without gridding (wgridder::plan,
// core/src/wgridder.rs:3525-3549; with nout > 0, gridding_footprint,
// :3551-3586, on float64 weights).
template <class T>
int32_t plan_host(const WgridderOptions* opts, std::size_t nrow, std::size_t nchan, const double* uvw,
                  const double* freq, const void* vis, int32_t vis_kind, const void* wgt, int32_t wgt_kind,
                  const std::uint8_t* mask, int nout, double* info, std::size_t* bytes) {
    try {
        if (!opts) return fail(1, "null options");
        if (!uvw || !freq || !info || (nout > 0 && !bytes)) return fail(1, "null array");
        std::string err;
        if (!kinds_ok(vis, vis_kind, wgt, nullptr, wgt_kind, err)) return fail(1, err);
        const Options o = to_options(*opts);
        if (!host_checks(o, nrow, nchan, freq, err)) return fail(1, err);
        Staging* s = staging();
        if (!s) return fail(2, "no HIP device");
        std::lock_guard<std::mutex> lk(s->mu);
        const std::size_t n = nrow * nchan;
        Stage st{s};
        DevInputs in;
        in.nrow = nrow;
        in.nchan = nchan;
        in.uvw = st.up(0, uvw, sizeof(double) * 3 * nrow);
        in.freq = st.up(1, freq, sizeof(double) * nchan);
        in.vis = st.up(2, static_cast<const char*>(vis), vis_bytes(vis_kind) * n);
        in.vis_kind = vis_kind;
        in.wgt = st.up(3, static_cast<const char*>(wgt), wgt_bytes(wgt_kind) * n);
        in.wgt_kind = wgt_kind;
        in.mask = st.up(5, mask, n);
        if (!st.ok) return fail(2, st.err);
        const int32_t status = device_plan<T>(o, in, freq, vis_kind != kVisUnit, nout, info, bytes,
                                              s->stream, err);
        return status ? fail(status, err) : 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

template <class T>
RangPrepared* prepare_dev(const WgridderOptions* opts, std::size_t nrow, std::size_t nchan,
                          const double* uvw, const double* freq, const void* wgt, int32_t wgt_kind,
                          void* stream) {
    try {
        if (!opts) return fail(1, "null options"), nullptr;
        if (!uvw || !freq) return fail(1, "null array"), nullptr;
        std::string err;
        if (!kinds_ok(nullptr, kVisUnit, wgt, nullptr, wgt_kind, err)) return fail(1, err), nullptr;
        std::vector<double> fh;
        if (!freq_to_host(freq, nchan, fh, stream, err)) return fail(2, err), nullptr;
        DevInputs in;
        in.nrow = nrow;
        in.nchan = nchan;
        in.uvw = uvw;
        in.freq = freq;
        in.wgt = wgt;
        in.wgt_kind = wgt_kind;
        int status = 0;
        PreparedBase* p = prepared_new<T>(to_options(*opts), in, fh.data(),
                                          static_cast<hipStream_t>(stream), status, err);
        if (!p) fail(status ? status : 2, err);
        return reinterpret_cast<RangPrepared*>(p);
    } catch (const std::exception& e) {
        fail(2, std::string("failure in the gridder: ") + e.what());
        return nullptr;
    }
}

template <class T>
PreparedBase* handle(RangPrepared* h, std::string& err) {
    PreparedBase* p = reinterpret_cast<PreparedBase*>(h);
    if (!p) {
        err = "null prepared plan";
        return nullptr;
    }
    if (p->single() != (sizeof(T) == 4)) {
        err = "prepared plan of the other precision";
        return nullptr;
    }
    return p;
}

template <class T, class F>
int32_t prepared_call(RangPrepared* h, F&& f) {
    try {
        std::string err;
        PreparedBase* p = handle<T>(h, err);
        if (!p) return fail(1, err);
        const int32_t status = f(p, err);
        return status ? fail(status, err) : 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

bool vis_ok(const void* vis, int32_t kind, std::string& err) {
    if (!vis || !(kind == kVisC64 || kind == kVisC128)) {
        err = "invalid visibility array or kind";
        return false;
    }
    return true;
}

}  // namespace

extern "C" {


int32_t rang_core_wgridder_vis2dirty_f64(const WgridderOptions* opts, size_t nrow, size_t nchan,
                                         const double* uvw, const double* freq, const double* vis,
                                         const double* wgt, const uint8_t* mask, double* dirty,
                                         double* info) {
    if (!vis || !dirty) return fail(1, "null array");
    return v2d_host<double>(opts, nrow, nchan, uvw, freq, vis, kVisC128, wgt, nullptr,
                            wgt ? kWgtF64 : kWgtNone, mask, dirty, nullptr, info);
}

int32_t rang_core_wgridder_vis2dirty_f32(const WgridderOptions* opts, size_t nrow, size_t nchan,
                                         const double* uvw, const double* freq, const float* vis,
                                         const float* wgt, const uint8_t* mask, float* dirty,
                                         double* info) {
    if (!vis || !dirty) return fail(1, "null array");
    return v2d_host<float>(opts, nrow, nchan, uvw, freq, vis, kVisC64, wgt, nullptr,
                           wgt ? kWgtF32 : kWgtNone, mask, dirty, nullptr, info);
}

int32_t rang_core_wgridder_dirty2vis_f64(const WgridderOptions* opts, size_t nrow, size_t nchan,
                                         const double* uvw, const double* freq, const double* dirty,
                                         const double* wgt, const uint8_t* mask, double* vis,
                                         double* info) {
    return d2v_host<double>(opts, nrow, nchan, uvw, freq, dirty, wgt, wgt ? kWgtF64 : kWgtNone, mask,
                            vis, info);
}

int32_t rang_core_wgridder_dirty2vis_f32(const WgridderOptions* opts, size_t nrow, size_t nchan,
                                         const double* uvw, const double* freq, const float* dirty,
                                         const float* wgt, const uint8_t* mask, float* vis,
                                         double* info) {
    return d2v_host<float>(opts, nrow, nchan, uvw, freq, dirty, wgt, wgt ? kWgtF32 : kWgtNone, mask,
                           vis, info);
}

int32_t rang_core_wgridder_vis2dirty_ex_f64(const WgridderOptions* opts, size_t nrow, size_t nchan,
                                            const double* uvw, const double* freq, const void* vis,
                                            int32_t vis_kind, const void* wgt, int32_t wgt_kind,
                                            const uint8_t* mask, double* dirty, double* psf,
                                            double* info) {
    return v2d_host<double>(opts, nrow, nchan, uvw, freq, vis, vis_kind, wgt, nullptr, wgt_kind,
                            mask, dirty, psf, info);
}

int32_t rang_core_wgridder_vis2dirty_ex_f32(const WgridderOptions* opts, size_t nrow, size_t nchan,
                                            const double* uvw, const double* freq, const void* vis,
                                            int32_t vis_kind, const void* wgt, int32_t wgt_kind,
                                            const uint8_t* mask, float* dirty, float* psf,
                                            double* info) {
    return v2d_host<float>(opts, nrow, nchan, uvw, freq, vis, vis_kind, wgt, nullptr, wgt_kind,
                           mask, dirty, psf, info);
}

int32_t rang_core_wgridder_vis2dirty_ex2_f64(const WgridderOptions* opts, size_t nrow,
                                             size_t nchan, const double* uvw, const double* freq,
                                             const void* vis, int32_t vis_kind, const void* wgt,
                                             const void* wgt2, int32_t wgt_kind,
                                             const uint8_t* mask, double* dirty, double* psf,
                                             double* info) {
    return v2d_host<double>(opts, nrow, nchan, uvw, freq, vis, vis_kind, wgt, wgt2, wgt_kind,
                            mask, dirty, psf, info);
}

int32_t rang_core_wgridder_vis2dirty_ex2_f32(const WgridderOptions* opts, size_t nrow,
                                             size_t nchan, const double* uvw, const double* freq,
                                             const void* vis, int32_t vis_kind, const void* wgt,
                                             const void* wgt2, int32_t wgt_kind,
                                             const uint8_t* mask, float* dirty, float* psf,
                                             double* info) {
    return v2d_host<float>(opts, nrow, nchan, uvw, freq, vis, vis_kind, wgt, wgt2, wgt_kind, mask,
                           dirty, psf, info);
}

int32_t rang_core_wgridder_dirty2vis_ex_f64(const WgridderOptions* opts, size_t nrow, size_t nchan,
                                            const double* uvw, const double* freq,
                                            const double* dirty, const void* wgt, int32_t wgt_kind,
                                            const uint8_t* mask, double* vis, double* info) {
    return d2v_host<double>(opts, nrow, nchan, uvw, freq, dirty, wgt, wgt_kind, mask, vis, info);
}

int32_t rang_core_wgridder_dirty2vis_ex_f32(const WgridderOptions* opts, size_t nrow, size_t nchan,
                                            const double* uvw, const double* freq,
                                            const float* dirty, const void* wgt, int32_t wgt_kind,
                                            const uint8_t* mask, float* vis, double* info) {
    return d2v_host<float>(opts, nrow, nchan, uvw, freq, dirty, wgt, wgt_kind, mask, vis, info);
}

int32_t rang_core_wgridder_kernel_row(size_t index, double* out) {
    const KernelRow* r = kernel_row(index);
    if (!r || !out) return 1;
    const double v[7] = {double(r->w), r->ofactor, r->epsilon, r->beta, r->e0, double(r->ndim),
                         double(r->single)};
    std::memcpy(out, v, sizeof v);
    return 0;
}

size_t rang_core_wgridder_last_error(uint8_t* buf, size_t len) {
    if (buf) std::memcpy(buf, t_last_error.data(), std::min(len, t_last_error.size()));
    return t_last_error.size();
}

int32_t rang_amd_wgridder_vis2dirty_dev_f64(const WgridderOptions* opts, size_t nrow,
                                            size_t nchan, const double* uvw, const double* freq,
                                            const void* vis, int32_t vis_kind, const void* wgt,
                                            const void* wgt2, int32_t wgt_kind,
                                            const uint8_t* mask, double* dirty, double* psf,
                                            double* info, void* stream) {
    return v2d_dev<double>(opts, nrow, nchan, uvw, freq, vis, vis_kind, wgt, wgt2, wgt_kind, mask,
                           dirty, psf, info, stream);
}

int32_t rang_amd_wgridder_vis2dirty_dev_f32(const WgridderOptions* opts, size_t nrow,
                                            size_t nchan, const double* uvw, const double* freq,
                                            const void* vis, int32_t vis_kind, const void* wgt,
                                            const void* wgt2, int32_t wgt_kind,
                                            const uint8_t* mask, float* dirty, float* psf,
                                            double* info, void* stream) {
    return v2d_dev<float>(opts, nrow, nchan, uvw, freq, vis, vis_kind, wgt, wgt2, wgt_kind, mask,
                          dirty, psf, info, stream);
}

int32_t rang_amd_wgridder_vis2dirty_noalloc_f32(const WgridderOptions* opts, size_t nrow,
                                                size_t nchan, const double* uvw, const double* freq,
                                                const void* vis, int32_t vis_kind, const void* wgt,
                                                const void* wgt2, int32_t wgt_kind,
                                                const uint8_t* mask, float* dirty, float* psf,
                                                double* info, void* stream) {
    return v2d_dev<float>(opts, nrow, nchan, uvw, freq, vis, vis_kind, wgt, wgt2, wgt_kind, mask,
                          dirty, psf, info, stream, true);
}

int32_t rang_amd_wgridder_dirty2vis_dev_f64(const WgridderOptions* opts, size_t nrow,
                                            size_t nchan, const double* uvw, const double* freq,
                                            const double* dirty, const void* wgt,
                                            int32_t wgt_kind, const uint8_t* mask, double* vis,
                                            double* info, void* stream) {
    return d2v_dev<double>(opts, nrow, nchan, uvw, freq, dirty, wgt, wgt_kind, mask, vis, info,
                           stream);
}

int32_t rang_amd_wgridder_dirty2vis_dev_f32(const WgridderOptions* opts, size_t nrow,
                                            size_t nchan, const double* uvw, const double* freq,
                                            const float* dirty, const void* wgt, int32_t wgt_kind,
                                            const uint8_t* mask, float* vis, double* info,
                                            void* stream) {
    return d2v_dev<float>(opts, nrow, nchan, uvw, freq, dirty, wgt, wgt_kind, mask, vis, info,
                          stream);
}


#define RANG_PREPARED_ABI(SUF, T)                                                                   \
    RangPrepared* rang_amd_wgridder_prepare_dev_##SUF(const WgridderOptions* opts, size_t nrow,    \
                                                      size_t nchan, const double* uvw,            \
                                                      const double* freq, const void* wgt,        \
                                                      int32_t wgt_kind, void* stream) {           \
        return prepare_dev<T>(opts, nrow, nchan, uvw, freq, wgt, wgt_kind, stream);                \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_vis2dirty_##SUF(RangPrepared* h, const void* vis,           \
                                                       int32_t vis_kind, T* dirty, double* info,   \
                                                       void* stream) {                             \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!vis_ok(vis, vis_kind, err)) return 1;                                             \
            if (!dirty || !info) return err = "null array", 1;                                     \
            return prepared_vis2dirty<T>(p, vis, vis_kind, dirty, info,                            \
                                         static_cast<hipStream_t>(stream), err);                   \
        });                                                                                         \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_dirty2vis_##SUF(RangPrepared* h, const T* dirty,            \
                                                       int32_t weighted, T* vis, void* stream) {   \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!dirty || !vis) return err = "null array", 1;                                      \
            return prepared_dirty2vis<T>(p, dirty, weighted != 0, vis,                             \
                                         static_cast<hipStream_t>(stream), err);                   \
        });                                                                                         \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_residual_##SUF(RangPrepared* h, const void* vis,            \
                                                      int32_t vis_kind, const T* model, T* dirty,  \
                                                      double* misfit, void* stream) {              \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!vis_ok(vis, vis_kind, err)) return 1;                                             \
            if (!dirty || !misfit) return err = "null array", 1;                                   \
            return prepared_residual<T>(p, vis, vis_kind, model, dirty, misfit,                    \
                                        static_cast<hipStream_t>(stream), err);                    \
        });                                                                                         \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_residual_sorted_##SUF(RangPrepared* h, const void* vis,     \
                                                             int32_t vis_kind, const T* model,     \
                                                             T* dirty, double* misfit,             \
                                                             void* stream) {                       \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!vis_ok(vis, vis_kind, err)) return 1;                                             \
            if (!dirty || !misfit) return err = "null array", 1;                                   \
            return prepared_residual_sorted<T>(p, vis, vis_kind, model, dirty, misfit,             \
                                               static_cast<hipStream_t>(stream), err);             \
        });                                                                                         \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_sorted_##SUF(RangPrepared* h, const void* dense,            \
                                                    size_t elem_bytes, void* out, void* stream) {  \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!dense || !out) return err = "null array", 1;                                      \
            if (elem_bytes != 4 && elem_bytes != 8 && elem_bytes != 16)                             \
                return err = "elements of 4, 8 or 16 bytes", 1;                                     \
            return prepared_sorted<T>(p, dense, int(elem_bytes), out,                              \
                                      static_cast<hipStream_t>(stream), err);                      \
        });                                                                                         \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_gram_##SUF(RangPrepared* h, const T* x, T* out,             \
                                                  void* stream) {                                  \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!x || !out) return err = "null array", 1;                                          \
            return prepared_gram<T>(p, x, out, static_cast<hipStream_t>(stream), err);             \
        });                                                                                         \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_cg_##SUF(RangPrepared* h, const T* rhs, T* x,               \
                                                int32_t iterations, double tolerance,              \
                                                int32_t* used, double* rel, void* stream) {        \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!rhs || !x || !used || !rel) return err = "null array", 1;                         \
            int u = 0;                                                                              \
            const int st = prepared_cg<T>(p, rhs, x, iterations, tolerance, &u, rel,               \
                                          static_cast<hipStream_t>(stream), err);                  \
            *used = u;                                                                              \
            return st;                                                                              \
        });                                                                                         \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_fista_##SUF(RangPrepared* h, const T* dirty, double q0,     \
                                                   double lambda, const T* factors, int32_t sgn,   \
                                                   const RangFistaSettings* settings, T* x, T* gx, \
                                                   RangFistaRecord* records, size_t nrecords,      \
                                                   RangFistaOutcome* out, void* stream) {          \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!dirty || !settings || !x || !out) return err = "null array", 1;                   \
            if (!(settings->step > 0.0)) return err = "the step must be positive", 1;              \
            if (!(lambda >= 0.0)) return err = "the regularisation weight must be non-negative", 1; \
            return prepared_fista<T>(p, dirty, q0, lambda, factors, sgn != 0, *settings, x, gx,    \
                                     records, nrecords, out, static_cast<hipStream_t>(stream), err); \
        });                                                                                         \
    }                                                                                               \
    int32_t rang_amd_wgridder_prepared_power_##SUF(RangPrepared* h, uint64_t seed, double tolerance, \
                                                   uint64_t max_iterations, double* value,          \
                                                   uint64_t* used, void* stream) {                  \
        return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {                        \
            if (!value || !used) return err = "null array", 1;                                     \
            if (max_iterations == 0) return err = "at least one iteration", 1;                     \
            return prepared_power<T>(p, seed, tolerance, max_iterations, value, used,              \
                                     static_cast<hipStream_t>(stream), err);                       \
        });                                                                                         \
    }
RANG_PREPARED_ABI(f64, double)
RANG_PREPARED_ABI(f32, float)
#undef RANG_PREPARED_ABI

#define RANG_PLAN_ABI(SUF, T)                                                                       \
    int32_t rang_amd_wgridder_plan_##SUF(const WgridderOptions* opts, size_t nrow, size_t nchan,       \
                                         const double* uvw, const double* freq, const void* vis,       \
                                         int32_t vis_kind, const void* wgt, int32_t wgt_kind,          \
                                         const uint8_t* mask, double* info) {                          \
        return plan_host<T>(opts, nrow, nchan, uvw, freq, vis, vis_kind, wgt, wgt_kind, mask, 0, info, \
                            nullptr);                                                                  \
    }                                                                                                  \
    int32_t rang_amd_wgridder_footprint_##SUF(const WgridderOptions* opts, size_t nrow, size_t nchan,  \
                                              const double* uvw, const double* freq,                   \
                                              const double* wgt, size_t nout, double* info,            \
                                              size_t* bytes) {                                         \
        if (!wgt) return fail(1, "null array");                                                        \
        if (nout == 0 || nout > 2) return fail(1, "one or two images");                                \
        return plan_host<T>(opts, nrow, nchan, uvw, freq, nullptr, kVisUnit, wgt, kWgtF64, nullptr,    \
                            int(nout), info, bytes);                                                   \
    }
RANG_PLAN_ABI(f64, double)
RANG_PLAN_ABI(f32, float)
#undef RANG_PLAN_ABI

// ------------------------------------------------------------ Gram operator

int32_t rang_amd_wexpansion(int32_t stencil, double theta, double ridge, double* sup_error, double* lebesgue,
                            double* table) {
    try {
        if (!sup_error || !lebesgue) return fail(1, "null array");
        WExpansion ex;
        std::string err;
        if (!wexpansion_build(stencil, theta, ridge, ex, err)) return fail(1, err);
        *sup_error = ex.sup_error;
        *lebesgue = ex.lebesgue;
        if (table) std::memcpy(table, ex.table.data(), sizeof(double) * ex.table.size());
        return 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the Gram operator: ") + e.what());
    }
}

}  // extern "C"

namespace {

template <class T>
RangGram* gram_make(const WgridderOptions* opts, std::size_t nrow, std::size_t nchan, const double* uvw,
                    const double* freq, const void* wgt, int32_t wgt_kind, bool with_w, int32_t stencil,
                    double theta, double ridge, double epsilon, void* stream) {
    try {
        if (!opts) return fail(1, "null options"), nullptr;
        if (!uvw || !freq || !wgt) return fail(1, "null array"), nullptr;
        if (wgt_kind != kWgtF32 && wgt_kind != kWgtF64) return fail(1, "invalid weight array or kind"), nullptr;
        std::string err;
        WExpansion ex;
        if (with_w && !wexpansion_build(stencil, theta, ridge, ex, err)) return fail(1, err), nullptr;
        std::vector<double> fh;
        if (!freq_to_host(freq, nchan, fh, stream, err)) return fail(2, err), nullptr;
        DevInputs in;
        in.nrow = nrow;
        in.nchan = nchan;
        in.uvw = uvw;
        in.freq = freq;
        in.wgt = wgt;
        in.wgt_kind = wgt_kind;
        int status = 0;
        GramBase* g = gram_new<T>(to_options(*opts), in, fh.data(), with_w ? &ex : nullptr, epsilon,
                                  static_cast<hipStream_t>(stream), status, err);
        if (!g) fail(status ? status : 2, err);
        return reinterpret_cast<RangGram*>(g);
    } catch (const std::exception& e) {
        fail(2, std::string("failure in the Gram operator: ") + e.what());
        return nullptr;
    }
}

template <class T>
int32_t gram_apply_abi(RangGram* h, const T* x, T* out, void* stream) {
    try {
        GramBase* g = reinterpret_cast<GramBase*>(h);
        if (!g || !x || !out) return fail(1, "null array");
        if (g->single() != (sizeof(T) == 4)) return fail(1, "Gram operator of the other precision");
        std::string err;
        const int st = gram_apply_dev<T>(g, x, out, static_cast<hipStream_t>(stream), err);
        return st ? fail(st, err) : 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the Gram operator: ") + e.what());
    }
}

bool same_side(const PreparedBase* p, const GramBase* g) {
    return p->image_nx() == p->image_ny() && p->image_nx() == std::size_t(g->side());
}

template <class T>
int32_t defect_abi(RangPrepared* h, RangGram* m, const T* dirty, double q0, const T* factors, int32_t sgn,
                   const RangFistaSettings* inner, const RangOuterSchedule* schedule, T* x, T* gx,
                   RangOuterRecord* records, size_t nrecords, RangDefectOutcome* out, void* stream) {
    return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {
        GramBase* g = reinterpret_cast<GramBase*>(m);
        if (!dirty || !inner || !schedule || !x || !out || !g) return err = "null array", 1;
        if (g->single() != (sizeof(T) == 4)) return err = "Gram operator of the other precision", 1;
        if (!same_side(p, g)) return err = "the Gram operator's image side differs from the plan's", 1;
        if (!(inner->step > 0.0)) return err = "the step must be positive", 1;
        if (!(schedule->lambda >= 0.0)) return err = "the regularisation weight must be non-negative", 1;
        if (!(schedule->gamma >= 0.0)) return err = "gamma must be non-negative", 1;
        std::function<void(const T*, T*, hipStream_t)> cheap = [g](const T* in, T* o, hipStream_t s) {
            std::string e;
            if (gram_apply_dev<T>(g, in, o, s, e)) throw std::runtime_error(e);
        };
        return prepared_defect<T>(p, cheap, dirty, q0, factors, sgn != 0, *inner, *schedule, x, gx, records,
                                  nrecords, out, static_cast<hipStream_t>(stream), err);
    });
}

template <class T>
int32_t pcg_abi(RangPrepared* h, RangGram* m, const T* rhs, T* x, int32_t iterations, double tolerance,
                int32_t* used, double* rel, void* stream) {
    return prepared_call<T>(h, [&](PreparedBase* p, std::string& err) {
        GramBase* g = reinterpret_cast<GramBase*>(m);
        if (!rhs || !x || !used || !rel || !g) return err = "null array", 1;
        if (g->single() != (sizeof(T) == 4)) return err = "Gram operator of the other precision", 1;
        if (!same_side(p, g)) return err = "the Gram operator's image side differs from the plan's", 1;
        int u = 0;
        std::function<void(const T*, T*, hipStream_t)> pre = [g](const T* r, T* z, hipStream_t s) {
            std::string e;
            if (gram_apply_dev<T>(g, r, z, s, e)) throw std::runtime_error(e);
        };
        const int st = prepared_pcg<T>(p, pre, rhs, x, iterations, tolerance, &u, rel,
                                       static_cast<hipStream_t>(stream), err);
        *used = u;
        return st;
    });
}

}  // namespace

extern "C" {

#define RANG_GRAM_ABI(SUF, T)                                                                          \
    RangGram* rang_amd_gram_new_##SUF(const WgridderOptions* opts, size_t nrow, size_t nchan,          \
                                      const double* uvw, const double* freq, const void* wgt,          \
                                      int32_t wgt_kind, int32_t stencil, double theta, double ridge,   \
                                      double epsilon, void* stream) {                                  \
        return gram_make<T>(opts, nrow, nchan, uvw, freq, wgt, wgt_kind, true, stencil, theta, ridge,   \
                            epsilon, stream);                                                          \
    }                                                                                                   \
    RangGram* rang_amd_gram_no_w_##SUF(const WgridderOptions* opts, size_t nrow, size_t nchan,         \
                                       const double* uvw, const double* freq, const void* wgt,         \
                                       int32_t wgt_kind, double epsilon, void* stream) {               \
        return gram_make<T>(opts, nrow, nchan, uvw, freq, wgt, wgt_kind, false, 0, 0.0, 0.0, epsilon,  \
                            stream);                                                                   \
    }                                                                                                   \
    int32_t rang_amd_gram_apply_##SUF(RangGram* g, const T* x, T* out, void* stream) {                 \
        return gram_apply_abi<T>(g, x, out, stream);                                                    \
    }                                                                                                   \
    int32_t rang_amd_wgridder_prepared_pcg_##SUF(RangPrepared* p, RangGram* m, const T* rhs, T* x,     \
                                                 int32_t iterations, double tolerance, int32_t* used,  \
                                                 double* rel, void* stream) {                          \
        return pcg_abi<T>(p, m, rhs, x, iterations, tolerance, used, rel, stream);                      \
    }                                                                                                   \
    int32_t rang_amd_wgridder_prepared_defect_##SUF(                                                   \
        RangPrepared* p, RangGram* cheap, const T* dirty, double q0, const T* factors, int32_t sgn,     \
        const RangFistaSettings* inner, const RangOuterSchedule* schedule, T* x, T* gx,                 \
        RangOuterRecord* records, size_t nrecords, RangDefectOutcome* out, void* stream) {              \
        return defect_abi<T>(p, cheap, dirty, q0, factors, sgn, inner, schedule, x, gx, records,       \
                             nrecords, out, stream);                                                    \
    }
RANG_GRAM_ABI(f64, double)
RANG_GRAM_ABI(f32, float)
#undef RANG_GRAM_ABI

RangGram* rang_amd_gram_inverse_filter(const RangGram* h, double rho, void* stream) {
    try {
        const GramBase* g = reinterpret_cast<const GramBase*>(h);
        if (!g) return fail(1, "null Gram operator"), nullptr;
        int status = 0;
        std::string err;
        GramBase* out = gram_inverse_filter(g, rho, static_cast<hipStream_t>(stream), status, err);
        if (!out) fail(status ? status : 2, err);
        return reinterpret_cast<RangGram*>(out);
    } catch (const std::exception& e) {
        fail(2, std::string("failure in the Gram operator: ") + e.what());
        return nullptr;
    }
}

int32_t rang_amd_gram_stats(const RangGram* h, double* out) {
    const GramBase* g = reinterpret_cast<const GramBase*>(h);
    if (!g || !out) return fail(1, "null array");
    g->stats(out);
    return 0;
}

size_t rang_amd_gram_side(const RangGram* h) {
    const GramBase* g = reinterpret_cast<const GramBase*>(h);
    return g ? size_t(g->side()) : 0;
}

void rang_amd_gram_free(RangGram* h) { delete reinterpret_cast<GramBase*>(h); }

size_t rang_amd_wgridder_prepared_active(const RangPrepared* h) {
    const PreparedBase* p = reinterpret_cast<const PreparedBase*>(h);
    return p ? p->active() : 0;
}
size_t rang_amd_wgridder_prepared_memory(const RangPrepared* h) {
    const PreparedBase* p = reinterpret_cast<const PreparedBase*>(h);
    return p ? p->memory() : 0;
}
void rang_amd_wgridder_prepared_release_grids(RangPrepared* h) {
    PreparedBase* p = reinterpret_cast<PreparedBase*>(h);
    if (p) p->release_grids();
}
void rang_amd_wgridder_prepared_free(RangPrepared* h) { delete reinterpret_cast<PreparedBase*>(h); }

void rang_amd_stage_timing(int32_t on) { set_stage_timing(on != 0); }
}  // extern "C"

namespace {

template <class T>
bool batch_image_of(const RangBatchImage& r, std::size_t k, BatchImage& b, std::string& err) {
    const std::string at = "image " + std::to_string(k) + ": ";
    if (!r.uvw || !r.freq || !r.vis || !r.dirty) {
        err = at + "null array";
        return false;
    }
    if (!(r.vis_kind == kVisC64 || r.vis_kind == kVisC128)) {
        err = at + "invalid visibility array or kind";
        return false;
    }
    if (!kinds_ok(r.vis, r.vis_kind, r.wgt, nullptr, r.wgt_kind, err)) {
        err = at + err;
        return false;
    }
    b.o = to_options(r.opts);
    b.in.nrow = r.nrow;
    b.in.nchan = r.nchan;
    b.in.uvw = r.uvw;
    b.in.vis = r.vis;
    b.in.vis_kind = r.vis_kind;
    b.in.wgt = r.wgt;
    b.in.wgt_kind = r.wgt_kind;
    b.in.mask = r.mask;
    b.freq = r.freq;
    b.dirty = r.dirty;
    b.tile = r.tile;
    return true;
}

template <class T>
int32_t batch_abi(RangBatchImage* images, std::size_t n, std::size_t budget, void* stream) {
    try {
        if (!images && n) return fail(1, "null array");
        if (n > (std::size_t(1) << 20)) return fail(1, "too many images in one batch");
        std::vector<BatchImage> b(n);
        std::string err;
        for (std::size_t k = 0; k < n; ++k)
            if (!batch_image_of<T>(images[k], k, b[k], err)) return fail(1, err);
        const int status = device_batch_vis2dirty<T>(b.data(), int(n), budget, static_cast<hipStream_t>(stream), err);
        if (status) return fail(status, err);
        for (std::size_t k = 0; k < n; ++k) std::copy(b[k].info, b[k].info + kNInfo, images[k].info);
        return 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

template <class T>
int32_t tune_abi(RangBatchImage* image, const char* path, int32_t kernels, RangTuneReport* report, void* stream) {
    try {
        if (!image) return fail(1, "null array");
        if (kernels < 1) return fail(1, "kernels must be at least 1");
        BatchImage b;
        std::string err;
        if (!batch_image_of<T>(*image, 0, b, err)) return fail(1, err);
        double ms = 0, dms = 0;
        int cached = 0;
        const int status =
            device_tune<T>(b, path, kernels, &ms, &dms, &cached, static_cast<hipStream_t>(stream), err);
        if (status) return fail(status, err);
        image->opts.kernel_index = b.o.kernel_index;
        image->opts.nu = b.o.nu;
        image->opts.nv = b.o.nv;
        image->tile = b.tile;
        if (report) {
            report->ms = ms;
            report->default_ms = cached ? 0.0 : dms;
            report->cached = cached;
            report->reserved = 0;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

}  // namespace

extern "C" {

int32_t rang_amd_wgridder_batch_vis2dirty_f32(RangBatchImage* images, size_t n, size_t budget_bytes, void* stream) {
    return batch_abi<float>(images, n, budget_bytes, stream);
}
int32_t rang_amd_wgridder_batch_vis2dirty_f64(RangBatchImage* images, size_t n, size_t budget_bytes, void* stream) {
    return batch_abi<double>(images, n, budget_bytes, stream);
}
int32_t rang_amd_wgridder_tune_f32(RangBatchImage* image, const char* cache_path, int32_t kernels,
                                   RangTuneReport* report, void* stream) {
    return tune_abi<float>(image, cache_path, kernels, report, stream);
}
int32_t rang_amd_wgridder_tune_f64(RangBatchImage* image, const char* cache_path, int32_t kernels,
                                   RangTuneReport* report, void* stream) {
    return tune_abi<double>(image, cache_path, kernels, report, stream);
}

int32_t rang_amd_release_workspaces(void) {
    try {
        release_workspaces();
        return 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

int32_t rang_amd_stage_count(void) { return stage_timing_count(); }
const char* rang_amd_stage_name(int32_t i) { return stage_timing_name(i); }
double rang_amd_stage_ms(int32_t i) { return stage_timing_ms(i); }

}  // extern "C"
