// Device pipeline of the exact w-gridder: vis2dirty and dirty2vis on
// device-resident inputs, on one HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>

#include "../../include/rang_core_wgridder.h"
#include "wg_plan.h"

namespace rang {

// Kinds of the reference's `_ex` entry points (core/src/ffi_wgridder.rs:296-310).
enum VisKind : int { kVisUnit = 0, kVisC64 = 1, kVisC128 = 2 };
enum WgtKind : int { kWgtNone = 0, kWgtF32 = 1, kWgtF64 = 2, kWgtF32Product = 3 };

// Device pointers of one call's inputs.
struct DevInputs {
    std::size_t nrow = 0, nchan = 0;
    const double* uvw = nullptr;   // nrow * 3
    const double* freq = nullptr;  // nchan
    const void* vis = nullptr;
    int vis_kind = kVisUnit;
    const void* wgt = nullptr;
    const void* wgt2 = nullptr;
    int wgt_kind = kWgtNone;
    const std::uint8_t* mask = nullptr;
};

// Status as the reference's C ABI: 0 ok, 1 invalid arguments (message in
// err), 2 failure of the device (message in err).
template <class T>
int device_vis2dirty(const Options& o, const DevInputs& in, const double* freq_host, T* dirty,
                     T* psf, double* info, hipStream_t stream, std::string& err);

// device_vis2dirty with every device buffer of the call carved from the
// fixed scratch block of the current device (fixed_scratch_bytes() bytes)
// and no allocation; a call that needs more returns 1 with a message.
template <class T>
int device_vis2dirty_fixed(const Options& o, const DevInputs& in, const double* freq_host,
                           T* dirty, T* psf, double* info, hipStream_t stream, std::string& err);
std::size_t fixed_scratch_bytes();

// One image of a sub-band batch: options, device inputs, the frequencies
// on the host and the device image (nx * ny of the precision); `tile` the
// gridding tile side to aim for (0: the default); info written per image.
struct BatchImage {
    Options o;
    DevInputs in;
    const double* freq = nullptr;
    void* dirty = nullptr;
    int tile = 0;
    double info[kNInfo] = {};
};

// vis2dirty of every image (weighted visibilities, one output each): the
// scans and plans of all images, then, per group of images whose buffers
// fit `budget` bytes (largest first), their records sorted, ONE gridding
// launch over every image's work items, and their transforms.
template <class T>
int device_batch_vis2dirty(BatchImage* imgs, int n, std::size_t budget, hipStream_t stream, std::string& err);

// The tuner: the fastest of the cost model's `kernels` cheapest kernel rows
// and of the gridding tile sides for one image, timed on the device and
// kept per signature (in `path` when not null); sets im.o.kernel_index, nu,
// nv and im.tile. ms: the chosen time, default_ms: the cost model's own
// choice, cached: 1 when the signature was known.
template <class T>
int device_tune(BatchImage& im, const char* path, int kernels, double* ms, double* default_ms, int* cached,
                hipStream_t stream, std::string& err);

template <class T>
int device_dirty2vis(const Options& o, const DevInputs& in, const double* freq_host,
                     const T* dirty, T* vis, double* info, hipStream_t stream, std::string& err);

// The plan without gridding (core/src/wgridder.rs:3525-3586); footprint: the
// gridding pass's device bytes for nout images (null or nout 0: none).
template <class T>
int device_plan(const Options& o, const DevInputs& in, const double* freq_host, bool use_vis, int nout,
                double* info, std::size_t* footprint, hipStream_t stream, std::string& err);

// A prepared set of visibilities (core/src/wgridder.rs:3165-3499): plan and
// sorted layouts kept on the device between calls.
struct PreparedBase {
    virtual ~PreparedBase() = default;
    virtual bool single() const = 0;
    virtual std::size_t active() const = 0;
    virtual std::size_t memory() const = 0;
    virtual void release_grids() = 0;
    virtual std::size_t image_nx() const = 0;
    virtual std::size_t image_ny() const = 0;
};

// Device pointers in `in` (uvw, weights f32/f64 or none); freq on the host.
// Returns null with status 1 (invalid) or 2 (device failure) and a message.
template <class T>
PreparedBase* prepared_new(const Options& o, const DevInputs& in, const double* freq_host,
                           hipStream_t stream, int& status, std::string& err);
template <class T>
int prepared_vis2dirty(PreparedBase* p, const void* vis, int vis_kind, T* dirty, double* info,
                       hipStream_t stream, std::string& err);
template <class T>
int prepared_dirty2vis(PreparedBase* p, const T* dirty, bool weighted, T* vis, hipStream_t stream,
                       std::string& err);
template <class T>
int prepared_residual(PreparedBase* p, const void* vis, int vis_kind, const T* model, T* dirty,
                      double* misfit, hipStream_t stream, std::string& err);
// Prepared::residual_sorted: the data in the order of prepared_sorted (one
// value per active visibility).
template <class T>
int prepared_residual_sorted(PreparedBase* p, const void* vis, int vis_kind, const T* model, T* dirty,
                             double* misfit, hipStream_t stream, std::string& err);
// Prepared::sorted: out = a dense nrow x nchan array (elements of `bytes`:
// 4, 8 or 16) at the active visibilities in the residual's order.
template <class T>
int prepared_sorted(PreparedBase* p, const void* dense, int bytes, void* out, hipStream_t stream, std::string& err);
template <class T>
int prepared_gram(PreparedBase* p, const T* x, T* out, hipStream_t stream, std::string& err);
template <class T>
int prepared_cg(PreparedBase* p, const T* rhs, T* x, int iterations, double tolerance, int* used,
                double* rel, hipStream_t stream, std::string& err);
template <class T>
int prepared_fista(PreparedBase* p, const T* dirty, double q0, double lam, const T* factors, int sgn,
                   const RangFistaSettings& settings, T* x, T* gx, RangFistaRecord* records,
                   std::size_t nrecords, RangFistaOutcome* out, hipStream_t stream, std::string& err);
template <class T>
int prepared_defect(PreparedBase* p, const std::function<void(const T*, T*, hipStream_t)>& cheap,
                    const T* dirty, double q0, const T* factors, int sgn, const RangFistaSettings& inner,
                    const RangOuterSchedule& schedule, T* x, T* gx, RangOuterRecord* records,
                    std::size_t nrecords, RangDefectOutcome* out, hipStream_t stream, std::string& err);
template <class T>
int prepared_power(PreparedBase* p, std::uint64_t seed, double tolerance, std::uint64_t max_iterations,
                   double* value, std::uint64_t* used, hipStream_t stream, std::string& err);

// Frees the device buffers kept between calls (workspaces, batch pools).
void release_workspaces();

// Per-kernel timing of the last call on this thread (hipEvents on the call's
// stream; enabled by set_stage_timing). Names and milliseconds.
void set_stage_timing(bool on);
int stage_timing_count();
const char* stage_timing_name(int i);
double stage_timing_ms(int i);

}  // namespace rang
