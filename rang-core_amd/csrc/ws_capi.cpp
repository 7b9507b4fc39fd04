// The C ABI of include/rang_core_grid.h: rang-core's ffi_grid.rs and
// gpu/ffi.rs entry points on the MI355X w-stacking path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "../../include/rang_core_grid.h"
#include "ws_device.h"

namespace {

thread_local std::string t_grid_error;

int32_t fail(int32_t status, const std::string& msg) {
    t_grid_error = msg;
    return status;
}

bool mul_ok(std::size_t a, std::size_t b, std::size_t& out) {
    if (a != 0 && b > std::numeric_limits<std::size_t>::max() / a) return false;
    out = a * b;
    return true;
}

// The checks of rang_core_grid_invert and grid::bucket that need no GPU
// (core/src/ffi_grid.rs:66-96, core/src/grid.rs:729-751).
int32_t checks(const GridInputs* i, const GridOutputs* o) {
    if (!i || !o) return fail(1, "null inputs or outputs");
    if (!i->uvw || !i->frequencies || !i->data || !i->weights || !i->imaging_weights || !o->dirty ||
        !o->psf || !o->stats)
        return fail(1, "null array");
    std::size_t cells, uvw_len, data_len, pixels;
    if (!mul_ok(i->nrow, i->nchan, cells) || !mul_ok(i->nrow, 3, uvw_len) ||
        !mul_ok(cells, 2, data_len) || !mul_ok(i->width, i->height, pixels))
        return fail(1, "array sizes overflow");
    if (i->width == 0 || i->height == 0 || i->kernel_size == 0 || i->oversampling == 0 ||
        i->n_w_layers == 0 || i->kernel_size % 2 == 0 || std::isnan(i->padding) || i->padding < 1.0 ||
        i->oversampling > 65535)
        return fail(1, "invalid image or kernel settings");
    if (cells >= (std::size_t(1) << 31)) return fail(1, "too many samples for one call (at most 2^31 - 1)");
    const double pw = std::ceil(double(i->width) * i->padding), ph = std::ceil(double(i->height) * i->padding);
    if (pw > 2147483647.0 || ph > 2147483647.0) return fail(1, "image too large");
    return 0;
}

// WSClean's padded side (grid.rs:80-89) and whether the reference's GPU path
// transforms it (radices up to 11 there: core/src/gpu/tests.rs:274-288).
std::size_t padded(std::size_t n, double padding) {
    const std::size_t p = std::size_t(std::ceil(double(n) * padding));
    return p + (4 - p % 4) % 4;
}
bool smooth(std::size_t n) {
    for (std::size_t q : {2, 3, 5, 7, 11})
        while (n % q == 0) n /= q;
    return n == 1;
}

int32_t run(const GridInputs* i, GridOutputs* o, bool single, double* timings, bool any_length) {
    if (const int32_t c = checks(i, o)) return c;
    if (!any_length && (!smooth(padded(i->width, i->padding)) || !smooth(padded(i->height, i->padding))))
        return fail(1, "padded image sides must be products of 2, 3, 5, 7 and 11 for the GPU transforms");
    try {
        std::string err;
        const int st = rang::ws::invert_gpu(*i, *o, single, timings, err);
        return st ? fail(st, err) : 0;
    } catch (const std::exception& e) {
        return fail(2, std::string("failure in the gridder: ") + e.what());
    }
}

}  // namespace

extern "C" {

int32_t rang_core_grid_invert(const GridInputs* inputs, GridOutputs* outputs) {
    return run(inputs, outputs, false, nullptr, true);
}

int32_t rang_core_grid_invert_gpu(const GridInputs* inputs, GridOutputs* outputs, int32_t precision,
                                  double* timings) {
    if (!inputs || !outputs) return fail(1, "null inputs or outputs");
    if (precision != 0 && precision != 1) return fail(1, "precision must be 0 (f64) or 1 (f32)");
    if (!timings) return fail(3, "null buffer");
    std::fill(timings, timings + RANG_GRID_N_TIMINGS, 0.0);
    // as core/src/gpu/ffi.rs:125-136: every other failure (arguments the
    // path rejects included) is 3
    return run(inputs, outputs, precision == 1, timings, false) ? 3 : 0;
}

int32_t rang_core_grid_kernel(size_t kernel_size, size_t oversampling, double* kernel) {
    std::size_t len;
    if (!mul_ok(kernel_size, oversampling, len) || !kernel || len == 0) return fail(1, "invalid kernel size");
    return rang::ws::kernel(kernel_size, oversampling, kernel);
}

size_t rang_amd_device_name(uint8_t* buf, size_t len) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 0;
    const std::string name = std::string(p.name) + " (" + p.gcnArchName + ")";
    if (buf && len) std::memcpy(buf, name.data(), std::min(len, name.size()));
    return name.size();
}

size_t rang_amd_grid_last_error(uint8_t* buf, size_t len) {
    const std::string& m = t_grid_error;
    if (buf && len) std::memcpy(buf, m.data(), std::min(len, m.size()));
    return m.size();
}

}  // extern "C"
