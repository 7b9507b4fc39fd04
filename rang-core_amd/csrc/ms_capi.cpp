// C ABI of the table reader (include/rang_core_ms.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>

#include "../../include/rang_core_ms.h"
#include "ms_table.h"

struct RangMsTable {
    rang::ms::Table t;
    explicit RangMsTable(const std::string& p) : t(p) {}
};

namespace {

thread_local std::string t_ms_error;

int32_t ms_fail(const std::string& m) {
    t_ms_error = m;
    return 1;
}

// The cells of a read in the C struct; the buffer is owned by `out`.
int32_t fill(rang::ms::Cells&& c, MsColumn* out) {
    if (c.shape.size() > 8) return ms_fail("more than 8 cell axes");
    std::size_t nb = c.bytes.size();
    if (c.dtype == rang::ms::kString) {
        nb = 0;
        for (const auto& s : c.strings) nb += 4 + s.size();
    }
    uint8_t* buf = static_cast<uint8_t*>(::operator new(nb ? nb : 1));
    if (c.dtype == rang::ms::kString) {
        std::size_t at = 0;
        for (const auto& s : c.strings) {
            const uint32_t n = uint32_t(s.size());
            for (int k = 0; k < 4; ++k) buf[at + k] = uint8_t(n >> (8 * k));
            std::memcpy(buf + at + 4, s.data(), s.size());
            at += 4 + s.size();
        }
    } else if (nb) {
        std::memcpy(buf, c.bytes.data(), nb);
    }
    std::memset(out, 0, sizeof *out);
    out->data = buf;
    out->nbytes = nb;
    out->nrows = c.nrows;
    out->ndim = c.shape.size();
    for (std::size_t i = 0; i < c.shape.size(); ++i) out->shape[i] = c.shape[i];
    out->dtype = int32_t(c.dtype);
    out->nstrings = c.strings.size();
    return 0;
}

bool slice_of(const int64_t* s, size_t n, rang::ms::Slice& out) {
    out.clear();
    if (n == 0) return false;
    for (size_t i = 0; i < n; ++i) out.push_back({s[2 * i], s[2 * i + 1]});
    return true;
}

template <class F>
int32_t guarded(F&& f) {
    try {
        return f();
    } catch (const rang::ms::Error& e) {
        return ms_fail(e.what());
    } catch (const std::bad_alloc&) {
        return ms_fail("out of host memory");
    } catch (const std::exception& e) {
        t_ms_error = e.what();
        return 2;
    } catch (...) {
        t_ms_error = "unknown failure";
        return 2;
    }
}

}  // namespace

extern "C" {

size_t rang_core_ms_last_error(uint8_t* buf, size_t len) {
    if (buf && len) {
        const size_t n = std::min(len, t_ms_error.size());
        std::memcpy(buf, t_ms_error.data(), n);
    }
    return t_ms_error.size();
}

int32_t rang_core_ms_open(const char* path, RangMsTable** out) {
    if (!path || !out) return ms_fail("null argument");
    return guarded([&] {
        *out = new RangMsTable(path);
        return 0;
    });
}

void rang_core_ms_close(RangMsTable* table) { delete table; }

size_t rang_core_ms_describe(const RangMsTable* table, uint8_t* buf, size_t len) {
    if (!table) return 0;
    try {
        const std::string j = table->t.describe_json();
        if (buf && len) std::memcpy(buf, j.data(), std::min(len, j.size()));
        return j.size();
    } catch (...) {
        return 0;
    }
}

int32_t rang_core_ms_read(const RangMsTable* table, const char* name, uint64_t start, uint64_t count,
                          MsColumn* out) {
    return rang_amd_ms_read_slice(table, name, start, count, nullptr, 0, out);
}

void rang_core_ms_free(MsColumn* out) {
    if (!out) return;
    ::operator delete(out->data);
    out->data = nullptr;
    out->nbytes = 0;
}

int32_t rang_amd_ms_read_slice(const RangMsTable* table, const char* name, uint64_t start, uint64_t count,
                               const int64_t* slice, size_t nslice, MsColumn* out) {
    if (!table || !name || !out || (nslice && !slice)) return ms_fail("null argument");
    return guarded([&] {
        rang::ms::Slice s;
        const bool cut = slice_of(slice, nslice, s);
        return fill(table->t.read(name, start, count, cut ? &s : nullptr), out);
    });
}

int32_t rang_amd_ms_read_rows(const RangMsTable* table, const char* name, const uint64_t* rows, size_t nrows,
                              const int64_t* slice, size_t nslice, MsColumn* out) {
    if (!table || !name || !out || (nrows && !rows) || (nslice && !slice)) return ms_fail("null argument");
    return guarded([&] {
        rang::ms::Slice s;
        const bool cut = slice_of(slice, nslice, s);
        std::vector<uint64_t> r(rows, rows + nrows);
        return fill(table->t.read_rows(name, r, cut ? &s : nullptr), out);
    });
}

int32_t rang_amd_ms_read_dev(const RangMsTable* table, const char* name, uint64_t start, uint64_t count,
                             const int64_t* slice, size_t nslice, void* dst, size_t cap, void* stream,
                             MsColumn* out) {
    if (!table || !name || !out || !dst || (nslice && !slice)) return ms_fail("null argument");
    return guarded([&] {
        rang::ms::Slice s;
        const bool cut = slice_of(slice, nslice, s);
        rang::ms::Cells c = table->t.read(name, start, count, cut ? &s : nullptr);
        if (c.dtype == rang::ms::kString) return ms_fail(std::string(name) + ": string columns do not go to the device");
        if (c.bytes.size() > cap)
            return ms_fail(std::string(name) + ": " + std::to_string(c.bytes.size()) + " bytes exceed the " +
                           std::to_string(cap) + "-byte device buffer");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const size_t nb = c.bytes.size();
        if (!c.bytes.empty()) {
            if (hipMemcpyAsync(dst, c.bytes.data(), c.bytes.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                t_ms_error = "device copy failed";
                return 2;
            }
        }
        c.bytes.clear();
        c.bytes.shrink_to_fit();
        const int32_t rc = fill(std::move(c), out);
        if (rc == 0) {
            ::operator delete(out->data);
            out->data = nullptr;
            out->nbytes = nb;
        }
        return rc;
    });
}

}  // extern "C"
