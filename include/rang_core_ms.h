/*
 * C ABI of the Measurement Set (casacore table) reader, MI355X build: the
 * host side of ingest to the device.
 *
 * The rang_core_ms_* symbols replace rang-core's core/src/ffi_ms.rs with
 * the same names, arguments, struct layout and status codes: 0 ok, 1 an
 * error (message kept per thread for rang_core_ms_last_error), 2 a failure
 * inside the library. A table is an opaque handle; a read returns a buffer
 * the library allocates and rang_core_ms_free releases.
 *
 * The rang_amd_* symbols have no C counterpart in the reference (it reaches
 * them through its Python extension's MsTable.read / read_rows,
 * core/py/src/lib.rs): a read with a cell slice, a read of listed rows, and
 * a read straight into device memory on a HIP stream.
 */
#ifndef RANG_CORE_MS_H
#define RANG_CORE_MS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RangMsTable RangMsTable;

/* core/src/ffi_ms.rs:24-33 (#[repr(C)] MsColumn): `data` holds `nbytes`
 * bytes, for numeric and boolean columns `nrows` cells of `shape` (casacore
 * order, booleans one byte each), for string columns `nstrings` strings,
 * each a little-endian uint32 byte count then UTF-8 bytes. `dtype` is the
 * casacore DataType code. */
typedef struct MsColumn {
    uint8_t *data;
    size_t nbytes;
    uint64_t nrows;
    size_t ndim;
    int64_t shape[8];
    int32_t dtype;
    size_t nstrings;
} MsColumn;

/* core/src/ffi_ms.rs:47 */
size_t rang_core_ms_last_error(uint8_t *buf, size_t len);
/* core/src/ffi_ms.rs:67: open the table in directory `path` */
int32_t rang_core_ms_open(const char *path, RangMsTable **out);
/* core/src/ffi_ms.rs:94 */
void rang_core_ms_close(RangMsTable *table);
/* core/src/ffi_ms.rs:110: JSON description (path, nrows, columns with
 * dtype, shape, manager, supported) into `buf`; returns the length needed */
size_t rang_core_ms_describe(const RangMsTable *table, uint8_t *buf, size_t len);
/* core/src/ffi_ms.rs:137: rows start .. start + count (count 0: to the end) */
int32_t rang_core_ms_read(const RangMsTable *table, const char *name, uint64_t start, uint64_t count,
                          MsColumn *out);
/* core/src/ffi_ms.rs:197 */
void rang_core_ms_free(MsColumn *out);

/* As rang_core_ms_read, each cell cut to `slice`: `nslice` (first pixel,
 * length) pairs for the last casacore axes (a length < 0 runs to the end);
 * the tiled managers read only the tiles the slice meets
 * (core/src/ms.rs:2111-2160). nslice 0 reads whole cells. */
int32_t rang_amd_ms_read_slice(const RangMsTable *table, const char *name, uint64_t start, uint64_t count,
                               const int64_t *slice, size_t nslice, MsColumn *out);
/* The rows `rows` (any order) in that order, cut as above
 * (core/src/ms.rs:2162-2222). */
int32_t rang_amd_ms_read_rows(const RangMsTable *table, const char *name, const uint64_t *rows, size_t nrows,
                              const int64_t *slice, size_t nslice, MsColumn *out);
/* Rows start .. start + count of a numeric or boolean column, cut as above,
 * written to device memory `dst` (capacity `cap` bytes) on HIP stream
 * `stream` (null: the default stream); complete on return. `out` gets the
 * description of what was written (its `data` is null, nothing to free). */
int32_t rang_amd_ms_read_dev(const RangMsTable *table, const char *name, uint64_t start, uint64_t count,
                             const int64_t *slice, size_t nslice, void *dst, size_t cap, void *stream,
                             MsColumn *out);

#ifdef __cplusplus
}
#endif

#endif
