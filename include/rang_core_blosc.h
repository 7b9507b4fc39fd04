/*
 * C ABI of the blosc-1 chunk decoder, MI355X build: zarr chunks expanded on
 * the GPU.
 *
 * rang_core_blosc_size / _header / _decode replace the functions of the same
 * names in rang-core's core/src/ffi.rs:537-660 (same arguments; status 0 ok,
 * 1 refused or corrupt, 2 a failure inside the library). Only LZ4-compressed
 * or stored chunks of blosc format 1 and 2, with or without byte shuffle, are
 * read (core/src/blosc.rs:1-12); bit shuffle and other codecs are refused.
 * rang_core_blosc_decode takes host buffers and runs on the GPU.
 *
 * rang_amd_blosc_decode_batch has no counterpart in the reference: many
 * chunks (host bytes) expanded into device buffers on `stream` in one
 * upload and one launch, complete on return; the message of a refusal is
 * kept per thread for rang_amd_blosc_last_error.
 */
#ifndef RANG_CORE_BLOSC_H
#define RANG_CORE_BLOSC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RangBloscChunk {
    const uint8_t *src; /* host: the compressed chunk */
    size_t len;         /* its length (the header's cbytes) */
    void *dst;          /* device: the decoded chunk */
    size_t dst_len;     /* its length (the header's nbytes) */
} RangBloscChunk;

/* core/src/ffi.rs:537: the decoded length of a whole chunk */
int32_t rang_core_blosc_size(const uint8_t *src, size_t len, size_t *nbytes);
/* core/src/ffi.rs:634: nbytes, blocksize, number of blocks, typesize from the first bytes */
int32_t rang_core_blosc_header(const uint8_t *head, size_t len, size_t *fields);
/* core/src/ffi.rs:565: a whole chunk into `out` (exactly its decoded length) */
int32_t rang_core_blosc_decode(const uint8_t *src, size_t len, uint8_t *out, size_t out_len);

int32_t rang_amd_blosc_decode_batch(const RangBloscChunk *chunks, size_t n, void *stream);
/* As rang_amd_blosc_decode_batch with each chunk's bytes already in device
 * memory at dev_src[i] (the host copy at chunks[i].src is read for the plan
 * only: headers, block starts and stream lengths); nothing is uploaded. */
int32_t rang_amd_blosc_decode_resident(const RangBloscChunk *chunks, const void *const *dev_src, size_t n,
                                       void *stream);
size_t rang_amd_blosc_last_error(uint8_t *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif
