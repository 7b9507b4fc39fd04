/* ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A scalar CPU restatement of blosc-1 chunk decoding as rang-core reads zarr
 * chunks (core/src/blosc.rs): the 16-byte header and its refusals
 * (:53-79), the LZ4 block format (:81-170; the textbook LZ4 sequence
 * format: token, literal run, 16-bit offset, match of length >= 4 that may
 * overlap its own output), the split rule and the byte unshuffle of a block
 * (:301-351), and whole chunks block by block (:452-465). Only tests/ and
 * bench.py's CPU-baseline leg call it, as the checker of the GPU decoder.
 */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum {
    F_SHUFFLE = 0x01,
    F_MEMCPYED = 0x02,
    F_BITSHUFFLE = 0x04,
    F_DONT_SPLIT = 0x10,
    HEAD = 16,
    MIN_BUF = 128,
    MAX_SPLITS = 16
};

static size_t le32(const uint8_t* p) {
    return (size_t)p[0] | ((size_t)p[1] << 8) | ((size_t)p[2] << 16) | ((size_t)p[3] << 24);
}

/* f: nbytes, blocksize, nblocks, typesize, flags, cbytes. 0 ok, 1 refused. */
int or_blosc_header(const uint8_t* src, size_t len, size_t* f) {
    if (len < HEAD) return 1;
    const unsigned flags = src[2];
    const size_t typesize = src[3], nbytes = le32(src + 4), blocksize = le32(src + 8), cbytes = le32(src + 12);
    if (src[0] != 1 && src[0] != 2) return 1;
    if (typesize == 0 || (nbytes > 0 && blocksize == 0)) return 1;
    if ((flags & 0x08) || (flags & F_BITSHUFFLE)) return 1;
    if (!(flags & F_MEMCPYED) && ((flags & 0xe0) >> 5) != 1) return 1; /* LZ4 only */
    f[0] = nbytes;
    f[1] = blocksize;
    f[2] = blocksize ? (nbytes + blocksize - 1) / blocksize : 0;
    f[3] = typesize;
    f[4] = flags;
    f[5] = cbytes;
    return 0;
}

/* One LZ4 block into exactly n bytes. */
int or_lz4_block(const uint8_t* src, size_t len, uint8_t* out, size_t n) {
    size_t s = 0, d = 0;
    for (;;) {
        if (s >= len) return 1;
        const unsigned token = src[s++];
        size_t lit = token >> 4;
        if (lit == 15) {
            unsigned b;
            do {
                if (s >= len) return 1;
                b = src[s++];
                lit += b;
            } while (b == 255);
        }
        if (lit > len - s || lit > n - d) return 1;
        for (size_t i = 0; i < lit; ++i) out[d + i] = src[s + i];
        s += lit;
        d += lit;
        if (s == len) break;
        if (len - s < 2) return 1;
        const size_t off = (size_t)src[s] | ((size_t)src[s + 1] << 8);
        s += 2;
        if (off == 0 || off > d) return 1;
        size_t m = token & 15;
        if (m == 15) {
            unsigned b;
            do {
                if (s >= len) return 1;
                b = src[s++];
                m += b;
            } while (b == 255);
        }
        m += 4;
        if (m > n - d) return 1;
        /* byte by byte: a match may read what it has just written */
        for (size_t i = 0; i < m; ++i) out[d + i] = out[d + i - off];
        d += m;
    }
    return d == n ? 0 : 1;
}

/* Block j of a chunk (header fields f) into dest (its decoded length). */
static int block(const uint8_t* src, size_t len, const size_t* f, size_t j, uint8_t* dest, uint8_t* scratch) {
    const size_t nbytes = f[0], bs = f[1], nblocks = f[2], ts = f[3], flags = f[4];
    const size_t left = nbytes % bs;
    const int last = (j + 1 == nblocks) && left > 0;
    const size_t bsize = last ? left : bs;
    if (flags & F_MEMCPYED) {
        const size_t at = HEAD + j * bs;
        if (at + bsize > len) return 1;
        memcpy(dest, src + at, bsize);
        return 0;
    }
    if (HEAD + 4 * (j + 1) > len) return 1;
    const size_t start = le32(src + HEAD + 4 * j);
    /* the block ends at the nearest larger block start (blocks may be stored
     * out of order) or at the end of the chunk */
    size_t end = f[5];
    for (size_t k = 0; k < nblocks; ++k) {
        if (HEAD + 4 * (k + 1) > len) return 1;
        const size_t o = le32(src + HEAD + 4 * k);
        if (o > start && o < end) end = o;
    }
    if (end > len || start > end) return 1;
    const int shuffled = (flags & F_SHUFFLE) && ts > 1;
    const size_t nsplit = (!(flags & F_DONT_SPLIT) && ts <= MAX_SPLITS && bsize / ts >= MIN_BUF && !last) ? ts : 1;
    const size_t ne = bsize / nsplit;
    uint8_t* target = shuffled ? scratch : dest;
    size_t at = start;
    for (size_t k = 0; k < nsplit; ++k) {
        if (at + 4 > end) return 1;
        const size_t cb = le32(src + at);
        at += 4;
        if (cb > end - at) return 1;
        if (cb == ne)
            memcpy(target + k * ne, src + at, ne);
        else if (or_lz4_block(src + at, cb, target + k * ne, ne))
            return 1;
        at += cb;
    }
    if (shuffled) {
        /* stream b holds byte b of every element; a tail shorter than one
         * element stays in place */
        const size_t q = bsize / ts;
        for (size_t i = 0; i < q; ++i)
            for (size_t b = 0; b < ts; ++b) dest[i * ts + b] = scratch[b * q + i];
        memcpy(dest + q * ts, scratch + q * ts, bsize - q * ts);
    }
    return 0;
}

/* A whole chunk into out (exactly its decoded length). 0 ok, 1 refused. */
int or_blosc_decode(const uint8_t* src, size_t len, uint8_t* out, size_t out_len) {
    size_t f[6];
    if (or_blosc_header(src, len, f)) return 1;
    if (f[5] != len || f[0] != out_len) return 1;
    if (!(f[4] & F_MEMCPYED) && f[2] > (len - HEAD) / 4) return 1;
    uint8_t* scratch = (uint8_t*)malloc(f[1] ? f[1] : 1);
    if (!scratch) return 1;
    int rc = 0;
    for (size_t j = 0; j < f[2] && !rc; ++j) {
        const size_t lo = j * f[1];
        rc = block(src, len, f, j, out + lo, scratch);
    }
    free(scratch);
    return rc;
}
