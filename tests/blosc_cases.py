"""Test-side writers of LZ4 blocks and blosc-1 chunks, and the chunk cases
the decoder tests share.

`lz4_compress` is a plain greedy LZ4 block encoder (4-byte hash matches,
offsets up to 65535, matches that overlap their own output, the format's
end rules: the last five bytes literal, no match starting in the last
twelve). `blosc_chunk` lays out a blosc-1 chunk as c-blosc 1.21 does
(16-byte header, block starts, per block one stream or one per byte of the
type, optional byte shuffle, streams that would not shrink stored raw), so
that every path of the decoder is reached with known contents.
"""

import struct
import zlib

import numpy as np


def _length(out, n):
    while n >= 255:
        out.append(255)
        n -= 255
    out.append(n)


def _sequence(out, literals, offset=None, match=0):
    lit = len(literals)
    token = (min(lit, 15) << 4) | (min(match - 4, 15) if offset is not None else 0)
    out.append(token)
    if lit >= 15:
        _length(out, lit - 15)
    out += literals
    if offset is not None:
        out += struct.pack("<H", offset)
        if match - 4 >= 15:
            _length(out, match - 4 - 15)


def lz4_compress(data):
    data = bytes(data)
    n = len(data)
    out = bytearray()
    table = {}
    i = anchor = 0
    limit = n - 12
    while i < limit:
        key = data[i : i + 4]
        j = table.get(key)
        table[key] = i
        if j is not None and i - j <= 65535:
            m = 4
            while i + m < n - 5 and data[j + m] == data[i + m]:
                m += 1
            _sequence(out, data[anchor:i], i - j, m)
            for k in range(i + 1, min(i + m, limit)):
                table[data[k : k + 4]] = k
            i += m
            anchor = i
        else:
            i += 1
    _sequence(out, data[anchor:])
    return bytes(out)


def shuffle(block, typesize):
    """c-blosc's byte shuffle: stream b holds byte b of every element; a
    tail shorter than one element stays at the end."""
    block = np.frombuffer(bytes(block), np.uint8)
    q = block.size // typesize
    body = block[: q * typesize].reshape(q, typesize).T.reshape(-1)
    return body.tobytes() + block[q * typesize :].tobytes()


def blosc_chunk(data, typesize, blocksize, *, shuffled=False, split=True, memcpyed=False, version=2,
                reverse_blocks=False, compress=True):
    data = bytes(data)
    nbytes = len(data)
    flags = (1 << 5) | (0x01 if shuffled else 0) | (0 if split else 0x10)
    if memcpyed:
        flags |= 0x02
        body = data
        head = struct.pack("<BBBBIII", version, 1, flags, typesize, nbytes, blocksize, 16 + nbytes)
        return head + body
    nblocks = -(-nbytes // blocksize) if nbytes else 0
    left = nbytes % blocksize
    blocks = []
    for j in range(nblocks):
        last = j + 1 == nblocks and left > 0
        bsize = left if last else blocksize
        block = data[j * blocksize : j * blocksize + bsize]
        if shuffled and typesize > 1:
            block = shuffle(block, typesize)
        nsplit = typesize if (split and typesize <= 16 and bsize // typesize >= 128 and not last) else 1
        ne = bsize // nsplit
        enc = bytearray()
        for k in range(nsplit):
            stream = block[k * ne : (k + 1) * ne]
            comp = lz4_compress(stream) if compress else stream
            if len(comp) >= ne:
                comp = stream
            enc += struct.pack("<I", len(comp)) + comp
        blocks.append(bytes(enc))
    order = list(range(nblocks))[::-1] if reverse_blocks else list(range(nblocks))
    starts = [0] * nblocks
    at = 16 + 4 * nblocks
    body = bytearray()
    for j in order:
        starts[j] = at
        body += blocks[j]
        at += len(blocks[j])
    head = struct.pack("<BBBBIII", version, 1, flags, typesize, nbytes, blocksize, at)
    return head + b"".join(struct.pack("<I", s) for s in starts) + bytes(body)


def contents(kind, n, seed=0):
    """Bytes of a chunk: visibility-like floats, flags, runs, periods."""
    rng = np.random.default_rng(seed)
    if kind == "vis":
        v = (rng.normal(size=n // 8) + 1j * rng.normal(size=n // 8)).astype(np.complex64)
        v[rng.random(v.size) < 0.2] = 0
        return v.tobytes()[:n].ljust(n, b"\0")
    if kind == "flags":
        f = (rng.random(n) < 0.05).astype(np.uint8)
        f[n // 3 : n // 2] = 1
        return f.tobytes()
    if kind == "weights":
        w = np.repeat(rng.random(max(1, n // 64)).astype(np.float32), 16)
        return w.tobytes()[:n].ljust(n, b"\0")
    if kind == "period":
        pat = bytes(rng.integers(0, 256, 7, dtype=np.uint8))
        return (pat * (n // 7 + 1))[:n]
    if kind == "random":
        return rng.integers(0, 256, n, dtype=np.uint8).tobytes()
    raise ValueError(kind)


# (name, contents, nbytes, typesize, blocksize, options)
CASES = [
    ("vis-shuffle-split", "vis", 3 * 65536 + 4096, 8, 65536, dict(shuffled=True)),
    ("vis-noshuffle", "vis", 70000, 8, 32768, dict()),
    ("vis-dont-split", "vis", 50000, 8, 16384, dict(shuffled=True, split=False)),
    ("flags-typesize1", "flags", 100000, 1, 32768, dict()),
    ("weights-shuffle4", "weights", 90004, 4, 16384, dict(shuffled=True)),
    ("period-long-matches", "period", 200000, 1, 131072, dict(split=False)),
    ("random-stored-raw", "random", 40000, 8, 16384, dict(shuffled=True)),
    ("memcpyed", "vis", 30000, 8, 8192, dict(memcpyed=True)),
    ("reversed-blocks", "vis", 5 * 4096 + 100, 2, 4096, dict(shuffled=True, reverse_blocks=True)),
    ("typesize16", "vis", 64 * 1024, 16, 32768, dict(shuffled=True)),
    ("typesize3-tail", "weights", 3 * 2000 + 2, 3, 3 * 500, dict(shuffled=True)),
    ("version1", "flags", 20000, 1, 4096, dict(version=1)),
    ("empty", "vis", 0, 8, 4096, dict()),
]


def build(case):
    name, kind, n, ts, bs, opts = case
    data = contents(kind, n, seed=zlib.crc32(name.encode()) % 1000)
    return data, blosc_chunk(data, ts, bs, **opts)
