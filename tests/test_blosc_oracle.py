"""The blosc-1 / LZ4 oracle (oracle/blosc.c) pinned before it checks the GPU
decoder: the LZ4 vector of rang-core's own unit test
(core/src/blosc.rs `lz4_literals_and_overlapping_match`), round trips
through an independent LZ4 and blosc writer (tests/blosc_cases.py) over
every layout the decoder meets (split and whole blocks, byte shuffle with
typesizes 1-16 and a tail, stored-raw streams, memcpyed chunks, blocks
stored out of order, format 1, empty), the closed form of the byte
shuffle, the reference's refusals, and real c-blosc chunks: the MSv4
store of m05.zip decoded equal, bit for bit, to the Measurement Set it was
converted from (read by the table reader), as tests/test_msv4.py of
rang-core checks.
"""

import struct
import zipfile
from pathlib import Path

import numpy as np
import pytest

import blosc_cases as B
from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden" / "ms"


def test_reference_lz4_vector():
    src = bytes([0x35, ord("a"), ord("b"), ord("c"), 3, 0, 0x00])
    assert O.lz4_block(src, 12) == b"abcabcabcabc"
    assert O.lz4_block(src, 13) is None  # the wrong size is an error


@pytest.mark.parametrize("case", B.CASES, ids=[c[0] for c in B.CASES])
def test_round_trips(case):
    data, chunk = B.build(case)
    assert O.blosc_header(chunk)[0] == len(data)
    assert O.blosc_decode(chunk) == data


def test_shuffle_closed_form():
    # element i = i as uint32: stream b holds byte b of each
    data = np.arange(300, dtype=np.uint32).tobytes()
    s = np.frombuffer(B.shuffle(data, 4), np.uint8).reshape(4, 300)
    assert np.array_equal(s[0], np.arange(300) & 255) and np.array_equal(s[1], np.arange(300) >> 8)
    assert not s[2].any() and not s[3].any()
    chunk = B.blosc_chunk(data, 4, 1200, shuffled=True)
    assert O.blosc_decode(chunk) == data


def test_overlapping_matches_of_every_short_period():
    for period in range(1, 20):
        data = bytes(range(period)) * (4000 // period)
        comp = B.lz4_compress(data)
        assert len(comp) < len(data) // 4
        assert O.lz4_block(comp, len(data)) == data


def test_refusals():
    data, chunk = B.build(B.CASES[0])
    bad = bytearray(chunk)
    bad[2] |= 0x04  # bit shuffle
    assert O.blosc_header(bytes(bad)) is None
    bad = bytearray(chunk)
    bad[2] = (bad[2] & 0x1F) | (2 << 5)  # another codec
    assert O.blosc_header(bytes(bad)) is None
    bad = bytearray(chunk)
    bad[0] = 3  # format version
    assert O.blosc_header(bytes(bad)) is None
    assert O.blosc_decode(chunk[:-1]) is None  # cbytes disagrees
    assert O.blosc_header(chunk[:10]) is None
    # a match reaching before the start of the output
    assert O.lz4_block(bytes([0x14, ord("a"), 2, 0, 0x00]), 9) is None
    # an offset of zero
    assert O.lz4_block(bytes([0x14, ord("a"), 0, 0, 0x10, ord("b")]), 10) is None
    # a literal run longer than the stream
    assert O.lz4_block(bytes([0x50, 1, 2]), 5) is None


def _m05(tmp_path):
    with zipfile.ZipFile(GOLD / "msv4" / "m05.zip") as z:
        z.extractall(tmp_path)
    return tmp_path


def test_real_chunks_equal_the_measurement_set(tmp_path):
    from rang import core_ms, msv4

    root = _m05(tmp_path)
    ps = msv4.open_processing_set(root / "m05.ps.zarr")
    assert len(ps.names) == 2 and ps.refused == {}
    ms = root / "m05.ms"
    with core_ms.open_table(ms) as t:
        ref = {c: t.getcol(c) for c in ("TIME", "ANTENNA1", "ANTENNA2", "DATA_DESC_ID")}
        ref["UVW"] = t.getcol_rows("UVW", np.arange(t.nrows))
        cells = {c: [t.getcol(c, r, 1)[0] for r in range(t.nrows)] for c in ("DATA", "FLAG", "WEIGHT_SPECTRUM")}
    with core_ms.open_table(ms / "ANTENNA") as t:
        names = [str(n) for n in t.getcol("NAME")]
    with core_ms.open_table(ms / "DATA_DESCRIPTION") as t:
        dd_spw = t.getcol("SPECTRAL_WINDOW_ID")
    with core_ms.open_table(ms / "SPECTRAL_WINDOW") as t:
        freqs = [t.getcol("CHAN_FREQ", r, 1)[0] for r in range(t.nrows)]
    seen = 0
    for part in ps.partitions():
        dec = {}
        for role in ("correlated_data", "flag", "weight", "uvw"):
            a = part.arrays[role]
            raw = (a.path / ("c/" + "/".join("0" * len(a.shape)))).read_bytes()
            assert O.blosc_header(raw) is not None
            dec[role] = np.frombuffer(O.blosc_decode(raw), a.dtype).reshape(a.chunks)
        (ddid,) = [d for d, s in enumerate(dd_spw) if np.array_equal(freqs[s], part.frequency)]
        key = {}
        for ti in range(len(part.time)):
            for bi in range(part.nbaseline):
                key[(part.time[ti], part.antenna_names[part.antenna1[bi]],
                     part.antenna_names[part.antenna2[bi]])] = (ti, bi)
        for r in np.flatnonzero(ref["DATA_DESC_ID"] == ddid):
            ti, bi = key[(ref["TIME"][r], names[ref["ANTENNA1"][r]], names[ref["ANTENNA2"][r]])]
            assert dec["correlated_data"][ti, bi].tobytes() == cells["DATA"][r].tobytes()
            assert dec["flag"][ti, bi].tobytes() == cells["FLAG"][r].tobytes()
            assert dec["weight"][ti, bi].tobytes() == cells["WEIGHT_SPECTRUM"][r].tobytes()
            assert dec["uvw"][ti, bi].tobytes() == ref["UVW"][r].tobytes()
            seen += 1
    assert seen == len(ref["TIME"])


def test_header_layout():
    h = struct.pack("<BBBBIII", 2, 1, 0x21, 8, 1000, 256, 16 + 4 * 4)
    assert O.blosc_header(h) == (1000, 256, 4, 8, 0x21, 32)
