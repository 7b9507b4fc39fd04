"""Ingest to the device: the GPU blosc-1 decoder (rang_amd_blosc_decode_batch,
rang_core_blosc_decode) against the oracle (oracle/blosc.c, pinned in
tests/test_blosc_oracle.py) on every chunk layout, bit for bit; the MSv4
reader, whose data chunks expand on the GPU, against the Measurement Set it
was converted from, bit for bit (tests/test_msv4.py of rang-core,
`test_xradio_store_equals_msv2`); the Stokes I forming on the device against
the oracle's rule; and the ingested visibilities gridded by the device
vis2dirty against the oracle's vis2dirty of the same host arrays (2e-5 of the
peak, the single-precision bar of the gridder's tests).
"""

import ctypes
import struct
import zipfile
from pathlib import Path

import numpy as np
import pytest

import blosc_cases as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLD = Path(__file__).resolve().parent / "golden" / "ms"


@pytest.fixture(scope="module")
def m05(tmp_path_factory):
    d = tmp_path_factory.mktemp("m05")
    with zipfile.ZipFile(GOLD / "msv4" / "m05.zip") as z:
        z.extractall(d)
    return d


def _decode(chunks):
    from rang import msv4

    sizes = [O.blosc_header(c)[0] for c in chunks]
    outs = [torch.empty(n, dtype=torch.uint8, device="cuda") for n in sizes]
    msv4.blosc_decode_device(chunks, outs)
    torch.cuda.synchronize()
    return [o.cpu().numpy().tobytes() for o in outs]


def test_every_layout_in_one_launch():
    built = [B.build(c) for c in B.CASES]
    got = _decode([c for _, c in built])
    for case, (data, chunk), g in zip(B.CASES, built, got):
        assert g == data == O.blosc_decode(chunk), case[0]


def test_one_chunk_and_the_host_entry_point():
    from rang import _core

    lib = _core.library()
    for case in B.CASES[:4]:
        data, chunk = B.build(case)
        assert _decode([chunk])[0] == data
        out = np.zeros(max(1, len(data)), np.uint8)
        assert lib.rang_core_blosc_decode(chunk, len(chunk), out.ctypes.data, len(data)) == 0
        assert out[: len(data)].tobytes() == data


def test_corrupt_streams_are_refused():
    from rang import msv4

    def chunk_of(stream, n):
        # one block of n bytes, typesize 1, one LZ4 stream
        body = struct.pack("<I", len(stream)) + stream
        cb = 16 + 4 + len(body)
        return struct.pack("<BBBBIII", 2, 1, 0x21, 1, n, n, cb) + struct.pack("<I", 20) + body

    good = chunk_of(B.lz4_compress(b"abcd" * 64), 256)
    assert _decode([good])[0] == b"abcd" * 64
    bad = [chunk_of(bytes([0x14, ord("a"), 9, 0, 0x00]), 9),      # a match before the output's start
           chunk_of(bytes([0x14, ord("a"), 0, 0, 0x00]), 9),      # offset zero
           chunk_of(bytes([0x10, ord("a"), 1, 0, 0x00]), 9),      # decodes to 5 bytes, 9 declared
           chunk_of(bytes([0xF0, 200]), 300)]                     # literal run past the stream
    for c in bad:
        out = torch.empty(O.blosc_header(c)[0], dtype=torch.uint8, device="cuda")
        with pytest.raises(ValueError, match="corrupt LZ4 stream"):
            msv4.blosc_decode_device([c], [out])
    # a corrupt chunk beside good ones fails the batch; the next call is clean
    out = [torch.empty(256, dtype=torch.uint8, device="cuda"), torch.empty(9, dtype=torch.uint8, device="cuda")]
    with pytest.raises(ValueError):
        msv4.blosc_decode_device([good, bad[0]], out)
    assert _decode([good])[0] == b"abcd" * 64
    # refusals on the host: truncated chunk, output of the wrong size
    with pytest.raises(ValueError, match="compressed size"):
        msv4.blosc_decode_device([good[:-1]], [torch.empty(256, dtype=torch.uint8, device="cuda")])
    with pytest.raises(ValueError, match="wrong size"):
        msv4.blosc_decode_device([good], [torch.empty(255, dtype=torch.uint8, device="cuda")])


def test_msv4_rows_equal_the_measurement_set(m05):
    from rang import core_ms, msv4

    ms = m05 / "m05.ms"
    ps = msv4.open_processing_set(m05 / "m05.ps.zarr")
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
    assert len(ps.names) == 2
    for part in ps.partitions():
        rows = part.rows()
        assert rows["PADS"] == 0
        (ddid,) = [d for d, s in enumerate(dd_spw) if np.array_equal(freqs[s], part.frequency)]
        key = {(rows["TIME"][i], part.antenna_names[a], part.antenna_names[b]): i
               for i, (a, b) in enumerate(zip(rows["ANTENNA1"], rows["ANTENNA2"]))}
        for r in np.flatnonzero(ref["DATA_DESC_ID"] == ddid):
            i = key[(ref["TIME"][r], names[ref["ANTENNA1"][r]], names[ref["ANTENNA2"][r]])]
            for c in ("DATA", "FLAG", "WEIGHT_SPECTRUM"):
                assert rows[c][i].dtype == cells[c][r].dtype
                assert rows[c][i].tobytes() == cells[c][r].tobytes()
            assert rows["UVW"][i].tobytes() == ref["UVW"][r].tobytes()
            seen += 1
        assert len(rows["TIME"]) == int(np.sum(ref["DATA_DESC_ID"] == ddid))
        pieces = list(part.chunks(time_rows=2))
        assert len(pieces) == 2
        assert np.concatenate([p["DATA"] for p in pieces]).tobytes() == rows["DATA"].tobytes()
    assert seen == len(ref["TIME"])


def test_stokes_i_and_gridding_of_ingested_rows(m05):
    from rang import core_ms, msv4
    from rang import device as D

    part = msv4.open_processing_set(m05 / "m05.ps.zarr").partition("m05-var-spw-tssm_1")
    dev = part.stokes_i_device()
    rows = part.rows()
    want_vis, want_wgt = O.stokes_i(rows["DATA"], rows["WEIGHT_SPECTRUM"], rows["FLAG"])
    assert dev["vis"].cpu().numpy().tobytes() == want_vis.tobytes()
    assert dev["wgt"].cpu().numpy().tobytes() == want_wgt.tobytes()
    # the same rows through the Measurement Set reader
    ms = core_ms.stokes_i_to_device(m05 / "m05.ms", ddid=1)
    assert ms["vis"].cpu().numpy().tobytes() == want_vis.tobytes()
    assert ms["wgt"].cpu().numpy().tobytes() == want_wgt.tobytes()
    assert ms["uvw"].cpu().numpy().tobytes() == rows["UVW"].tobytes()
    assert np.array_equal(ms["freq"], part.frequency)
    # gridded on the device from HBM, against the oracle on the host arrays
    nx = ny = 64
    dx = 2e-4
    o = [nx, ny, dx, dx, 1e-5, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    freq = torch.from_numpy(part.frequency).cuda()
    dirty, _, _ = D.vis2dirty(o, dev["uvw"], freq, dev["vis"], dev["wgt"], single=True)
    imgs, _ = O.vis2dirty(rows["UVW"], part.frequency, want_vis, want_wgt, npix_x=nx, npix_y=ny, pixsize_x=dx,
                          pixsize_y=dx, epsilon=1e-5, single=True, do_wgridding=True)
    ref = imgs["dirty"]
    got = dirty.cpu().numpy()
    assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max()


def test_columns_straight_to_the_device(m05):
    from rang import core_ms

    with core_ms.open_table(m05 / "m05.ms") as t:
        uvw = t.getcol_device("UVW")
        assert uvw.dtype == torch.float64 and uvw.cpu().numpy().tobytes() == t.getcol("UVW").tobytes()
        part = t.getcol_device("TIME", 6, 12)
        assert part.cpu().numpy().tobytes() == t.getcol("TIME", 6, 12).tobytes()
        with pytest.raises(OSError, match="fixed cell shape"):
            t.getcol_device("DATA")
