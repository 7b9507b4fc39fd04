"""The C-ABI library loads and exports every symbol its header declares, and
the checks that need no GPU behave as the reference's (CPU only)."""

import re

import numpy as np
import pytest
from conftest import ROOT, ref_case

from oracle import oracle as O
from rang import _core, core_wgridder

HEADERS = [ROOT / "include" / "rang_core_wgridder.h", ROOT / "include" / "rang_core_grid.h",
           ROOT / "include" / "rang_core_precond.h", ROOT / "include" / "rang_core_gram.h",
           ROOT / "include" / "rang_core_weights.h", ROOT / "include" / "rang_core_batch.h",
           ROOT / "include" / "rang_core_ms.h", ROOT / "include" / "rang_core_blosc.h"]


def declared():
    text = "".join(h.read_text() for h in HEADERS)
    return sorted(set(re.findall(r"\b(rang_(?:core|amd)_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) == 119
    lib = _core.library()
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_core.SIGNATURES)


def test_fista_structs_match_the_header():
    # RangFistaSettings / Record / Outcome as laid out by a C compiler
    import ctypes

    assert ctypes.sizeof(_core.FistaSettings) == 48
    assert ctypes.sizeof(_core.FistaRecord) == 88
    assert ctypes.sizeof(_core.FistaOutcome) == 64 + 88
    assert _core.FistaOutcome.last.offset == 64


def test_options_struct_layout():
    # 9 size/double fields, 6 int32, 1 int64, 2 size_t: 120 bytes
    import ctypes

    assert ctypes.sizeof(_core.WgridderOptions) == 120
    assert _core.WgridderOptions.kernel_index.offset == 96


def test_batch_structs_match_a_c_compiler(tmp_path):
    # RangBatchImage / RangTuneReport as gcc lays them out from the header
    import ctypes
    import subprocess

    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "rang_core_batch.h"\n'
        "int main(void) { printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(RangBatchImage),"
        " offsetof(RangBatchImage, nrow), offsetof(RangBatchImage, vis_kind), offsetof(RangBatchImage, tile),"
        " offsetof(RangBatchImage, info), sizeof(RangTuneReport)); return 0; }\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    B = _core.BatchImage
    assert got == [ctypes.sizeof(B), B.nrow.offset, B.vis_kind.offset, B.tile.offset, B.info.offset,
                   ctypes.sizeof(_core.TuneReport)]


def test_ingest_structs_match_a_c_compiler(tmp_path):
    # MsColumn (core/src/ffi_ms.rs:24-33) and RangBloscChunk as gcc lays them out
    import ctypes
    import subprocess

    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "rang_core_ms.h"\n#include "rang_core_blosc.h"\n'
        "int main(void) { printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(MsColumn), offsetof(MsColumn, shape),"
        " offsetof(MsColumn, dtype), offsetof(MsColumn, nstrings), sizeof(RangBloscChunk)); return 0; }\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    M = _core.MsColumn
    assert got == [ctypes.sizeof(M), M.shape.offset, M.dtype.offset, M.nstrings.offset,
                   ctypes.sizeof(_core.BloscChunk)]


def test_blosc_header_refusals_without_a_gpu():
    # header parsing is host work: the reference's refusals (core/src/blosc.rs:53-79)
    import ctypes
    import struct

    lib = _core.library()
    n = ctypes.c_size_t()
    good = struct.pack("<BBBBIII", 2, 1, 0x21, 8, 0, 256, 16)
    assert lib.rang_core_blosc_size(good, len(good), ctypes.byref(n)) == 0 and n.value == 0
    fields = (ctypes.c_size_t * 4)()
    assert lib.rang_core_blosc_header(good, len(good), fields) == 0 and list(fields) == [0, 256, 0, 8]
    for bad in (struct.pack("<BBBBIII", 2, 1, 0x25, 8, 0, 256, 16),   # bit shuffle
                struct.pack("<BBBBIII", 2, 1, 0x41, 8, 0, 256, 16),   # not LZ4
                struct.pack("<BBBBIII", 9, 1, 0x21, 8, 0, 256, 16),   # version
                good[:12]):
        assert lib.rang_core_blosc_size(bad, len(bad), ctypes.byref(n)) == 1


def test_kernel_rows_through_the_abi_are_the_table():
    table = core_wgridder.kernel_table()
    assert len(table) == 939 == O.kernel_count()
    for i in range(0, 939, 7):
        assert table[i] == O.kernel_row(i)
    assert core_wgridder.kernel_row(939) is None


def test_host_side_errors_match_the_reference():
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nrow=6)
    kw = dict(uvw=uvw, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx, vis=vis, wgt=wgt)
    with pytest.raises(ValueError, match="ascending"):
        core_wgridder.vis2dirty(freq=freq[::-1], epsilon=1e-6, **kw)
    with pytest.raises(ValueError, match="positive"):
        core_wgridder.vis2dirty(freq=freq, epsilon=0.0, **kw)
    with pytest.raises(ValueError, match="positive"):
        core_wgridder.vis2dirty(freq=-freq, epsilon=1e-6, **kw)
    with pytest.raises(ValueError, match="sizes disagree"):
        _core.wgridder_vis2dirty_f64([nx, ny, dx, dx, 1e-6, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0],
                                     uvw, freq, vis[:-1], wgt, None, None,
                                     np.empty((nx, ny)), None, np.empty(14))
    with pytest.raises(ValueError, match="C-contiguous"):
        _core.wgridder_vis2dirty_f64([nx, ny, dx, dx, 1e-6, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0],
                                     uvw, freq, np.asfortranarray(vis), wgt, None, None,
                                     np.empty((nx, ny)), None, np.empty(14))
    with pytest.raises(ValueError, match="two weights"):
        _core.wgridder_vis2dirty_f64([nx, ny, dx, dx, 1e-6, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0],
                                     uvw, freq, vis, wgt, wgt.astype(np.float32), None,
                                     np.empty((nx, ny)), None, np.empty(14))
    with pytest.raises(ValueError, match="18 gridder options"):
        _core.wgridder_vis2dirty_f64([1.0] * 17, uvw, freq, vis, wgt, None, None,
                                     np.empty((nx, ny)), None, np.empty(14))


def test_grid_kernel_and_padding_without_a_gpu():
    from rang import core_grid

    assert np.array_equal(core_grid.kernel(7, 1023), O.ws_kernel(7, 1023))
    # core/src/grid.rs:954-970
    assert (core_grid.padded_size(512), core_grid.padded_size(100)) == (616, 120)
    with pytest.raises(ValueError, match="invalid kernel settings"):
        core_grid.kernel(0, 1023)
