"""The casacore table reader (include/rang_core_ms.h, rang.core_ms) against
rang-core's own fixtures, on the host: the table reader is host code (files
in, cells out), so these run without a GPU.

* `tests/golden/ms/msv2-edge`: three tables written by python-casacore
  3.8.1 (TiledCellStMan; IncrementalStMan scalars, strings and arrays;
  StandardStMan and IncrementalStMan string arrays) with arcae 0.5.9's cells
  in oracle.json, as tests/test_core_ms_edge.py of rang-core reads them;
* `tests/golden/ms/strarray`: a table with python-casacore's cells in
  oracle.json (tests/test_core_ms_strarray.py of rang-core);
* `tests/golden/ms/msv4/m05.zip`: the two-window Measurement Set (m05.ms)
  beside its MSv4 conversion (tests/test_msv4.py of rang-core).

Every cell is compared exactly (bytes and dtype, or the string lists).
"""

import hashlib
import json
import zipfile
from pathlib import Path

import numpy as np
import pytest

from rang import core_ms

GOLD = Path(__file__).resolve().parent / "golden" / "ms"
EDGE = GOLD / "msv2-edge"
STR = GOLD / "strarray"
ORACLE = json.loads((EDGE / "oracle.json").read_text(encoding="utf-8"))["tables"]
SORACLE = json.loads((STR / "oracle.json").read_text(encoding="utf-8"))
CASES = [(t, c) for t, e in ORACLE.items() for c in e["columns"]]


def _want(entry):
    if entry["dtype"] == "str":
        return np.asarray(entry["values"], dtype=object).reshape(entry["shape"])
    return np.asarray(entry["values"], dtype=entry["dtype"]).reshape(entry["shape"])


def _same(got, want):
    """Exact equality; arcae's booleans are uint8 0 and 1."""
    got = np.asarray(got)
    if want.dtype == np.uint8 and got.dtype == bool:
        got = got.astype(np.uint8)
    if want.dtype == object:
        return got.shape == want.shape and got.tolist() == want.tolist()
    return got.shape == want.shape and got.dtype == want.dtype and got.tobytes() == want.tobytes()


def test_fixtures_are_the_references():
    for line in (GOLD / "SHA256SUMS").read_text().splitlines():
        digest, rel = line.split("  ", 1)
        assert hashlib.sha256((GOLD / rel).read_bytes()).hexdigest() == digest, rel


@pytest.mark.parametrize("case,name", CASES)
def test_every_cell_equals_arcae(case, name):
    cells = ORACLE[case]["columns"][name]
    with core_ms.open_table(EDGE / case) as t:
        assert t.nrows == ORACLE[case]["nrows"]
        assert t.columns[name]["supported"]
        for row, entry in enumerate(cells):
            if entry is None:
                with pytest.raises(OSError, match="has no (tiled )?array"):
                    t.getcol(name, row, 1)
                continue
            assert _same(t.getcol(name, row, 1)[0], _want(entry)), (case, name, row)
        defined = [r for r, e in enumerate(cells) if e is not None][::-1]
        shapes = {tuple(cells[r]["shape"]) for r in defined}
        if len(shapes) == 1:
            got = t.getcol_rows(name, defined)
            for i, r in enumerate(defined):
                assert _same(got[i], _want(cells[r])), (case, name, r)
        else:
            with pytest.raises(OSError, match="differ in shape"):
                t.getcol_rows(name, defined)


def test_incremental_scalars_read_beside_arrays():
    with core_ms.open_table(EDGE / "t06-ism") as t:
        assert t.getcol("I32").tolist() == [1, 1, 1, 1, 2, 2, 2, 3, 3, 3]
        assert t.getcol("B", 3, 3).tolist() == [True, False, False]
        assert t.getcol("S", 6, 2).tolist() == ["value 2", "value 3"]
        assert t.getcol("ARR", 2, 3).tolist() == [[1.0] * 3] * 2 + [[2.0] * 3]


def test_tiled_cell_sliced_read():
    entry = ORACLE["t05-tcellsm"]["columns"]["F32"][2]
    with core_ms.open_table(EDGE / "t05-tcellsm") as t:
        got = t.getcol("F32", 2, 1, chans=slice(1, 3))[0]
        assert np.array_equal(got, _want(entry)[1:3])
        # a stepped slice is cut on the host, with the same values
        assert np.array_equal(t.getcol("F32", 2, 1, chans=slice(0, 4, 2))[0], _want(entry)[0:4:2])


@pytest.fixture
def strtable():
    with core_ms.open_table(STR / "table") as t:
        yield t


@pytest.mark.parametrize("name", list(SORACLE["columns"]))
def test_every_string_cell_equals_casacore(strtable, name):
    assert strtable.nrows == SORACLE["nrows"]
    for row, entry in enumerate(SORACLE["columns"][name]):
        if entry is None:
            with pytest.raises(OSError, match="has no array"):
                strtable.getcol(name, row, 1)
            continue
        got = strtable.getcol(name, row, 1)
        want = np.asarray(entry["values"], dtype=object).reshape(entry["shape"])
        assert got.shape == (1,) + want.shape, (name, row)
        assert got.tolist() == [want.tolist()], (name, row)


@pytest.mark.parametrize("name", ["POLARIZATION_TYPE", "FIXED", "DIRECT", "NAME", "UNSET_FIXED"])
def test_whole_string_columns(strtable, name):
    want = np.stack([np.asarray(e["values"], dtype=object).reshape(e["shape"]) for e in SORACLE["columns"][name]])
    got = strtable.getcol(name)
    assert got.shape == want.shape and got.tolist() == want.tolist()
    rows = [7, 2, 3, 11, 0]
    assert strtable.getcol_rows(name, rows).tolist() == want[rows].tolist()


def test_string_edge_cases(strtable):
    got = strtable.getcol("POLARIZATION_TYPE")
    assert got.shape == (12, 2) and got[1].tolist() == ["X", "Y"] and got[0].tolist() == ["R", "L"]
    # LONG's first strings (300 to 707 characters) exceed the 512-byte bucket
    got = strtable.getcol("LONG", 11, 1)[0]
    assert len(got[0]) == 300 + 37 * 11 and got[1] == "t11"
    with pytest.raises(OSError, match="differ in shape"):
        strtable.getcol("VAR", 0, 2)
    assert strtable.columns["INCR"]["supported"]
    assert strtable.getcol("INCR").tolist() == [["X", "Y"]] * SORACLE["nrows"]
    assert strtable.getcol_rows("INCR", [5, 0]).tolist() == [["X", "Y"]] * 2


@pytest.fixture(scope="module")
def m05(tmp_path_factory):
    d = tmp_path_factory.mktemp("m05")
    with zipfile.ZipFile(GOLD / "msv4" / "m05.zip") as z:
        z.extractall(d)
    return d


def test_measurement_set_main_table(m05):
    with core_ms.open_table(m05 / "m05.ms") as t:
        assert t.nrows == 36
        cols = t.columns
        assert cols["DATA"]["manager_type"] == "TiledShapeStMan" and cols["UVW"]["manager_type"] == "StandardStMan"
        assert all(c["supported"] for c in cols.values())
        ddid = t.getcol("DATA_DESC_ID")
        assert ddid.dtype == np.int32 and sorted(set(ddid.tolist())) == [0, 1]
        # rows sorted by (time, data description, baseline), ANTENNA1 < ANTENNA2
        a1, a2 = t.getcol("ANTENNA1"), t.getcol("ANTENNA2")
        assert (a1 < a2).all()
        uvw = t.getcol("UVW")
        assert uvw.shape == (36, 3) and uvw.dtype == np.float64
        assert np.array_equal(t.getcol_rows("UVW", [35, 0]), uvw[[35, 0]])
        # two windows of 4 and 8 channels: cells of DATA differ in shape
        assert t.getcol("DATA", 0, 1).shape == (1, 4, 4) and t.getcol("DATA", 35, 1).shape == (1, 8, 4)
        with pytest.raises(OSError, match="differ in shape"):
            t.getcol("DATA")
        rows = np.flatnonzero(ddid == 1)
        data = t.getcol_rows("DATA", rows)
        assert data.shape == (rows.size, 8, 4) and data.dtype == np.complex64
        for i, r in enumerate(rows[:5]):
            assert data[i].tobytes() == t.getcol("DATA", int(r), 1)[0].tobytes()
        # a channel slice reads only those channels
        part = t.getcol_rows("DATA", rows, chans=slice(2, 5))
        assert part.tobytes() == np.ascontiguousarray(data[:, 2:5]).tobytes()
        assert t.getcol("FLAG", 0, 1).dtype == np.bool_
        with pytest.raises(OSError, match="does not exist"):
            t.getcol("NO_SUCH_COLUMN")
        with pytest.raises(OSError, match="exceed"):
            t.getcol("TIME", 30, 10)
    with core_ms.open_table(m05 / "m05.ms" / "SPECTRAL_WINDOW") as s:
        assert s.getcol("CHAN_FREQ", 1, 1)[0].tolist() == [1.1e9 + 1e6 * k for k in range(8)]
    with core_ms.open_table(m05 / "m05.ms" / "POLARIZATION") as p:
        assert p.getcol("CORR_TYPE", 0, 1)[0].tolist() == [9, 10, 11, 12]
    with core_ms.open_table(m05 / "m05.ms" / "ANTENNA") as a:
        assert a.getcol("NAME").tolist() == ["A0", "A1", "A2", "A3"]


def test_missing_table_is_an_error(tmp_path):
    with pytest.raises(OSError, match="table.dat"):
        core_ms.open_table(tmp_path / "nothing")
