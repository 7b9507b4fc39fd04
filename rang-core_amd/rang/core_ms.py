"""Measurement Set reads at the edge of the gridder, and ingest to the device.

Mirrors python/rang/core_ms.py of rang-core: `open_table(path)` returns a
`Table` whose `getcol` / `getcol_rows` return columns as python-casacore's
`getcol` does (a leading row axis, then the cell axes in reversed casacore
order; NumPy dtypes bool, int32, float32, float64, complex64 and friends;
string columns as object arrays), reading through this library's casacore
reader (C ABI rang_core_ms_*, include/rang_core_ms.h). Errors of the reader
raise OSError with its message, as the reference's extension does.

`Table.getcol_device` and `stokes_i_to_device` are the ingest: columns
copied into device memory and, for four-correlation data, the Stokes I
visibilities and weights formed on the GPU (WSClean's rule,
python/rang/core_grid.py:154-185), ready for the gridder's device calls.
"""

import ctypes
import json
from pathlib import Path

import numpy as np

from . import _core

DTYPES = {
    0: np.dtype("bool"),
    1: np.dtype("int8"),
    2: np.dtype("uint8"),
    3: np.dtype("int16"),
    4: np.dtype("uint16"),
    5: np.dtype("int32"),
    6: np.dtype("uint32"),
    7: np.dtype("float32"),
    8: np.dtype("float64"),
    9: np.dtype("complex64"),
    10: np.dtype("complex128"),
    29: np.dtype("int64"),
}
STRING = 11


def has_reader():
    """True when the library exposes the Measurement Set reader."""
    try:
        return hasattr(_core.library(), "rang_core_ms_open")
    except RuntimeError:
        return False


def _error():
    lib = _core.library()
    n = lib.rang_core_ms_last_error(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib.rang_core_ms_last_error(buf, n)
    return buf.raw[:n].decode("utf-8", "replace")


def _check(code):
    if code == 1:
        raise OSError(_error())
    if code != 0:
        raise RuntimeError(f"Measurement Set reader failure: {_error()}")


def _take(col):
    """A filled MsColumn as (raw bytes, nrows, cell shape, dtype, nstrings),
    the buffer released."""
    lib = _core.library()
    try:
        raw = ctypes.string_at(col.data, col.nbytes) if col.nbytes else b""
    finally:
        lib.rang_core_ms_free(ctypes.byref(col))
    return raw, int(col.nrows), tuple(int(col.shape[i]) for i in range(col.ndim)), int(col.dtype), int(col.nstrings)


def _decode(raw, nrows, cell, dtype, nstrings):
    """A column's packed bytes as python-casacore's `getcol` returns them."""
    shape = (int(nrows),) + tuple(cell)[::-1]
    if dtype == STRING:
        strings, at = [], 0
        for _ in range(nstrings):
            n = int.from_bytes(raw[at : at + 4], "little")
            strings.append(raw[at + 4 : at + 4 + n].decode("utf-8"))
            at += 4 + n
        return np.asarray(strings, dtype=object).reshape(shape)
    return np.frombuffer(bytearray(raw), dtype=DTYPES[dtype]).reshape(shape)


def _slice_args(cut):
    if cut is None:
        return None, 0
    flat = (ctypes.c_int64 * (2 * len(cut)))(*[int(v) for pair in cut for v in pair])
    return flat, len(cut)


class Table:
    """An open casacore table; use as a context manager or call `close`."""

    def __init__(self, path):
        self.path = Path(path)
        lib = _core.library()
        handle = ctypes.c_void_p()
        _check(lib.rang_core_ms_open(str(self.path).encode(), ctypes.byref(handle)))
        self._table = handle
        n = lib.rang_core_ms_describe(handle, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib.rang_core_ms_describe(handle, buf, n)
        self.description = json.loads(buf.raw[:n].decode("utf-8"))
        self.nrows = int(self.description["nrows"])
        self.columns = {c["name"]: c for c in self.description["columns"]}

    def close(self):
        if self._table is not None:
            _core.library().rang_core_ms_close(self._table)
        self._table = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def colnames(self):
        return list(self.columns)

    def subtable(self, name):
        return Table(self.path / name)

    def _open(self):
        if self._table is None:
            raise ValueError("table is closed")
        return self._table

    def getcol(self, name, startrow=0, nrow=None, chans=None):
        """Rows `startrow` onwards (all, or `nrow` of them) of one column.
        `chans` (a slice of the first cell axis, the channels of a
        visibility column) reads only those channels; the tiled managers
        read only the tiles they meet."""
        handle = self._open()
        count = 0 if nrow is None else int(nrow)
        if nrow is not None and count <= 0:
            raise ValueError("nrow must be positive")
        cut = reader_slice(chans)
        if cut is False:
            return self.getcol(name, startrow, nrow)[:, chans]
        flat, n = _slice_args(cut)
        col = _core.MsColumn()
        _check(_core.library().rang_amd_ms_read_slice(handle, name.encode(), int(startrow), count, flat, n,
                                                       ctypes.byref(col)))
        return _decode(*_take(col))

    def getcol_rows(self, name, rows, chans=None):
        """The rows `rows` (any order) of one column, in that order, read in
        one call; `chans` as `getcol`."""
        handle = self._open()
        cut = reader_slice(chans)
        if cut is False:
            return self.getcol_rows(name, rows)[:, chans]
        rows = np.ascontiguousarray(rows, np.uint64)
        flat, n = _slice_args(cut)
        col = _core.MsColumn()
        _check(_core.library().rang_amd_ms_read_rows(
            handle, name.encode(), rows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), rows.size, flat, n,
            ctypes.byref(col)))
        return _decode(*_take(col))

    def getcol_device(self, name, startrow=0, nrow=None, chans=None, device="cuda"):
        """As `getcol` for a numeric or boolean column of fixed cell shape,
        the values written straight into a new torch tensor on `device`
        (rang_amd_ms_read_dev; booleans as uint8)."""
        import torch

        handle = self._open()
        desc = self.columns[name]
        if desc["dtype"] == "str":
            raise OSError(f"{name}: string columns do not go to the device")
        count = (self.nrows - int(startrow)) if nrow is None else int(nrow)
        cell = [int(v) for v in desc["shape"]]
        if desc["is_array"] and not cell:
            raise OSError(f"{name}: a column without a fixed cell shape has no size before it is read")
        cut = reader_slice(chans)
        if cut is False:
            raise ValueError("chans must be a plain slice for a device read")
        if cut is not None:
            region = _resolve(cut, cell)
            cell = [n for _, n in region]
        dt = {v.name: k for k, v in DTYPES.items()}[("uint8" if desc["dtype"] == "bool" else desc["dtype"])]
        npd = DTYPES[dt]
        tdt = {"uint8": torch.uint8, "int8": torch.int8, "int16": torch.int16, "int32": torch.int32,
               "int64": torch.int64, "float32": torch.float32, "float64": torch.float64,
               "complex64": torch.complex64, "complex128": torch.complex128}[npd.name]
        shape = (count,) + tuple(cell[::-1])
        out = torch.empty(shape, dtype=tdt, device=device)
        flat, n = _slice_args(cut)
        col = _core.MsColumn()
        stream = torch.cuda.current_stream(out.device).cuda_stream
        _check(_core.library().rang_amd_ms_read_dev(handle, name.encode(), int(startrow), count if nrow else 0,
                                                     flat, n, ctypes.c_void_p(out.data_ptr()),
                                                     out.numel() * out.element_size(), ctypes.c_void_p(stream),
                                                     ctypes.byref(col)))
        got = (int(col.nrows),) + tuple(int(col.shape[i]) for i in range(col.ndim))[::-1]
        if got != shape:
            raise OSError(f"{name}: read {got}, expected {shape}")
        return out


def _resolve(cut, shape):
    """A reader slice against a cell of `shape` (casacore order)."""
    full = [(0, -1)] * (len(shape) - len(cut)) + list(cut)
    out = []
    for (a, n), size in zip(full, shape):
        n = size - a if n < 0 else n
        out.append((a, n))
    return out


def reader_slice(chans):
    """`chans` (None or a slice of NumPy's first cell axis) as the reader's
    cell slice (the last casacore axis: first pixel and length, -1 to the
    end); None for whole cells, False when the reader cannot take it (a
    step, a negative or empty bound: the caller cuts in NumPy).
    (python/rang/core_ms.py:108-128 of rang-core.)"""
    if chans is None:
        return None
    if not isinstance(chans, slice) or chans.step not in (None, 1):
        return False
    start = 0 if chans.start is None else int(chans.start)
    if start < 0:
        return False
    if chans.stop is None:
        return [(start, -1)]
    stop = int(chans.stop)
    if stop <= start:
        return False
    return [(start, stop - start)]


def open_table(path):
    return Table(path)


def stokes_i_to_device(path, ddid=0, data_column="DATA", device="cuda"):
    """The rows of data description `ddid` of the Measurement Set at `path`
    as the gridder's device inputs: `uvw` (nrow, 3) float64, `freq` (nchan,)
    float64 (the spectral window's CHAN_FREQ), `vis` (nrow, nchan) complex64
    Stokes I and `wgt` (nrow, nchan) float32, zero where either of XX and YY
    is flagged or not finite and on flagged rows (WSClean's reader rule;
    WEIGHT_SPECTRUM, or WEIGHT spread over the channels). The columns go
    to the device as read; the Stokes I forming runs there
    (rang_amd_stokes_i_dev)."""
    import torch

    path = Path(path)
    with open_table(path) as t:
        ddids = t.getcol("DATA_DESC_ID")
        rows = np.flatnonzero(ddids == ddid).astype(np.uint64)
        if rows.size == 0:
            raise ValueError(f"no rows of data description {ddid}")
        with t.subtable("DATA_DESCRIPTION") as dd:
            spw = int(dd.getcol("SPECTRAL_WINDOW_ID")[ddid])
            pol = int(dd.getcol("POLARIZATION_ID")[ddid])
        with t.subtable("SPECTRAL_WINDOW") as sw:
            freq = np.ascontiguousarray(sw.getcol("CHAN_FREQ", spw, 1)[0], np.float64)
        with t.subtable("POLARIZATION") as pt:
            corr = [int(c) for c in pt.getcol("CORR_TYPE", pol, 1)[0]]
        if corr != [9, 10, 11, 12]:
            raise ValueError(f"Stokes I needs correlations XX, XY, YX, YY (types 9-12), found {corr}")
        uvw = np.ascontiguousarray(t.getcol_rows("UVW", rows), np.float64)
        data = np.ascontiguousarray(t.getcol_rows(data_column, rows), np.complex64)
        flag = np.ascontiguousarray(t.getcol_rows("FLAG", rows), np.uint8)
        flag_row = t.getcol_rows("FLAG_ROW", rows).astype(bool)
        if "WEIGHT_SPECTRUM" in t.columns and t.columns["WEIGHT_SPECTRUM"]["supported"]:
            try:
                weight = np.ascontiguousarray(t.getcol_rows("WEIGHT_SPECTRUM", rows), np.float32)
            except OSError:
                weight = None
        else:
            weight = None
        if weight is None:
            w = np.ascontiguousarray(t.getcol_rows("WEIGHT", rows), np.float32)
            weight = np.ascontiguousarray(np.broadcast_to(w[:, None, :], data.shape), np.float32)
    flag[flag_row] = 1
    dev = torch.device(device)
    tdata = torch.from_numpy(data.view(np.float32)).to(dev)
    tweight = torch.from_numpy(weight).to(dev)
    tflag = torch.from_numpy(flag).to(dev)
    nrow, nchan = data.shape[:2]
    vis = torch.empty((nrow, nchan), dtype=torch.complex64, device=dev)
    wgt = torch.empty((nrow, nchan), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    inp = _core.StokesIInputs(nrow * nchan, tdata.data_ptr(), tweight.data_ptr(), tflag.data_ptr())
    out = _core.StokesIOutputs(vis.data_ptr(), wgt.data_ptr())
    code = _core.library().rang_amd_stokes_i_dev(ctypes.byref(inp), ctypes.byref(out), ctypes.c_void_p(stream))
    if code:
        raise RuntimeError(f"Stokes I on the device failed with code {code}")
    return {
        "uvw": torch.from_numpy(uvw).to(dev),
        "freq": freq,
        "vis": vis,
        "wgt": wgt,
        "rows": rows,
    }
