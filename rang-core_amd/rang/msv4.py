"""MSv4 processing sets (zarr v3 stores) read into device memory.

Mirrors python/rang/msv4.py of rang-core (`open_processing_set`,
`ProcessingSet.partition` / `partitions`, `Partition.rows` / `chunks`, the
refusals and the row rule: rows time-major in baseline order, a
(time, baseline) cell whose UVW and VISIBILITY are all NaN a pad and
dropped, TIME in MJD seconds) without xarray or zarr: the store's
zarr.json files are read directly, and the chunks of the data variables
(blosc-1 with LZ4, as xradio writes them) are expanded on the GPU
(rang_amd_blosc_decode_batch) straight into device buffers.
`Partition.rows_device` keeps them there; `Partition.rows` copies them back
for the host API. Small coordinate arrays compressed with zstd are decoded
on the host (pyarrow's codec).
"""

import ctypes
import json
import math
from pathlib import Path

import numpy as np

from . import _core

UNIX_TO_MJD_SECONDS = 3506716800.0
GRID = ("time", "baseline_id", "frequency", "polarization")
DTYPES = {
    "correlated_data": (np.dtype("complex64"),),
    "flag": (np.dtype("bool"), np.dtype("uint8")),
    "weight": (np.dtype("float32"),),
    "uvw": (np.dtype("float64"),),
}
_NUMERIC = {"bool": "bool", "int8": "int8", "uint8": "uint8", "int16": "int16", "int32": "int32",
            "int64": "int64", "uint32": "uint32", "uint64": "uint64", "float32": "float32",
            "float64": "float64", "complex64": "complex64", "complex128": "complex128"}


def blosc_decode_device(chunks, outs, resident=None):
    """Expand blosc-1 chunks (bytes) into device tensors `outs` (uint8 views
    of each chunk's decoded length) in one upload and one launch; with
    `resident` (device tensors holding the same bytes) nothing is uploaded."""
    import torch

    lib = _core.library()
    keep = [np.frombuffer(c, np.uint8) for c in chunks]
    arr = (_core.BloscChunk * max(1, len(chunks)))()
    for i, (c, o) in enumerate(zip(keep, outs)):
        arr[i] = _core.BloscChunk(c.ctypes.data if c.size else None, c.size, o.data_ptr() if o.numel() else None,
                                  o.numel() * o.element_size())
    stream = torch.cuda.current_stream().cuda_stream
    if resident is not None:
        ptrs = (ctypes.c_void_p * max(1, len(chunks)))(*[r.data_ptr() for r in resident])
        code = lib.rang_amd_blosc_decode_resident(arr, ptrs, len(chunks), ctypes.c_void_p(stream))
    else:
        code = lib.rang_amd_blosc_decode_batch(arr, len(chunks), ctypes.c_void_p(stream))
    if code:
        n = lib.rang_amd_blosc_last_error(None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib.rang_amd_blosc_last_error(buf, n)
        raise ValueError(f"blosc chunk refused: {buf.raw[:n].decode()}")


def _zstd(raw, size):
    import pyarrow as pa

    return pa.Codec("zstd").decompress(raw, decompressed_size=size).to_pybytes()


class ZarrArray:
    """One zarr v3 array of a store: its metadata and its chunks."""

    def __init__(self, path):
        self.path = Path(path)
        meta = json.loads((self.path / "zarr.json").read_text())
        if meta.get("zarr_format") != 3 or meta.get("node_type") != "array":
            raise ValueError(f"{self.path}: not a zarr v3 array")
        self.meta = meta
        self.shape = tuple(int(s) for s in meta["shape"])
        self.dims = tuple(meta.get("dimension_names") or ())
        self.attrs = meta.get("attributes", {})
        grid = meta["chunk_grid"]
        if grid["name"] != "regular":
            raise ValueError(f"{self.path}: chunk grid {grid['name']} is not supported")
        self.chunks = tuple(int(s) for s in grid["configuration"]["chunk_shape"])
        enc = meta["chunk_key_encoding"]
        self.sep = enc.get("configuration", {}).get("separator", "/")
        self.key_kind = enc["name"]
        dt = meta["data_type"]
        if isinstance(dt, dict) and dt["name"] == "fixed_length_utf32":
            self.dtype = np.dtype(f"<U{int(dt['configuration']['length_bytes']) // 4}")
        elif isinstance(dt, str) and dt in _NUMERIC:
            self.dtype = np.dtype(_NUMERIC[dt])
        else:
            raise ValueError(f"{self.path}: data type {dt} is not supported")
        self.codec = None
        for c in meta["codecs"]:
            if c["name"] == "bytes":
                if c.get("configuration", {}).get("endian", "little") != "little" and self.dtype.itemsize > 1:
                    raise ValueError(f"{self.path}: big-endian chunks are not supported")
            elif c["name"] in ("blosc", "zstd") and self.codec is None:
                self.codec = c
            else:
                raise ValueError(f"{self.path}: codec {c['name']} is not supported")
        if self.codec and self.codec["name"] == "blosc":
            cfg = self.codec["configuration"]
            if cfg.get("cname") != "lz4" or cfg.get("shuffle") == "bitshuffle":
                raise ValueError(f"{self.path}: blosc {cfg.get('cname')}/{cfg.get('shuffle')} is not supported")
        self.fill = self._fill(meta.get("fill_value"))

    def _fill(self, v):
        if self.dtype.kind == "U":
            return v or ""
        if self.dtype.kind == "c":
            re, im = (v if isinstance(v, list) else (v, 0))
            return complex(float(re), float(im))
        if v in ("NaN", "Infinity", "-Infinity"):
            return float(v.replace("Infinity", "inf"))
        return 0 if v is None else v

    def _key(self, idx):
        if self.key_kind == "default":
            return self.path / ("c" + self.sep + self.sep.join(str(i) for i in idx)) if idx else self.path / "c"
        return self.path / self.sep.join(str(i) for i in idx)

    def _grid(self):
        counts = [max(1, math.ceil(s / c)) if s else 0 for s, c in zip(self.shape, self.chunks)]
        return np.ndindex(*counts) if self.shape else [()]

    def read(self):
        """The whole array on the host (NumPy)."""
        if self.codec and self.codec["name"] == "blosc":
            return self.read_device().cpu().numpy().astype(self.dtype, copy=False)
        out = np.full(self.shape, self.fill, dtype=self.dtype)
        nchunk = int(np.prod(self.chunks)) if self.chunks else 1
        for idx in self._grid():
            f = self._key(idx)
            if not f.exists():
                continue
            raw = f.read_bytes()
            if self.codec:
                raw = _zstd(raw, nchunk * self.dtype.itemsize)
            block = np.frombuffer(raw, self.dtype).reshape(self.chunks)
            sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, self.chunks, self.shape))
            out[sl] = block[tuple(slice(0, s.stop - s.start) for s in sl)]
        return out

    def read_device(self, device="cuda", index=None):
        """The array (or the chunks along axis 0 covering `index`, a slice of
        axis 0) in device memory, its blosc chunks expanded on the GPU."""
        import torch

        tdt = {"bool": torch.bool, "uint8": torch.uint8, "int32": torch.int32, "int64": torch.int64,
               "float32": torch.float32, "float64": torch.float64, "complex64": torch.complex64,
               "complex128": torch.complex128}.get(self.dtype.name)
        if tdt is None:
            raise ValueError(f"{self.path}: {self.dtype} does not go to the device")
        lo, hi = (0, self.shape[0]) if index is None else (index.start, index.stop)
        shape = (hi - lo,) + self.shape[1:]
        out = torch.empty(shape, dtype=tdt, device=device)
        if isinstance(self.fill, float) and math.isnan(self.fill) or self.fill:
            out.fill_(self.fill)
        else:
            out.zero_()
        nchunk = int(np.prod(self.chunks))
        raws, bufs, places = [], [], []
        for idx in self._grid():
            t0 = idx[0] * self.chunks[0]
            if t0 >= hi or t0 + self.chunks[0] <= lo:
                continue
            f = self._key(idx)
            if not f.exists():
                continue
            raw = f.read_bytes()
            buf = torch.empty(self.chunks, dtype=tdt, device=device)
            if self.codec and self.codec["name"] == "zstd":
                host = np.frombuffer(_zstd(raw, nchunk * self.dtype.itemsize), self.dtype).reshape(self.chunks)
                buf.copy_(torch.from_numpy(host.copy()))
            elif self.codec:
                raws.append(raw)
                bufs.append(buf.view(torch.uint8).reshape(-1) if buf.numel() else buf.view(torch.uint8))
            else:
                host = np.frombuffer(raw, self.dtype).reshape(self.chunks)
                buf.copy_(torch.from_numpy(host.copy()))
            places.append((idx, buf))
        if raws:
            blosc_decode_device(raws, bufs)
        for idx, buf in places:
            src, dst = [], []
            for a, (i, c, s) in enumerate(zip(idx, self.chunks, self.shape)):
                g0, g1 = i * c, min((i + 1) * c, s)
                if a == 0:
                    g0, g1 = max(g0, lo), min(g1, hi)
                    src.append(slice(g0 - i * c, g1 - i * c))
                    dst.append(slice(g0 - lo, g1 - lo))
                else:
                    src.append(slice(0, g1 - g0))
                    dst.append(slice(g0, g1))
            out[tuple(dst)] = buf[tuple(src)]
        return out


def _group(path):
    meta = json.loads((Path(path) / "zarr.json").read_text())
    if meta.get("node_type") != "group":
        raise ValueError(f"{path}: not a zarr group")
    return meta.get("attributes", {})


def _schema_major(attrs):
    version = str(attrs.get("schema_version", ""))
    return version.split(".")[0] if version else ""


def open_processing_set(source):
    return ProcessingSet(Path(source))


class ProcessingSet:
    """The partitions of a processing set; `refused` names those not read."""

    def __init__(self, root):
        self.root = Path(root)
        self.names = []
        self.refused = {}
        for child in sorted(p for p in self.root.iterdir() if (p / "zarr.json").is_file()):
            try:
                attrs = _group(child)
            except ValueError:
                continue
            if "data_groups" not in attrs:
                continue
            if _schema_major(attrs) != "4":
                self.refused[child.name] = f"schema_version {attrs.get('schema_version')!r}"
            elif attrs.get("type") != "visibility":
                self.refused[child.name] = f"type {attrs.get('type')!r}, not visibility"
            else:
                self.names.append(child.name)

    def partition(self, name, data_group="base"):
        if name in self.refused:
            raise ValueError(f"{name}: {self.refused[name]}")
        return Partition(self.root / name, name, data_group)

    def partitions(self, data_group="base"):
        return [self.partition(name, data_group) for name in self.names]


class Partition:
    """One MSv4 of a processing set: its grid and its rows."""

    def __init__(self, path, name, data_group="base"):
        self.path = Path(path)
        self.name = name
        attrs = _group(self.path)
        groups = attrs["data_groups"]
        if data_group not in groups:
            raise ValueError(f"{name}: no data group {data_group!r} in {list(groups)}")
        group = groups[data_group]
        self.variables = {}
        self.arrays = {}
        for role, dtypes in DTYPES.items():
            var = group.get(role)
            if var is None or not (self.path / var / "zarr.json").is_file():
                raise ValueError(f"{name}: data group {data_group!r} has no {role}")
            array = ZarrArray(self.path / var)
            want = GRID if role != "uvw" else ("time", "baseline_id", "uvw_label")
            if array.dims != want:
                raise ValueError(f"{name}: {var} has dimensions {array.dims}, not {want}")
            if array.dtype not in dtypes:
                want_dtype = " or ".join(str(d) for d in dtypes)
                raise ValueError(f"{name}: {var} is {array.dtype}, not {want_dtype}")
            self.variables[role] = var
            self.arrays[role] = array
        self.time_unix = np.asarray(ZarrArray(self.path / "time").read(), np.float64)
        self.time = self.time_unix + UNIX_TO_MJD_SECONDS
        self.frequency = np.asarray(ZarrArray(self.path / "frequency").read(), np.float64)
        self.polarization = tuple(str(p) for p in ZarrArray(self.path / "polarization").read())
        names = [str(a) for a in ZarrArray(self.path / "antenna_xds" / "antenna_name").read()]
        self.antenna_names = names
        index = {a: i for i, a in enumerate(names)}
        n1 = [str(a) for a in ZarrArray(self.path / "baseline_antenna1_name").read()]
        n2 = [str(a) for a in ZarrArray(self.path / "baseline_antenna2_name").read()]
        try:
            self.antenna1 = np.asarray([index[a] for a in n1], np.int32)
            self.antenna2 = np.asarray([index[a] for a in n2], np.int32)
        except KeyError as exc:
            raise ValueError(f"{name}: baseline antenna {exc} not in antenna_xds") from exc
        self.field_name = self._per_time("field_name")
        self.scan_name = self._per_time("scan_name")

    def _per_time(self, coord):
        if not (self.path / coord / "zarr.json").is_file():
            return np.full(len(self.time), "", dtype=object)
        return np.asarray([str(v) for v in ZarrArray(self.path / coord).read()], dtype=object)

    @property
    def nbaseline(self):
        return len(self.antenna1)

    def chunks(self, time_rows=None):
        """Row chunks of `time_rows` times each (all times by default)."""
        ntime = len(self.time)
        step = ntime if not time_rows else int(time_rows)
        for start in range(0, ntime, max(step, 1)):
            yield self.rows(start, min(start + step, ntime))

    def rows_device(self, start=0, stop=None, device="cuda"):
        """The rows of times `start` to `stop` (exclusive), pads dropped, as
        device tensors (DATA, FLAG, WEIGHT_SPECTRUM, UVW) beside host arrays
        (ANTENNA1, ANTENNA2, TIME, FIELD_NAME, SCAN_NAME, PADS)."""
        import torch

        stop = len(self.time) if stop is None else stop
        sel = slice(start, stop)
        vis = self.arrays["correlated_data"].read_device(device, sel)
        flag = self.arrays["flag"].read_device(device, sel)
        weight = self.arrays["weight"].read_device(device, sel)
        uvw = self.arrays["uvw"].read_device(device, sel)
        if flag.dtype == torch.uint8:
            bad = flag > 1
            if bool(bad.any()):
                values = torch.unique(flag[bad])[:8].tolist()
                raise ValueError(f"{self.name}: {self.variables['flag']} is uint8 with values {values} other "
                                 "than 0 and 1")
            flag = flag != 0
        ntime, nbl, nchan, ncorr = vis.shape
        uvw = uvw.reshape(ntime * nbl, 3)
        vis = vis.reshape(ntime * nbl, nchan, ncorr)
        # a pad: UVW all NaN and every sample NaN in either part
        pad = torch.isnan(uvw).all(dim=1) & torch.isnan(vis).reshape(ntime * nbl, -1).all(dim=1)
        keep = ~pad
        keep_host = keep.cpu().numpy()
        time = np.repeat(self.time[start:stop], nbl)
        return {
            "ANTENNA1": np.tile(self.antenna1, ntime)[keep_host],
            "ANTENNA2": np.tile(self.antenna2, ntime)[keep_host],
            "TIME": time[keep_host],
            "FLAG_ROW": np.zeros(int(keep_host.sum()), bool),
            "DATA": vis[keep],
            "FLAG": flag.reshape(ntime * nbl, nchan, ncorr)[keep],
            "WEIGHT_SPECTRUM": weight.reshape(ntime * nbl, nchan, ncorr)[keep],
            "UVW": uvw[keep],
            "FIELD_NAME": np.repeat(self.field_name[start:stop], nbl)[keep_host],
            "SCAN_NAME": np.repeat(self.scan_name[start:stop], nbl)[keep_host],
            "PADS": int(pad.sum().item()),
        }

    def rows(self, start=0, stop=None):
        """As `rows_device`, every array on the host (NumPy)."""
        out = self.rows_device(start, stop)
        for k in ("DATA", "FLAG", "WEIGHT_SPECTRUM", "UVW"):
            out[k] = out[k].cpu().numpy()
        return out

    def stokes_i_device(self, start=0, stop=None, device="cuda"):
        """The gridder's device inputs of the rows: `uvw` (nrow, 3) float64,
        `freq` (nchan,), `vis` (nrow, nchan) complex64 Stokes I and `wgt`
        (nrow, nchan) float32, WSClean's rule on XX, XY, YX, YY
        (python/rang/core_grid.py:154-185), formed on the GPU."""
        import torch

        if self.polarization != ("XX", "XY", "YX", "YY"):
            raise ValueError(f"Stokes I needs correlations XX, XY, YX, YY, found {self.polarization}")
        r = self.rows_device(start, stop, device)
        data = r["DATA"].contiguous()
        nrow, nchan = data.shape[:2]
        weight = r["WEIGHT_SPECTRUM"].contiguous()
        flag = r["FLAG"].to(torch.uint8).contiguous()
        vis = torch.empty((nrow, nchan), dtype=torch.complex64, device=data.device)
        wgt = torch.empty((nrow, nchan), dtype=torch.float32, device=data.device)
        stream = torch.cuda.current_stream(data.device).cuda_stream
        inp = _core.StokesIInputs(nrow * nchan, data.data_ptr(), weight.data_ptr(), flag.data_ptr())
        out = _core.StokesIOutputs(vis.data_ptr(), wgt.data_ptr())
        code = _core.library().rang_amd_stokes_i_dev(ctypes.byref(inp), ctypes.byref(out), ctypes.c_void_p(stream))
        if code:
            raise RuntimeError(f"Stokes I on the device failed with code {code}")
        return {"uvw": r["UVW"].contiguous(), "freq": self.frequency, "vis": vis, "wgt": wgt}
