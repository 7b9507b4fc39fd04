"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes access to the CPU restatement of rang-core's exact w-gridder
(`oracle/wg_*.c`, following core/src/wgridder.rs) and to the FP64 direct
DFT (`oracle/dft.c`). Only tests/, `__graft_entry__.smoke()` and bench.py's
`cpu_baseline` leg use this module, as the checker or the timed CPU
baseline; the product library (rang-core_amd/) never imports it.

The calls mirror the reference's `_ex2` / `dirty2vis_ex` entry points
(core/src/ffi_wgridder.rs:542-573, 668-720): the same option struct, the
same visibility and weight kinds, the same status codes.
"""

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
C_LIGHT = 299792458.0
N_INFO = 14
VIS_KIND = {np.dtype(np.complex64): 1, np.dtype(np.complex128): 2}
WGT_KIND = {np.dtype(np.float32): 1, np.dtype(np.float64): 2}


class Options(ctypes.Structure):
    """`WgridderOptions` (core/src/ffi_wgridder.rs:22-43)."""

    _fields_ = [
        ("nx", ctypes.c_size_t),
        ("ny", ctypes.c_size_t),
        ("pixsize_x", ctypes.c_double),
        ("pixsize_y", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("center_x", ctypes.c_double),
        ("center_y", ctypes.c_double),
        ("sigma_min", ctypes.c_double),
        ("sigma_max", ctypes.c_double),
        ("do_wgridding", ctypes.c_int32),
        ("divide_by_n", ctypes.c_int32),
        ("flip_u", ctypes.c_int32),
        ("flip_v", ctypes.c_int32),
        ("flip_w", ctypes.c_int32),
        ("allow_nshift", ctypes.c_int32),
        ("kernel_index", ctypes.c_int64),
        ("nu", ctypes.c_size_t),
        ("nv", ctypes.c_size_t),
    ]


def options(nx, ny, pixsize_x, pixsize_y, epsilon, **kw):
    kernel = kw.pop("kernel", None)
    o = Options(
        nx=int(nx),
        ny=int(ny),
        pixsize_x=float(pixsize_x),
        pixsize_y=float(pixsize_y),
        epsilon=float(epsilon),
        center_x=float(kw.pop("center_x", 0.0)),
        center_y=float(kw.pop("center_y", 0.0)),
        sigma_min=float(kw.pop("sigma_min", 1.1)),
        sigma_max=float(kw.pop("sigma_max", 2.6)),
        do_wgridding=int(bool(kw.pop("do_wgridding", True))),
        divide_by_n=int(bool(kw.pop("divide_by_n", False))),
        flip_u=int(bool(kw.pop("flip_u", False))),
        flip_v=int(bool(kw.pop("flip_v", False))),
        flip_w=int(bool(kw.pop("flip_w", False))),
        allow_nshift=int(bool(kw.pop("allow_nshift", True))),
        kernel_index=-1 if kernel is None else int(kernel[0]),
        nu=0 if kernel is None else int(kernel[1]),
        nv=0 if kernel is None else int(kernel[2]),
    )
    if kw:
        raise TypeError(f"unexpected options {sorted(kw)}")
    return o


_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle library missing: run make -C {HERE}")
        L = ctypes.CDLL(str(LIB))
        vp, sz, i32, dp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
        for suf in ("f32", "f64"):
            f = getattr(L, f"or_vis2dirty_{suf}")
            f.argtypes = [ctypes.POINTER(Options), sz, sz, vp, vp, vp, i32, vp, vp, i32, vp, vp, vp, vp]
            f.restype = i32
            f = getattr(L, f"or_dirty2vis_{suf}")
            f.argtypes = [ctypes.POINTER(Options), sz, sz, vp, vp, vp, vp, i32, vp, vp, vp]
            f.restype = i32
            f = getattr(L, f"or_fft_{suf}")
            f.argtypes = [sz, vp, i32]
            f.restype = None
        L.or_last_error.restype = ctypes.c_char_p
        L.or_kernel_count.restype = sz
        L.or_kernel_row.argtypes = [sz, vp]
        L.or_good_size_complex.argtypes = [sz]
        L.or_good_size_complex.restype = sz
        L.or_gauss_legendre_symmetric.argtypes = [i32, vp, vp]
        L.or_kernel_eval_many.argtypes = [i32, sz, vp, vp, vp, vp]
        L.or_plan.argtypes = [ctypes.POINTER(Options), i32, sz, dp, dp, vp]
        L.or_dft_dirty.argtypes = [sz, vp, vp, vp, vp, vp, sz, vp, vp, i32, vp]
        L.or_dft_vis.argtypes = [sz, vp, vp, vp, sz, vp, vp, vp, i32, vp, vp]
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads of the oracle's later calls (OMP_NUM_THREADS before its
    first use, omp_set_num_threads after)."""
    os.environ["OMP_NUM_THREADS"] = str(int(n))
    if _lib is not None:
        _lib.or_set_threads(int(n))


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(status):
    if status != 0:
        raise ValueError(lib().or_last_error().decode())


class KernelRow(ctypes.Structure):
    _fields_ = [
        ("w", ctypes.c_int),
        ("ofactor", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("e0", ctypes.c_double),
        ("ndim", ctypes.c_int),
        ("single", ctypes.c_int),
    ]


def kernel_row(index):
    r = KernelRow()
    if lib().or_kernel_row(int(index), ctypes.byref(r)):
        return None
    return (r.w, r.ofactor, r.epsilon, r.beta, r.e0, r.ndim, bool(r.single))


def kernel_count():
    return lib().or_kernel_count()


def good_size_complex(n):
    return lib().or_good_size_complex(int(n))


def gauss_legendre(n):
    m = (n + 1) // 2
    x = np.zeros(m)
    w = np.zeros(m)
    lib().or_gauss_legendre_symmetric(int(n), _p(x), _p(w))
    return x, w


def kernel_eval(index, v):
    """(polynomial kernel, ES kernel, correction function) of row `index` at v."""
    v = np.ascontiguousarray(v, np.float64)
    poly, phi, corr = (np.zeros_like(v) for _ in range(3))
    _check(lib().or_kernel_eval_many(int(index), v.size, _p(v), _p(poly), _p(phi), _p(corr)))
    return poly, phi, corr


def plan(nvis, wmin_d, wmax_d, single, **opt):
    o = options(**opt)
    info = np.zeros(N_INFO)
    _check(lib().or_plan(ctypes.byref(o), int(single), int(nvis), float(wmin_d), float(wmax_d), _p(info)))
    return info


def fft(x, inverse):
    """The oracle's unnormalised FFT of a complex vector (sign + for inverse)."""
    x = np.ascontiguousarray(x)
    single = x.dtype == np.complex64
    x = x.astype(np.complex64 if single else np.complex128).copy()
    getattr(lib(), f"or_fft_{'f32' if single else 'f64'}")(x.size, _p(x), 1 if inverse else -1)
    return x


def _coords(uvw, freq):
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    return uvw, freq, uvw.shape[0], freq.shape[0]


def vis2dirty(uvw, freq, vis, wgt=None, mask=None, *, npix_x, npix_y, pixsize_x, pixsize_y,
              epsilon, single, outputs=("dirty",), wgt2=None, **kw):
    """The oracle's gridding: `outputs` is a subset of ("dirty", "psf");
    returns (images dict, info array)."""
    uvw, freq, nrow, nchan = _coords(uvw, freq)
    o = options(npix_x, npix_y, pixsize_x, pixsize_y, epsilon, **kw)
    vkind = 0
    if vis is not None:
        vis = np.ascontiguousarray(vis)
        vkind = VIS_KIND[vis.dtype]
    wkind = 0
    if wgt is not None:
        wgt = np.ascontiguousarray(wgt)
        wkind = WGT_KIND[wgt.dtype]
    if wgt2 is not None:
        wgt2 = np.ascontiguousarray(wgt2, np.float32)
        wkind = 3
    if mask is not None:
        mask = np.ascontiguousarray(mask, np.uint8)
    real = np.float32 if single else np.float64
    imgs = {k: np.zeros((int(npix_x), int(npix_y)), real) for k in outputs}
    info = np.zeros(N_INFO)
    fn = getattr(lib(), f"or_vis2dirty_{'f32' if single else 'f64'}")
    st = fn(ctypes.byref(o), nrow, nchan, _p(uvw), _p(freq), _p(vis), vkind, _p(wgt), _p(wgt2),
            wkind, _p(mask), _p(imgs.get("dirty")), _p(imgs.get("psf")), _p(info))
    _check(st)
    return imgs, info


def dirty2vis(uvw, freq, dirty, wgt=None, mask=None, *, pixsize_x, pixsize_y, epsilon, **kw):
    """The oracle's degridding; precision follows `dirty`."""
    uvw, freq, nrow, nchan = _coords(uvw, freq)
    dirty = np.ascontiguousarray(dirty)
    single = dirty.dtype == np.float32
    if not single:
        dirty = dirty.astype(np.float64)
    o = options(dirty.shape[0], dirty.shape[1], pixsize_x, pixsize_y, epsilon, **kw)
    wkind = 0
    if wgt is not None:
        wgt = np.ascontiguousarray(wgt)
        wkind = WGT_KIND[wgt.dtype]
    if mask is not None:
        mask = np.ascontiguousarray(mask, np.uint8)
    vis = np.zeros((nrow, nchan), np.complex64 if single else np.complex128)
    info = np.zeros(N_INFO)
    fn = getattr(lib(), f"or_dirty2vis_{'f32' if single else 'f64'}")
    st = fn(ctypes.byref(o), nrow, nchan, _p(uvw), _p(freq), _p(dirty), _p(wgt), wkind, _p(mask),
            _p(vis), _p(info))
    _check(st)
    return vis, info


def _uvw_lambda(uvw, freq, flip_u=False, flip_v=False, flip_w=False):
    uvw = np.asarray(uvw, np.float64)
    f = np.asarray(freq, np.float64) / C_LIGHT
    s = np.array([-1.0 if flip_u else 1.0, -1.0 if flip_v else 1.0, -1.0 if flip_w else 1.0])
    lam = (uvw * s)[:, None, :] * f[None, :, None]
    return [np.ascontiguousarray(lam[..., i].ravel()) for i in range(3)]


def pixel_lm(nx, ny, dx, dy, pixels=None, center_x=0.0, center_y=0.0, flip_u=False, flip_v=False):
    """l, m of pixels (i, j) (all when None): l = lshift + (i - nx/2) dx with
    lshift = -center_x under flip_u (core/src/wgridder.rs:30-36, 615-624)."""
    if pixels is None:
        ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        pixels = (ii.ravel(), jj.ravel())
    i, j = (np.asarray(a) for a in pixels)
    l0 = -center_x if flip_u else center_x
    m0 = -center_y if flip_v else center_y
    return (np.ascontiguousarray(l0 + (i - nx // 2) * dx, np.float64),
            np.ascontiguousarray(m0 + (j - ny // 2) * dy, np.float64))


def dft_dirty(uvw, freq, vis, wgt, nx, ny, dx, dy, pixels=None, divide_by_n=False, mask=None,
              **flips):
    """FP64 direct sum over every visibility for the given pixels (all by
    default); returns a flat array over `pixels` (or an (nx, ny) image)."""
    center = {k: flips.pop(k) for k in ("center_x", "center_y") if k in flips}
    u, v, w = _uvw_lambda(uvw, freq, **flips)
    vis = np.ascontiguousarray(np.asarray(vis, np.complex128).ravel())
    wt = None if wgt is None else np.asarray(wgt, np.float64).ravel()
    if mask is not None:
        m = np.asarray(mask).ravel().astype(np.float64)
        wt = m if wt is None else wt * m
    wt = None if wt is None else np.ascontiguousarray(wt)
    l, m = pixel_lm(nx, ny, dx, dy, pixels, flip_u=flips.get("flip_u", False),
                    flip_v=flips.get("flip_v", False), **center)
    out = np.zeros(l.size)
    lib().or_dft_dirty(u.size, _p(u), _p(v), _p(w), _p(vis), _p(wt), l.size, _p(l), _p(m),
                       int(divide_by_n), _p(out))
    return out.reshape(nx, ny) if pixels is None else out


def dft_vis(uvw, freq, image, wgt, dx, dy, rows=None, divide_by_n=False, **flips):
    """FP64 direct degridding of `image` (nx, ny) for every visibility (or
    the flat visibility indices `rows`); complex128."""
    center = {k: flips.pop(k) for k in ("center_x", "center_y") if k in flips}
    nrow, nchan = np.asarray(uvw).shape[0], np.asarray(freq).shape[0]
    u, v, w = _uvw_lambda(uvw, freq, **flips)
    wt = None if wgt is None else np.asarray(wgt, np.float64).ravel()
    if rows is not None:
        u, v, w = (np.ascontiguousarray(a[rows]) for a in (u, v, w))
        wt = None if wt is None else wt[rows]
    wt = None if wt is None else np.ascontiguousarray(wt)
    nx, ny = image.shape
    l, m = pixel_lm(nx, ny, dx, dy, None, flip_u=flips.get("flip_u", False),
                    flip_v=flips.get("flip_v", False), **center)
    img = np.ascontiguousarray(np.asarray(image, np.float64).ravel())
    out = np.zeros(u.size, np.complex128)
    lib().or_dft_vis(u.size, _p(u), _p(v), _p(w), l.size, _p(l), _p(m), _p(img),
                     int(divide_by_n), _p(wt), _p(out))
    return out.reshape(nrow, nchan) if rows is None else out


# ---------------------------------------------------------------- w-stacking


class WsInputs(ctypes.Structure):
    """core/src/ffi_grid.rs:7-29."""
    _fields_ = [(n, ctypes.c_size_t) for n in ("nrow", "nchan", "width", "height", "kernel_size",
                                                "oversampling", "n_w_layers")] + [
        ("pixel_size_x", ctypes.c_double), ("pixel_size_y", ctypes.c_double),
        ("padding", ctypes.c_double), ("uvw", ctypes.c_void_p), ("frequencies", ctypes.c_void_p),
        ("data", ctypes.c_void_p), ("weights", ctypes.c_void_p), ("imaging_weights", ctypes.c_void_p)]


class WsOutputs(ctypes.Structure):
    _fields_ = [("dirty", ctypes.c_void_p), ("psf", ctypes.c_void_p), ("stats", ctypes.c_void_p)]


WS_STAT_NAMES = ("image_weight", "gridded_count", "visibility_weight_sum", "max_gridded_weight",
                 "min_w", "max_w", "psf_normalization", "n_w_layers")


def ws_kernel(kernel_size=7, oversampling=1023):
    out = np.zeros(kernel_size * oversampling)
    if lib().ws_kernel(ctypes.c_size_t(kernel_size), ctypes.c_size_t(oversampling), _p(out)):
        raise ValueError("invalid kernel size")
    return out


def ws_fft2_backward(grid):
    """The oracle's unnormalised backward 2-D DFT of a (height, width) complex
    array (rows of width, then columns)."""
    g = np.ascontiguousarray(grid, np.complex128).copy()
    lib().ws_fft2_backward(_p(g), ctypes.c_size_t(g.shape[1]), ctypes.c_size_t(g.shape[0]))
    return g


def ws_invert(uvw, frequencies, data, weights, imaging_weights, *, width, height, pixel_size,
              w_layers, kernel_size=7, oversampling=1023, padding=1.2):
    """The oracle's w-stacking inversion: (height, width) dirty and PSF in
    FITS order and the eight counters."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    frequencies = np.ascontiguousarray(frequencies, np.float64)
    data = np.ascontiguousarray(data, np.complex64)
    weights = np.ascontiguousarray(weights, np.float32)
    imaging_weights = np.ascontiguousarray(imaging_weights, np.float32)
    nrow, nchan = data.shape
    inp = WsInputs(nrow, nchan, width, height, kernel_size, oversampling, w_layers, pixel_size,
                   pixel_size, padding, uvw.ctypes.data, frequencies.ctypes.data, data.ctypes.data,
                   weights.ctypes.data, imaging_weights.ctypes.data)
    dirty = np.zeros((height, width))
    psf = np.zeros((height, width))
    stats = np.zeros(8)
    out = WsOutputs(dirty.ctypes.data, psf.ctypes.data, stats.ctypes.data)
    if lib().ws_invert(ctypes.byref(inp), ctypes.byref(out)):
        raise ValueError("invalid w-stacking inputs")
    rec = dict(zip(WS_STAT_NAMES, stats.tolist()))
    return {"dirty": dirty, "psf": psf, "stats": rec}


# --------------------------------------------------------------- weights
# The serial restatement of core/src/weights.rs (oracle/weights.c).

WEIGHT_CLASSES = {"natural": 0, "uniform": 1, "briggs": 2}


def _weights_lib():
    L = lib()
    if not getattr(L, "_weights_bound", False):
        vp, sz, i32, dp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
        L.or_padded_size.argtypes = [sz, dp]
        L.or_padded_size.restype = sz
        L.or_grid_shape.argtypes = [sz, sz, dp, vp]
        L.or_grid_shape.restype = i32
        L.or_stokes_i.argtypes = [sz, vp, vp, vp, vp, vp]
        L.or_stokes_i.restype = i32
        L.or_imaging_weights.argtypes = [sz, sz, vp, vp, vp, vp, i32, dp, sz, sz, dp, dp, dp, dp, sz, vp, vp, vp]
        L.or_imaging_weights.restype = i32
        L._weights_bound = True
    return L


def padded_size(trimmed, padding=1.2):
    return int(_weights_lib().or_padded_size(int(trimmed), float(padding)))


def grid_shape(width, height, super_weight=1.0):
    out = np.zeros(2, np.uintp)
    if _weights_lib().or_grid_shape(int(width), int(height), float(super_weight), _p(out)):
        raise ValueError("positive super weight required")
    return int(out[0]), int(out[1])


def stokes_i(data, weights, flags):
    data = np.ascontiguousarray(data, np.complex64)
    weights = np.ascontiguousarray(weights, np.float32)
    flags = np.ascontiguousarray(flags, np.uint8)
    shape = data.shape[:-1]
    n = int(np.prod(shape))
    vis = np.zeros(shape, np.complex64)
    w = np.zeros(shape, np.float32)
    _weights_lib().or_stokes_i(n, _p(data), _p(weights), _p(flags), _p(vis), _p(w))
    return vis, w


def imaging_weights(uvw, wavelength, weights, selected, *, mode="briggs", robust=0.5, width=616, height=616,
                    pixel_scale=8.0 * np.pi / (180.0 * 3600.0), super_weight=1.0, rank_filter=(3.0, 16)):
    """(weights f32 (nrow, nchan), grid (h/2, w), sums [vis weight sum, total, max, count])."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    wavelength = np.ascontiguousarray(wavelength, np.float64)
    weights = np.ascontiguousarray(weights, np.float32)
    selected = np.ascontiguousarray(selected, np.uint8)
    nrow, nchan = weights.shape
    gw, gh = grid_shape(width, height, super_weight)
    out = np.zeros((nrow, nchan), np.float32)
    grid = np.zeros((gh // 2, gw))
    sums = np.zeros(4)
    level, size = rank_filter if rank_filter is not None else (0.0, 0)
    st = _weights_lib().or_imaging_weights(nrow, nchan, _p(uvw), _p(wavelength), _p(weights), _p(selected),
                                           WEIGHT_CLASSES[mode], float(robust), int(width), int(height),
                                           float(pixel_scale), float(pixel_scale), float(super_weight),
                                           float(level), int(size), _p(out), _p(grid), _p(sums))
    if st:
        raise ValueError("invalid imaging-weight arguments")
    return out, grid, sums


def _blosc_lib():
    L = lib()
    if not getattr(L, "_blosc_bound", False):
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.or_blosc_header.argtypes = [vp, sz, vp]
        L.or_blosc_header.restype = i32
        L.or_blosc_decode.argtypes = [vp, sz, vp, sz]
        L.or_blosc_decode.restype = i32
        L.or_lz4_block.argtypes = [vp, sz, vp, sz]
        L.or_lz4_block.restype = i32
        L._blosc_bound = True
    return L


def blosc_header(chunk):
    """(nbytes, blocksize, nblocks, typesize, flags, cbytes) of a blosc-1
    chunk, or None when it is refused (oracle/blosc.c)."""
    b = np.frombuffer(bytes(chunk), np.uint8)
    f = np.zeros(6, np.uint64)
    if _blosc_lib().or_blosc_header(_p(b), b.size, _p(f)):
        return None
    return tuple(int(v) for v in f)


def blosc_decode(chunk):
    """The decoded bytes of a whole blosc-1 chunk, or None when refused."""
    h = blosc_header(chunk)
    if h is None:
        return None
    b = np.frombuffer(bytes(chunk), np.uint8)
    out = np.zeros(max(h[0], 1), np.uint8)
    if _blosc_lib().or_blosc_decode(_p(b), b.size, _p(out), h[0]):
        return None
    return out[: h[0]].tobytes()


def lz4_block(src, n):
    """One LZ4 block expanded to exactly `n` bytes, or None when corrupt."""
    b = np.frombuffer(bytes(src), np.uint8)
    out = np.zeros(max(n, 1), np.uint8)
    if _blosc_lib().or_lz4_block(_p(b) if b.size else None, b.size, _p(out), n):
        return None
    return out[:n].tobytes()
