"""Exact gridder and degridder (ducc0-style w-gridding) on the MI355X.

The Python surface of rang-core's `rang.core_wgridder`
(python/rang/core_wgridder.py): `vis2dirty`, `dirty_psf`, `dirty2vis`,
`fits_image`, `kernel_row`, `kernel_table` and `info_dict`, with the same
keywords, defaults, dtype rules and errors. Every call runs through
`rang._core` (the C ABI of include/rang_core_wgridder.h) on the GPU.

Conventions: the image is (nx, ny) with the first axis l, pixel (i, j) at
l = (i - nx/2) dx, m = (j - ny/2) dy; vis2dirty computes
sum wgt Re[V exp(+2 pi i (u l + v m + w (n-1)))] and dirty2vis its adjoint
(core/src/wgridder.rs:30-36). The precision of a call follows its data:
complex64 visibilities or float32 images work in single precision.
"""

import numpy as np

from . import _core

N_INFO = _core.N_INFO
INFO_NAMES = (
    "kernel_index",
    "supp",
    "nu",
    "nv",
    "nplanes",
    "nvis",
    "wmin_d",
    "wmax_d",
    "nm1min",
    "nm1max",
    "nshift",
    "dw",
    "wmin",
    "ofactor",
)
ARCSEC = np.pi / (180.0 * 3600.0)
VIS_DTYPES = (np.dtype(np.complex64), np.dtype(np.complex128))
WGT_DTYPES = (np.dtype(np.float32), np.dtype(np.float64))

# Option order of WgridderOptions and the reference's defaults
# (core/src/wgridder.rs:446-467, python/rang/core_wgridder.py:77-103).
_DEFAULTS = (
    ("center_x", 0.0),
    ("center_y", 0.0),
    ("sigma_min", 1.1),
    ("sigma_max", 2.6),
    ("do_wgridding", True),
    ("divide_by_n", False),
    ("flip_u", False),
    ("flip_v", False),
    ("flip_w", False),
    ("allow_nshift", True),
)


def extension():
    """The module that runs the gridder (`rang._core`)."""
    return _core


def backend():
    return "extension"


def library_path():
    return _core.library_path()


def _options(nx, ny, px, py, epsilon, kw):
    kw = dict(kw)
    kernel = kw.pop("kernel", None)
    vals = [int(nx), int(ny), float(px), float(py), float(epsilon)]
    for name, default in _DEFAULTS:
        v = kw.pop(name, default)
        vals.append(int(bool(v)) if isinstance(default, bool) else float(v))
    if kernel is None:
        vals += [-1, 0, 0]
    else:
        vals += [int(kernel[0]), int(kernel[1]), int(kernel[2])]
    if kw:
        raise TypeError(f"unexpected options {sorted(kw)}")
    return [float(v) for v in vals]


def _check(status):
    if status == 0:
        return
    if status == 1:
        raise ValueError(_core.wgridder_last_error())
    raise RuntimeError(f"core gridder failed with status {status}: {_core.wgridder_last_error()}")


def _coords(uvw, freq):
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    if uvw.ndim != 2 or uvw.shape[1] != 3 or freq.ndim != 1:
        raise ValueError("uvw must be (nrow, 3) and freq (nchan,)")
    return uvw, freq, (uvw.shape[0], freq.shape[0])


def _weights(wgt, mask, shape):
    """float32/float64 weights are read in place (other dtypes become
    float64); the mask becomes uint8."""
    if wgt is not None:
        wgt = np.asarray(wgt)
        if wgt.dtype not in WGT_DTYPES:
            wgt = wgt.astype(np.float64)
        wgt = np.ascontiguousarray(wgt)
        if wgt.shape != shape:
            raise ValueError("wgt must be (nrow, nchan)")
    if mask is not None:
        mask = np.ascontiguousarray(mask, np.uint8)
        if mask.shape != shape:
            raise ValueError("mask must be (nrow, nchan)")
    return wgt, mask


def _visibilities(vis, shape):
    vis = np.asarray(vis)
    if vis.dtype not in VIS_DTYPES:
        vis = vis.astype(np.complex128)
    vis = np.ascontiguousarray(vis)
    if vis.shape != shape:
        raise ValueError("visibility arrays must be (nrow, nchan)")
    return vis


def _second_weight(wgt, wgt2, shape):
    """A float32 imaging weight beside a float32 weight is multiplied in the
    library; any other pair is multiplied here."""
    if wgt2 is None:
        return wgt, None
    if wgt is None:
        raise ValueError("wgt2 needs wgt")
    wgt2 = np.asarray(wgt2)
    if wgt.dtype == np.float32 and wgt2.dtype == np.float32:
        wgt2 = np.ascontiguousarray(wgt2)
        if wgt2.shape != shape:
            raise ValueError("wgt2 must be (nrow, nchan)")
        return wgt, wgt2
    return np.ascontiguousarray(wgt * wgt2), None


def _grid(uvw, freq, vis, wgt, mask, npix, pix, epsilon, single, outputs, kw, wgt2=None):
    uvw, freq, shape = _coords(uvw, freq)
    vis = _visibilities(vis, shape)
    wgt, mask = _weights(wgt, mask, shape)
    wgt, wgt2 = _second_weight(wgt, wgt2, shape)
    opts = _options(npix[0], npix[1], pix[0], pix[1], epsilon, kw)
    real = np.float32 if single else np.float64
    images = {k: np.empty((int(npix[0]), int(npix[1])), real) for k in outputs}
    info = np.empty(N_INFO, np.float64)
    fn = _core.wgridder_vis2dirty_f32 if single else _core.wgridder_vis2dirty_f64
    _check(fn(opts, uvw, freq, vis, wgt, wgt2, mask, images.get("dirty"), images.get("psf"), info))
    return images, info_dict(info)


def vis2dirty(*, uvw, freq, vis, wgt=None, mask=None, npix_x, npix_y, pixsize_x, pixsize_y,
              epsilon, return_info=False, **kw):
    """Dirty image (npix_x, npix_y) of `vis` (nrow, nchan); single precision
    for complex64 data, double otherwise. Keywords as ducc0's
    (do_wgridding, divide_by_n, flip_u/v/w, center_x/y, allow_nshift,
    sigma_min/max) plus kernel=(row, nu, nv)."""
    single = np.asarray(vis).dtype == np.complex64
    images, info = _grid(uvw, freq, vis, wgt, mask, (npix_x, npix_y), (pixsize_x, pixsize_y),
                         epsilon, single, ("dirty",), kw)
    return (images["dirty"], info) if return_info else images["dirty"]


def dirty_psf(*, uvw, freq, vis, wgt=None, mask=None, npix_x, npix_y, pixsize_x, pixsize_y,
              epsilon, single=False, return_info=False, wgt2=None, **kw):
    """Dirty image and PSF in one pass, both on the plan of the visibilities
    with nonzero weight times mask (python/rang/core_wgridder.py:234-274)."""
    images, info = _grid(uvw, freq, vis, wgt, mask, (npix_x, npix_y), (pixsize_x, pixsize_y),
                         epsilon, single, ("dirty", "psf"), kw, wgt2=wgt2)
    if return_info:
        return images["dirty"], images["psf"], info
    return images["dirty"], images["psf"]


def dirty2vis(*, uvw, freq, dirty, wgt=None, mask=None, pixsize_x, pixsize_y, epsilon,
              return_info=False, **kw):
    """Visibilities (nrow, nchan) of the image `dirty` (nx, ny); single
    precision for a float32 image."""
    dirty = np.asarray(dirty)
    single = dirty.dtype == np.float32
    real = np.float32 if single else np.float64
    dirty = np.ascontiguousarray(dirty, real)
    uvw, freq, shape = _coords(uvw, freq)
    wgt, mask = _weights(wgt, mask, shape)
    opts = _options(dirty.shape[0], dirty.shape[1], pixsize_x, pixsize_y, epsilon, kw)
    vis = np.empty(shape, np.complex64 if single else np.complex128)
    info = np.empty(N_INFO, np.float64)
    fn = _core.wgridder_dirty2vis_f32 if single else _core.wgridder_dirty2vis_f64
    _check(fn(opts, uvw, freq, dirty, wgt, mask, vis.view(real), info))
    return (vis, info_dict(info)) if return_info else vis


def plan(*, uvw, freq, npix_x, npix_y, pixsize_x, pixsize_y, epsilon, vis=None, wgt=None,
         mask=None, single=False, **kw):
    """The plan (info dict) a call on these arrays would choose, without
    gridding (core/src/wgridder.rs:3525-3549); `vis` None plans the PSF or
    degridding pass."""
    import ctypes

    uvw, freq, shape = _coords(uvw, freq)
    wgt, mask = _weights(wgt, mask, shape)
    vis_kind = _core.VIS_UNIT
    if vis is not None:
        vis = _visibilities(vis, shape)
        vis_kind = _core.VIS_C64 if vis.dtype == np.complex64 else _core.VIS_C128
    wgt_kind = _core.WGT_NONE if wgt is None else (
        _core.WGT_F32 if wgt.dtype == np.float32 else _core.WGT_F64)
    o = _core.options(_options(npix_x, npix_y, pixsize_x, pixsize_y, epsilon, kw))
    info = np.zeros(N_INFO)
    lib = _core.library()
    fn = lib.rang_amd_wgridder_plan_f32 if single else lib.rang_amd_wgridder_plan_f64
    p = lambda a: None if a is None else ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    _check(fn(ctypes.byref(o), shape[0], shape[1], p(uvw), p(freq), p(vis), vis_kind, p(wgt),
              wgt_kind, p(mask), p(info)))
    return info_dict(info)


def gridding_footprint(*, uvw, freq, wgt, npix_x, npix_y, pixsize_x, pixsize_y, epsilon, nout=1,
                       single=False, **kw):
    """(plan, bytes): the plan of the gridding pass on float64 weights and the
    device bytes that pass holds (core/src/wgridder.rs:3551-3586)."""
    import ctypes

    uvw, freq, shape = _coords(uvw, freq)
    wgt = np.ascontiguousarray(wgt, np.float64)
    if wgt.shape != shape:
        raise ValueError("array lengths disagree")
    o = _core.options(_options(npix_x, npix_y, pixsize_x, pixsize_y, epsilon, kw))
    info = np.zeros(N_INFO)
    nbytes = ctypes.c_size_t(0)
    lib = _core.library()
    fn = lib.rang_amd_wgridder_footprint_f32 if single else lib.rang_amd_wgridder_footprint_f64
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    _check(fn(ctypes.byref(o), shape[0], shape[1], p(uvw), p(freq), p(wgt), int(nout), p(info),
              ctypes.byref(nbytes)))
    return info_dict(info), int(nbytes.value)


def info_dict(info):
    rec = {k: float(v) for k, v in zip(INFO_NAMES, info)}
    for k in ("kernel_index", "supp", "nu", "nv", "nplanes", "nvis"):
        rec[k] = int(rec[k])
    return rec


def kernel_row(index):
    """(W, ofactor, epsilon, beta, e0, ndim, singleprec) of table row
    `index`, None past the end."""
    out = np.empty(7, np.float64)
    if _core.wgridder_kernel_row(int(index), out):
        return None
    return (int(out[0]), float(out[1]), float(out[2]), float(out[3]), float(out[4]),
            int(out[5]), bool(out[6]))


def kernel_table():
    rows = []
    while (row := kernel_row(len(rows))) is not None:
        rows.append(row)
    return rows


def fits_image(uvw, freq, data, weights, width, height, scale_arcsec, epsilon=1e-6, single=False,
               imaging_weights=None, **kw):
    """Dirty image and PSF as (height, width) arrays in WSClean's FITS pixel
    order, scaled so that the PSF centre is one: the image of the data with v
    and w flipped, transposed (python/rang/core_wgridder.py:335-382)."""
    dirty, psf, info = dirty_psf(
        uvw=uvw, freq=freq, vis=data, wgt=weights, wgt2=imaging_weights, **kw,
        npix_x=width, npix_y=height, pixsize_x=scale_arcsec * ARCSEC,
        pixsize_y=scale_arcsec * ARCSEC, epsilon=epsilon, single=single, return_info=True,
        flip_v=True, flip_w=True,
    )
    norm = psf[width // 2, height // 2]
    return {
        "dirty": (dirty.T / norm).astype(np.float64),
        "psf": (psf.T / norm).astype(np.float64),
        "info": info,
        "info_psf": info,
        "psf_centre_raw": float(norm),
    }
