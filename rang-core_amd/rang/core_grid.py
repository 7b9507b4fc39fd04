"""The w-stacking gridder (WSClean 3.6's inversion, rang-core's
core/src/grid.rs) on the MI355X, with the surface of rang-core's
python/rang/core_grid.py: `stokes_i_reader` (what WSClean's reader hands the
gridder), `invert` (gridder "wstack", or "wgrid" for the exact gridder),
`kernel`, `wcs_header` and `write_fits`, the device and precision resolved
from the same environment variables.

Both device choices run on the GPU: "cpu" (the reference's CPU path,
rang_core_grid_invert) grids and transforms in double precision for any
padded side; "gpu" (the reference's GPU path, rang_core_grid_invert_gpu)
runs in the requested precision ("f64" by default, or "f32"), refuses
padded sides with a prime factor above 11 as the reference's does, and
returns the timing breakdown under "gpu". Without the built library every
call raises; nothing falls back to a CPU implementation.
"""

import math
import os

import numpy as np

from . import _core

ARCSEC = math.pi / 648000.0
GRIDDERS = ("wstack", "wgrid")
# the exact gridder's accuracy in the reference's drivers (python/rang/core_grid.py:36)
WGRID_EPSILON = 1e-6
N_STATS = 8
STAT_NAMES = ("image_weight", "gridded_count", "visibility_weight_sum", "max_gridded_weight",
              "min_w", "max_w", "psf_normalization", "n_w_layers")
DEVICE_ENV = "RANG_GRIDDER_DEVICE"
PRECISION_ENV = "RANG_GRIDDER_GPU_PRECISION"
DEVICES = ("cpu", "gpu")
PRECISIONS = ("f64", "f32")
N_GPU_TIMINGS = 14
# the order of Timings::as_array (core/src/gpu/invert.rs:57-74)
GPU_TIMING_NAMES = ("context_s", "compile_s", "prepare_s", "upload_s", "compute_s", "download_s",
                    "finish_s", "gpu_bytes", "batches", "layers_per_batch", "grid_s", "fft_s",
                    "project_s", "bucket_s")


def resolve_device(value=None):
    """The gridder device: `value`, else RANG_GRIDDER_DEVICE, else "cpu"
    (python/rang/core_grid.py:70-77)."""
    value = value if value is not None else os.environ.get(DEVICE_ENV, "cpu")
    value = value.strip().lower() or "cpu"
    if value not in DEVICES:
        raise ValueError(f"{DEVICE_ENV} must be one of {DEVICES}, not {value!r}")
    return value


def resolve_precision(value=None):
    """The GPU precision: `value`, else RANG_GRIDDER_GPU_PRECISION, else
    "f64" (python/rang/core_grid.py:80-87)."""
    value = value if value is not None else os.environ.get(PRECISION_ENV, "f64")
    value = value.strip().lower() or "f64"
    if value not in PRECISIONS:
        raise ValueError(f"{PRECISION_ENV} must be one of {PRECISIONS}, not {value!r}")
    return value


def gpu_compiled():
    """Whether the library carries the GPU gridder (always: it is a GPU
    library)."""
    return hasattr(_core.library(), "rang_core_grid_invert_gpu")


def gpu_adapter():
    """Description of the device the GPU gridder uses; RuntimeError when no
    GPU is usable (python/rang/core_grid.py:95-106)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("no usable GPU for the gridder (HIP reports no device)")
    props = torch.cuda.get_device_properties(0)
    arch = getattr(props, "gcnArchName", "")
    return f"{props.name} ({arch}, {props.total_memory >> 30} GiB)" if arch else props.name


def gpu_available():
    """(True, adapter) or (False, reason), without raising."""
    try:
        return True, gpu_adapter()
    except RuntimeError as error:
        return False, str(error)


def padded_size(n, padding=1.2):
    """WSClean's padded side (grid.rs:80-89): ceil(n * padding) rounded up to
    a multiple of four."""
    p = int(math.ceil(n * padding))
    return p + (4 - p % 4) % 4


def kernel(kernel_size=7, oversampling=1023):
    """The oversampled Kaiser-Bessel kernel (grid.rs:165-188)."""
    out = np.zeros(kernel_size * oversampling)
    code = _core.grid_kernel(kernel_size, oversampling, out)
    if code:
        raise ValueError("invalid kernel settings")
    return out


def kernel_reference(kernel_size=7, oversampling=1023, alpha=8.6):
    """The same kernel from its formula in NumPy (python/rang/core_grid.py:
    126-150): the windowed sinc sin(pi x / oversampling) / (pi x) times
    I0(alpha sqrt(1 - (x / mid)^2)) oversampling / I0(alpha), I0 by its power
    series to a relative term of 1e-8, mirrored about the centre."""

    def i0(x, precision=1e-8):
        term, total, d = 1.0, 1.0, 0.0
        while True:
            d += 2.0
            term *= x * x / (d * d)
            total += term
            if term <= total * precision:
                return total

    n = kernel_size * oversampling
    mid = n // 2
    ratio = 1.0 / oversampling
    x = np.arange(mid + 1, dtype=float)
    sinc = np.empty(mid + 1)
    sinc[0] = ratio
    sinc[1:] = np.sin(math.pi * ratio * x[1:]) / (math.pi * x[1:])
    scale = oversampling / i0(alpha)
    out = np.empty(n)
    for k in range(mid + 1):
        out[mid + k] = sinc[k] * i0(alpha * math.sqrt(1.0 - (k / mid) ** 2)) * scale
    out[:mid] = out[n - 1:n - 1 - mid:-1]
    return out


def stokes_i_reader(data, weight_spectrum, flags, flag_rows=None):
    """What WSClean's reader hands the gridder for `-pol I` on XX, XY, YX, YY
    (python/rang/core_grid.py:153-185): `data`, `weight_spectrum` and `flags`
    are (nrow, nchan, 4). Returns the sample (XX + YY) / 2 as complex64 (zero
    where not finite) and the weight 4 min(w_XX, w_YY) as float32, zero where
    either correlation is flagged or not finite, and on rows with
    `flag_rows` set. (rang.core_weights.stokes_i forms the same on the GPU.)"""
    data = np.asarray(data, np.complex64)
    weight_spectrum = np.asarray(weight_spectrum, np.float32)
    flags = np.asarray(flags, bool)
    if data.shape != weight_spectrum.shape or data.shape != flags.shape or data.shape[-1] != 4:
        raise ValueError("four-correlation data, weights and flags of one shape required")
    xx, yy = data[..., 0], data[..., 3]
    sample = (xx + yy) * np.float32(0.5)
    sample = np.where(np.isfinite(sample), sample, np.complex64(0)).astype(np.complex64)
    ok = ~flags[..., 0] & np.isfinite(xx) & ~flags[..., 3] & np.isfinite(yy)
    w = np.minimum(weight_spectrum[..., 0] * np.float32(4.0), weight_spectrum[..., 3] * np.float32(4.0))
    weight = np.where(ok, w, np.float32(0)).astype(np.float32)
    if flag_rows is not None:
        weight[np.asarray(flag_rows, bool)] = 0.0
    return sample, weight


def _record(stats):
    rec = dict(zip(STAT_NAMES, (float(v) for v in stats)))
    rec["gridded_count"] = int(rec["gridded_count"])
    rec["n_w_layers"] = int(rec["n_w_layers"])
    return rec


def invert(uvw, frequencies, data, weights, imaging_weights, width=512, height=512,
           scale_arcsec=8.0, w_layers=64, kernel_size=7, oversampling=1023, padding=1.2,
           gridder="wstack", epsilon=WGRID_EPSILON, device=None, precision=None):
    """Dirty image and PSF, (height, width) arrays in FITS pixel order, with
    WSClean's counters (python/rang/core_grid.py:188-313). `gridder`
    "wstack" (the default) or "wgrid" (the exact gridder at `epsilon`, one
    pass for dirty and PSF in double precision; w_layers, kernel_size,
    oversampling and padding do not apply to it, and as in the reference it
    refuses device "gpu"). `device` and `precision` default to the
    environment (resolve_device, resolve_precision)."""
    if gridder == "wgrid":
        if resolve_device(device) != "cpu":
            raise RuntimeError(
                "the wgrid gridder runs on the CPU only; the GPU path is the "
                "wstack gridder's (unset RANG_GRIDDER_DEVICE or pass device='cpu')"
            )
        return _invert_wgrid(uvw, frequencies, data, weights, imaging_weights, width, height, scale_arcsec,
                             epsilon)
    if gridder != "wstack":
        raise ValueError(f"gridder must be one of {GRIDDERS}, not {gridder!r}")
    device = resolve_device(device)
    uvw = np.ascontiguousarray(uvw, np.float64)
    frequencies = np.ascontiguousarray(frequencies, np.float64)
    data = np.ascontiguousarray(data, np.complex64)
    weights = np.ascontiguousarray(weights, np.float32)
    imaging_weights = np.ascontiguousarray(imaging_weights, np.float32)
    nrow, nchan = data.shape
    if uvw.shape != (nrow, 3) or frequencies.shape != (nchan,):
        raise ValueError("uvw must be (nrow, 3) and frequencies (nchan,)")
    if weights.shape != (nrow, nchan) or imaging_weights.shape != (nrow, nchan):
        raise ValueError("weights must match data")
    dirty = np.empty((height, width), np.float64)
    psf = np.empty((height, width), np.float64)
    stats = np.empty(_core.GRID_N_STATS, np.float64)
    args = (nrow, nchan, width, height, kernel_size, oversampling, w_layers, scale_arcsec * ARCSEC,
            scale_arcsec * ARCSEC, float(padding), uvw, frequencies, data.view(np.float32), weights,
            imaging_weights)
    if device == "gpu":
        precision = resolve_precision(precision)
        gpu_adapter()
        timings = np.zeros(_core.GRID_N_TIMINGS)
        code = _core.grid_invert_gpu(*args, precision, dirty, psf, stats, timings)
        gpu = dict(zip(GPU_TIMING_NAMES, (float(v) for v in timings)))
        gpu["precision"] = precision
        for key in ("gpu_bytes", "batches", "layers_per_batch"):
            gpu[key] = int(gpu[key])
        if code:
            raise RuntimeError(f"core GPU gridder failed with code {code}: {_core.grid_last_error()}")
        return {"dirty": dirty, "psf": psf, "stats": _record(stats), "gpu": gpu}
    code = _core.grid_invert(*args, dirty, psf, stats)
    if code:
        raise RuntimeError(f"core gridder failed with code {code}: {_core.grid_last_error()}")
    return {"dirty": dirty, "psf": psf, "stats": _record(stats), "gridder": "wstack"}


def _weight_counters(weights, imaging_weights, block=4096):
    """WSClean's counters of the gridded samples (w iw != 0, the product in
    float32 as the core forms it), in row blocks: image weight, count,
    visibility weight sum and the largest gridded weight
    (python/rang/core_grid.py:316-334)."""
    total = vsum = 0.0
    count = 0
    largest = 0.0
    for start in range(0, len(weights), block):
        w = weights[start:start + block]
        cum = w * imaging_weights[start:start + block]
        live = cum != 0
        total += float(cum.sum(dtype=np.float64))
        vsum += float(w[live].sum(dtype=np.float64))
        count += int(np.count_nonzero(live))
        if cum.size:
            largest = max(largest, float(cum.max()))
    return total, count, vsum, largest


def _invert_wgrid(uvw, frequencies, data, weights, imaging_weights, width, height, scale_arcsec, epsilon):
    """invert through the exact gridder (python/rang/core_grid.py:337-405)."""
    from . import core_wgridder

    data = np.asarray(data)
    if data.dtype not in (np.complex64, np.complex128):
        data = data.astype(np.complex64)
    data = np.ascontiguousarray(data)
    weights = np.ascontiguousarray(weights, np.float32)
    imaging_weights = np.ascontiguousarray(imaging_weights, np.float32)
    nrow, nchan = data.shape
    if np.shape(uvw) != (nrow, 3) or np.shape(frequencies) != (nchan,):
        raise ValueError("uvw must be (nrow, 3) and frequencies (nchan,)")
    if weights.shape != (nrow, nchan) or imaging_weights.shape != (nrow, nchan):
        raise ValueError("weights must match data")
    image_weight, count, vsum, largest = _weight_counters(weights, imaging_weights)
    if image_weight == 0.0:
        raise RuntimeError("core gridder failed: no visibility carries weight")
    out = core_wgridder.fits_image(np.ascontiguousarray(uvw, np.float64),
                                   np.ascontiguousarray(frequencies, np.float64), data, weights, width,
                                   height, scale_arcsec, epsilon=epsilon, imaging_weights=imaging_weights)
    info = out["info"]
    centre = out["psf_centre_raw"]
    stats = {
        "image_weight": image_weight,
        "gridded_count": count,
        "visibility_weight_sum": vsum,
        "max_gridded_weight": largest,
        "min_w": info["wmin_d"],
        "max_w": info["wmax_d"],
        "psf_normalization": image_weight / centre if centre else 0.0,
        "n_w_layers": info["nplanes"],
    }
    return {
        "dirty": out["dirty"],
        "psf": out["psf"],
        "stats": stats,
        "gridder": "wgrid",
        "wgridder": {"epsilon": float(epsilon), "backend": core_wgridder.backend(),
                     "psf_centre_raw": float(centre), **info},
    }


def wcs_header(phase_dir_rad, width, height, scale_arcsec, frequencies):
    """FITS keywords in WSClean's convention for a SIN image at the phase
    centre; the frequency axis spans the channel edges
    (python/rang/core_grid.py:408-446)."""
    ra = math.degrees(phase_dir_rad[0])
    if ra > 180.0:
        ra -= 360.0
    dec = math.degrees(phase_dir_rad[1])
    frequencies = np.asarray(frequencies, float)
    step = frequencies[1] - frequencies[0] if len(frequencies) > 1 else 0.0
    low, high = frequencies[0] - step / 2, frequencies[-1] + step / 2
    scale = scale_arcsec / 3600.0
    return {
        "BUNIT": "JY/BEAM", "BTYPE": "Intensity", "EQUINOX": 2000.0, "LONPOLE": 180.0,
        "CTYPE1": "RA---SIN", "CRPIX1": width / 2 + 1, "CRVAL1": ra, "CDELT1": -scale, "CUNIT1": "deg",
        "CTYPE2": "DEC--SIN", "CRPIX2": height / 2 + 1, "CRVAL2": dec, "CDELT2": scale, "CUNIT2": "deg",
        "CTYPE3": "FREQ", "CRPIX3": 1.0, "CRVAL3": (low + high) / 2, "CDELT3": high - low, "CUNIT3": "Hz",
        "CTYPE4": "STOKES", "CRPIX4": 1.0, "CRVAL4": 1.0, "CDELT4": 1.0, "SPECSYS": "TOPOCENT",
    }


def write_fits(path, image, header, extra=None):
    """Write a (height, width) image as a 4-axis float32 FITS file in
    WSClean's pixel order (python/rang/core_grid.py:449-461); needs astropy,
    as the reference's does."""
    from astropy.io import fits

    image = np.asarray(image, np.float32)
    hdu = fits.PrimaryHDU(image[np.newaxis, np.newaxis, :, :])
    for key, value in header.items():
        hdu.header[key] = value
    hdu.header["ORIGIN"] = "Rang core gridder"
    for key, value in (extra or {}).items():
        hdu.header[key] = value
    hdu.writeto(str(path), overwrite=False)
