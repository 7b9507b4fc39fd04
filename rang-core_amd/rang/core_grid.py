"""The w-stacking gridder (WSClean 3.6's inversion, rang-core's
core/src/grid.rs) on the MI355X: the `invert` and `kernel` of rang-core's
python/rang/core_grid.py:116-405 for gridder "wstack", over
`rang._core.grid_invert` / `grid_invert_gpu`.

Both device choices run on the GPU: "cpu" (the reference's CPU path,
rang_core_grid_invert) grids and transforms in double precision, "gpu" in
the requested precision ("f64" by default, or "f32") and returns the timing
breakdown under "gpu".
"""

import math

import numpy as np

from . import _core

ARCSEC = math.pi / 648000.0
STAT_NAMES = ("image_weight", "gridded_count", "visibility_weight_sum", "max_gridded_weight",
              "min_w", "max_w", "psf_normalization", "n_w_layers")
GPU_TIMING_NAMES = ("context_s", "compile_s", "prepare_s", "bucket_s", "upload_s", "compute_s",
                    "grid_s", "fft_s", "project_s", "download_s", "finish_s", "gpu_bytes", "batches",
                    "layers_per_batch")


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
        raise RuntimeError(f"core kernel failed with code {code}")
    return out


def _record(stats):
    rec = dict(zip(STAT_NAMES, (float(v) for v in stats)))
    rec["gridded_count"] = int(rec["gridded_count"])
    rec["n_w_layers"] = int(rec["n_w_layers"])
    return rec


def invert(uvw, frequencies, data, weights, imaging_weights, width=512, height=512,
           scale_arcsec=8.0, w_layers=64, kernel_size=7, oversampling=1023, padding=1.2,
           device="gpu", precision="f64"):
    """Dirty image and PSF, (height, width) arrays in FITS pixel order, with
    WSClean's counters (python/rang/core_grid.py:188-313)."""
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
        timings = np.zeros(_core.GRID_N_TIMINGS)
        code = _core.grid_invert_gpu(*args, precision, dirty, psf, stats, timings)
        if code:
            raise RuntimeError(f"core GPU gridder failed with code {code}: {_core.grid_last_error()}")
        gpu = dict(zip(GPU_TIMING_NAMES, (float(v) for v in timings)))
        gpu["precision"] = precision
        for key in ("gpu_bytes", "batches", "layers_per_batch"):
            gpu[key] = int(gpu[key])
        return {"dirty": dirty, "psf": psf, "stats": _record(stats), "gpu": gpu}
    if device != "cpu":
        raise ValueError(f"device must be 'cpu' or 'gpu', not {device!r}")
    code = _core.grid_invert(*args, dirty, psf, stats)
    if code:
        raise RuntimeError(f"core gridder failed with code {code}: {_core.grid_last_error()}")
    return {"dirty": dirty, "psf": psf, "stats": _record(stats), "gridder": "wstack"}
