"""Sub-band batches: many independent images of uneven size and visibility
count imaged in one call, with every image's gridding in one persistent
launch, and the tuner that picks each image's kernel row and gridding tile
(include/rang_core_batch.h). No counterpart in the reference, which images
one call at a time (core/src/wgridder.rs:2430-2696).

    images = [batch.Image(opts, uvw, freq, vis, wgt) for ...]
    for im in images:
        batch.tune(im, cache_path="tuning.txt")
    dirty = batch.vis2dirty(images, single=True)

Arrays are torch tensors on the current GPU; `freq` may be a tensor or a
numpy array (the plan is made on the host from it).
"""

import ctypes

import numpy as np
import torch

from . import _core
from .device import _VIS_KIND, _WGT_KIND, _check, _ptr, _stream


class Image:
    """One image of a batch: the 18-float option list of `rang._core`, device
    uvw (nrow x 3), frequencies, device visibilities (nrow x nchan complex),
    optional device weights and mask. `tile` (0: the default) and the
    option list's kernel_index, nu and nv are what the tuner sets."""

    def __init__(self, opts, uvw, freq, vis, wgt=None, mask=None, tile=0):
        self.opts = list(opts)
        self.uvw, self.vis, self.wgt, self.mask = uvw, vis, wgt, mask
        f = freq.detach().cpu().numpy() if isinstance(freq, torch.Tensor) else np.asarray(freq)
        self.freq = np.ascontiguousarray(f, dtype=np.float64)
        self.tile = int(tile)
        self.info = np.zeros(_core.N_INFO)

    def _struct(self, dirty):
        b = _core.BatchImage()
        b.opts = _core.options(self.opts)
        b.nrow, b.nchan = self.uvw.shape[0], self.freq.shape[0]
        if self.vis.shape != (b.nrow, b.nchan):
            raise ValueError("visibilities must be nrow x nchan")
        b.uvw = _ptr(self.uvw)
        b.freq = ctypes.c_void_p(self.freq.ctypes.data)
        b.vis = _ptr(self.vis)
        b.wgt = _ptr(self.wgt)
        b.mask = _ptr(self.mask)
        b.dirty = _ptr(dirty)
        b.vis_kind = _VIS_KIND[self.vis.dtype]
        b.wgt_kind = _core.WGT_NONE if self.wgt is None else _WGT_KIND[self.wgt.dtype]
        b.tile = self.tile
        return b

    def _take(self, b):
        o = b.opts
        # the option list's kernel_index, nu, nv (positions 15-17)
        self.opts[15], self.opts[16], self.opts[17] = int(o.kernel_index), int(o.nu), int(o.nv)
        self.tile = int(b.tile)
        self.info = np.array(b.info[:], dtype=np.float64)


def vis2dirty(images, *, single, budget=None, out=None):
    """Dirty images (torch tensors, nx x ny) of every image in `images`, the
    gridding of all of them in one launch per group of images whose buffers
    fit `budget` bytes (default: half the device's free memory). Each
    image's plan is written to its `info`."""
    real = torch.float32 if single else torch.float64
    if out is None:
        out = [torch.empty((int(im.opts[0]), int(im.opts[1])), dtype=real, device=im.uvw.device) for im in images]
    if budget is None:
        free, _ = torch.cuda.mem_get_info()
        budget = free // 2
    arr = (_core.BatchImage * max(len(images), 1))()
    for k, im in enumerate(images):
        arr[k] = im._struct(out[k])
    lib = _core.library()
    fn = lib.rang_amd_wgridder_batch_vis2dirty_f32 if single else lib.rang_amd_wgridder_batch_vis2dirty_f64
    _check(fn(arr, len(images), int(budget), _stream()))
    for k, im in enumerate(images):
        im.info = np.array(arr[k].info[:], dtype=np.float64)
    return out


def tune(image, *, single, cache_path=None, kernels=3, out=None):
    """Times the image's vis2dirty over the cost model's `kernels` cheapest
    kernel rows and over gridding tile sides, and sets the fastest into the
    image. Returns a dict: ms (chosen), default_ms (the cost model's choice;
    0 when cached), cached."""
    real = torch.float32 if single else torch.float64
    d = out if out is not None else torch.empty((int(image.opts[0]), int(image.opts[1])), dtype=real,
                                                device=image.uvw.device)
    b = image._struct(d)
    rep = _core.TuneReport()
    lib = _core.library()
    fn = lib.rang_amd_wgridder_tune_f32 if single else lib.rang_amd_wgridder_tune_f64
    path = None if cache_path is None else str(cache_path).encode()
    _check(fn(ctypes.byref(b), path, int(kernels), ctypes.byref(rep), _stream()))
    image._take(b)
    return {"ms": rep.ms, "default_ms": rep.default_ms, "cached": bool(rep.cached)}
