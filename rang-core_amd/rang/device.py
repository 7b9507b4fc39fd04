"""Device-resident calls of the gridder on torch tensors in HBM.

The rang_amd_wgridder_*_dev entry points of include/rang_core_wgridder.h:
the same transforms as `rang._core` with every array already on the GPU, run
on torch's current stream. No counterpart in the reference; this is what
solvers and benchmarks use to keep visibilities resident between calls.
"""

import ctypes

import numpy as np
import torch

from . import _core

_VIS_KIND = {torch.complex64: _core.VIS_C64, torch.complex128: _core.VIS_C128}
_WGT_KIND = {torch.float32: _core.WGT_F32, torch.float64: _core.WGT_F64}


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("device arrays must be contiguous GPU tensors")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(status):
    if status == 0:
        return
    msg = _core.wgridder_last_error()
    if status == 1:
        raise ValueError(msg)
    raise RuntimeError(f"core gridder failed with status {status}: {msg}")


def vis2dirty(opts, uvw, freq, vis=None, wgt=None, wgt2=None, mask=None, *, single,
              dirty=True, psf=False, out=None, noalloc=False):
    """Dirty image and/or PSF (torch tensors) of device visibilities. `opts`
    is the 18-float option list of `rang._core`; returns (dirty, psf, info).
    noalloc (single precision): the entry point that allocates no device
    memory and works in the library's fixed scratch block instead."""
    if noalloc and not single:
        raise ValueError("the allocation-free entry point is single precision")
    o = _core.options(opts)
    nrow, nchan = uvw.shape[0], freq.shape[0]
    vis_kind = _core.VIS_UNIT if vis is None else _VIS_KIND[vis.dtype]
    wgt_kind = _core.WGT_NONE if wgt is None else _WGT_KIND[wgt.dtype]
    if wgt2 is not None:
        wgt_kind = _core.WGT_F32_PRODUCT
    real = torch.float32 if single else torch.float64
    dev = uvw.device
    d = p = None
    if dirty:
        d = out[0] if out is not None else torch.empty((o.nx, o.ny), dtype=real, device=dev)
    if psf:
        p = out[1] if out is not None else torch.empty((o.nx, o.ny), dtype=real, device=dev)
    info = np.zeros(_core.N_INFO)
    lib = _core.library()
    fn = lib.rang_amd_wgridder_vis2dirty_dev_f32 if single else lib.rang_amd_wgridder_vis2dirty_dev_f64
    if noalloc:
        fn = lib.rang_amd_wgridder_vis2dirty_noalloc_f32
    st = fn(ctypes.byref(o), nrow, nchan, _ptr(uvw), _ptr(freq), _ptr(vis), vis_kind, _ptr(wgt),
            _ptr(wgt2), wgt_kind, _ptr(mask), _ptr(d), _ptr(p),
            ctypes.c_void_p(info.ctypes.data), _stream())
    _check(st)
    return d, p, info


def dirty2vis(opts, uvw, freq, dirty, wgt=None, mask=None, out=None):
    """Visibilities (complex torch tensor) of a device image; the precision
    follows the image. Returns (vis, info)."""
    o = _core.options(opts)
    nrow, nchan = uvw.shape[0], freq.shape[0]
    single = dirty.dtype == torch.float32
    wgt_kind = _core.WGT_NONE if wgt is None else _WGT_KIND[wgt.dtype]
    cdt = torch.complex64 if single else torch.complex128
    vis = out if out is not None else torch.empty((nrow, nchan), dtype=cdt, device=uvw.device)
    info = np.zeros(_core.N_INFO)
    lib = _core.library()
    fn = lib.rang_amd_wgridder_dirty2vis_dev_f32 if single else lib.rang_amd_wgridder_dirty2vis_dev_f64
    st = fn(ctypes.byref(o), nrow, nchan, _ptr(uvw), _ptr(freq), _ptr(dirty), _ptr(wgt), wgt_kind,
            _ptr(mask), _ptr(vis), ctypes.c_void_p(info.ctypes.data), _stream())
    _check(st)
    return vis, info


def stage_timing(on=True):
    _core.library().rang_amd_stage_timing(int(bool(on)))


def stage_times():
    """[(stage, ms)] of this thread's last call with stage timing on."""
    lib = _core.library()
    return [(lib.rang_amd_stage_name(i).decode(), lib.rang_amd_stage_ms(i))
            for i in range(lib.rang_amd_stage_count())]


def release():
    """Frees the device memory the library keeps between calls (its
    workspaces and sub-band batch pools; not live prepared plans)."""
    _check(_core.library().rang_amd_release_workspaces())
