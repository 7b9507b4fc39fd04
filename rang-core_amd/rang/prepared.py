"""Prepared plans on the GPU: the plan and sorted layouts of a fixed set of
visibilities kept between calls, as rang-core's `wgridder::Prepared`
(core/src/wgridder.rs:3165-3499), on torch tensors in HBM.

    p = Prepared(opts, uvw, freq, wgt, single=True)
    dirty = p.vis2dirty(vis)            # image of vis with the prepared weights
    vis = p.dirty2vis(model)            # A model (weighted=True: times the weights)
    dirty, q = p.residual(vis, model)   # image of vis - A model, sum w |vis - A model|^2
    out = p.gram(x)                     # A^H W A x
    x, used, rel = p.conjugate_gradients(rhs, iterations, tolerance)

The active set is the nonzero weights. All arrays are contiguous tensors on
the current GPU; calls run on torch's current stream.
"""

import ctypes

import numpy as np
import torch

from . import _core
from .device import _VIS_KIND, _WGT_KIND, _check, _ptr, _stream


class Prepared:
    def __init__(self, opts, uvw, freq, wgt=None, *, single):
        o = _core.options(opts)
        self.nx, self.ny = o.nx, o.ny
        self.nrow, self.nchan = int(uvw.shape[0]), int(freq.shape[0])
        self.single = bool(single)
        self.real = torch.float32 if single else torch.float64
        self.complex = torch.complex64 if single else torch.complex128
        self.device = uvw.device
        lib = _core.library()
        fn = lib.rang_amd_wgridder_prepare_dev_f32 if single else lib.rang_amd_wgridder_prepare_dev_f64
        kind = _core.WGT_NONE if wgt is None else _WGT_KIND[wgt.dtype]
        if wgt is not None and wgt.numel() != self.nrow * self.nchan:
            raise ValueError("array lengths disagree")
        h = fn(ctypes.byref(o), self.nrow, self.nchan, _ptr(uvw), _ptr(freq), _ptr(wgt), kind, _stream())
        if not h:
            raise ValueError(_core.wgridder_last_error())
        self._h = ctypes.c_void_p(h)
        self._suf = "f32" if single else "f64"

    def _fn(self, name):
        return getattr(_core.library(), f"rang_amd_wgridder_prepared_{name}_{self._suf}")

    def close(self):
        if getattr(self, "_h", None):
            _core.library().rang_amd_wgridder_prepared_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _image(self):
        return torch.empty((self.nx, self.ny), dtype=self.real, device=self.device)

    def _vis(self, vis):
        if vis.numel() != self.nrow * self.nchan:
            raise ValueError("array lengths disagree")
        return _VIS_KIND[vis.dtype]

    def active(self):
        """Active visibilities (nonzero weight)."""
        return int(_core.library().rang_amd_wgridder_prepared_active(self._h))

    def memory(self):
        """Device bytes the plan holds."""
        return int(_core.library().rang_amd_wgridder_prepared_memory(self._h))

    def release_grids(self):
        """Free the work grids; the next call allocates them again."""
        _core.library().rang_amd_wgridder_prepared_release_grids(self._h)

    def vis2dirty(self, vis, out=None):
        kind = self._vis(vis)
        d = out if out is not None else self._image()
        info = np.zeros(_core.N_INFO)
        _check(self._fn("vis2dirty")(self._h, _ptr(vis), kind, _ptr(d),
                                      ctypes.c_void_p(info.ctypes.data), _stream()))
        return d

    def dirty2vis(self, dirty, weighted=False, out=None):
        v = out if out is not None else torch.empty((self.nrow, self.nchan), dtype=self.complex,
                                                    device=self.device)
        _check(self._fn("dirty2vis")(self._h, _ptr(dirty), int(bool(weighted)), _ptr(v), _stream()))
        return v

    def residual(self, vis, model=None, out=None):
        """(image of vis - A model, sum w |vis - A model|^2)."""
        kind = self._vis(vis)
        d = out if out is not None else self._image()
        q = ctypes.c_double(0.0)
        _check(self._fn("residual")(self._h, _ptr(vis), kind, _ptr(model), _ptr(d), ctypes.byref(q),
                                     _stream()))
        return d, q.value

    def gram(self, x, out=None):
        """A^H W A x."""
        d = out if out is not None else self._image()
        _check(self._fn("gram")(self._h, _ptr(x), _ptr(d), _stream()))
        return d

    def fista(self, dirty, q0, lam, x, *, factors=None, signed=False, settings, gx=None, nrecords=0):
        """prox_solver::solve (core/src/prox_solver.rs:283-445) with the L1 prior
        on this plan's normal operator. x (the start) is overwritten with the
        result; returns (outcome, records[:min(iterations + 1, nrecords)])."""
        recs = (_core.FistaRecord * max(int(nrecords), 1))()
        out = _core.FistaOutcome()
        _check(self._fn("fista")(self._h, _ptr(dirty), float(q0), float(lam), _ptr(factors),
                                  int(bool(signed)), ctypes.byref(settings), _ptr(x), _ptr(gx), recs,
                                  int(nrecords), ctypes.byref(out), _stream()))
        n = min(int(out.iterations) + 1, int(nrecords))
        return out, [recs[k] for k in range(n)]

    def power_iteration(self, seed, tolerance, max_iterations):
        """prox_solver::power_iteration (core/src/prox_solver.rs:886-908):
        (largest eigenvalue of A^H W A, applications used)."""
        value = ctypes.c_double(0.0)
        used = ctypes.c_uint64(0)
        _check(self._fn("power")(self._h, int(seed), float(tolerance), int(max_iterations),
                                  ctypes.byref(value), ctypes.byref(used), _stream()))
        return value.value, int(used.value)

    def conjugate_gradients(self, rhs, iterations, tolerance, out=None):
        """CG on A^H W A x = rhs from zero (core/src/clean.rs:263-306):
        (x, Gram applications used, |rhs - A^H W A x| / |rhs|)."""
        x = out if out is not None else self._image()
        used = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        _check(self._fn("cg")(self._h, _ptr(rhs), _ptr(x), int(iterations), float(tolerance),
                               ctypes.byref(used), ctypes.byref(rel), _stream()))
        return x, used.value, rel.value

    def preconditioned_cg(self, precond, rhs, iterations, tolerance, out=None):
        """CG on A^H W A x = rhs from zero preconditioned by a symmetric
        positive definite `precond` (a rang.gram.GramOperator, e.g. an inverse
        filter), stopping on the same |r| / |rhs| as conjugate_gradients:
        (x, Gram applications used, relative residual)."""
        x = out if out is not None else self._image()
        used = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        _check(self._fn("pcg")(self._h, precond._h, _ptr(rhs), _ptr(x), int(iterations), float(tolerance),
                                ctypes.byref(used), ctypes.byref(rel), _stream()))
        return x, used.value, rel.value

    def defect(self, cheap, dirty, q0, x, *, factors=None, signed=False, inner, schedule, gx=None, nrecords=0):
        """prox_solver::defect_correction (core/src/prox_solver.rs:778-875) with
        the L1 prior: major cycles on this plan's normal operator, each inner
        FISTA solve on `cheap` (a rang.gram.GramOperator of the same side and
        precision). x (the start) is overwritten with the result; returns
        (outcome, records[:min(cycles + 1, nrecords)])."""
        recs = (_core.OuterRecord * max(int(nrecords), 1))()
        out = _core.DefectOutcome()
        _check(self._fn("defect")(self._h, cheap._h, _ptr(dirty), float(q0), _ptr(factors), int(bool(signed)),
                                   ctypes.byref(inner), ctypes.byref(schedule), _ptr(x), _ptr(gx), recs,
                                   int(nrecords), ctypes.byref(out), _stream()))
        n = min(int(out.cycles) + 1, int(nrecords))
        return out, [recs[k] for k in range(n)]

    def sorted(self, dense):
        """Prepared::sorted (core/src/wgridder.rs:3269-3286): `dense` (nrow x
        nchan, 4-, 8- or 16-byte elements) at the active visibilities, in the
        order residual_sorted reads its data."""
        if dense.numel() != self.nrow * self.nchan or not dense.is_contiguous():
            raise ValueError("array lengths disagree")
        out = torch.empty(self.active(), dtype=dense.dtype, device=dense.device)
        _check(self._fn("sorted")(self._h, _ptr(dense), dense.element_size(), _ptr(out), _stream()))
        return out

    def residual_sorted(self, vis, model=None, out=None):
        """Prepared::residual_sorted (:3257-3267): residual() of data held only
        for the active visibilities, in the order of sorted()."""
        if vis.numel() != self.active():
            raise ValueError("array lengths disagree")
        d = out if out is not None else self._image()
        q = ctypes.c_double(0.0)
        _check(self._fn("residual_sorted")(self._h, _ptr(vis), _VIS_KIND[vis.dtype], _ptr(model), _ptr(d),
                                            ctypes.byref(q), _stream()))
        return d, q.value
