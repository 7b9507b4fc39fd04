"""The normal operator of the measurement without visibilities, on the GPU:
rang-core's `gridder_gram` (core/src/gridder_gram.rs).

    ex = WExpansion(10, 1.0, 1e-13)                  # :84-183
    G = GramOperator.new(opts, uvw, freq, wgt, ex, epsilon, single=True)   # :226-407
    P = GramOperator.no_w(opts, uvw, freq, wgt, epsilon, single=True)      # :414-478
    M = P.inverse_filter(rho)                         # :488-509, a preconditioner
    y = G.apply(x)                                    # :519-647

G m ~ Re sum_k e_k (K_k (*) (conj(e_k) m)) over w-nodes k, each Toeplitz
product by FFTs on the doubled grid. Arrays are contiguous torch tensors on
the current GPU; calls run on torch's current stream.
"""

import ctypes

import numpy as np
import torch

from . import _core
from .device import _WGT_KIND, _check, _ptr, _stream

BINS = 8192
STATS = ("planes", "stencil", "theta", "wmin", "wmax", "x_range", "sup_error", "lebesgue",
         "active_pairs", "cube_bytes", "hermitian_check", "spectral_sparsity", "spectral_peak")


class WExpansion:
    """The compact w-expansion's fit (gridder_gram.rs:84-141): its measured
    sup_error, Lebesgue constant and coefficient table
    ((BINS + 1) x stencil, psi_j at s = -1/2 + b / BINS)."""

    def __init__(self, stencil, theta, ridge):
        self.stencil, self.theta, self.ridge = int(stencil), float(theta), float(ridge)
        sup, leb = ctypes.c_double(0.0), ctypes.c_double(0.0)
        table = np.zeros((BINS + 1, max(self.stencil, 1)))
        _check(_core.library().rang_amd_wexpansion(self.stencil, self.theta, self.ridge, ctypes.byref(sup),
                                                    ctypes.byref(leb), ctypes.c_void_p(table.ctypes.data)))
        self.sup_error, self.lebesgue, self.table = sup.value, leb.value, table

    def coefficient(self, s, j):
        """psi_j(s) by linear interpolation in the table (gridder_gram.rs:167-175)."""
        f = (s + 0.5) * BINS
        b = min(max(int(f), 0), BINS - 1)
        a = f - b
        return self.table[b, j] + a * (self.table[b + 1, j] - self.table[b, j])


class GramOperator:
    def __init__(self, handle, single):
        if not handle:
            raise ValueError(_core.wgridder_last_error())
        self._h = ctypes.c_void_p(handle)
        self.single = bool(single)
        self.real = torch.float32 if single else torch.float64
        self.n = int(_core.library().rang_amd_gram_side(self._h))

    @staticmethod
    def _args(opts, uvw, freq, wgt):
        if wgt is None or wgt.dtype not in _WGT_KIND:
            raise ValueError("the Gram operator needs float32 or float64 weights")
        if wgt.numel() != uvw.shape[0] * freq.shape[0]:
            raise ValueError("array lengths disagree")
        o = _core.options(opts)
        return (ctypes.byref(o), int(uvw.shape[0]), int(freq.shape[0]), _ptr(uvw), _ptr(freq), _ptr(wgt),
                _WGT_KIND[wgt.dtype]), o

    @classmethod
    def new(cls, opts, uvw, freq, wgt, expansion, epsilon, *, single):
        """The w-aware operator of weights `wgt` (used as given)."""
        args, keep = cls._args(opts, uvw, freq, wgt)
        lib = _core.library()
        fn = lib.rang_amd_gram_new_f32 if single else lib.rang_amd_gram_new_f64
        h = fn(*args, int(expansion.stencil), float(expansion.theta), float(expansion.ridge),
               float(epsilon), _stream())
        return cls(h, single)

    @classmethod
    def no_w(cls, opts, uvw, freq, wgt, epsilon, *, single):
        """One node without the w-term: the translation-invariant PSF."""
        args, keep = cls._args(opts, uvw, freq, wgt)
        lib = _core.library()
        fn = lib.rang_amd_gram_no_w_f32 if single else lib.rang_amd_gram_no_w_f64
        return cls(fn(*args, float(epsilon), _stream()), single)

    def inverse_filter(self, rho):
        """The circulant operator with spectrum 1 / (max(K, 0) + rho)."""
        return GramOperator(_core.library().rang_amd_gram_inverse_filter(self._h, float(rho), _stream()),
                            self.single)

    def apply(self, x, out=None):
        y = out if out is not None else torch.empty((self.n, self.n), dtype=self.real, device=x.device)
        fn = _core.library().rang_amd_gram_apply_f32 if self.single else _core.library().rang_amd_gram_apply_f64
        _check(fn(self._h, _ptr(x), _ptr(y), _stream()))
        return y

    __call__ = apply

    def stats(self):
        out = np.zeros(len(STATS))
        _check(_core.library().rang_amd_gram_stats(self._h, ctypes.c_void_p(out.ctypes.data)))
        return dict(zip(STATS, out.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            _core.library().rang_amd_gram_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
