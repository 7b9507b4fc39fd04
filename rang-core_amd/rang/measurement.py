"""The exact measurement operator of rang-core's certified imaging
(core/src/clean_certified.rs:54-236) on the GPU: `WgridMeasurement` with its
predict / image / residual / gram_apply, the conjugate-gradient solve of
`clean::conjugate_gradients` (core/src/clean.rs:263-306) and `misfit_floor`
(core/src/clean_certified.rs:609-630), over a prepared plan.
"""

import time
from dataclasses import dataclass

import torch

from .prepared import Prepared


def _opts(side, cell, epsilon, **kw):
    return [side, side, cell, cell, epsilon, kw.get("center_x", 0.0), kw.get("center_y", 0.0),
            kw.get("sigma_min", 1.1), kw.get("sigma_max", 2.6), int(kw.get("do_wgridding", True)),
            int(kw.get("divide_by_n", False)), int(kw.get("flip_u", False)),
            int(kw.get("flip_v", False)), int(kw.get("flip_w", False)),
            int(kw.get("allow_nshift", True)), -1, 0, 0]


class WgridMeasurement:
    """Visibilities `vis` and weights `wgt` (nrow x nchan, device tensors);
    the weights normalised to sum to one (clean_certified.rs:117-150).
    `single` selects the float32 gridder (the reference uses float64)."""

    def __init__(self, uvw, freq, vis, wgt, side, cell, epsilon, *, single=False, **opts):
        if vis.shape != wgt.shape or vis.shape[0] != uvw.shape[0] or vis.shape[1] != freq.shape[0]:
            raise ValueError("vis and wgt must be nrow * nchan")
        real = torch.float32 if single else torch.float64
        wsum = float(wgt.to(torch.float64).sum())
        if not wsum > 0.0:
            raise ValueError("weights must have a positive sum")
        self.weight_sum = wsum
        self.wgt = (wgt.to(torch.float64) / wsum).to(real).contiguous()
        self.vis = vis.to(torch.complex64 if single else torch.complex128).contiguous()
        self.side = int(side)
        self.plan = Prepared(_opts(side, cell, epsilon, **opts), uvw, freq, self.wgt, single=single)

    def predict(self, model):
        """The model visibilities A m (no weights)."""
        return self.plan.dirty2vis(model)

    def image(self, vis):
        """A^H W v."""
        return self.plan.vis2dirty(vis)

    def residual(self, model):
        """(A^H W (v - A m), half the weighted misfit) (clean_certified.rs:216-231)."""
        img, q = self.plan.residual(self.vis, model)
        return img, 0.5 * q

    def gram_apply(self, x):
        return self.plan.gram(x)


def conjugate_gradients(measurement, rhs, iterations, tolerance):
    """(x, iterations used, relative residual) of CG on the measurement's
    normal operator, from zero, on the GPU."""
    return measurement.plan.conjugate_gradients(rhs, iterations, tolerance)


@dataclass
class Floor:
    q_floor: float
    iterations: int
    relative_residual: float
    seconds: float


def misfit_floor(measurement, dirty, iterations, tolerance):
    """Q_inf by CG on A^H W A m = dirty and the exact misfit of its model."""
    t0 = time.perf_counter()
    x, used, rel = conjugate_gradients(measurement, dirty, iterations, tolerance)
    _, q = measurement.residual(x)
    torch.cuda.synchronize()
    return Floor(q, used, rel, time.perf_counter() - t0), x
