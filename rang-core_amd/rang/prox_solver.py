"""Forward-backward and FISTA for 1/2 |y - Phi m|^2_W + R(m) on the GPU, as
rang-core's `prox_solver` (core/src/prox_solver.rs): `solve` (:283-445) with
the `L1Prior` (:55-132) over a `WgridMeasurement`, `defect_correction`
(:760-875: major cycles judged by the exact operator, each solved on a cheap
one such as rang.gram's visibility-free operator), and `power_iteration`
(:886-908) for the step 1 / L.

The iteration runs natively on the device (one Gram application per
iteration, G x by linearity, sums in double); the per-iterate records come
back at the end instead of through an observer callback.

    m = WgridMeasurement(uvw, freq, vis, wgt, side, cell, epsilon)
    dirty, q0 = m.residual(zeros)                  # d = Phi^H W y, Q0
    lmax, _ = power_iteration(m, 3, 1e-8, 2000)
    out = solve(m, Problem(dirty, q0), L1Prior(0.02, 1.0),
                Settings(step=1 / (1.01 * lmax), gap_tolerance=1e-9), x0)
"""

import math
from dataclasses import dataclass, field
from typing import List, Optional

import torch

from . import _core


class L1Prior:
    """lambda |x|_1 with x >= 0 unless signed, optionally reweighted by
    delta / (delta + |x_i|) (prox_solver.rs:55-132). Factors live on the
    device as a tensor of the image's shape."""

    def __init__(self, lam, delta, *, signed=False):
        self.lam = float(lam)
        self.delta = float(delta)
        self.signed = bool(signed)
        self.factors = None

    @classmethod
    def new_signed(cls, lam, delta):
        return cls(lam, delta, signed=True)

    def prox(self, z, step):
        tau = step * self.lam
        t = tau if self.factors is None else tau * self.factors
        if self.signed:
            return torch.sign(z) * torch.clamp(z.abs() - t, min=0.0)
        return torch.clamp(z - t, min=0.0)

    def value(self, x):
        if self.factors is None:
            return self.lam * float(x.abs().sum())
        return self.lam * float((x.abs() * self.factors).sum())

    def reweight(self, x):
        self.factors = (self.delta / (self.delta + x.abs())).contiguous()

    def set_lambda(self, lam):
        self.lam = float(lam)


@dataclass
class Problem:
    """The data side (prox_solver.rs:167-172): d = Phi^H W y and Q0 = 1/2 |y|^2_W."""

    dirty: torch.Tensor
    q0: float


@dataclass
class Settings:
    """prox_solver.rs:175-208 (objective_every does not apply to the L1 prior,
    whose per-pixel thresholds give the objective every iteration)."""

    step: float = 1.0
    max_iterations: int = 100
    momentum: bool = True
    restart: bool = True
    gap_tolerance: Optional[float] = None
    kkt_tolerance: Optional[float] = None
    step_tolerance: Optional[float] = None


@dataclass
class Record:
    """The state of one iterate (prox_solver.rs:211-231)."""

    iteration: int
    objective: float
    misfit: float
    penalty: float
    gap: float
    kkt: float
    nonzero: int
    flux: float
    curvature: float
    relative_step: float
    restarted: bool


@dataclass
class Outcome:
    """prox_solver.rs:234-249."""

    x: torch.Tensor
    gx: torch.Tensor
    last: Record
    iterations: int
    stop: str
    restarts: int
    objective_rises: int
    largest_rise: float
    negative_curvatures: int
    min_curvature: float
    max_curvature: float
    records: List[Record] = field(default_factory=list)


def _record(r):
    return Record(int(r.iteration), r.objective, r.misfit, r.penalty, r.gap, r.kkt, int(r.nonzero), r.flux,
                  r.curvature, r.relative_step, bool(r.restarted))


def _tol(v):
    return math.nan if v is None else float(v)


def solve(measurement, problem, prior, settings, x0, *, keep_records=True):
    """FISTA (momentum) or forward-backward from x0 on the measurement's
    normal operator (prox_solver.rs:283-445). Returns an Outcome whose
    `records` hold every iterate's Record when keep_records."""
    if not isinstance(prior, L1Prior):
        raise TypeError("the device solver takes the L1 prior")
    plan = measurement.plan
    real = torch.float32 if plan.single else torch.float64
    x = x0.to(real).contiguous().clone()
    gx = torch.empty_like(x)
    dirty = problem.dirty.to(real).contiguous()
    factors = None if prior.factors is None else prior.factors.to(real).contiguous()
    s = _core.FistaSettings(float(settings.step), int(settings.max_iterations), int(bool(settings.momentum)),
                            int(bool(settings.restart)), _tol(settings.gap_tolerance),
                            _tol(settings.kkt_tolerance), _tol(settings.step_tolerance))
    nrec = int(settings.max_iterations) + 1 if keep_records else 0
    out, recs = plan.fista(dirty, problem.q0, prior.lam, x, factors=factors, signed=prior.signed, settings=s,
                           gx=gx, nrecords=nrec)
    return Outcome(x=x, gx=gx, last=_record(out.last), iterations=int(out.iterations),
                   stop="converged" if out.converged else "iterations", restarts=int(out.restarts),
                   objective_rises=int(out.objective_rises), largest_rise=out.largest_rise,
                   negative_curvatures=int(out.negative_curvatures), min_curvature=out.min_curvature,
                   max_curvature=out.max_curvature, records=[_record(r) for r in recs])


@dataclass
class OuterSchedule:
    """prox_solver.rs:760-772: at most `cycles` cycles; stop when the exact
    optimality residual at the target `lam` is <= kkt; a cycle solves at
    max(lam, gamma * peak residual) (0 keeps lam)."""

    cycles: int
    kkt: float
    lam: float
    gamma: float = 0.0


@dataclass
class OuterRecord:
    """One major cycle (prox_solver.rs:447-462)."""

    outer: int
    objective: float
    misfit: float
    gap: float
    kkt: float
    inner_iterations: int
    defect_sup: float


@dataclass
class DefectOutcome:
    """What defect_correction returns (prox_solver.rs:778-875): the judged
    iterate, its exact G x, the last cycle's record, the inner iterations
    summed, the stop ("converged" or "cycles"), and the applications of each
    operator."""

    x: torch.Tensor
    gx: torch.Tensor
    last: OuterRecord
    iterations: int
    stop: str
    restarts: int
    negative_curvatures: int
    exact_applies: int
    cheap_applies: int
    records: List[OuterRecord] = field(default_factory=list)


def _outer(r):
    return OuterRecord(int(r.outer), r.objective, r.misfit, r.gap, r.kkt, int(r.inner_iterations), r.defect_sup)


def defect_correction(exact, cheap, problem, prior, inner, schedule, x0):
    """Solve on a cheap operator with its error at x_j moved into the data:
    min 1/2 <x, G~ x> - <d - (G - G~) x_j, x> + R(x), cycle after cycle, each
    cycle judged by the exact operator's certificate at schedule.lam
    (prox_solver.rs:778-875). `exact` is a WgridMeasurement, `cheap` a
    rang.gram.GramOperator of the same side and precision (for instance
    GramOperator.new of the same uvw and normalised weights); `inner` the
    FISTA settings of each inner solve (its gap tolerance is not used). The
    whole loop runs natively on the device."""
    if not isinstance(prior, L1Prior):
        raise TypeError("the device solver takes the L1 prior")
    plan = exact.plan
    real = torch.float32 if plan.single else torch.float64
    if cheap.single != plan.single:
        raise ValueError("the cheap operator's precision differs from the measurement's")
    x = x0.to(real).contiguous().clone()
    gx = torch.empty_like(x)
    dirty = problem.dirty.to(real).contiguous()
    factors = None if prior.factors is None else prior.factors.to(real).contiguous()
    s = _core.FistaSettings(float(inner.step), int(inner.max_iterations), int(bool(inner.momentum)),
                            int(bool(inner.restart)), math.nan, _tol(inner.kkt_tolerance),
                            _tol(inner.step_tolerance))
    sched = _core.OuterSchedule(int(schedule.cycles), float(schedule.kkt), float(schedule.lam),
                                float(schedule.gamma))
    out, recs = plan.defect(cheap, dirty, problem.q0, x, factors=factors, signed=prior.signed, inner=s,
                            schedule=sched, gx=gx, nrecords=int(schedule.cycles) + 1)
    return DefectOutcome(x=x, gx=gx, last=_outer(out.last), iterations=int(out.iterations),
                         stop="converged" if out.converged else "cycles", restarts=int(out.restarts),
                         negative_curvatures=int(out.negative_curvatures), exact_applies=int(out.exact_applies),
                         cheap_applies=int(out.cheap_applies), records=[_outer(r) for r in recs])


def power_iteration(measurement, seed, tolerance, max_iterations):
    """(largest eigenvalue of the normal operator, applications used)
    (prox_solver.rs:886-908)."""
    return measurement.plan.power_iteration(seed, tolerance, max_iterations)
