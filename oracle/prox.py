"""TEST INFRASTRUCTURE ONLY: a CPU restatement of rang-core's forward-backward
/ FISTA solver (core/src/prox_solver.rs) for the parity tests of the device
solver. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
may import it; the product never does.

It follows prox_solver.rs:
  * seeded            - :871-882 (xorshift64 start vectors)
  * L1Prior           - :55-132 (prox, value, thresholds, reweight)
  * certificate       - :236-281 (misfit, penalty, duality gap, KKT residual)
  * solve             - :283-445 (FISTA with the linearity trick and gradient
                                  restart; forward-backward without momentum)
  * defect_correction - :760-875 (major cycles: exact certificate, the cheap
                                  operator's defect moved into the data, FISTA
                                  on the cheap operator)
  * power_iteration   - :886-908

The operator is any callable gram(x) -> G x on flat float64 arrays. Sums run
serially in float64 (the reference sums in parallel; both agree to rounding).
Pinned by the reference's own known answers (tests/test_prox_oracle.py): the
proximal map of prox_solver.rs:1066-1076, the dense problem of :1000-1063 with
its gap, linearity, KKT, curvature and forward-backward checks of :1078-1133
and :1241-1276, the diagonal spectrum of :1278-1300, and the defect
correction from an operator 20 % off of :1301-1376.
"""

import math

import numpy as np


def seeded(n, seed):
    # prox_solver.rs:871-882
    state = max(int(seed), 1) & 0xFFFFFFFFFFFFFFFF
    out = np.empty(n, dtype=np.float64)
    mask = 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        out[i] = (state >> 11) / float(1 << 53) - 0.5
    return out


class L1Prior:
    """lambda |x|_1 with x >= 0 unless signed, optionally reweighted by
    delta / (delta + |x_i|) (prox_solver.rs:55-132)."""

    def __init__(self, lam, delta, signed=False):
        self.lam = float(lam)
        self.delta = float(delta)
        self.signed = bool(signed)
        self.factors = None

    def _shrink(self, v, t):
        if self.signed:
            return np.sign(v) * np.maximum(np.abs(v) - t, 0.0)
        return np.maximum(v - t, 0.0)

    def prox(self, z, step):
        tau = step * self.lam
        if self.factors is None:
            return self._shrink(z, tau)
        return self._shrink(z, tau * self.factors)

    def value(self, x):
        if self.factors is None:
            return self.lam * float(np.sum(np.abs(x)))
        return self.lam * float(np.sum(np.abs(x) * self.factors))

    def threshold(self, i=None):
        if self.factors is None:
            return self.lam if i is None else self.lam
        return self.lam * (self.factors if i is None else self.factors[i])

    def reweight(self, x):
        self.factors = self.delta / (self.delta + np.abs(x))

    def set_lambda(self, lam):
        self.lam = float(lam)


def certificate(dirty, q0, x, gx, prior):
    """(misfit, penalty, gap, kkt) of x from G x (prox_solver.rs:236-281)."""
    d_x = float(np.dot(dirty, x))
    misfit = q0 - d_x + 0.5 * float(np.dot(x, gx))
    lam = prior.threshold()
    lam = np.broadcast_to(lam, x.shape)
    r = dirty - gx
    pos, neg = x > 0.0, x < 0.0
    if prior.signed:
        zero_viol = np.maximum(np.abs(r) - lam, 0.0)
    else:
        zero_viol = np.maximum(r - lam, 0.0)
    viol = np.where(pos, np.abs(r - lam), np.where(neg, np.abs(r + lam), zero_viol))
    excess = np.abs(r) if prior.signed else r
    with np.errstate(divide="ignore", invalid="ignore"):
        bound = np.where(excess > 0.0, lam / np.where(excess > 0.0, excess, 1.0), np.inf)
    pen = float(np.sum(lam * np.abs(x)))
    kkt = float(np.max(viol)) if viol.size else 0.0
    scale_max = float(np.min(bound)) if bound.size else math.inf
    y_e = 2.0 * q0 - d_x
    e_e = 2.0 * misfit
    s = min(max(y_e / max(e_e, 1e-300), 0.0), scale_max)
    dual = s * y_e - 0.5 * s * s * e_e
    gap = misfit + pen - dual
    return misfit, pen, gap, kkt


def solve(gram, dirty, q0, prior, x0, *, step, max_iterations=100, momentum=True, restart=True,
          gap_tolerance=None, kkt_tolerance=None, step_tolerance=None, observer=None):
    """FISTA (momentum) or forward-backward from x0 (prox_solver.rs:283-445).
    Returns a dict with x, gx, last (the last record), iterations, stop,
    restarts, objective_rises, largest_rise, negative_curvatures,
    min_curvature, max_curvature, applies."""
    dirty = np.asarray(dirty, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    n2 = x.size
    x_prev = x.copy()
    y = x.copy()
    y_prev = np.zeros(n2)
    g_y_prev = np.zeros(n2)
    g_x_prev = np.zeros(n2)
    t = 1.0
    beta = 0.0
    out = dict(restarts=0, objective_rises=0, largest_rise=0.0, negative_curvatures=0,
               min_curvature=math.inf, max_curvature=-math.inf, applies=0)
    last_objective = math.inf
    restarted = False
    for k in range(max_iterations + 1):
        g_y = np.asarray(gram(y), dtype=np.float64)
        out["applies"] += 1
        g_x = g_y.copy() if beta == 0.0 else (g_y + beta * g_x_prev) / (1.0 + beta)
        if k == 0:
            curvature = math.nan
        else:
            dy = y - y_prev
            num = float(np.dot(dy, g_y - g_y_prev))
            den = float(np.dot(dy, dy))
            curvature = num / den if den > 0.0 else math.nan
        if math.isfinite(curvature):
            out["min_curvature"] = min(out["min_curvature"], curvature)
            out["max_curvature"] = max(out["max_curvature"], curvature)
            if curvature < 0.0:
                out["negative_curvatures"] += 1
        misfit, pen, gap, kkt = certificate(dirty, q0, x, g_x, prior)
        change = float(np.sum((x - x_prev) ** 2))
        norm = float(np.sum(x * x))
        relative_step = math.sqrt(change) / max(math.sqrt(norm), 1e-300)
        record = dict(iteration=k, objective=misfit + pen, misfit=misfit, penalty=pen, gap=gap, kkt=kkt,
                      nonzero=int(np.count_nonzero(x)), flux=float(np.sum(x)), curvature=curvature,
                      relative_step=relative_step, restarted=restarted)
        if observer is not None:
            observer(record, x)
        if math.isfinite(record["objective"]):
            if record["objective"] > last_objective:
                out["objective_rises"] += 1
                out["largest_rise"] = max(out["largest_rise"], record["objective"] - last_objective)
            last_objective = record["objective"]
        out["last"] = record
        out["iterations"] = k
        converged = ((gap_tolerance is not None and gap <= gap_tolerance)
                     or (kkt_tolerance is not None and kkt <= kkt_tolerance)
                     or (k > 0 and step_tolerance is not None and relative_step <= step_tolerance))
        if converged or k == max_iterations:
            out["stop"] = "converged" if converged else "iterations"
            out["gx"] = g_x
            out["x"] = x
            return out
        z = y + step * (dirty - g_y)
        x_next = prior.prox(z, step)
        next_beta = 0.0
        restarted = False
        if momentum:
            inner = float(np.dot(y - x_next, x_next - x))
            if restart and inner > 0.0:
                t = 1.0
                restarted = True
                out["restarts"] += 1
            else:
                t_next = 0.5 * (1.0 + math.sqrt(1.0 + 4.0 * t * t))
                next_beta = (t - 1.0) / t_next
                t = t_next
        y_prev = y
        g_y_prev = g_y
        g_x_prev = g_x
        x_prev = x
        x = x_next
        y = x + next_beta * (x - x_prev)
        beta = next_beta
    raise AssertionError("the loop returns at max_iterations")


def defect_correction(exact, cheap, dirty, q0, prior, x0, *, step, max_iterations, cycles, kkt, lam,
                      gamma=0.0, momentum=True, restart=True, kkt_tolerance=None, step_tolerance=None,
                      observer=None):
    """Major cycles on a cheap operator with its error at x_j moved into the
    data (prox_solver.rs:760-875): each cycle applies both operators at x_j,
    judges x_j by the exact certificate at `lam`, and solves
    min 1/2 <x, G~ x> - <d - (G - G~) x_j, x> + R(x) with FISTA on the cheap
    operator from x_j at max(lam, gamma * peak residual), no gap test. Returns
    a dict with x, gx (exact G x), last (objective, misfit, gap, kkt),
    iterations (inner, summed), stop, restarts, negative_curvatures,
    records (one per cycle), exact_applies and cheap_applies."""
    dirty = np.asarray(dirty, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    out = dict(restarts=0, negative_curvatures=0, records=[], exact_applies=0, cheap_applies=0)
    inner_iterations = 0
    for j in range(cycles + 1):
        g = np.asarray(exact(x), dtype=np.float64)
        out["exact_applies"] += 1
        prior.set_lambda(lam)
        misfit, pen, gap, k_res = certificate(dirty, q0, x, g, prior)
        g_cheap = np.asarray(cheap(x), dtype=np.float64)
        out["cheap_applies"] += 1
        c = g - g_cheap
        defect_sup = float(np.max(np.abs(c))) if c.size else 0.0
        corrected = dirty - c
        record = dict(outer=j, objective=misfit + pen, misfit=misfit, gap=gap, kkt=k_res,
                      inner_iterations=inner_iterations, defect_sup=defect_sup)
        out["records"].append(record)
        if observer is not None:
            observer(record, x)
        out["last"] = dict(objective=misfit + pen, misfit=misfit, gap=gap, kkt=k_res)
        if k_res <= kkt or j == cycles:
            out["stop"] = "converged" if k_res <= kkt else "cycles"
            out["gx"] = g
            out["x"] = x
            out["iterations"] = inner_iterations
            return out
        r = dirty - g
        peak = max(0.0, float(np.max(np.abs(r) if prior.signed else r))) if r.size else 0.0
        prior.set_lambda(max(lam, gamma * peak))
        solved = solve(cheap, corrected, q0, prior, x, step=step, max_iterations=max_iterations,
                       momentum=momentum, restart=restart, gap_tolerance=None, kkt_tolerance=kkt_tolerance,
                       step_tolerance=step_tolerance)
        inner_iterations += solved["iterations"]
        out["restarts"] += solved["restarts"]
        out["negative_curvatures"] += solved["negative_curvatures"]
        out["cheap_applies"] += solved["applies"]
        x = solved["x"]
    raise AssertionError("the loop returns at the last cycle")


def power_iteration(gram, n2, seed, tolerance, max_iterations):
    """(largest eigenvalue, applies) from a seeded start (prox_solver.rs:886-908)."""
    v = seeded(n2, seed)
    v /= math.sqrt(float(np.dot(v, v)))
    value = 0.0
    for k in range(1, max_iterations + 1):
        g = np.asarray(gram(v), dtype=np.float64)
        nxt = math.sqrt(float(np.dot(g, g)))
        v = g / max(nxt, 1e-300)
        if k > 1 and abs(nxt - value) <= tolerance * nxt:
            return nxt, k
        value = nxt
    return value, max_iterations
