"""The FISTA oracle (oracle/prox.py) against the known answers of rang-core's
own solver tests (core/src/prox_solver.rs:1000-1300): the proximal map, the
dense problem's gap, linearity, KKT and curvature checks, forward-backward
reaching the FISTA objective, the signed prior, and power iteration on a
known spectrum. CPU only."""

import math

import numpy as np
import pytest

from oracle import prox as X


def dense_problem(n, rows):
    # prox_solver.rs:1000-1063: G = A^T A for a seeded dense A
    a = X.seeded(rows * n * n, 21).reshape(rows, n * n)
    truth = np.zeros(n * n)
    truth[3] = 2.0
    truth[n * n // 2 + 1] = 1.0
    truth[n * n - 2] = 0.5
    noise = X.seeded(rows, 99)
    y = a @ truth + 0.02 * noise
    d = a.T @ y
    q0 = 0.5 * float(np.sum(y * y))

    def gram(x):
        return a.T @ (a @ x)

    return gram, d, q0


def test_seeded_is_xorshift64():
    # the first values of the reference's generator from seed 1, by hand
    state = 1
    mask = (1 << 64) - 1
    want = []
    for _ in range(3):
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        want.append((state >> 11) / float(1 << 53) - 0.5)
    assert np.array_equal(X.seeded(3, 1), np.array(want))
    assert np.array_equal(X.seeded(3, 0), X.seeded(3, 1))  # seed.max(1)
    v = X.seeded(100000, 7)
    assert v.min() >= -0.5 and v.max() < 0.5 and abs(v.mean()) < 0.01


def test_the_proximal_operator_is_the_shifted_positive_part():
    # prox_solver.rs:1066-1076
    p = X.L1Prior(2.0, 1.0)
    assert np.array_equal(p.prox(np.array([3.0, 1.0, -1.0]), 0.5), [2.0, 0.0, 0.0])
    assert p.value(np.array([1.0, 0.0, 2.0])) == 6.0
    p.reweight(np.array([1.0, 0.0, 3.0]))
    assert np.array_equal(p.prox(np.array([3.0, 1.0, -1.0]), 0.5), [2.5, 0.0, 0.0])


def test_fista_closes_the_gap_and_the_linear_trick_is_exact():
    # prox_solver.rs:1078-1133
    gram, d, q0 = dense_problem(6, 20)
    lmax, _ = X.power_iteration(gram, 36, 3, 1e-8, 2000)
    prior = X.L1Prior(0.02, 1.0)
    worst = [0.0]

    def observer(r, x):
        g = gram(x)
        want = q0 - float(np.dot(d, x)) + 0.5 * float(np.dot(x, g))
        worst[0] = max(worst[0], abs(want - r["misfit"]))

    out = X.solve(gram, d, q0, prior, np.zeros(36), step=1.0 / (1.01 * lmax), max_iterations=20000,
                  gap_tolerance=1e-11, observer=observer)
    assert out["stop"] == "converged"
    assert worst[0] < 1e-10
    assert out["last"]["kkt"] < 1e-3
    assert out["last"]["gap"] >= -1e-12
    assert out["min_curvature"] >= -1e-10
    assert out["max_curvature"] <= lmax * (1.0 + 1e-6)
    fb = X.solve(gram, d, q0, prior, np.zeros(36), step=1.0 / (1.01 * lmax), max_iterations=200000,
                 gap_tolerance=1e-11, momentum=False)
    assert fb["largest_rise"] < 1e-12
    assert abs(fb["last"]["objective"] - out["last"]["objective"]) < 1e-9
    assert fb["iterations"] > out["iterations"]
    # one Gram application per iteration (the linearity trick)
    assert out["applies"] == out["iterations"] + 1


def test_a_plausible_mistake_breaks_the_linear_trick():
    # G x taken as G y (the extrapolated point) instead of by linearity: the
    # misfit the solver reports then departs from the direct one
    gram, d, q0 = dense_problem(6, 20)
    lmax, _ = X.power_iteration(gram, 36, 3, 1e-8, 2000)
    prior = X.L1Prior(0.02, 1.0)
    x = np.abs(X.seeded(36, 4))
    x_prev = 0.5 * x
    beta = 0.4
    y = (1 + beta) * x - beta * x_prev
    gx_lin = (gram(y) + beta * gram(x_prev)) / (1 + beta)
    assert np.allclose(gx_lin, gram(x), atol=1e-12)
    m_good = X.certificate(d, q0, x, gx_lin, prior)[0]
    m_bad = X.certificate(d, q0, x, gram(y), prior)[0]
    m_true = q0 - float(np.dot(d, x)) + 0.5 * float(np.dot(x, gram(x)))
    assert abs(m_good - m_true) < 1e-10
    assert abs(m_bad - m_true) > 1e-3


def test_signed_l1_closes_its_gap_below_the_nonnegative_objective():
    # prox_solver.rs:1241-1276
    gram, d, q0 = dense_problem(6, 20)
    lmax, _ = X.power_iteration(gram, 36, 3, 1e-8, 2000)
    kw = dict(step=1.0 / (1.01 * lmax), max_iterations=50000, gap_tolerance=1e-11)
    a = X.solve(gram, d, q0, X.L1Prior(0.02, 1.0, signed=True), np.zeros(36), **kw)
    assert a["stop"] == "converged"
    assert a["last"]["kkt"] < 1e-3
    b = X.solve(gram, d, q0, X.L1Prior(0.02, 1.0), np.zeros(36), **kw)
    assert a["last"]["objective"] <= b["last"]["objective"] + 1e-10
    assert np.all(b["x"] >= 0.0)


def test_power_iteration_finds_a_known_spectrum():
    # prox_solver.rs:1278-1300 (the power-iteration half)
    diag = np.arange(1, 17) / 16.0
    diag[0] = -0.25
    lam, _ = X.power_iteration(lambda x: diag * x, 16, 5, 1e-10, 5000)
    assert abs(lam - 1.0) < 1e-6


def test_certificate_gap_vanishes_at_the_exact_minimiser():
    # a separable problem, G = I: the minimiser of 1/2|x - d|^2 + lam |x|_1,
    # x >= 0, is max(d - lam, 0); there the KKT residual and the gap vanish
    rng = np.random.default_rng(3)
    d = rng.normal(size=50)
    q0 = 0.5 * float(np.sum(d * d))
    prior = X.L1Prior(0.3, 1.0)
    x = np.maximum(d - 0.3, 0.0)
    misfit, pen, gap, kkt = X.certificate(d, q0, x, x, prior)
    assert kkt < 1e-15
    assert abs(gap) < 1e-12
    assert math.isclose(misfit + pen, 0.5 * float(np.sum((x - d) ** 2)) + 0.3 * float(np.sum(x)) + q0
                        - 0.5 * float(np.sum(d * d)), rel_tol=1e-12)
    # and FISTA with step 1 reaches it in one step from zero
    out = X.solve(lambda v: v, d, q0, prior, np.zeros(50), step=1.0, max_iterations=5, gap_tolerance=1e-12)
    assert out["stop"] == "converged"
    assert np.allclose(out["x"], x, atol=1e-15)


def test_defect_correction_reaches_the_exact_problem_from_a_wrong_operator():
    # prox_solver.rs:1301-1376: the cheap operator is the exact one scaled by
    # 1.2; the cycles converge to the exact problem's objective, while the
    # wrong operator alone gives a clearly different solution
    gram, d, q0 = dense_problem(6, 20)
    lmax, _ = X.power_iteration(gram, 36, 3, 1e-8, 2000)

    def wrong(x):
        return 1.2 * gram(x)

    prior = X.L1Prior(0.02, 1.0)
    seen = []
    out = X.defect_correction(gram, wrong, d, q0, prior, np.zeros(36), step=1.0 / (1.3 * lmax),
                              max_iterations=400, kkt_tolerance=1e-7, cycles=60, kkt=1e-6, lam=0.02,
                              gamma=0.0, observer=lambda r, x: seen.append(r))
    assert out["stop"] == "converged"
    direct = X.solve(gram, d, q0, prior, np.zeros(36), step=1.0 / (1.01 * lmax), max_iterations=20000,
                     gap_tolerance=1e-12)
    assert abs(out["last"]["objective"] - direct["last"]["objective"]) < 1e-7
    alone = X.solve(wrong, d, q0, prior, np.zeros(36), step=1.0 / (1.3 * lmax), max_iterations=20000,
                    kkt_tolerance=1e-7)
    assert float(np.max(np.abs(alone["x"] - direct["x"]))) > 1e-2
    assert len(seen) > 1
    # one exact and one cheap application per cycle, plus the inner solves'
    assert out["exact_applies"] == len(seen)
    assert out["records"][-1]["kkt"] <= 1e-6
    # the defect at the fixed point is the operators' difference at x: 0.2 G x
    assert math.isclose(out["records"][-1]["defect_sup"], float(np.max(np.abs(0.2 * gram(out["x"])))),
                        rel_tol=1e-12)


def test_defect_correction_with_the_exact_operator_is_one_solve():
    # G~ = G: the correction vanishes, so the first cycle's inner solve is the
    # plain solve from x0, and the next certificate judges its iterate
    gram, d, q0 = dense_problem(6, 20)
    lmax, _ = X.power_iteration(gram, 36, 3, 1e-8, 2000)
    kw = dict(step=1.0 / (1.01 * lmax), max_iterations=150)
    out = X.defect_correction(gram, gram, d, q0, X.L1Prior(0.02, 1.0), np.zeros(36), cycles=1, kkt=0.0,
                              lam=0.02, **kw)
    plain = X.solve(gram, d, q0, X.L1Prior(0.02, 1.0), np.zeros(36), **kw)
    assert out["stop"] == "cycles"
    assert np.array_equal(out["x"], plain["x"])
    assert out["records"][0]["defect_sup"] == 0.0
    assert out["iterations"] == plain["iterations"] == 150
    m, p, _, k = X.certificate(d, q0, plain["x"], gram(plain["x"]), X.L1Prior(0.02, 1.0))
    assert math.isclose(out["last"]["objective"], m + p, rel_tol=1e-13)
    assert math.isclose(out["last"]["kkt"], k, rel_tol=1e-13)


def test_defect_correction_with_the_sign_of_the_defect_flipped_misses_the_exact_problem(monkeypatch):
    # a planted mistake in the oracle: d + (G - G~) x_j instead of d - (...);
    # the cycles then settle away from the exact problem's objective
    gram, d, q0 = dense_problem(6, 20)
    lmax, _ = X.power_iteration(gram, 36, 3, 1e-8, 2000)
    orig = X.solve

    def flipped(cheap, corrected, q0_, prior, x, **kw):
        # corrected = d - c arrives here; hand the inner solve d + c
        return orig(cheap, 2.0 * d - corrected, q0_, prior, x, **kw)

    monkeypatch.setattr(X, "solve", flipped)
    out = X.defect_correction(gram, lambda x: 1.2 * gram(x), d, q0, X.L1Prior(0.02, 1.0), np.zeros(36),
                              step=1.0 / (1.3 * lmax), max_iterations=400, kkt_tolerance=1e-7, cycles=60,
                              kkt=1e-6, lam=0.02)
    monkeypatch.setattr(X, "solve", orig)
    direct = X.solve(gram, d, q0, X.L1Prior(0.02, 1.0), np.zeros(36), step=1.0 / (1.01 * lmax),
                     max_iterations=20000, gap_tolerance=1e-12)
    assert out["stop"] == "cycles"
    assert abs(out["last"]["objective"] - direct["last"]["objective"]) > 1e-4
