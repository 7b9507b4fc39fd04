"""Major cycles on the GPU (rang.prox_solver.defect_correction, the device
loop of rang_amd_wgridder_prepared_defect_*) against the oracle's
`defect_correction` (core/src/prox_solver.rs:760-875, restated in
oracle/prox.py and pinned there by the reference's own test of :1301-1376).

The exact operator is the prepared plan's normal operator (the oracle's is
its own vis2dirty of dirty2vis, as in test_fista_gpu.py); the cheap one is
the visibility-free operator of rang.gram (its accuracy against the direct
sum is test_gram_gpu.py's subject). Both runs use that same cheap operator,
the oracle through host copies, so what is compared is the loop: per cycle
the exact certificate (objective, misfit, gap, optimality residual), the
defect's size and the inner iterations. And the point of the method, as the
reference's test puts it: from an operator 20 % off (the visibility-free one
of weights 1.2 W), the cycles reach the exact problem's objective, which FISTA
on that operator alone misses.
"""

import numpy as np
import pytest
from conftest import ref_case

from oracle import oracle as O
from oracle import prox as X

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def host(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def problem():
    from rang import gram
    from rang.measurement import WgridMeasurement
    from rang.prox_solver import power_iteration

    print("[defect] 32^2 problem, 40 rows x 3 channels, f64", flush=True)
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=32, ny=32, nrow=40, seed=17, nchan=3)
    eps = 1e-9
    m = WgridMeasurement(dev(uvw), dev(freq), dev(vis), dev(wgt), nx, dx, eps)
    wn = wgt / wgt.sum()

    def gram_exact(x):
        img = x.reshape(nx, nx)
        v, _ = O.dirty2vis(uvw, freq, img, None, (wn != 0).astype(np.uint8), pixsize_x=dx, pixsize_y=dx,
                           epsilon=eps)
        out, _ = O.vis2dirty(uvw, freq, v, wn, None, npix_x=nx, npix_y=nx, pixsize_x=dx, pixsize_y=dx,
                             epsilon=eps, single=False)
        return out["dirty"].reshape(-1)

    opts = [nx, nx, dx, dx, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    cheap = gram.GramOperator.new(opts, dev(uvw), dev(freq), m.wgt, gram.WExpansion(8, 1.2, 1e-13), eps,
                                  single=False)

    def gram_cheap(x):
        return host(cheap.apply(dev(x.reshape(nx, nx)))).reshape(-1)

    dirty, q0 = m.residual(torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    lmax, _ = power_iteration(m, 3, 1e-8, 500)
    return dict(m=m, cheap=cheap, gram_exact=gram_exact, gram_cheap=gram_cheap, dirty=dirty, q0=q0,
                d=host(dirty).reshape(-1), nx=nx, lmax=lmax, opts=opts, uvw=uvw, freq=freq)


def test_cycles_match_the_oracle(problem):
    from rang.prox_solver import L1Prior, OuterSchedule, Problem, Settings, defect_correction

    p, nx = problem, problem["nx"]
    lam = 0.05 * float(np.max(p["d"]))
    step = 1.0 / (1.3 * p["lmax"])
    scale = float(np.max(np.abs(p["d"])))
    inner = Settings(step=step, max_iterations=60, kkt_tolerance=1e-9 * scale)
    sched = OuterSchedule(cycles=6, kkt=1e-12 * scale, lam=lam, gamma=0.0)
    out = defect_correction(p["m"], p["cheap"], Problem(p["dirty"], p["q0"]), L1Prior(lam, 1.0), inner, sched,
                            torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    ref = X.defect_correction(p["gram_exact"], p["gram_cheap"], p["d"], p["q0"], X.L1Prior(lam, 1.0),
                              np.zeros(nx * nx), step=step, max_iterations=60, kkt_tolerance=1e-9 * scale,
                              cycles=6, kkt=1e-12 * scale, lam=lam)
    print(f"[defect] device: {len(out.records)} cycles, {out.iterations} inner iterations, stop {out.stop}; "
          f"oracle: {len(ref['records'])} cycles, {ref['iterations']} inner", flush=True)
    assert out.stop == ref["stop"] == "cycles"
    assert len(out.records) == len(ref["records"]) == 7
    q = abs(out.records[0].objective)
    for g, r in zip(out.records, ref["records"]):
        assert g.outer == r["outer"]
        for k in ("objective", "misfit", "gap"):
            assert abs(getattr(g, k) - r[k]) <= 1e-8 * q, (g.outer, k, getattr(g, k), r[k])
        assert abs(g.kkt - r["kkt"]) <= 1e-6 * scale, (g.outer, g.kkt, r["kkt"])
        assert abs(g.defect_sup - r["defect_sup"]) <= 1e-8 * max(r["defect_sup"], 1e-300) + 1e-12 * scale
        assert abs(g.inner_iterations - r["inner_iterations"]) <= 2
    assert out.exact_applies == ref["exact_applies"] == 7
    assert abs(out.cheap_applies - ref["cheap_applies"]) <= 8
    x = host(out.x).reshape(-1)
    assert np.max(np.abs(x - ref["x"])) <= 1e-6 * np.max(np.abs(ref["x"]))
    # gx is the exact G x of the returned x
    gx = host(p["m"].gram_apply(out.x)).reshape(-1)
    assert np.max(np.abs(host(out.gx).reshape(-1) - gx)) <= 1e-12 * np.max(np.abs(gx))


def test_cycles_reach_the_exact_problem_from_an_operator_20_percent_off(problem):
    # prox_solver.rs:1301-1376 on the gridder: the cheap operator is the
    # visibility-free one of weights 1.2 W, i.e. 1.2 G; alone it gives a
    # clearly different solution, while the cycles reach the exact problem's
    # objective
    from rang import gram
    from rang.prox_solver import L1Prior, OuterSchedule, Problem, Settings, defect_correction, solve

    p, nx = problem, problem["nx"]
    lam = 0.05 * float(np.max(p["d"]))
    scale = float(np.max(np.abs(p["d"])))
    zeros = torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0")
    wrong = gram.GramOperator.new(p["opts"], dev(p["uvw"]), dev(p["freq"]), (1.2 * p["m"].wgt).contiguous(),
                                  gram.WExpansion(8, 1.2, 1e-13), 1e-9, single=False)

    def gram_wrong(x):
        return host(wrong.apply(dev(x.reshape(nx, nx)))).reshape(-1)

    x = dev(np.random.default_rng(5).standard_normal((nx, nx)))
    ratio = float((wrong.apply(x) - 1.2 * p["m"].gram_apply(x)).abs().max() / (1.2 * p["m"].gram_apply(x)).abs().max())
    assert ratio < 1e-3
    direct = solve(p["m"], Problem(p["dirty"], p["q0"]), L1Prior(lam, 1.0),
                   Settings(step=1.0 / (1.01 * p["lmax"]), max_iterations=20000, gap_tolerance=1e-13 * p["q0"]),
                   zeros)
    assert direct.stop == "converged"
    out = defect_correction(p["m"], wrong, Problem(p["dirty"], p["q0"]), L1Prior(lam, 1.0),
                            Settings(step=1.0 / (1.3 * p["lmax"]), max_iterations=400, kkt_tolerance=1e-9 * scale),
                            OuterSchedule(cycles=60, kkt=1e-7 * scale, lam=lam), zeros)
    print(f"[defect] 1.2 G: converged in {len(out.records)} cycles ({out.exact_applies} exact, "
          f"{out.cheap_applies} cheap applications); objective {out.last.objective:.12e} vs direct "
          f"{direct.last.objective:.12e}", flush=True)
    assert out.stop == "converged"
    assert out.last.kkt <= 1e-7 * scale
    assert abs(out.last.objective - direct.last.objective) <= 1e-9 * p["q0"]
    assert float(out.x.min()) >= 0.0
    assert out.exact_applies == len(out.records) > 1
    alone = X.solve(gram_wrong, p["d"], p["q0"], X.L1Prior(lam, 1.0), np.zeros(nx * nx),
                    step=1.0 / (1.3 * p["lmax"]), max_iterations=20000, kkt_tolerance=1e-9 * scale)
    diff = float(np.max(np.abs(alone["x"] - host(direct.x).reshape(-1))))
    assert diff > 1e-2 * float(np.max(np.abs(host(direct.x))))


def test_defect_errors(problem):
    from rang import gram
    from rang.prox_solver import L1Prior, OuterSchedule, Problem, Settings, defect_correction

    p, nx = problem, problem["nx"]
    zeros = torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0")
    pr = Problem(p["dirty"], p["q0"])
    with pytest.raises(Exception, match="step must be positive"):
        defect_correction(p["m"], p["cheap"], pr, L1Prior(0.1, 1.0), Settings(step=0.0),
                          OuterSchedule(cycles=2, kkt=0.0, lam=0.1), zeros)
    # a cheap operator of another side is refused before any work
    o2 = list(p["opts"])
    o2[0] = o2[1] = 2 * nx
    other = gram.GramOperator.no_w(o2, dev(p["uvw"]), dev(p["freq"]), p["m"].wgt, 1e-9, single=False)
    with pytest.raises(Exception, match="side differs"):
        defect_correction(p["m"], other, pr, L1Prior(0.1, 1.0), Settings(step=1.0),
                          OuterSchedule(cycles=2, kkt=0.0, lam=0.1), zeros)
    single = gram.GramOperator.no_w(p["opts"], dev(p["uvw"]), dev(p["freq"]), p["m"].wgt.float(), 1e-5,
                                    single=True)
    with pytest.raises(ValueError, match="precision"):
        defect_correction(p["m"], single, pr, L1Prior(0.1, 1.0), Settings(step=1.0),
                          OuterSchedule(cycles=2, kkt=0.0, lam=0.1), zeros)
