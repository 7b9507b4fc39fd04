"""FISTA and power iteration on the GPU's prepared normal operator against
the oracle (rang-core's `prox_solver::solve` and `power_iteration`,
core/src/prox_solver.rs:283-445, 886-908, restated in oracle/prox.py) driven
by the oracle's own Gram operator (vis2dirty of dirty2vis with the weights
normalised to sum to one, core/src/clean_certified.rs:117-150, 233-236).

The two operators agree to about 1e-11 in double precision, so: the power
iteration's estimate to 1e-9 relative; per-iterate objectives, misfits and
gaps of a fixed-length FISTA run to 1e-8 relative (of the first objective);
the final images to 1e-6 of their peak; and the solver's own checks of the
reference's tests (:1078-1133): the misfit it reports equals the one computed
directly from the final image, the gap is not negative, forward-backward
descends.
"""

import math

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
    from rang.measurement import WgridMeasurement

    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=32, ny=32, nrow=40, seed=17, nchan=3)
    eps = 1e-9
    m = WgridMeasurement(dev(uvw), dev(freq), dev(vis), dev(wgt), nx, dx, eps)
    wn = wgt / wgt.sum()

    def gram(x):
        img = x.reshape(nx, nx)
        v, _ = O.dirty2vis(uvw, freq, img, None, (wn != 0).astype(np.uint8), pixsize_x=dx, pixsize_y=dx,
                           epsilon=eps)
        out, _ = O.vis2dirty(uvw, freq, v, wn, None, npix_x=nx, npix_y=nx, pixsize_x=dx, pixsize_y=dx,
                             epsilon=eps, single=False)
        return out["dirty"].reshape(-1)

    d_img, _ = O.vis2dirty(uvw, freq, vis, wn, None, npix_x=nx, npix_y=nx, pixsize_x=dx, pixsize_y=dx,
                           epsilon=eps, single=False)
    d = d_img["dirty"].reshape(-1)
    q0 = 0.5 * float(np.sum(wn * np.abs(vis) ** 2))
    return dict(m=m, gram=gram, d=d, q0=q0, nx=nx)


def test_power_iteration_matches_the_oracle(problem):
    from rang.prox_solver import power_iteration

    lam, used = power_iteration(problem["m"], 3, 1e-8, 500)
    lam_r, used_r = X.power_iteration(problem["gram"], problem["nx"] ** 2, 3, 1e-8, 500)
    assert abs(lam - lam_r) <= 1e-9 * lam_r
    assert abs(used - used_r) <= 1


def test_fista_iterates_match_the_oracle(problem):
    from rang.prox_solver import L1Prior, Problem, Settings, solve

    p, nx = problem, problem["nx"]
    lam_r, _ = X.power_iteration(p["gram"], nx * nx, 3, 1e-8, 500)
    step = 1.0 / (1.01 * lam_r)
    lam1 = 0.05 * float(np.max(p["d"]))
    dirty, q0 = p["m"].residual(torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    assert abs(q0 - p["q0"]) <= 1e-10 * p["q0"]
    assert np.max(np.abs(host(dirty).reshape(-1) - p["d"])) <= 1e-11 * np.max(np.abs(p["d"]))
    its = 40
    out = solve(p["m"], Problem(dirty, q0), L1Prior(lam1, 1.0), Settings(step=step, max_iterations=its),
                torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    ref = X.solve(p["gram"], p["d"], p["q0"], X.L1Prior(lam1, 1.0), np.zeros(nx * nx), step=step,
                  max_iterations=its)
    assert out.stop == ref["stop"] == "iterations"
    assert out.iterations == ref["iterations"] == its
    assert len(out.records) == its + 1
    # the reference's records, iterate by iterate
    scale = abs(out.records[0].objective)
    got_rec = []
    X.solve(p["gram"], p["d"], p["q0"], X.L1Prior(lam1, 1.0), np.zeros(nx * nx), step=step,
            max_iterations=its, observer=lambda r, x: got_rec.append(r))
    for g, r in zip(out.records, got_rec):
        assert g.iteration == r["iteration"]
        for k in ("objective", "misfit", "penalty", "gap"):
            assert abs(getattr(g, k) - r[k]) <= 1e-8 * scale, (g.iteration, k)
        assert g.restarted == r["restarted"]
        assert g.nonzero == r["nonzero"]
    assert out.restarts == ref["restarts"]
    x = host(out.x).reshape(-1)
    assert np.max(np.abs(x - ref["x"])) <= 1e-6 * np.max(np.abs(ref["x"]))
    # the linearity trick: G x from the run equals the operator applied to x
    gx_direct = host(p["m"].gram_apply(out.x)).reshape(-1)
    assert np.max(np.abs(host(out.gx).reshape(-1) - gx_direct)) <= 1e-9 * np.max(np.abs(gx_direct))


def test_fista_closes_the_gap_and_forward_backward_descends(problem):
    from rang.prox_solver import L1Prior, Problem, Settings, power_iteration, solve

    p, nx = problem, problem["nx"]
    lmax, _ = power_iteration(p["m"], 3, 1e-8, 500)
    dirty, q0 = p["m"].residual(torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    lam1 = 0.05 * float(torch.max(dirty))
    st = Settings(step=1.0 / (1.01 * lmax), max_iterations=3000, gap_tolerance=1e-10 * q0)
    out = solve(p["m"], Problem(dirty, q0), L1Prior(lam1, 1.0), st,
                torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    assert out.stop == "converged"
    assert out.last.gap >= -1e-12 * q0
    # the reported misfit is the one computed directly at the result
    x = out.x
    gx = p["m"].gram_apply(x)
    direct = q0 - float((dirty * x).sum()) + 0.5 * float((x * gx).sum())
    assert abs(direct - out.last.misfit) <= 1e-10 * q0
    assert float(x.min()) >= 0.0
    assert out.min_curvature >= -1e-10 * lmax and out.max_curvature <= lmax * (1 + 1e-6)
    fb = solve(p["m"], Problem(dirty, q0), L1Prior(lam1, 1.0),
               Settings(step=st.step, max_iterations=3000, momentum=False, gap_tolerance=1e-8 * q0),
               torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    assert fb.largest_rise <= 1e-12 * q0
    assert fb.last.objective >= out.last.objective - 1e-8 * q0
    # signed prior: its objective is at most the non-negative one
    sg = solve(p["m"], Problem(dirty, q0), L1Prior(lam1, 1.0, signed=True), st,
               torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    assert sg.last.objective <= out.last.objective + 1e-10 * q0


def test_fista_errors(problem):
    from rang.prox_solver import L1Prior, Problem, Settings, solve

    nx = problem["nx"]
    dirty, q0 = problem["m"].residual(torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    with pytest.raises(Exception, match="step must be positive"):
        solve(problem["m"], Problem(dirty, q0), L1Prior(0.1, 1.0), Settings(step=0.0),
              torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    assert math.isfinite(q0)
