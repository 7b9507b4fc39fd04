"""The normal operator without visibilities on the GPU (rang.gram, the
C ABI of include/rang_core_gram.h) against the oracle's direct-sum Gram
and the exact normal operator, and its inverse filter as a CG
preconditioner.

* The reference's test (core/src/gridder_gram.rs:825-846): on its xorshift
  case (32^2, 800 rows, w up to 400 and 2000 wavelengths, with and without
  flips) the operator of each expansion (6, 1.2), (8, 1.2), (10, 1.0),
  (12, 0.9) is within (sup_error + Lebesgue 1e-9) |m|_1 of the direct sum,
  its node spectra real (imaginary / real below 1e-9), and the finest
  within 1e-5.
* Symmetric (<a, G b> = <b, G a> to 1e-9) and the single-precision operator
  within 1e-3 (:848-864).
* Without w (w = 0), the one-node operator equals the exact normal operator
  A^H W A of a prepared plan to 1e-9 of its peak (f64, epsilon 1e-12).
* Preconditioned CG with the inverse filter reaches the tolerance of plain
  CG in fewer Gram applications, and its residual, recomputed by the exact
  operator, meets the tolerance."""

import numpy as np
import pytest
from gram_cases import gram_case, model_of

from oracle import gram as OG

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def opts(n, cell, eps, flip=False, wgrid=True):
    return [n, n, cell, cell, eps, 0, 0, 1.1, 2.6, int(wgrid), 0, 0, int(flip), int(flip), 1, -1, 0, 0]


def dev(c, dtype=torch.float64):
    d = torch.device("cuda:0")
    return (torch.from_numpy(c["uvw"]).to(d), torch.from_numpy(c["freq"]).to(d),
            torch.from_numpy(c["wgt"]).to(d).to(dtype))


@pytest.mark.parametrize("flip,wmax", [(False, 400.0), (True, 2000.0)])
def test_matches_the_direct_sum_within_the_bound(flip, wmax):
    from rang import gram

    print(f"[gram] direct sum, flip {flip}, wmax {wmax}", flush=True)
    n = 32
    c = gram_case(n, 800, wmax, flip)
    m = model_of(n)
    want = OG.gram_direct(c["uvw"], c["freq"], c["wgt"], m, c["cell"], flip_v=flip, flip_w=flip)
    uvw, freq, wgt = dev(c)
    x = torch.from_numpy(m).cuda()
    l1 = np.abs(m).sum()
    last = None
    for s, theta in [(6, 1.2), (8, 1.2), (10, 1.0), (12, 0.9)]:
        ex = gram.WExpansion(s, theta, 1e-13)
        op = gram.GramOperator.new(opts(n, c["cell"], 1e-10, flip), uvw, freq, wgt, ex, 1e-10, single=False)
        got = op.apply(x).cpu().numpy()
        err = np.abs(got - want).max()
        bound = (ex.sup_error + ex.lebesgue * 1e-9) * l1
        st = op.stats()
        print(f"[gram] s {s} theta {theta}: error {err:.2e} bound {bound:.2e} planes {st['planes']:.0f}",
              flush=True)
        assert err <= bound
        assert st["hermitian_check"] < 1e-9
        last = err
    assert last < 1e-5


def test_symmetric_and_single_precision_close():
    from rang import gram

    n = 32
    c = gram_case(n, 500, 800.0, True)
    uvw, freq, wgt = dev(c)
    ex = gram.WExpansion(10, 1.0, 1e-13)
    op = gram.GramOperator.new(opts(n, c["cell"], 1e-10, True), uvw, freq, wgt, ex, 1e-10, single=False)
    a = torch.from_numpy(model_of(n)).cuda()
    b = torch.zeros((n, n), dtype=torch.float64, device="cuda:0")
    b.view(-1)[100], b.view(-1)[700] = 2.0, -1.0
    ga, gb = op.apply(a), op.apply(b)
    assert abs(float((a * gb).sum() - (b * ga).sum())) < 1e-9
    op32 = gram.GramOperator.new(opts(n, c["cell"], 1e-5, True), uvw, freq, wgt.float(), ex, 1e-5, single=True)
    g32 = op32.apply(a.float()).double()
    assert float((ga - g32).abs().max()) < 1e-3


def test_without_w_equals_the_exact_normal_operator():
    from rang import gram
    from rang.prepared import Prepared

    n = 48
    c = gram_case(n, 700, 0.0, False)
    uvw, freq, wgt = dev(c)
    o = opts(n, c["cell"], 1e-12, wgrid=False)
    op = gram.GramOperator.no_w(o, uvw, freq, wgt, 1e-12, single=False)
    plan = Prepared(o, uvw, freq, wgt, single=False)
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal((n, n))).cuda()
    got, ref = op.apply(x), plan.gram(x)
    err = float((got - ref).abs().max() / ref.abs().max())
    print(f"[gram] no-w operator vs exact normal operator: {err:.2e}", flush=True)
    assert err < 1e-9


@pytest.mark.parametrize("single", [False, True])
def test_inverse_filter_preconditions_cg(single):
    from rang import gram, synth
    from rang.measurement import WgridMeasurement

    case = synth.imaging_case(32, 8, 1.3e9, 1.45e9, 128, 8.0, seed=77, dtype=np.complex128)
    d = torch.device("cuda:0")
    uvw = torch.from_numpy(case["uvw"]).to(d)
    freq = torch.from_numpy(case["freq"]).to(d)
    vis = torch.from_numpy(case["vis"]).to(d)
    wgt = torch.ones(vis.shape, dtype=torch.float64, device=d)
    eps = 1e-5 if single else 1e-9
    meas = WgridMeasurement(uvw, freq, vis, wgt, 128, case["pixsize"], eps, single=single)
    dirty, _ = meas.residual(torch.zeros((128, 128), dtype=meas.wgt.dtype, device=d))
    o = opts(128, case["pixsize"], eps)
    psf = gram.GramOperator.no_w(o, uvw, freq, meas.wgt, eps, single=single)
    peak = psf.stats()["spectral_peak"]
    tol = 1e-3
    x_cg, n_cg, rel_cg = meas.plan.conjugate_gradients(dirty, 3000, tol)
    best = None
    for rho_rel in (1e-2, 1e-3, 1e-4):
        m = psf.inverse_filter(rho_rel * peak)
        x, n_p, rel = meas.plan.preconditioned_cg(m, dirty, 3000, tol)
        true_rel = float((dirty - meas.plan.gram(x)).norm() / dirty.norm())
        print(f"[gram] single {single}: CG {n_cg} ({rel_cg:.1e}); PCG rho {rho_rel:g}: {n_p} ({rel:.1e}, "
              f"exact {true_rel:.1e})", flush=True)
        assert true_rel < 2 * tol
        best = n_p if best is None else min(best, n_p)
    assert rel_cg <= tol and best < n_cg


def test_nodes_in_several_passes_give_the_same_operator(monkeypatch):
    # the applications process the w-nodes in passes whose scratch fits a
    # budget; a budget of one node per pass must give the one-pass result
    from rang import gram

    n = 32
    c = gram_case(n, 800, 2000.0, False)
    uvw, freq, wgt = dev(c)
    x = torch.from_numpy(model_of(n)).cuda()
    ex = gram.WExpansion(10, 1.0, 1e-13)
    one = gram.GramOperator.new(opts(n, c["cell"], 1e-10), uvw, freq, wgt, ex, 1e-10, single=False)
    want = one.apply(x).cpu().numpy()
    assert one.stats()["planes"] > 2
    monkeypatch.setenv("RANG_GRAM_PASS_BYTES", "1")
    many = gram.GramOperator.new(opts(n, c["cell"], 1e-10), uvw, freq, wgt, ex, 1e-10, single=False)
    got = many.apply(x).cpu().numpy()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("single", [False, True])
def test_transposed_column_pass_gives_the_same_operator(monkeypatch, single):
    # large images transpose the half-transformed slices around the column
    # pass (whole columns read and written); forced on a small one, in one and
    # in several passes, it must give the in-place form's result
    from rang import gram

    n = 64
    c = gram_case(n, 800, 2000.0, False)
    dtype = torch.float32 if single else torch.float64
    uvw, freq, wgt = dev(c, dtype)
    x = torch.from_numpy(model_of(n)).cuda().to(dtype)
    ex = gram.WExpansion(10, 1.0, 1e-13)
    eps = 1e-5 if single else 1e-10
    monkeypatch.setenv("RANG_GRAM_TRANSPOSE", "0")
    plain = gram.GramOperator.new(opts(n, c["cell"], eps), uvw, freq, wgt, ex, eps, single=single)
    want = plain.apply(x).cpu().numpy()
    monkeypatch.setenv("RANG_GRAM_TRANSPOSE", "1")
    for budget in (None, "1"):
        if budget:
            monkeypatch.setenv("RANG_GRAM_PASS_BYTES", budget)
        op = gram.GramOperator.new(opts(n, c["cell"], eps), uvw, freq, wgt, ex, eps, single=single)
        got = op.apply(x).cpu().numpy()
        print(f"[gram] transposed column pass, single {single}, budget {budget}: "
              f"{np.abs(got - want).max() / np.abs(want).max():.2e}", flush=True)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
