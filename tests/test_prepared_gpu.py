"""Prepared plans, the Gram operator and conjugate gradients on the GPU
against the oracle (rang-core's `wgridder::Prepared`, core/src/wgridder.rs:
3165-3499; `WgridMeasurement`, core/src/clean_certified.rs:97-236;
`clean::conjugate_gradients`, core/src/clean.rs:263-306).

The oracle side composes the oracle's one-shot calls as the reference's
Prepared does: prediction = dirty2vis without weights at the active
(nonzero-weight) visibilities, images = vis2dirty with the weights. Image
tolerances as the one-shot parity tests (max |difference| over the peak):
1e-11 in double, 2e-5 in single precision; the misfit is a sum of squares
of those visibilities, relative 1e-9 in double and 1e-4 in single.
"""

import numpy as np
import pytest
from conftest import ref_case, rel_linf

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = {False: 1e-11, True: 2e-5}
TOL_Q = {False: 1e-9, True: 1e-4}


def case(single, seed=3, zero_frac=0.2, nrow=60, nchan=3):
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=48, ny=40, nrow=nrow, seed=seed, nchan=nchan)
    rng = np.random.default_rng(seed + 11)
    wgt = wgt * (rng.uniform(size=wgt.shape) >= zero_frac)
    cdt, rdt = (np.complex64, np.float32) if single else (np.complex128, np.float64)
    return uvw, freq, vis.astype(cdt), wgt.astype(rdt), nx, ny, dx


def opts(nx, ny, dx, eps):
    return [nx, ny, dx, dx, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def host(t):
    return t.detach().cpu().numpy()


def oracle_predict(uvw, freq, img, wgt, dx, eps):
    # A img at the active visibilities, no weights
    v, _ = O.dirty2vis(uvw, freq, img, None, (wgt != 0).astype(np.uint8), pixsize_x=dx, pixsize_y=dx,
                       epsilon=eps)
    return v


def oracle_image(uvw, freq, vis, wgt, nx, ny, dx, eps, single):
    imgs, _ = O.vis2dirty(uvw, freq, vis, wgt, None, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx,
                          epsilon=eps, single=single)
    return imgs["dirty"]


@pytest.mark.parametrize("single", [False, True])
def test_prepared_images_and_predictions(single):
    from rang.prepared import Prepared

    uvw, freq, vis, wgt, nx, ny, dx = case(single)
    eps = 1e-5 if single else 1e-9
    p = Prepared(opts(nx, ny, dx, eps), dev(uvw), dev(freq), dev(wgt), single=single)
    assert p.active() == int((wgt != 0).sum())
    # Prepared::vis2dirty: the image with the prepared weights; zero-weight
    # visibilities add nothing whatever their value
    got = host(p.vis2dirty(dev(vis)))
    ref = oracle_image(uvw, freq, vis, wgt, nx, ny, dx, eps, single)
    assert rel_linf(got, ref) < TOL[single]
    # Prepared::dirty2vis, without and with the weights
    rng = np.random.default_rng(5)
    img = rng.standard_normal((nx, ny)).astype(np.float32 if single else np.float64)
    pv = host(p.dirty2vis(dev(img)))
    rv = oracle_predict(uvw, freq, img, wgt, dx, eps)
    assert rel_linf(pv, rv) < TOL[single]
    assert np.all(pv[wgt == 0] == 0)
    pw = host(p.dirty2vis(dev(img), weighted=True))
    assert rel_linf(pw, rv * wgt) < TOL[single]
    # the same plan again after the grids were released
    p.release_grids()
    assert rel_linf(host(p.vis2dirty(dev(vis))), ref) < TOL[single]
    assert p.memory() > 0
    p.close()


@pytest.mark.parametrize("single", [False, True])
def test_residual_and_misfit(single):
    from rang.prepared import Prepared

    uvw, freq, vis, wgt, nx, ny, dx = case(single, seed=7)
    eps = 1e-5 if single else 1e-9
    rng = np.random.default_rng(9)
    model = (rng.standard_normal((nx, ny)) * (rng.uniform(size=(nx, ny)) < 0.05)).astype(
        np.float32 if single else np.float64)
    p = Prepared(opts(nx, ny, dx, eps), dev(uvw), dev(freq), dev(wgt), single=single)
    img, q = p.residual(dev(vis), dev(model))
    r = vis - oracle_predict(uvw, freq, model, wgt, dx, eps).astype(vis.dtype)
    q_ref = float(np.sum(wgt.astype(np.float64) * np.abs(r.astype(np.complex128)) ** 2))
    ref = oracle_image(uvw, freq, r, wgt, nx, ny, dx, eps, single)
    assert rel_linf(host(img), ref) < TOL[single]
    assert abs(q - q_ref) <= TOL_Q[single] * q_ref
    # without a model: the data alone
    img0, q0 = p.residual(dev(vis))
    q0_ref = float(np.sum(wgt.astype(np.float64) * np.abs(vis.astype(np.complex128)) ** 2))
    assert rel_linf(host(img0), oracle_image(uvw, freq, vis, wgt, nx, ny, dx, eps, single)) < TOL[single]
    assert abs(q0 - q0_ref) <= TOL_Q[single] * q0_ref


@pytest.mark.parametrize("single", [False, True])
def test_gram_apply_is_image_of_prediction(single):
    from rang.prepared import Prepared

    uvw, freq, vis, wgt, nx, ny, dx = case(single, seed=13)
    eps = 1e-5 if single else 1e-9
    x = np.random.default_rng(1).standard_normal((nx, ny)).astype(np.float32 if single else np.float64)
    p = Prepared(opts(nx, ny, dx, eps), dev(uvw), dev(freq), dev(wgt), single=single)
    got = host(p.gram(dev(x)))
    pred = oracle_predict(uvw, freq, x, wgt, dx, eps)
    ref = oracle_image(uvw, freq, pred.astype(vis.dtype), wgt, nx, ny, dx, eps, single)
    assert rel_linf(got, ref) < 10 * TOL[single]
    # symmetric positive semidefinite: <y, B x> = <B y, x>, <x, B x> >= 0
    y = np.random.default_rng(2).standard_normal((nx, ny)).astype(x.dtype)
    bx, by = got.astype(np.float64), host(p.gram(dev(y))).astype(np.float64)
    lhs, rhs = float(np.sum(y * bx)), float(np.sum(by * x))
    assert abs(lhs - rhs) <= (1e-10 if not single else 1e-4) * np.sqrt(np.sum(bx**2) * np.sum(y**2))
    assert float(np.sum(x * bx)) >= 0.0


def numpy_cg(apply, rhs, iterations, tolerance):
    # clean::conjugate_gradients (core/src/clean.rs:263-306) on the oracle
    x = np.zeros_like(rhs)
    r = rhs.copy()
    p = r.copy()
    rr = float(np.sum(r * r))
    nd = np.sqrt(rr)
    used = 0
    while used < iterations and np.sqrt(rr) > tolerance * nd:
        ap = apply(p)
        pap = float(np.sum(p * ap))
        if pap <= 0:
            break
        a = rr / pap
        x += a * p
        r -= a * ap
        rn = float(np.sum(r * r))
        p = r + (rn / rr) * p
        rr = rn
        used += 1
    return x, used, np.sqrt(rr) / nd


def test_conjugate_gradients_match_the_oracle_solve():
    from rang.measurement import WgridMeasurement, conjugate_gradients, misfit_floor

    uvw, freq, vis, wgt, nx, ny, dx = case(False, seed=21, nrow=40)
    eps = 1e-9
    m = WgridMeasurement(dev(uvw), dev(freq), dev(vis), dev(wgt), nx, dx, eps)
    assert abs(m.weight_sum - float(wgt.sum())) < 1e-9 * m.weight_sum
    wn = wgt / wgt.sum()
    d0, q0 = m.residual(torch.zeros((nx, nx), dtype=torch.float64, device="cuda:0"))
    ref_d0 = oracle_image(uvw, freq, vis, wn, nx, nx, dx, eps, False)
    assert rel_linf(host(d0), ref_d0) < TOL[False]
    assert abs(q0 - 0.5 * float(np.sum(wn * np.abs(vis) ** 2))) < 1e-9 * q0

    def apply(v):
        return oracle_image(uvw, freq, oracle_predict(uvw, freq, v, wn, dx, eps), wn, nx, nx, dx, eps, False)

    for tol in (1e-2, 1e-4):
        x, used, rel = conjugate_gradients(m, d0, 200, tol)
        xr, used_r, rel_r = numpy_cg(apply, ref_d0, 200, tol)
        # the same recurrence on operators equal to 1e-11: the same count of
        # applications (within one where the stopping test is close)
        assert abs(used - used_r) <= 1
        assert rel <= tol or used == 200
        # the solution's residual measured with the oracle's operator
        xh = host(x)
        true_rel = np.linalg.norm(ref_d0 - apply(xh)) / np.linalg.norm(ref_d0)
        assert true_rel <= 1.01 * tol
        if used == used_r:
            assert np.linalg.norm(xh - xr) <= 1e-6 * np.linalg.norm(xr)
    floor, xf = misfit_floor(m, d0, 50, 1e-4)
    assert floor.q_floor <= q0 and floor.iterations >= 1


def test_no_active_visibilities():
    from rang.prepared import Prepared

    uvw, freq, vis, wgt, nx, ny, dx = case(True)
    p = Prepared(opts(nx, ny, dx, 1e-5), dev(uvw), dev(freq), dev(np.zeros_like(wgt)), single=True)
    assert p.active() == 0
    assert not host(p.vis2dirty(dev(vis))).any()
    assert not host(p.dirty2vis(dev(np.ones((nx, ny), np.float32)))).any()
    img, q = p.residual(dev(vis))
    assert q == 0.0 and not host(img).any()


def test_prepared_errors():
    from rang.prepared import Prepared

    uvw, freq, vis, wgt, nx, ny, dx = case(True)
    with pytest.raises(ValueError, match="ascending"):
        Prepared(opts(nx, ny, dx, 1e-5), dev(uvw), dev(freq[::-1].copy()), dev(wgt), single=True)
    with pytest.raises(ValueError, match="lengths disagree"):
        Prepared(opts(nx, ny, dx, 1e-5), dev(uvw), dev(freq), dev(wgt[:-1]), single=True)


@pytest.mark.parametrize("single", [False, True])
def test_residual_sorted_and_sorted(single):
    # Prepared::sorted and ::residual_sorted (core/src/wgridder.rs:3257-3286):
    # the data held only for the active visibilities, in the plan's order,
    # give the residual and misfit of the dense call; 9,000 visibilities on a
    # 48 x 40 image put more than one gridding item in a tile (split tiles)
    from rang.prepared import Prepared

    uvw, freq, vis, wgt, nx, ny, dx = case(single, seed=17, nrow=3000, nchan=3)
    eps = 1e-5 if single else 1e-9
    rng = np.random.default_rng(19)
    model = (rng.standard_normal((nx, ny)) * (rng.uniform(size=(nx, ny)) < 0.05)).astype(
        np.float32 if single else np.float64)
    p = Prepared(opts(nx, ny, dx, eps), dev(uvw), dev(freq), dev(wgt), single=single)
    sw = host(p.sorted(dev(wgt)))
    assert sw.size == p.active() == int((wgt != 0).sum())
    assert np.array_equal(np.sort(sw), np.sort(wgt[wgt != 0]))
    sv = p.sorted(dev(vis))
    assert np.array_equal(np.sort_complex(host(sv)), np.sort_complex(vis[wgt != 0]))
    for m in (dev(model), None):
        img, q = p.residual(dev(vis), m)
        img_s, q_s = p.residual_sorted(sv, m)
        assert rel_linf(host(img_s), host(img)) < TOL[single]
        assert abs(q_s - q) <= 1e-12 * q
    r = vis - oracle_predict(uvw, freq, model, wgt, dx, eps).astype(vis.dtype)
    ref = oracle_image(uvw, freq, r, wgt, nx, ny, dx, eps, single)
    img, q = p.residual(dev(vis), dev(model))
    assert rel_linf(host(img), ref) < TOL[single]
    q_ref = float(np.sum(wgt.astype(np.float64) * np.abs(r.astype(np.complex128)) ** 2))
    assert abs(q - q_ref) <= TOL_Q[single] * q_ref
    with pytest.raises(ValueError, match="disagree"):
        p.residual_sorted(dev(vis), None)


@pytest.mark.parametrize("single", [False, True])
def test_graph_replays_match_eager_calls(single):
    # A prepared call's first run is eager, its second is captured into a
    # graph and launched, later ones replay it; new contents of the same
    # buffers, and other eager calls in between, must show in every replay.
    from rang.prepared import Prepared

    uvw, freq, vis, wgt, nx, ny, dx = case(single, seed=23, nrow=400)
    eps = 1e-5 if single else 1e-9
    rdt = np.float32 if single else np.float64
    rng = np.random.default_rng(4)
    img1, img2 = (rng.standard_normal((nx, ny)).astype(rdt) for _ in range(2))
    p = Prepared(opts(nx, ny, dx, eps), dev(uvw), dev(freq), dev(wgt), single=single)
    x = dev(img1)
    out = torch.empty((nx, ny), dtype=x.dtype, device=x.device)

    def ref_gram(img):
        pred = oracle_predict(uvw, freq, img, wgt, dx, eps)
        return oracle_image(uvw, freq, pred.astype(vis.dtype), wgt, nx, ny, dx, eps, single)

    g1, g2 = ref_gram(img1), ref_gram(img2)
    got = []
    for k in range(6):
        x.copy_(dev(img1 if k % 2 == 0 else img2))
        got.append(host(p.gram(x, out=out)).copy())
        # other work between the replays: an eager call of another key and torch's own
        p.vis2dirty(dev(vis))
        torch.cuda.synchronize()
    for k, gk in enumerate(got):
        assert rel_linf(gk, g1 if k % 2 == 0 else g2) < 10 * TOL[single], k
    # the image direction and the residual the same way
    v = dev(vis)
    d = torch.empty((nx, ny), dtype=x.dtype, device=x.device)
    ref = oracle_image(uvw, freq, vis, wgt, nx, ny, dx, eps, single)
    for k in range(4):
        assert rel_linf(host(p.vis2dirty(v, out=d)), ref) < TOL[single], k
    for k in range(4):
        img, q = p.residual(v, x, out=d)
        img0, q0 = p.residual(v, x)
        assert rel_linf(host(img), host(img0)) < TOL[single] and abs(q - q0) <= 1e-12 * q0, k
