"""Parity of the HIP gridder with the oracle and the FP64 direct transform,
through the C ABI (rang._core / rang.core_wgridder) on the GPU.

Tolerances (max |difference| over the peak of the reference): against the
oracle, which grids with the same kernel, grid and planes, 1e-11 in double
and 2e-5 in single precision (the two differ by summation order and FFT
rounding only, amplified towards the image edges by the correction
factor); against the direct transform, the reference's own gates
(1e-8 at epsilon 1e-9 in double, 1e-3 at 1e-4 in single) and 10 epsilon at
the workload's epsilon 1e-5."""

import numpy as np
import pytest
from conftest import C_LIGHT, lcg_case, ref_case, rel_linf

from oracle import oracle as O
from rang import core_wgridder as W

pytestmark = pytest.mark.gpu

TOL_ORACLE = {False: 1e-11, True: 2e-5}


def v2d(uvw, freq, vis, wgt, nx, ny, dx, eps, single, mask=None, **kw):
    vis = vis.astype(np.complex64 if single else np.complex128)
    wgt = None if wgt is None else wgt.astype(np.float32 if single else np.float64)
    got, info = W.vis2dirty(uvw=uvw, freq=freq, vis=vis, wgt=wgt, mask=mask, npix_x=nx, npix_y=ny,
                            pixsize_x=dx, pixsize_y=dx, epsilon=eps, return_info=True, **kw)
    ref, rinfo = O.vis2dirty(uvw, freq, vis, wgt, mask, npix_x=nx, npix_y=ny, pixsize_x=dx,
                             pixsize_y=dx, epsilon=eps, single=single, **kw)
    return got, ref["dirty"], info, rinfo


def d2v(uvw, freq, img, wgt, dx, eps, single, mask=None, **kw):
    img = img.astype(np.float32 if single else np.float64)
    wgt = None if wgt is None else wgt.astype(np.float32 if single else np.float64)
    got, info = W.dirty2vis(uvw=uvw, freq=freq, dirty=img, wgt=wgt, mask=mask, pixsize_x=dx,
                            pixsize_y=dx, epsilon=eps, return_info=True, **kw)
    ref, rinfo = O.dirty2vis(uvw, freq, img, wgt, mask, pixsize_x=dx, pixsize_y=dx, epsilon=eps, **kw)
    return got, ref, info, rinfo


def same_plan(info, rinfo):
    keys = ("kernel_index", "supp", "nu", "nv", "nplanes", "nvis")
    assert [info[k] for k in keys] == [int(v) for v in rinfo[:6]]
    assert np.allclose([info[k] for k in W.INFO_NAMES[6:]], rinfo[6:], rtol=0, atol=0)


def test_vis2dirty_matches_dft_both_precisions():
    # rang-core tests/test_core_wgridder.py:75-95
    uvw, freq, vis, wgt, nx, ny, dx = ref_case()
    ref = O.dft_dirty(uvw, freq, vis, wgt, nx, ny, dx, dx)
    d64, info = W.vis2dirty(uvw=uvw, freq=freq, vis=vis, wgt=wgt, npix_x=nx, npix_y=ny,
                            pixsize_x=dx, pixsize_y=dx, epsilon=1e-9, return_info=True)
    assert d64.dtype == np.float64 and rel_linf(d64, ref) < 1e-8
    assert info["nvis"] == vis.size and info["nplanes"] > info["supp"]
    d32 = W.vis2dirty(uvw=uvw, freq=freq, vis=vis.astype(np.complex64), wgt=wgt.astype(np.float32),
                      npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx, epsilon=1e-4)
    assert d32.dtype == np.float32 and rel_linf(d32, ref) < 1e-3


@pytest.mark.parametrize("nx,ny,field,wmax", [(32, 32, 0.2, 500.0), (32, 16, 0.05, 50.0)])
def test_lcg_cases_match_dft(nx, ny, field, wmax):
    # core/src/wgridder.rs:4667-4684, 4779-4796
    uvw, freq, vis, wgt, dx = lcg_case(nx, ny, field, wmax, 60, 7)
    ref = O.dft_dirty(uvw, freq, vis, wgt, nx, ny, dx, dx)
    got, oref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, 1e-9, False)
    assert rel_linf(got, ref) < 1e-8
    assert rel_linf(got, oref) < TOL_ORACLE[False]
    same_plan(info, rinfo)
    got, oref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, 1e-4, True)
    assert rel_linf(got, ref) < 1e-3
    assert rel_linf(got, oref) < TOL_ORACLE[True]


def test_adjoint_and_dirty2vis():
    # rang-core tests/test_core_wgridder.py:98-125
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(seed=8)
    x = np.random.default_rng(1).standard_normal((nx, ny))
    common = dict(uvw=uvw, freq=freq, wgt=wgt, pixsize_x=dx, pixsize_y=dx, epsilon=1e-9)
    ax = W.dirty2vis(dirty=x, **common)
    aty = W.vis2dirty(vis=vis, npix_x=nx, npix_y=ny, **common)
    lhs = np.real(np.vdot(ax, vis))
    rhs = np.sum(x * aty)
    assert abs(lhs - rhs) / (np.linalg.norm(ax) * np.linalg.norm(vis)) < 1e-12
    img = np.zeros((nx, ny))
    img[5, 20] = 2.0
    got = W.dirty2vis(dirty=img, **common)
    l, m = (5 - nx // 2) * dx, (20 - ny // 2) * dx
    n1 = np.sqrt(1 - l * l - m * m) - 1
    u, v, w = (uvw[:, i, None] * freq / C_LIGHT for i in range(3))
    ref = 2.0 * wgt * np.exp(-2j * np.pi * (u * l + v * m + w * n1))
    assert rel_linf(got, ref) < 1e-8


def test_no_kernel_in_a_narrow_sigma_range():
    # single precision reaches epsilon 1e-5 only above sigma 1.6: refused as
    # the reference refuses (core/src/wgridder.rs:154-181)
    uvw, freq, vis, wgt, nx, ny, dx = ref_case()
    with pytest.raises(ValueError, match="no appropriate kernel"):
        W.vis2dirty(uvw=uvw, freq=freq, vis=vis.astype(np.complex64), npix_x=nx, npix_y=ny,
                    pixsize_x=dx, pixsize_y=dx, epsilon=1e-5, sigma_min=1.5, sigma_max=1.6)
    with pytest.raises(ValueError, match="no appropriate kernel"):
        O.vis2dirty(uvw, freq, vis.astype(np.complex64), None, npix_x=nx, npix_y=ny, pixsize_x=dx,
                    pixsize_y=dx, epsilon=1e-5, single=True, sigma_min=1.5, sigma_max=1.6)


def test_zero_weights_and_bad_input():
    # rang-core tests/test_core_wgridder.py:128-147
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nrow=6)
    common = dict(uvw=uvw, freq=freq, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx)
    empty, info = W.vis2dirty(vis=vis, wgt=np.zeros_like(wgt), epsilon=1e-6, return_info=True, **common)
    assert not empty.any() and info["nvis"] == 0
    with pytest.raises(ValueError, match="ascending"):
        W.vis2dirty(vis=vis, wgt=wgt, epsilon=1e-6, **{**common, "freq": freq[::-1]})
    with pytest.raises(ValueError, match="kernel"):
        W.vis2dirty(vis=vis, wgt=wgt, epsilon=1e-14, **common)
    with pytest.raises(ValueError, match="even"):
        W.vis2dirty(vis=vis, wgt=wgt, epsilon=1e-6, **{**common, "npix_x": 31})


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("case", [
    dict(nx=32, ny=32, field=0.2, wmax=600.0, nrow=40, seed=3),
    dict(nx=64, ny=48, field=0.3, wmax=2000.0, nrow=200, seed=4, nchan=5),
    dict(nx=128, ny=96, field=0.05, wmax=20.0, nrow=300, seed=5, nchan=4),
    dict(nx=16, ny=16, field=0.6, wmax=3000.0, nrow=7, seed=6, nchan=3),
    # 12,000 visibilities on a few tiles: tiles split over several gridding
    # items, merged with global atomics into cells cleared first
    dict(nx=48, ny=40, field=0.15, wmax=300.0, nrow=4000, seed=21, nchan=3),
])
def test_matches_oracle(case, single):
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(**case)
    eps = 1e-5 if single else 1e-10
    got, ref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, eps, single)
    same_plan(info, rinfo)
    assert rel_linf(got, ref) < TOL_ORACLE[single]
    x = np.random.default_rng(7).standard_normal((nx, ny))
    got, ref, info, rinfo = d2v(uvw, freq, x, wgt, dx, eps, single)
    same_plan(info, rinfo)
    assert rel_linf(got, ref) < TOL_ORACLE[single]


OPTION_CASES = [
    dict(center_x=0.01, center_y=-0.02),
    dict(flip_u=True, flip_v=True),
    dict(flip_w=True),
    dict(divide_by_n=True),
    dict(allow_nshift=False),
    dict(do_wgridding=False),
    dict(center_x=0.03, flip_u=True, flip_w=True, divide_by_n=True),
    dict(sigma_min=1.8, sigma_max=2.0),
]


@pytest.mark.parametrize("opts", OPTION_CASES)
@pytest.mark.parametrize("single", [False, True])
def test_options_match_oracle(opts, single):
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=48, ny=40, field=0.15, wmax=300.0, nrow=60, seed=9)
    eps = 1e-5 if single else 1e-10
    got, ref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, eps, single, **opts)
    same_plan(info, rinfo)
    assert rel_linf(got, ref) < TOL_ORACLE[single]
    x = np.random.default_rng(8).standard_normal((nx, ny))
    got, ref, info, rinfo = d2v(uvw, freq, x, wgt, dx, eps, single, **opts)
    assert rel_linf(got, ref) < TOL_ORACLE[single]


def test_explicit_kernel_and_wide_support():
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=32, ny=32, field=0.3, wmax=1500.0, nrow=50, seed=10)
    # f64 at epsilon 1e-13: the widest supports of the table
    got, ref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, 1e-13, False)
    assert info["supp"] >= 14
    assert rel_linf(got, ref) < TOL_ORACLE[False]
    # an explicit kernel row and grid (kernel=(row, nu, nv)), grid sizes with
    # factors 7 and 11
    for kernel in [(879, 154, 98), (211, 110, 84)]:
        single = O.kernel_row(kernel[0])[6]
        got, ref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, 1e-5, single, kernel=kernel)
        assert (info["kernel_index"], info["nu"], info["nv"]) == kernel
        assert rel_linf(got, ref) < TOL_ORACLE[single]


def test_more_than_32_planes_in_lds_tiles_of_the_support():
    # f64, an explicit support-16 row on a 16 x 16 grid: the LDS gridder's
    # tile is the support itself (no padding), so more than 32 planes would
    # fit one item; an item holds at most 32 (a staged record keeps its item
    # plane in 5 bits), and every plane must land where the oracle puts it
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=8, ny=8, field=0.3, wmax=1500.0, nrow=20, seed=14)
    got, ref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, 1e-5, False, kernel=(773, 16, 16))
    assert info["supp"] == 16 and info["nplanes"] > 64
    same_plan(info, rinfo)
    assert rel_linf(got, ref) < TOL_ORACLE[False]


def test_masks_and_empty_inputs():
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nrow=30, seed=12)
    mask = (np.random.default_rng(1).uniform(size=vis.shape) > 0.4).astype(np.uint8)
    for single in (False, True):
        eps = 1e-5 if single else 1e-10
        got, ref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, eps, single, mask=mask)
        assert info["nvis"] == int(mask.sum())
        assert rel_linf(got, ref) < TOL_ORACLE[single]
        x = np.random.default_rng(2).standard_normal((nx, ny))
        got, ref, info, rinfo = d2v(uvw, freq, x, wgt, dx, eps, single, mask=mask)
        assert np.all(got[mask == 0] == 0)
        assert rel_linf(got, ref) < TOL_ORACLE[single]
    # all flagged, and no rows at all
    zero = W.vis2dirty(uvw=uvw, freq=freq, vis=vis, mask=np.zeros_like(mask), npix_x=nx, npix_y=ny,
                       pixsize_x=dx, pixsize_y=dx, epsilon=1e-6)
    assert not zero.any()
    none = W.vis2dirty(uvw=uvw[:0], freq=freq, vis=vis[:0], npix_x=nx, npix_y=ny, pixsize_x=dx,
                       pixsize_y=dx, epsilon=1e-6)
    assert none.shape == (nx, ny) and not none.any()
    v0 = W.dirty2vis(uvw=uvw[:0], freq=freq, dirty=np.ones((nx, ny)), pixsize_x=dx, pixsize_y=dx,
                     epsilon=1e-6)
    assert v0.shape == (0, 2)


def test_psf_pass_and_product_weights():
    # rang-core tests/test_core_wgridder.py:215-245, core/src/wgridder.rs:4812-4871
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(seed=21, nrow=50)
    vis64, wgt32 = vis.astype(np.complex64), wgt.astype(np.float32)
    wgt32[3, 1] = 0.0
    common = dict(uvw=uvw, freq=freq, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx, epsilon=1e-7)
    dirty, psf = W.dirty_psf(vis=vis64, wgt=wgt32, **common)
    sep = W.vis2dirty(vis=vis64.astype(np.complex128), wgt=wgt32.astype(np.float64), **common)
    ones = np.where(wgt32 != 0, 1.0, 0.0).astype(np.complex128)
    sep_psf = W.vis2dirty(vis=ones, wgt=wgt32.astype(np.float64), **common)
    # summation order differs between passes (LDS accumulation): within the
    # double-precision oracle tolerance, not bit for bit
    assert rel_linf(dirty, sep) < 1e-11 and rel_linf(psf, sep_psf) < 1e-11
    odir, _ = O.vis2dirty(uvw, freq, vis64, wgt32, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx,
                          epsilon=1e-7, single=False, outputs=("dirty", "psf"))
    assert rel_linf(dirty, odir["dirty"]) < 1e-11 and rel_linf(psf, odir["psf"]) < 1e-11
    b = (0.1 + np.random.default_rng(99).uniform(size=wgt32.shape)).astype(np.float32)
    b.ravel()[::7] = 0.0
    d1, p1 = W.dirty_psf(vis=vis64, wgt=wgt32 * b, **common)
    d2, p2 = W.dirty_psf(vis=vis64, wgt=wgt32, wgt2=b, **common)
    assert rel_linf(d1, d2) < 1e-11 and rel_linf(p1, p2) < 1e-11
    d32, p32 = W.dirty_psf(vis=vis, wgt=wgt, single=True, **{**common, "epsilon": 1e-4})
    assert d32.dtype == np.float32
    ref32 = W.vis2dirty(vis=vis64, wgt=wgt.astype(np.float32), **{**common, "epsilon": 1e-4})
    assert rel_linf(d32, ref32) < 2e-6


def test_fits_image_orientation():
    # fits_image is the rang-convention image with v and w flipped, transposed
    # and scaled to a unit PSF centre (python/rang/core_wgridder.py:335-382)
    rng = np.random.default_rng(4)
    width, scale = 48, 60.0
    freq = np.array([1.2e9, 1.4e9])
    cell = scale * np.pi / 648000.0
    lam = freq.max() / C_LIGHT
    uvw = np.stack([rng.uniform(-0.4, 0.4, 60) / cell / lam, rng.uniform(-0.4, 0.4, 60) / cell / lam,
                    rng.uniform(-100, 100, 60) / lam], axis=1)
    l0, m0 = 7 * cell, -4 * cell
    n0 = np.sqrt(1 - l0**2 - m0**2) - 1
    u, v, w = (uvw[:, i, None] * freq / C_LIGHT for i in range(3))
    data = np.exp(-2j * np.pi * (u * l0 + v * m0 + w * n0)).astype(np.complex64)
    weights = np.ones(data.shape, np.float32)
    got = W.fits_image(uvw, freq, data, weights, width, width, scale, epsilon=1e-7)
    assert got["psf"][width // 2, width // 2] == 1.0
    # the point source at l0, m0: rang pixel (i, j) = (24 + 7, 24 + 4) under
    # flip_v; FITS order transposes
    peak = np.unravel_index(np.argmax(got["dirty"]), got["dirty"].shape)
    assert peak == (width // 2 + 4, width // 2 + 7)


@pytest.mark.parametrize("single", [False, True])
def test_plan_and_footprint_without_gridding(single):
    # core/src/wgridder.rs:3525-3586: the plan a call would choose
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nrow=50, seed=8)
    eps = 1e-5 if single else 1e-9
    cdt, rdt = (np.complex64, np.float32) if single else (np.complex128, np.float64)
    p = W.plan(uvw=uvw, freq=freq, vis=vis.astype(cdt), wgt=wgt.astype(rdt), npix_x=nx, npix_y=ny,
               pixsize_x=dx, pixsize_y=dx, epsilon=eps, single=single)
    _, info = W.vis2dirty(uvw=uvw, freq=freq, vis=vis.astype(cdt), wgt=wgt.astype(rdt), npix_x=nx,
                          npix_y=ny, pixsize_x=dx, pixsize_y=dx, epsilon=eps, return_info=True)
    assert p == info
    pf, nbytes = W.gridding_footprint(uvw=uvw, freq=freq, wgt=wgt, npix_x=nx, npix_y=ny,
                                      pixsize_x=dx, pixsize_y=dx, epsilon=eps, nout=2, single=single)
    _, _, info2 = W.dirty_psf(uvw=uvw, freq=freq, vis=vis.astype(cdt), wgt=wgt.astype(rdt),
                              npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx, epsilon=eps,
                              single=single, return_info=True)
    assert pf == info2
    cx = 8 if single else 16
    assert nbytes >= 2 * pf["nplanes"] * pf["nu"] * pf["nv"] * cx
    # no active visibility: an empty plan
    e = W.plan(uvw=uvw, freq=freq, vis=None, wgt=np.zeros_like(wgt), npix_x=nx, npix_y=ny,
               pixsize_x=dx, pixsize_y=dx, epsilon=eps, single=single)
    assert e["nvis"] == 0 and e["nu"] == 0


@pytest.mark.parametrize("single", [False, True])
def test_two_pass_transforms_match_oracle(single, monkeypatch):
    # Grids longer than one LDS transform take two passes over memory; a low
    # limit sends these small grids through that path in both directions.
    monkeypatch.setenv("RANG_FFT_LDS_MAX", "48")
    for case in ({"nx": 64, "ny": 48, "nrow": 60}, {"nx": 40, "ny": 72, "nrow": 50}):
        uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=case["nx"], ny=case["ny"], nrow=case["nrow"], seed=17)
        eps = 1e-5 if single else 1e-9
        got, ref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, eps, single)
        assert max(info["nu"], info["nv"]) > 48
        same_plan(info, rinfo)
        assert rel_linf(got, ref) < TOL_ORACLE[single]
        img = np.random.default_rng(3).standard_normal((nx, ny))
        gv, rv, _, _ = d2v(uvw, freq, img, wgt, dx, eps, single)
        assert rel_linf(gv, rv) < TOL_ORACLE[single]
        # dirty and PSF in one pass (two outputs through the two-pass rows)
        d, p = W.dirty_psf(uvw=uvw, freq=freq, vis=vis.astype(np.complex64 if single else np.complex128),
                           wgt=wgt.astype(np.float32 if single else np.float64), npix_x=nx, npix_y=ny,
                           pixsize_x=dx, pixsize_y=dx, epsilon=eps, single=single)
        assert rel_linf(d, ref) < TOL_ORACLE[single]


@pytest.mark.parametrize("single", [True, False])
@pytest.mark.parametrize("mode", ["0", "1"])
@pytest.mark.parametrize("occ", ["0", "1"])
def test_column_transforms_strided_and_transposed(single, mode, occ, monkeypatch):
    # The column transforms read or write whole columns with a tiled
    # transpose beside them (the default for large planes), or strided
    # columns (small planes); both forms in both precisions (the paired
    # single-precision transpose and the plain one), dirty and PSF in one
    # pass. With `occ` the row transforms read only the tiles an item
    # reached and write only the rows degridding reads (the default for
    # large planes).
    monkeypatch.setenv("RANG_FFT_TCM", mode)
    monkeypatch.setenv("RANG_GRID_OCC", occ)
    for case in ({"nx": 64, "ny": 48, "nrow": 60}, {"nx": 40, "ny": 72, "nrow": 50}):
        uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=case["nx"], ny=case["ny"], nrow=case["nrow"], seed=23)
        eps = 1e-5 if single else 1e-9
        got, ref, info, rinfo = v2d(uvw, freq, vis, wgt, nx, ny, dx, eps, single)
        same_plan(info, rinfo)
        assert rel_linf(got, ref) < TOL_ORACLE[single]
        img = np.random.default_rng(5).standard_normal((nx, ny))
        gv, rv, _, _ = d2v(uvw, freq, img, wgt, dx, eps, single)
        assert rel_linf(gv, rv) < TOL_ORACLE[single]
        d, p = W.dirty_psf(uvw=uvw, freq=freq, vis=vis.astype(np.complex64 if single else np.complex128),
                           wgt=wgt.astype(np.float32 if single else np.float64), npix_x=nx, npix_y=ny,
                           pixsize_x=dx, pixsize_y=dx, epsilon=eps, single=single)
        assert rel_linf(d, ref) < TOL_ORACLE[single]


@pytest.mark.parametrize("occ", ["0", "1"])
def test_cells_no_visibility_reaches_are_cleared(occ, monkeypatch):
    """A call after a denser one: the grid's workspace still holds the first
    call's planes, and the (tile, plane group)s that none of the second
    call's visibilities reach are cleared rather than gridded (or, with
    `occ`, skipped by the row transform), so nothing of the first call may
    show in the second's image."""
    monkeypatch.setenv("RANG_GRID_OCC", occ)
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=64, ny=64, nrow=300, seed=21, nchan=4)
    keep = np.zeros(vis.shape, dtype=np.uint8)
    keep[:12] = 1
    for single in (True, False):
        eps = 1e-5 if single else 1e-10
        got, ref, _, _ = v2d(uvw, freq, vis, wgt, nx, ny, dx, eps, single)
        assert rel_linf(got, ref) < TOL_ORACLE[single]
        got, ref, info, _ = v2d(uvw, freq, vis, wgt, nx, ny, dx, eps, single, mask=keep)
        assert info["nvis"] == int(keep.sum())
        assert rel_linf(got, ref) < TOL_ORACLE[single]
