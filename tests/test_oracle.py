"""Pins the oracle (the CPU restatement of core/src/wgridder.rs) to what the
reference's own tests and published figures hold, and to the FP64 direct
Fourier transform. CPU only."""

import hashlib
import re

import numpy as np
import pytest
from conftest import C_LIGHT, ROOT, lcg_case, ref_case, rel_linf

from oracle import oracle as O

# SHA256 of ducc0 0.42.0's KernelDB rows as rang-core's test reads them from
# its source (tests/test_core_wgridder.py:12-14, 55-72).
KERNEL_ROWS_SHA256 = "0427f5940e0faa20efc615a1f4fb73eadbf754465f483f216c11b317f4ae3831"
ARCSEC = np.pi / 648000.0


@pytest.fixture(scope="module", autouse=True)
def built():
    O.build()


def table_rows():
    text = (ROOT / "rang-core_amd" / "csrc" / "kernel_db.inc").read_text()
    return re.findall(r"^\{(\d+), (\S+), (\S+), (\S+), (\S+), (\d), (\d)\},$", text, re.M)


def test_kernel_table_is_ducc0s():
    rows = table_rows()
    # The reference's row text: "(W, ofactor,%14s epsilon, beta, e0, ndim,%6s bool)"
    lines = ["(%s, %s,%14s, %s, %s, %s,%6s)" % (w, of, eps, beta, e0, nd, "true" if sp == "1" else "false")
             for w, of, eps, beta, e0, nd, sp in rows]
    text = "\n".join(lines) + "\n"
    assert hashlib.sha256(text.encode()).hexdigest() == KERNEL_ROWS_SHA256
    assert O.kernel_count() == len(rows) == 939
    assert O.kernel_row(0) == (4, 1.2, 0.027638471, 1.30709338717990464, 0.588199361165364421, 2, True)
    last = O.kernel_row(938)
    assert (last[0], last[5], last[6]) == (16, 3, False)
    assert O.kernel_row(939) is None
    # every row as parsed by the oracle equals the decimal tokens
    for i in (0, 17, 400, 938):
        w, of, eps, beta, e0, nd, sp = rows[i]
        assert O.kernel_row(i) == (int(w), float(of), float(eps), float(beta), float(e0), int(nd), sp == "1")


def test_good_sizes():
    # core/src/wgridder.rs:4635-4639
    assert [O.good_size_complex(n) for n in (13, 97, 181, 1025, 7)] == [14, 98, 189, 1029, 7]
    # brute force: the smallest 2^a 3^b 5^c 7^d 11^e >= n
    def smooth(m):
        for p in (2, 3, 5, 7, 11):
            while m % p == 0:
                m //= p
        return m == 1

    for n in list(range(13, 400)) + [1187, 2375, 4097]:
        m = n
        while not smooth(m):
            m += 1
        assert O.good_size_complex(n) == m


def test_gauss_legendre_integrates_polynomials():
    # core/src/wgridder.rs:4643-4650
    x, w = O.gauss_legendre(10)
    assert abs(np.sum(w * x**8) - 2.0 / 9.0) < 1e-14
    assert abs(w.sum() - 2.0) < 1e-14
    # odd n: the centre node once (weight not doubled)
    x, w = O.gauss_legendre(7)
    xs, ws = np.polynomial.legendre.leggauss(7)
    assert x[0] == 0.0 and abs(w[0] - ws[3]) < 1e-14
    assert np.allclose(x[1:], xs[4:], atol=1e-14) and np.allclose(w[1:], 2 * ws[4:], atol=1e-14)


def test_polynomial_kernel_matches_es():
    # core/src/wgridder.rs:4653-4664
    v = -1.0 + 2.0 * np.arange(2001) / 2000.0
    for index in (0, 200, 500, 938):
        poly, phi, _ = O.kernel_eval(index, v)
        eps = O.kernel_row(index)[2]
        assert np.abs(poly - phi).max() < max(eps, 1e-13), index


def test_correction_is_inverse_fourier_transform_of_kernel():
    # corfunc(v) = 1 / (W/2 int_{-1}^{1} phi(x) cos(pi W x v) dx), checked by
    # an independent quadrature
    index = 500
    row = O.kernel_row(index)
    wsup, beta, e0 = row[0], row[3] * row[0], row[4]
    x = np.linspace(-1, 1, 200001)
    phi = np.exp(beta * (np.sqrt(np.maximum(1 - x * x, 0)) ** (2 * e0) - 1))
    phi[np.abs(x) >= 1] = np.exp(-beta)
    vs = np.array([0.0, 0.1, 0.25, 0.4])
    _, _, corr = O.kernel_eval(index, vs)
    for v, c in zip(vs, corr):
        integral = np.trapezoid(phi * np.cos(np.pi * wsup * x * v), x) * wsup * 0.5
        assert abs(1.0 / integral - c) / c < 1e-6


@pytest.mark.parametrize("n", [2, 3, 5, 7, 11, 16, 22, 60, 2376, 1188, 900, 5 * 7 * 11])
def test_oracle_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref_f, ref_i = np.fft.fft(x), np.fft.ifft(x) * n
    scale = np.abs(x).sum()
    assert np.abs(O.fft(x, False) - ref_f).max() / scale < 1e-15 * np.log2(n + 1) * 10
    assert np.abs(O.fft(x, True) - ref_i).max() / scale < 1e-15 * np.log2(n + 1) * 10
    x32 = x.astype(np.complex64)
    assert np.abs(O.fft(x32, False) - np.fft.fft(x32.astype(np.complex128))).max() / scale < 2e-6


@pytest.mark.parametrize("nx,ny,field,wmax", [(32, 32, 0.2, 500.0), (32, 16, 0.05, 50.0)])
def test_vis2dirty_matches_dft(nx, ny, field, wmax):
    # core/src/wgridder.rs:4667-4684
    uvw, freq, vis, wgt, dx = lcg_case(nx, ny, field, wmax, 60, 7)
    imgs, info = O.vis2dirty(uvw, freq, vis, wgt, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx,
                             epsilon=1e-9, single=False)
    ref = O.dft_dirty(uvw, freq, vis, wgt, nx, ny, dx, dx)
    assert rel_linf(imgs["dirty"], ref) < 1e-8


def test_single_precision_matches_dft():
    # core/src/wgridder.rs:4779-4796
    uvw, freq, vis, wgt, dx = lcg_case(32, 32, 0.2, 400.0, 60, 5)
    imgs, _ = O.vis2dirty(uvw, freq, vis.astype(np.complex64), wgt.astype(np.float32), npix_x=32,
                          npix_y=32, pixsize_x=dx, pixsize_y=dx, epsilon=1e-4, single=True)
    assert imgs["dirty"].dtype == np.float32
    assert rel_linf(imgs["dirty"], O.dft_dirty(uvw, freq, vis, wgt, 32, 32, dx, dx)) < 1e-3


def test_python_case_both_precisions():
    # tests/test_core_wgridder.py:75-95
    uvw, freq, vis, wgt, nx, ny, dx = ref_case()
    ref = O.dft_dirty(uvw, freq, vis, wgt, nx, ny, dx, dx)
    kw = dict(npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx)
    d64, info = O.vis2dirty(uvw, freq, vis, wgt, epsilon=1e-9, single=False, **kw)
    assert rel_linf(d64["dirty"], ref) < 1e-8
    assert info[5] == vis.size and info[4] > info[1]
    d32, _ = O.vis2dirty(uvw, freq, vis.astype(np.complex64), wgt.astype(np.float32), epsilon=1e-4,
                         single=True, **kw)
    assert rel_linf(d32["dirty"], ref) < 1e-3


def test_adjoint_and_dirty2vis():
    # tests/test_core_wgridder.py:98-125 and core/src/wgridder.rs:4733-4754
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(seed=8)
    x = np.random.default_rng(1).standard_normal((nx, ny))
    kw = dict(pixsize_x=dx, pixsize_y=dx, epsilon=1e-9)
    ax, _ = O.dirty2vis(uvw, freq, x, wgt, **kw)
    aty, _ = O.vis2dirty(uvw, freq, vis, wgt, npix_x=nx, npix_y=ny, single=False, **kw)
    lhs = np.real(np.vdot(ax, vis))
    rhs = np.sum(x * aty["dirty"])
    assert abs(lhs - rhs) / (np.linalg.norm(ax) * np.linalg.norm(vis)) < 1e-12
    img = np.zeros((nx, ny))
    img[5, 20] = 2.0
    got, _ = O.dirty2vis(uvw, freq, img, wgt, **kw)
    ref = O.dft_vis(uvw, freq, img, wgt, dx, dx)
    assert rel_linf(got, ref) < 1e-8
    # and against the closed form of one pixel, computed here
    l, m = (5 - nx // 2) * dx, (20 - ny // 2) * dx
    n1 = np.sqrt(1 - l * l - m * m) - 1
    u, v, w = (uvw[:, i, None] * freq / C_LIGHT for i in range(3))
    closed = 2.0 * wgt * np.exp(-2j * np.pi * (u * l + v * m + w * n1))
    assert rel_linf(got, closed) < 1e-8
    uvw, freq, vis, wgt, dx = lcg_case(32, 32, 0.25, 800.0, 50, 11)
    s = np.random.default_rng(3).uniform(-0.5, 0.5, (32, 32))
    ax, _ = O.dirty2vis(uvw, freq, s, wgt, pixsize_x=dx, pixsize_y=dx, epsilon=1e-9)
    aty, _ = O.vis2dirty(uvw, freq, vis, wgt, npix_x=32, npix_y=32, pixsize_x=dx, pixsize_y=dx,
                         epsilon=1e-9, single=False)
    lhs = np.real(np.vdot(ax, vis))
    rhs = np.sum(s * aty["dirty"])
    assert abs(lhs - rhs) / (np.linalg.norm(ax) * np.linalg.norm(vis)) < 1e-13


def test_zero_weights_and_bad_input():
    # tests/test_core_wgridder.py:128-147, core/src/wgridder.rs:4799-4807
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nrow=6)
    kw = dict(npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx, single=False)
    empty, info = O.vis2dirty(uvw, freq, vis, np.zeros_like(wgt), epsilon=1e-6, **kw)
    assert not empty["dirty"].any() and not info.any()
    with pytest.raises(ValueError, match="ascending"):
        O.vis2dirty(uvw, freq[::-1], vis, wgt, epsilon=1e-6, **kw)
    with pytest.raises(ValueError, match="kernel"):
        O.vis2dirty(uvw, freq, vis, wgt, epsilon=1e-14, **kw)
    with pytest.raises(ValueError, match="even"):
        O.vis2dirty(uvw, freq, vis, wgt, epsilon=1e-6, **{**kw, "npix_x": 31})


def test_psf_in_one_pass_equals_separate_calls():
    # tests/test_core_wgridder.py:215-245: bit for bit
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(seed=21, nrow=50)
    vis64, wgt32 = vis.astype(np.complex64), wgt.astype(np.float32)
    wgt32[3, 1] = 0.0
    kw = dict(npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx, epsilon=1e-7, single=False)
    both, _ = O.vis2dirty(uvw, freq, vis64, wgt32, outputs=("dirty", "psf"), **kw)
    sep, _ = O.vis2dirty(uvw, freq, vis64.astype(np.complex128), wgt32.astype(np.float64), **kw)
    ones = np.where(wgt32 != 0, 1.0, 0.0).astype(np.complex128)
    sep_psf, _ = O.vis2dirty(uvw, freq, ones, wgt32.astype(np.float64), **kw)
    assert np.array_equal(both["dirty"], sep["dirty"])
    assert np.array_equal(both["psf"], sep_psf["dirty"])


def test_product_weights_match_their_product():
    # core/src/wgridder.rs:4812-4871
    uvw, freq, vis, wgt, dx = lcg_case(32, 24, 0.2, 400.0, 40, 7)
    rng = np.random.default_rng(99)
    a = wgt.astype(np.float32)
    b = (0.1 + rng.uniform(size=a.shape)).astype(np.float32)
    b.ravel()[::7] = 0.0
    kw = dict(npix_x=32, npix_y=24, pixsize_x=dx, pixsize_y=dx, epsilon=1e-6, single=False,
              outputs=("dirty", "psf"))
    one, i1 = O.vis2dirty(uvw, freq, vis.astype(np.complex64), a * b, **kw)
    two, i2 = O.vis2dirty(uvw, freq, vis.astype(np.complex64), a, wgt2=b, **kw)
    assert np.array_equal(one["dirty"], two["dirty"]) and np.array_equal(one["psf"], two["psf"])
    assert np.array_equal(i1, i2)


@pytest.mark.parametrize("name,nvis,npix,pix,supp,grid", [
    ("whole array", 23811985, 1024, 1.5, 7, 2376),
    ("K x 41", 4820000, 512, 8.0, 7, 1188),
    ("K arm", 117556, 512, 8.0, 8, 900),
])
def test_published_plan_choices(name, nvis, npix, pix, supp, grid):
    # research/df2-wgridder-speed.md:290-291: f64, epsilon 1e-6, planes = support
    # when the w range is well below dw
    info = O.plan(nvis, 0.0, 100.0, False, nx=npix, ny=npix, pixsize_x=pix * ARCSEC,
                  pixsize_y=pix * ARCSEC, epsilon=1e-6)
    assert (int(info[1]), int(info[2]), int(info[3]), int(info[4])) == (supp, grid, grid, supp)


OPTION_CASES = [
    dict(),
    dict(center_x=0.01, center_y=-0.02),
    dict(flip_u=True, flip_v=True),
    dict(flip_w=True),
    dict(divide_by_n=True),
    dict(allow_nshift=False),
    dict(center_x=0.03, flip_u=True, flip_w=True, divide_by_n=True),
]


@pytest.mark.parametrize("opts", OPTION_CASES)
def test_options_match_dft(opts):
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(nx=32, ny=24, field=0.15, wmax=300.0, nrow=30, seed=5)
    imgs, _ = O.vis2dirty(uvw, freq, vis, wgt, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx,
                          epsilon=1e-9, single=False, **opts)
    dft_opts = {k: v for k, v in opts.items() if k != "allow_nshift"}
    ref = O.dft_dirty(uvw, freq, vis, wgt, nx, ny, dx, dx, **dft_opts)
    assert rel_linf(imgs["dirty"], ref) < 1e-8
    x = np.random.default_rng(2).standard_normal((nx, ny))
    got, _ = O.dirty2vis(uvw, freq, x, wgt, pixsize_x=dx, pixsize_y=dx, epsilon=1e-9, **opts)
    refv = O.dft_vis(uvw, freq, x, wgt, dx, dx, **dft_opts)
    assert rel_linf(got, refv) < 1e-8


def test_without_w_gridding_ignores_w():
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(wmax=200.0)
    imgs, info = O.vis2dirty(uvw, freq, vis, wgt, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx,
                             epsilon=1e-9, single=False, do_wgridding=False)
    flat = uvw.copy()
    flat[:, 2] = 0.0
    assert info[4] == 0
    assert rel_linf(imgs["dirty"], O.dft_dirty(flat, freq, vis, wgt, nx, ny, dx, dx)) < 1e-8


def test_checks_can_fail():
    """The DFT comparison is sensitive to the conventions: gridding with w
    flipped against the unflipped transform is far off (the reference's
    precondition, research/df1-wgridder.md:117-120, found 1.01)."""
    uvw, freq, vis, wgt, nx, ny, dx = ref_case(field=0.4, wmax=2000.0)
    imgs, _ = O.vis2dirty(uvw, freq, vis, wgt, npix_x=nx, npix_y=ny, pixsize_x=dx, pixsize_y=dx,
                          epsilon=1e-9, single=False, flip_w=True)
    assert rel_linf(imgs["dirty"], O.dft_dirty(uvw, freq, vis, wgt, nx, ny, dx, dx)) > 0.05
