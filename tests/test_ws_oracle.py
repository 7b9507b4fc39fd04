"""The w-stacking oracle (CPU restatement of core/src/grid.rs) pinned to
rang-core's golden vectors and its own known-answer tests (CPU only)."""

import hashlib
import math

import numpy as np
from conftest import ROOT
from ws_cases import golden_case, point_source_case

from oracle import oracle as O

GOLD = ROOT / "tests" / "golden" / "ws_golden.npz"


def test_fixture_is_the_references_data():
    z = np.load(GOLD, allow_pickle=False)
    for key, sha in zip(("dirty", "psf", "kernel"), z["sha256"]):
        assert hashlib.sha256(np.ascontiguousarray(z[key]).tobytes()).hexdigest() == sha


def test_golden_dirty_psf_kernel_and_counters():
    z = np.load(GOLD, allow_pickle=False)
    c = golden_case()
    r = O.ws_invert(c["uvw"], c["frequencies"], c["data"], c["weights"], c["imaging_weights"],
                    width=48, height=48, pixel_size=c["pixel_size"], w_layers=4)
    # same algorithm, a different FFT (this one's against rustfft): rounding
    for key in ("dirty", "psf"):
        assert np.abs(r[key] - z[key]).max() <= 1e-14 * np.abs(z[key]).max()
    kern = O.ws_kernel(7, 1023)
    assert np.array_equal(kern, z["kernel"])
    stats = np.array([r["stats"][k] for k in O.WS_STAT_NAMES])
    assert np.array_equal(stats, z["stats"])


def test_fft2_matches_numpy():
    # core/src/grid.rs:931-951 sizes: factors 2, 3, 7, 11, primes 13, 17
    rng = np.random.default_rng(4)
    for w, h in [(14, 11), (12, 13), (22, 17), (1, 7), (8, 8)]:
        g = rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))
        ref = np.fft.ifft2(g) * (w * h)
        assert np.abs(O.ws_fft2_backward(g) - ref).max() < 1e-10


def test_kernel_is_symmetric_and_peaks_at_one():
    k = O.ws_kernel(7, 1023)
    mid = k.size // 2
    assert abs(k[mid] - 1.0) < 1e-12
    assert np.array_equal(k[:mid], k[::-1][:mid])


def test_point_source_lands_at_wsclean_pixel():
    # core/src/grid.rs:1036-1063
    c = point_source_case(4)
    r = O.ws_invert(c["uvw"], c["frequencies"], c["data"], c["weights"], c["imaging_weights"],
                    width=32, height=32, pixel_size=c["pixel_size"], w_layers=4)
    d = r["dirty"]
    y, x = np.unravel_index(np.argmax(d), d.shape)
    assert (x, y) == (16 - 5, 16 - 3)
    assert abs(r["psf"][16, 16] - 1.0) <= 2 * np.finfo(float).eps
    assert abs(d.max() - 1.0) < 0.05
    assert r["stats"]["gridded_count"] == 108


def test_dirty_matches_a_direct_sum_without_w():
    # one layer and no w: the dirty image is the weighted direct transform
    # divided by the kernel correction's ideal (the correction makes the
    # gridded image equal to the direct sum up to the kernel's accuracy)
    c = point_source_case(1)
    c["uvw"][:, 2] = 0.0
    r = O.ws_invert(c["uvw"], c["frequencies"], c["data"], c["weights"], c["imaging_weights"],
                    width=32, height=32, pixel_size=c["pixel_size"], w_layers=1)
    ps = c["pixel_size"]
    lam = 299792458.0 / c["frequencies"]
    u = c["uvw"][:, :1] / lam[None, :]
    v = c["uvw"][:, 1:2] / lam[None, :]
    wgt = c["weights"] * c["imaging_weights"]
    # WSClean's orientation (grid.rs:497-529): pixel (x, y) of the output is
    # the backward transform at l = (x - W/2) ps, m = -(y - H/2) ps; the
    # images are normalised so that the PSF centre (the weight sum) is one
    img = np.zeros((32, 32))
    for y in range(32):
        for x in range(32):
            l, m = (x - 16) * ps, -(y - 16) * ps
            img[y, x] = np.sum(wgt * np.real(c["data"] * np.exp(2j * math.pi * (u * l + v * m))))
    img /= wgt.sum()
    # the 7-tap Kaiser-Bessel kernel with 1.2 padding is accurate at the
    # centre and aliases towards the edges (measured: 6e-5 over the central
    # 8 x 8, 1.7e-3 over 16 x 16, 5.6e-2 at the border)
    err = np.abs(r["dirty"] - img)
    assert err[12:20, 12:20].max() < 2e-4
    assert err[8:24, 8:24].max() < 5e-3
    assert err.max() < 0.1
