"""The HIP w-stacking gridder (core/src/grid.rs through the C ABI of
core/src/ffi_grid.rs and core/src/gpu/ffi.rs) against rang-core's golden
vectors and the oracle.

Tolerances (max |difference| over the image peak): 1e-12 in double (the
GPU adds each layer's samples in a different order and transforms with its
own FFT), 2e-5 in single precision (grid and transforms in float, sums in
double). Counters that are sums of float32 weights in double: relative
1e-12; the gridded count and the w range: exact."""

import math

import numpy as np
import pytest
from conftest import ROOT
from ws_cases import golden_case, point_source_case, survey_case

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = ROOT / "tests" / "golden" / "ws_golden.npz"
TOL = {"f64": 1e-12, "f32": 2e-5}


def gpu(case, precision="f64", device="gpu", **kw):
    from rang import core_grid

    return core_grid.invert(case["uvw"], case["frequencies"], case["data"], case["weights"],
                            case["imaging_weights"], width=case["width"], height=case["height"],
                            scale_arcsec=case["pixel_size"] / core_grid.ARCSEC,
                            w_layers=case["w_layers"], device=device, precision=precision, **kw)


def oracle(case, **kw):
    return O.ws_invert(case["uvw"], case["frequencies"], case["data"], case["weights"],
                       case["imaging_weights"], width=case["width"], height=case["height"],
                       pixel_size=case["pixel_size"], w_layers=case["w_layers"], **kw)


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def same_stats(got, ref, precision="f64", exact_sums=False):
    for k in ("gridded_count", "min_w", "max_w", "n_w_layers"):
        assert got[k] == ref[k], k
    for k in ("image_weight", "visibility_weight_sum", "max_gridded_weight"):
        tol = 0.0 if exact_sums else 1e-12
        assert abs(got[k] - ref[k]) <= tol * abs(ref[k]), k
    # the PSF centre comes out of the grid and transforms: their precision
    assert abs(got["psf_normalization"] - ref["psf_normalization"]) <= TOL[precision] * ref["psf_normalization"]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_golden_case(precision):
    z = np.load(GOLD, allow_pickle=False)
    c = golden_case()
    r = gpu(c, precision)
    assert rel(r["dirty"], z["dirty"]) < TOL[precision]
    assert rel(r["psf"], z["psf"]) < TOL[precision]
    names = O.WS_STAT_NAMES
    ref = dict(zip(names, z["stats"].tolist()))
    same_stats(r["stats"], ref, precision, exact_sums=True)
    assert r["gpu"]["batches"] >= 1 and r["gpu"]["gpu_bytes"] > 0


def test_cpu_entry_is_the_double_path():
    z = np.load(GOLD, allow_pickle=False)
    r = gpu(golden_case(), device="cpu")
    assert rel(r["dirty"], z["dirty"]) < TOL["f64"]
    assert r["gridder"] == "wstack"


def test_kernel_through_the_abi():
    from rang import core_grid

    z = np.load(GOLD, allow_pickle=False)
    assert np.abs(core_grid.kernel(7, 1023) - z["kernel"]).max() <= 1e-15


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_point_source_lands_at_wsclean_pixel(precision):
    # core/src/grid.rs:1036-1063
    r = gpu(point_source_case(4), precision)
    d = r["dirty"]
    y, x = np.unravel_index(np.argmax(d), d.shape)
    assert (x, y) == (16 - 5, 16 - 3)
    assert abs(r["psf"][16, 16] - 1.0) <= 2 * np.finfo(float).eps
    assert abs(d.max() - 1.0) < 0.05
    assert r["stats"]["gridded_count"] == 108


@pytest.mark.parametrize("layers", [1, 8, 40])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_matches_oracle(layers, precision):
    c = survey_case(seed=layers, w_layers=layers)
    r, ref = gpu(c, precision), oracle(c)
    assert rel(r["dirty"], ref["dirty"]) < TOL[precision]
    assert rel(r["psf"], ref["psf"]) < TOL[precision]
    same_stats(r["stats"], ref["stats"], precision)


def test_many_layers_uneven_imaging_weights():
    # core/src/grid.rs:1065-1094's case: 40 layers, imaging weights 0.5..2.5
    c = point_source_case(40, imaging=(0.5 + np.arange(120) % 3).reshape(60, 2))
    r, ref = gpu(c), oracle(c)
    assert rel(r["dirty"], ref["dirty"]) < TOL["f64"]
    assert rel(r["psf"], ref["psf"]) < TOL["f64"]
    same_stats(r["stats"], ref["stats"])


def test_errors():
    from rang import core_grid

    c = golden_case()
    # no weight at all: invalid (1 from the CPU entry, 3 from the GPU entry)
    z = dict(c, weights=np.zeros_like(c["weights"]))
    with pytest.raises(RuntimeError, match="code 1"):
        gpu(z, device="cpu")
    with pytest.raises(RuntimeError, match="code 3"):
        gpu(z)
    # even kernel size
    with pytest.raises(RuntimeError, match="code 1"):
        gpu(c, device="cpu", kernel_size=6)
    # a padded side with a prime factor above 11 (42 -> 52 = 4 x 13): the GPU
    # entry refuses it as the reference's GPU path does (core/src/gpu/tests.rs:
    # 274-288); the CPU entry takes any side (test_any_padded_side)
    with pytest.raises(RuntimeError, match="code 3"):
        gpu(dict(c, width=42))
    with pytest.raises(ValueError):
        gpu(c, precision="f16")
    assert "gfx950" in __import__("rang._core", fromlist=["x"]).grid_gpu_adapter()


@pytest.mark.parametrize("width,height,padding,lds_max", [
    (42, 40, 1.2, None),    # 52 = 4 x 13 by 48: Bluestein in LDS along x
    (52, 68, 1.0, None),    # 52 x 68 (4 x 17): Bluestein along both axes
    (60, 52, 1.0, "32"),    # 60 in four steps (6 x 10), 52 by Bluestein over four steps (M = 105)
    (48, 48, 1.0, "16"),    # 48 in four steps both ways (transposed columns)
])
def test_any_padded_side(width, height, padding, lds_max, monkeypatch):
    # The reference's CPU entry transforms any side (rustfft, core/src/grid.rs:
    # 206-259); the oracle's FFT takes any length too. RANG_WS_FFT_LDS_MAX
    # lowers the in-LDS limit so that small sides take the four-step and chirp
    # paths that sides above 8192 (double) and 16384 (single) take.
    if lds_max:
        monkeypatch.setenv("RANG_WS_FFT_LDS_MAX", lds_max)
    print(f"[ws] {width} x {height}, padding {padding}, LDS limit {lds_max}", flush=True)
    c = survey_case(seed=width, width=width, height=height, w_layers=6)
    r = gpu(c, device="cpu", padding=padding)
    ref = oracle(c, padding=padding)
    assert rel(r["dirty"], ref["dirty"]) < TOL["f64"]
    assert rel(r["psf"], ref["psf"]) < TOL["f64"]
    same_stats(r["stats"], ref["stats"])


def _point(l_pix, m_pix, width=48, scale_arcsec=60.0, seed=20261029):
    # a unit point source (l_pix, m_pix) pixels from the centre through 80
    # random baselines at two frequencies (the reference's
    # tests/test_core_grid.py:37-55 case)
    from rang import core_grid

    scale = scale_arcsec * core_grid.ARCSEC
    l, m = l_pix * scale, m_pix * scale
    n = math.sqrt(1.0 - l * l - m * m)
    freq = np.array([1.0e9, 1.1e9])
    rng = np.random.default_rng(seed)
    radius, angle = rng.uniform(5.0, 120.0, 80), rng.uniform(0.0, 2 * math.pi, 80)
    uvw = np.stack([radius * np.cos(angle), radius * np.sin(angle), rng.normal(0, 4, 80)], 1)
    phase = 2 * math.pi * (uvw[:, :1] * l + uvw[:, 1:2] * m + uvw[:, 2:3] * (n - 1)) * freq / 299792458.0
    data = np.exp(1j * phase).astype(np.complex64)
    w = np.ones(data.shape, np.float32)
    return (uvw, freq, data, w, w), dict(width=width, height=width, scale_arcsec=scale_arcsec, w_layers=5)


@pytest.mark.parametrize("gridder", ["wstack", "wgrid"])
def test_point_source_at_the_wsclean_pixel_for_both_gridders(gridder, monkeypatch):
    from rang import core_grid

    monkeypatch.delenv(core_grid.DEVICE_ENV, raising=False)
    args, kw = _point(7, -4)
    r = core_grid.invert(*args, **kw, gridder=gridder)
    dirty, psf, w = r["dirty"], r["psf"], kw["width"]
    y, x = np.unravel_index(np.argmax(dirty), dirty.shape)
    assert (x, y) == (w // 2 - 7, w // 2 - 4)
    assert abs(psf[w // 2, w // 2] - 1.0) <= 1e-12
    assert abs(dirty[y, x] - 1.0) < 0.05
    assert r["gridder"] == gridder and "gpu" not in r
    assert r["stats"]["gridded_count"] == args[2].size
    if gridder == "wstack":
        # the default device is the reference's CPU path (run here on the GPU,
        # whose atomic sums may differ between calls in the last bits)
        explicit = core_grid.invert(*args, **kw, device="cpu")
        assert "gridder" in explicit and rel(explicit["dirty"], dirty) <= TOL["f64"]
    else:
        assert r["wgridder"]["epsilon"] == core_grid.WGRID_EPSILON


def test_wide_field_8192_pads_to_a_side_with_a_large_prime():
    # 8192^2 at 1.3 arcsec pads (1.2) to 9832 = 8 x 1229: the reference's CPU
    # path takes it through rustfft, here through Bluestein transforms in
    # four steps; a unit point source 300 x 200 pixels off centre lands on
    # its WSClean pixel with the PSF centre at one
    from rang import core_grid

    assert core_grid.padded_size(8192) == 9832
    args, kw = _point(300, -200, width=8192, scale_arcsec=1.3)
    kw["w_layers"] = 8
    r = core_grid.invert(*args, **kw, device="cpu")
    dirty, psf, w = r["dirty"], r["psf"], 8192
    y, x = np.unravel_index(np.argmax(dirty), dirty.shape)
    assert (x, y) == (w // 2 - 300, w // 2 - 200)
    assert abs(psf[w // 2, w // 2] - 1.0) <= 1e-12
    assert abs(dirty[y, x] - 1.0) < 0.05
