"""rang.core_grid's surface beside `invert` (rang-core's
python/rang/core_grid.py: the environment's device and precision, the
Stokes I reader rule, the kernel's formula, the WCS header), checked as the
reference's tests/test_core_grid.py:11-110 and tests/test_x2_gpu.py:40-90
check them. CPU only; `invert` itself is in test_ws_gpu.py."""

import math

import numpy as np
import pytest

from rang import core_grid


def test_kernel_equals_its_formula():
    k = core_grid.kernel(7, 1023)
    ref = core_grid.kernel_reference(7, 1023)
    assert k.shape == (7 * 1023,)
    assert np.abs(k - ref).max() <= 1e-15
    assert k[k.size // 2] == 1.0
    assert np.array_equal(k, k[::-1])


def test_stokes_i_reader_rule():
    # (XX + YY) / 2; weight 4 min(w_XX, w_YY), zero where XX or YY is flagged
    # or a row is flagged; XY and YX flags do not matter
    data = np.zeros((2, 3, 4), np.complex64)
    data[..., 0] = 2.0
    data[..., 3] = 4.0 + 2j
    w = np.ones((2, 3, 4), np.float32)
    w[0, 0, 0] = 0.25
    w[0, 1, 3] = 0.5
    flags = np.zeros((2, 3, 4), bool)
    flags[0, 2, 1] = True
    flags[1, 0, 3] = True
    sample, weight = core_grid.stokes_i_reader(data, w, flags, [False, False])
    assert np.array_equal(sample, np.full((2, 3), 3.0 + 1j, np.complex64))
    assert (weight[0, 0], weight[0, 1], weight[0, 2], weight[1, 0]) == (1.0, 2.0, 4.0, 0.0)
    _, rowless = core_grid.stokes_i_reader(data, w, flags, [True, False])
    assert rowless[0].sum() == 0.0 and rowless[1, 1] == 4.0
    bad = data.copy()
    bad[1, 2, 0] = np.nan
    s2, w2 = core_grid.stokes_i_reader(bad, w, flags)
    assert s2[1, 2] == 0 and w2[1, 2] == 0.0
    with pytest.raises(ValueError, match="four-correlation"):
        core_grid.stokes_i_reader(data[..., :3], w[..., :3], flags[..., :3])


def test_wcs_header():
    freq = 856e6 + 208984.375 * np.arange(1024, 3072)
    h = core_grid.wcs_header((5.1461782, -1.11199581), 512, 512, 8.0, freq)
    assert h["CRPIX1"] == 257.0 and h["CRPIX2"] == 257.0
    assert math.isclose(h["CRVAL1"], math.degrees(5.1461782) - 360.0, abs_tol=1e-9)
    assert math.isclose(h["CRVAL2"], math.degrees(-1.11199581), abs_tol=1e-9)
    assert h["CDELT1"] == -8 / 3600 and h["CDELT2"] == 8 / 3600
    # the channel edges: half a channel beyond the first and last centres
    assert h["CRVAL3"] == 1283895507.8125 and h["CDELT3"] == 428000000.0
    assert (h["CTYPE1"], h["CTYPE2"], h["CTYPE4"]) == ("RA---SIN", "DEC--SIN", "STOKES")


def test_device_and_precision_from_the_environment(monkeypatch):
    monkeypatch.delenv(core_grid.DEVICE_ENV, raising=False)
    monkeypatch.delenv(core_grid.PRECISION_ENV, raising=False)
    assert (core_grid.resolve_device(), core_grid.resolve_precision()) == ("cpu", "f64")
    monkeypatch.setenv(core_grid.DEVICE_ENV, " GPU ")
    monkeypatch.setenv(core_grid.PRECISION_ENV, "f32")
    assert (core_grid.resolve_device(), core_grid.resolve_precision()) == ("gpu", "f32")
    assert core_grid.resolve_device("cpu") == "cpu"
    monkeypatch.setenv(core_grid.DEVICE_ENV, "cuda")
    with pytest.raises(ValueError, match="RANG_GRIDDER_DEVICE"):
        core_grid.resolve_device()
    with pytest.raises(ValueError, match="RANG_GRIDDER_GPU_PRECISION"):
        core_grid.resolve_precision("f16")


def test_gridder_choice_is_checked():
    args = (np.zeros((1, 3)), np.array([1e9]), np.zeros((1, 1), np.complex64), np.ones((1, 1), np.float32),
            np.ones((1, 1), np.float32))
    with pytest.raises(ValueError, match="gridder"):
        core_grid.invert(*args, gridder="fft")
    with pytest.raises(RuntimeError, match="wgrid gridder runs on the CPU only"):
        core_grid.invert(*args, gridder="wgrid", device="gpu")
    assert core_grid.gpu_compiled()
    ok, why = core_grid.gpu_available()
    assert isinstance(ok, bool) and why
