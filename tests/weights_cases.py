"""The imaging-weights case of rang-core tests/test_core_backends.py:281-304:
numpy's default_rng(7) draws, in order, 5 x 3 four-correlation complex64
data, float32 weights in [0.5, 2), flags, uvw in +-300 m (5 x 3), the
wavelengths 0.2, 0.22, 0.25 m, float32 visibility weights in [0, 2), and
rows 0, 1, 3, 4 selected; the grid of rang's defaults (512 pixels of 8
arcsec, padding 1.2 -> 616), Briggs robust 0.5, rank filter 3 over 16."""

import numpy as np


def golden_weights_case():
    rng = np.random.default_rng(7)
    data = (rng.standard_normal((5, 3, 4)) + 1j * rng.standard_normal((5, 3, 4))).astype(np.complex64)
    weights = rng.uniform(0.5, 2.0, (5, 3, 4)).astype(np.float32)
    flags = rng.integers(0, 2, (5, 3, 4)).astype(bool)
    uvw = rng.uniform(-300, 300, (5, 3))
    wavelength = np.array([0.2, 0.22, 0.25])
    vis_weights = rng.uniform(0.0, 2.0, (5, 3)).astype(np.float32)
    selected = np.array([True, True, False, True, True])
    return dict(data=data, weights=weights, flags=flags, uvw=uvw, wavelength=wavelength,
                vis_weights=vis_weights, selected=selected)


PIXEL_SCALE = 8.0 * np.pi / (180.0 * 60.0 * 60.0)
