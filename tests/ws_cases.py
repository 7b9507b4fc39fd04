"""Inputs of the w-stacking gridder's reference test cases, regenerated
from their recipes (no reference code runs)."""

import math

import numpy as np

C_LIGHT = 299792458.0


def golden_case():
    """rang-core tests/test_core_backends.py:307-338: 80 rows, two channels,
    a unit point source at (5, -3) pixels of 60 arcsec, 48 x 48, 4 layers."""
    rng = np.random.default_rng(20261029)
    frequencies = np.array([1.0e9, 1.1e9])
    radius = rng.uniform(5.0, 120.0, 80)
    angle = rng.uniform(0.0, 2 * math.pi, 80)
    uvw = np.stack([radius * np.cos(angle), radius * np.sin(angle), rng.normal(0, 4, 80)], 1)
    scale = 60.0 * math.pi / 648000.0
    l, m = 5 * scale, -3 * scale
    n = math.sqrt(1.0 - l * l - m * m)
    phase = 2 * math.pi * (uvw[:, :1] * l + uvw[:, 1:2] * m + uvw[:, 2:3] * (n - 1)) / (
        C_LIGHT / frequencies)
    data = np.exp(1j * phase).astype(np.complex64)
    weights = np.ones(data.shape, np.float32)
    return dict(uvw=uvw, frequencies=frequencies, data=data, weights=weights,
                imaging_weights=weights.copy(), width=48, height=48, pixel_size=scale, w_layers=4)


def point_source_case(n_w_layers=4, imaging=None):
    """core/src/grid.rs:997-1034: 60 rows on a spiral with w terms, 1.0 and
    1.2 GHz, a source at (5, -3) pixels of 1 mrad, one flagged sample per
    fifth row; 32 x 32."""
    ps = 1e-3
    l, m = 5.0 * ps, -3.0 * ps
    n = math.sqrt(1.0 - l * l - m * m)
    freqs = np.array([1.0e9, 1.2e9])
    uvw, data, weights = [], [], []
    for i in range(60):
        a = i * 0.37
        r = 20.0 + (i % 7) * 15.0
        u, v, w = r * math.cos(a), r * math.sin(a), 3.0 * math.sin(a * 1.7)
        uvw.append([u, v, w])
        row_d, row_w = [], []
        for f in freqs:
            lam = C_LIGHT / f
            ph = 2.0 * math.pi * (u * l + v * m + w * (n - 1.0)) / lam
            row_d.append(complex(np.float32(math.cos(ph)), np.float32(math.sin(ph))))
            row_w.append(0.0 if (i % 5 == 0 and f > 1.1e9) else 1.0)
        data.append(row_d)
        weights.append(row_w)
    weights = np.array(weights, np.float32)
    if imaging is None:
        imaging = np.ones_like(weights)
    return dict(uvw=np.array(uvw), frequencies=freqs, data=np.array(data, np.complex64),
                weights=weights, imaging_weights=np.asarray(imaging, np.float32), width=32,
                height=32, pixel_size=ps, w_layers=n_w_layers)


def survey_case(seed, nrow=400, nchan=8, width=64, height=48, scale_arcsec=20.0, w_layers=8,
                wmax=200.0, flag_frac=0.2):
    """A random case with flags, uneven imaging weights and w terms."""
    rng = np.random.default_rng(seed)
    freqs = np.linspace(1.0e9, 1.4e9, nchan)
    scale = scale_arcsec * math.pi / 648000.0
    lam = C_LIGHT / freqs.max()
    umax = 0.45 / scale
    uvw = np.stack([rng.uniform(-umax, umax, nrow) * lam, rng.uniform(-umax, umax, nrow) * lam,
                    rng.uniform(-wmax, wmax, nrow) * lam], 1)
    data = (rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))).astype(np.complex64)
    weights = (rng.uniform(0.5, 1.5, (nrow, nchan)) * (rng.uniform(size=(nrow, nchan)) >= flag_frac)).astype(np.float32)
    imaging = rng.uniform(0.2, 2.0, (nrow, nchan)).astype(np.float32)
    return dict(uvw=uvw, frequencies=freqs, data=data, weights=weights, imaging_weights=imaging,
                width=width, height=height, pixel_size=scale, w_layers=w_layers)
