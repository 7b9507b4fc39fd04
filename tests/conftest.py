import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "rang-core_amd"))

C_LIGHT = 299792458.0


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


def ref_case(nx=32, ny=32, field=0.2, wmax=600.0, nrow=40, seed=3, nchan=2):
    """The synthetic case of rang-core tests/test_core_wgridder.py:17-34."""
    rng = np.random.default_rng(seed)
    freq = np.array([1.0e9, 1.3e9]) if nchan == 2 else np.linspace(1.0e9, 1.4e9, nchan)
    dx = field / nx
    lam = freq.max() / C_LIGHT
    uvw = np.stack(
        [
            rng.uniform(-0.45, 0.45, nrow) / dx / lam,
            rng.uniform(-0.45, 0.45, nrow) / dx / lam,
            rng.uniform(-wmax, wmax, nrow) / lam,
        ],
        axis=1,
    )
    vis = rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))
    wgt = rng.uniform(0.5, 1.5, (nrow, nchan))
    return uvw, freq, vis, wgt, nx, ny, dx


def lcg_case(nx, ny, field, wmax, nrow, seed):
    """The LCG case of core/src/wgridder.rs:4547-4593 (freq 1.0/1.2/1.4 GHz,
    weights 0.5 + U)."""
    s = [seed]

    def lcg():
        s[0] = (s[0] * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        return (s[0] >> 11) / float(1 << 53)

    freq = np.array([1.0e9, 1.2e9, 1.4e9])
    dx = field / nx
    lam = 1.4e9 / C_LIGHT
    uvw = []
    for _ in range(nrow):
        uvw.append((lcg() - 0.5) * 0.9 / dx / lam)
        uvw.append((lcg() - 0.5) * 0.9 / dx / lam)
        uvw.append((lcg() - 0.5) * 2.0 * wmax / lam)
    uvw = np.array(uvw).reshape(nrow, 3)
    n = nrow * 3
    vis = np.array([complex(lcg() - 0.5, lcg() - 0.5) for _ in range(n)]).reshape(nrow, 3)
    wgt = np.array([0.5 + lcg() for _ in range(n)]).reshape(nrow, 3)
    return uvw, freq, vis, wgt, dx


def rel_linf(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())
