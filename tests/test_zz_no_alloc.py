"""The allocation-free vis2dirty entry point
(rang_amd_wgridder_vis2dirty_noalloc_f32) against the allocating one
(rang_amd_wgridder_vis2dirty_dev_f32) through the C ABI, on two calls in a
row, the second smaller, and against the oracle.

Tolerances (max |difference| over the peak): 1e-6 between the two entry
points, which run the same kernels and differ only in the order of the
atomic merges of split tiles; 2e-5 against the oracle, the single-precision
parity bar of the gridder's other tests."""

import numpy as np
import pytest
from conftest import ref_case, rel_linf

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EPS = 1e-5


def opts(nx, ny, dx, eps=EPS):
    return [nx, ny, dx, dx, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]


def on_device(uvw, freq, vis, wgt):
    dev = torch.device("cuda:0")
    return (torch.from_numpy(uvw).to(dev), torch.from_numpy(freq).to(dev),
            torch.from_numpy(vis.astype(np.complex64)).to(dev),
            torch.from_numpy(wgt.astype(np.float32)).to(dev))


def both(case, psf=False):
    from rang import device as D

    uvw, freq, vis, wgt, nx, ny, dx = case
    tu, tf, tv, tw = on_device(uvw, freq, vis, wgt)
    o = opts(nx, ny, dx)
    ref = D.vis2dirty(o, tu, tf, tv, tw, single=True, psf=psf)
    got = D.vis2dirty(o, tu, tf, tv, tw, single=True, psf=psf, noalloc=True)
    torch.cuda.synchronize()
    return ref, got


def check(case, psf=False):
    (rd, rp, rinfo), (gd, gp, ginfo) = both(case, psf)
    assert np.array_equal(rinfo, ginfo)
    assert rel_linf(gd.cpu().numpy(), rd.cpu().numpy()) < 1e-6
    if psf:
        assert rel_linf(gp.cpu().numpy(), rp.cpu().numpy()) < 1e-6
    uvw, freq, vis, wgt, nx, ny, dx = case
    oref, _ = O.vis2dirty(uvw, freq, vis.astype(np.complex64), wgt.astype(np.float32), npix_x=nx,
                          npix_y=ny, pixsize_x=dx, pixsize_y=dx, epsilon=EPS, single=True)
    assert rel_linf(gd.cpu().numpy(), oref["dirty"]) < 2e-5
    return ginfo


def test_noalloc_matches_the_allocating_entry_point_on_two_calls():
    print("[noalloc] 128^2, 6000 rows x 4 channels, then 48x40, 60 rows x 2", flush=True)
    big = ref_case(nx=128, ny=128, field=0.3, wmax=800.0, nrow=6000, seed=11, nchan=4)
    info = check(big)
    assert info[5] == 24000
    small = ref_case(nx=48, ny=40, field=0.2, wmax=600.0, nrow=60, seed=12)
    info = check(small)
    assert info[5] == 120


def test_noalloc_dirty_and_psf_in_one_call():
    print("[noalloc] dirty + psf, 96^2, 800 rows x 3", flush=True)
    check(ref_case(nx=96, ny=96, field=0.25, wmax=500.0, nrow=800, seed=13, nchan=3), psf=True)


def test_noalloc_refuses_a_call_beyond_the_fixed_block():
    from rang import device as D

    print("[noalloc] 4096^2 refusal", flush=True)
    uvw, freq, vis, wgt, _, _, _ = ref_case(nx=4096, ny=4096, field=1.0, wmax=100.0, nrow=50, seed=14)
    tu, tf, tv, tw = on_device(uvw, freq, vis, wgt)
    with pytest.raises(ValueError, match="fixed scratch"):
        D.vis2dirty(opts(4096, 4096, 1.0 / 4096), tu, tf, tv, tw, single=True, noalloc=True)
    # the block is usable again after a refusal
    check(ref_case(nx=64, ny=64, field=0.2, wmax=300.0, nrow=100, seed=15))
