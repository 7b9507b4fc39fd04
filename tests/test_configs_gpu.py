"""Parity on the workloads themselves, through the C ABI on the GPU, with the
L2 and L-infinity errors stated.

* The small workload, whole (2016 baselines x 100 times x 16 channels =
  3,225,600 complex64 visibilities, 1024^2 at 2 arcsec, epsilon 1e-5,
  single precision): vis2dirty and dirty2vis against the oracle at the same
  plan on every pixel and visibility, and against the FP64 direct transform
  on 256 pixels (over every visibility) and on 100,000 visibilities of a
  20-source image.
* The scaled-up workload's plan (kernel row and 14784^2 grid of the
  464,486,400-visibility call) on a row subset that the oracle finishes in
  well under a minute: 6 time steps spread over the track (12,096 rows x
  256 channels), every w-plane of that subset, both directions against the
  oracle.

Tolerances: against the oracle, max error / peak 2e-5 (the single-precision
bar of the other parity tests; the two differ in summation order and FFT
rounding) and L2 1e-5; against the direct transform the reference's gate of
10 epsilon (max error / peak; research/df1-wgridder.md:184-186)."""

import numpy as np
import pytest
from direct import dirty_pixels, errors, vis_of_points

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EPS = 1e-5


def opts(npix, pix, kernel=None):
    k = (-1, 0, 0) if kernel is None else kernel
    return [npix, npix, pix, pix, EPS, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, k[0], k[1], k[2]]


def report(what, e):
    print(f"[configs] {what}: L2 {e[0]:.2e}  Linf/peak {e[1]:.2e}", flush=True)


@pytest.fixture(scope="module")
def small():
    from rang import synth

    print("[configs] forming the 3,225,600-visibility workload", flush=True)
    return synth.config1()


def test_small_workload_against_the_oracle_and_the_direct_transform(small):
    from rang import core_wgridder as W

    c = small
    n, pix = c["npix"], c["pixsize"]
    kw = dict(npix_x=n, npix_y=n, pixsize_x=pix, pixsize_y=pix, epsilon=EPS)
    dirty, info = W.vis2dirty(uvw=c["uvw"], freq=c["freq"], vis=c["vis"], wgt=c["wgt"],
                              return_info=True, **kw)
    ref, rinfo = O.vis2dirty(c["uvw"], c["freq"], c["vis"], c["wgt"], single=True, **kw)
    assert [info[k] for k in ("kernel_index", "supp", "nu", "nv", "nplanes", "nvis")] == \
        [int(v) for v in rinfo[:6]]
    e = errors(dirty, ref["dirty"])
    report(f"small vis2dirty vs oracle (grid {info['nu']}, {info['nplanes']} planes)", e)
    assert e[1] < 2e-5 and e[0] < 1e-5

    dev = torch.device("cuda:0")
    tu, tf = torch.from_numpy(c["uvw"]).to(dev), torch.from_numpy(c["freq"]).to(dev)
    tv, tw = torch.from_numpy(c["vis"]).to(dev), torch.from_numpy(c["wgt"]).to(dev)
    g = torch.Generator(device=dev).manual_seed(5)
    i = torch.randint(0, n, (256,), device=dev, generator=g)
    j = torch.randint(0, n, (256,), device=dev, generator=g)
    flux, sl, sm = c["sky"]
    b = int(np.argmax(flux))
    i[0], j[0] = int(round(n // 2 + sl[b] / pix)), int(round(n // 2 + sm[b] / pix))
    d256 = dirty_pixels(tu, tf, tv, tw, i, j, n, n, pix, pix).cpu().numpy()
    e = errors(dirty[i.cpu().numpy(), j.cpu().numpy()], d256)
    report("small vis2dirty vs direct transform, 256 pixels", e)
    assert e[1] < 10 * EPS

    model = np.zeros((n, n), np.float32)
    rng = np.random.default_rng(8)
    pi, pj = rng.integers(n // 8, 7 * n // 8, 20), rng.integers(n // 8, 7 * n // 8, 20)
    pf = rng.uniform(0.1, 1.0, 20)
    model[pi, pj] = pf
    vis = W.dirty2vis(uvw=c["uvw"], freq=c["freq"], dirty=model, wgt=c["wgt"], pixsize_x=pix,
                      pixsize_y=pix, epsilon=EPS)
    vref, _ = O.dirty2vis(c["uvw"], c["freq"], model, c["wgt"], pixsize_x=pix, pixsize_y=pix,
                          epsilon=EPS)
    e = errors(vis, vref)
    report("small dirty2vis vs oracle", e)
    assert e[1] < 2e-5 and e[0] < 1e-5
    idx = torch.randint(0, vis.size, (100_000,), device=dev, generator=g)
    nchan = c["freq"].size
    ref = vis_of_points(tu, tf, idx // nchan, idx % nchan, torch.from_numpy(pi).to(dev),
                        torch.from_numpy(pj).to(dev), torch.from_numpy(pf).to(dev), n, n, pix, pix)
    e = errors(vis.reshape(-1)[idx.cpu().numpy()], ref.cpu().numpy())
    report("small dirty2vis vs direct transform, 1e5 visibilities", e)
    assert e[1] < 10 * EPS


def test_large_plan_on_a_row_subset_against_the_oracle():
    from rang import core_wgridder as W
    from rang import synth

    # the plan of the whole 464M-visibility call (kernel row, grid), forced on
    # 6 of its 900 time steps: the w-planes follow the subset's w range
    kidx, nu = 226, 14784
    enu = synth.antennas(64, 8000.0, 1234)
    tracks = synth.uvw_tracks(synth.baselines_xyz(enu, -30.7), 900, 8.0, -45.0).reshape(900, -1, 3)
    uvw = np.ascontiguousarray(tracks[::150][:6].reshape(-1, 3))
    freq = np.linspace(0.9e9, 1.7e9, 256)
    sky = synth.point_sky(20, 0.5, 1e-3, 1.0, 1234 + 7)
    print(f"[configs] large plan on {uvw.shape[0]} rows x 256 channels", flush=True)
    vis = synth.visibilities(uvw, freq, sky, 1e-3, 1234)
    wgt = np.ones(vis.shape, np.float32)
    n, pix = 8192, synth.ARCSEC
    kw = dict(npix_x=n, npix_y=n, pixsize_x=pix, pixsize_y=pix, epsilon=EPS, kernel=(kidx, nu, nu))
    dirty, info = W.vis2dirty(uvw=uvw, freq=freq, vis=vis, wgt=wgt, return_info=True, **kw)
    assert (info["kernel_index"], info["nu"], info["nv"]) == (kidx, nu, nu)
    ref, rinfo = O.vis2dirty(uvw, freq, vis, wgt, single=True, **kw)
    assert int(rinfo[4]) == info["nplanes"]
    e = errors(dirty, ref["dirty"])
    report(f"large plan vis2dirty vs oracle ({info['nplanes']} planes)", e)
    assert e[1] < 2e-5 and e[0] < 1e-5
    img = dirty / np.abs(dirty).max()
    del ref
    got = W.dirty2vis(uvw=uvw, freq=freq, dirty=img, wgt=wgt, pixsize_x=pix, pixsize_y=pix,
                      epsilon=EPS, kernel=(kidx, nu, nu))
    vref, _ = O.dirty2vis(uvw, freq, img, wgt, pixsize_x=pix, pixsize_y=pix, epsilon=EPS,
                          kernel=(kidx, nu, nu))
    e = errors(got, vref)
    report("large plan dirty2vis vs oracle", e)
    assert e[1] < 2e-5 and e[0] < 1e-5
