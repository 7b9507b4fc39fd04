"""The exact gridder at the full size of the scaled-up workload (464,486,400
visibilities, 8192^2 at 1 arcsec, epsilon 1e-5, single precision; grid
14784^2, 26 w-planes) through size-independent checks, all on the GPU:

* 256 random pixels of vis2dirty against the FP64 direct transform over
  every visibility (torch, float64, on the device);
* 100,000 random visibilities of dirty2vis of a 20-source model image
  against the FP64 direct transform of that image;
* adjointness Re<A x, v> = <x, A^H v> for a random image and random data.

Tolerance: the reference's gate, 10 epsilon of the peak (L-infinity over
the sampled values / the peak), and 1e-4 relative for adjointness in single
precision."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

C_LIGHT = 299792458.0
EPS = 1e-5


@pytest.fixture(scope="module")
def large():
    from rang import synth

    dev = torch.device("cuda:0")
    case = synth.config2_device(dev)
    print(f"[large] visibilities {tuple(case['vis'].shape)}", flush=True)
    yield case
    del case
    torch.cuda.empty_cache()


def opts(case):
    n, pix = case["npix"], case["pixsize"]
    return [n, n, pix, pix, EPS, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]


def lm_of(i, j, n, pix):
    # pixel (i, j) at l = (i - n/2) pix, m = (j - n/2) pix (core/src/wgridder.rs:30-36)
    l = (i.to(torch.float64) - n // 2) * pix
    m = (j.to(torch.float64) - n // 2) * pix
    return l, m


@pytest.mark.parametrize("single", [True, False])
def test_vis2dirty_pixels_match_the_direct_transform(large, single):
    # the double-precision variant (grid 14000^2) takes the two-pass transforms
    from rang import device as D

    c = large
    wgt = c["wgt"] if single else c["wgt"].to(torch.float64)
    d, _, info = D.vis2dirty(opts(c), c["uvw"], c["freq"], c["vis"], wgt, single=single)
    assert int(info[5]) == c["vis"].numel()
    n, pix = c["npix"], c["pixsize"]
    g = torch.Generator(device="cuda:0").manual_seed(5)
    i = torch.randint(0, n, (256,), device="cuda:0", generator=g)
    j = torch.randint(0, n, (256,), device="cuda:0", generator=g)
    # the pixel nearest the brightest source, so the sampled peak is the
    # image's peak
    flux, sl, sm = c["sky"]
    b = int(np.argmax(flux))
    i[0] = int(round(n // 2 + sl[b] / pix))
    j[0] = int(round(n // 2 + sm[b] / pix))
    l, m = lm_of(i, j, n, pix)
    nm1 = torch.sqrt(1.0 - l * l - m * m) - 1.0
    foc = c["freq"] / C_LIGHT
    ref = torch.zeros(256, dtype=torch.float64, device="cuda:0")
    rows = c["uvw"].shape[0]
    for r0 in range(0, rows, 4096):
        u = c["uvw"][r0:r0 + 4096]
        v = c["vis"][r0:r0 + 4096].to(torch.complex128)
        # phase (rows, chan, pixel): dirty = sum Re[V exp(+2 pi i (u l + v m + w (n-1)) f/c)]
        path = (u[:, 0:1] * l[None, :] + u[:, 1:2] * m[None, :] + u[:, 2:3] * nm1[None, :])
        ph = 2.0 * math.pi * path[:, None, :] * foc[None, :, None]
        ref += (v[:, :, None].real * torch.cos(ph) - v[:, :, None].imag * torch.sin(ph)).sum(dim=(0, 1))
    got = d[i, j].to(torch.float64)
    peak = float(ref.abs().max())
    err = float((got - ref).abs().max()) / peak
    print(f"[large] vis2dirty ({'f32' if single else 'f64'}, grid {int(info[2])}) 256 pixels: "
          f"max error / peak {err:.2e}", flush=True)
    assert err < 10 * EPS


@pytest.mark.parametrize("single", [True, False])
def test_dirty2vis_of_point_sources_matches_the_direct_transform(large, single):
    from rang import device as D

    c = large
    n, pix = c["npix"], c["pixsize"]
    rng = np.random.default_rng(7)
    pi = rng.integers(n // 4, 3 * n // 4, 20)
    pj = rng.integers(n // 4, 3 * n // 4, 20)
    flux = rng.uniform(0.1, 1.0, 20)
    real = torch.float32 if single else torch.float64
    img = torch.zeros((n, n), dtype=real, device="cuda:0")
    img[torch.from_numpy(pi).cuda(), torch.from_numpy(pj).cuda()] = torch.from_numpy(flux).cuda().to(real)
    vis, _ = D.dirty2vis(opts(c), c["uvw"], c["freq"], img)
    g = torch.Generator(device="cuda:0").manual_seed(9)
    idx = torch.randint(0, vis.numel(), (100_000,), device="cuda:0", generator=g)
    nchan = c["freq"].numel()
    row, ch = idx // nchan, idx % nchan
    uvw = c["uvw"][row]
    foc = (c["freq"] / C_LIGHT)[ch]
    l, m = lm_of(torch.from_numpy(pi).cuda(), torch.from_numpy(pj).cuda(), n, pix)
    nm1 = torch.sqrt(1.0 - l * l - m * m) - 1.0
    path = uvw[:, 0:1] * l[None, :] + uvw[:, 1:2] * m[None, :] + uvw[:, 2:3] * nm1[None, :]
    ph = -2.0 * math.pi * path * foc[:, None]
    f = torch.from_numpy(flux).cuda()[None, :]
    ref = torch.complex((f * torch.cos(ph)).sum(1), (f * torch.sin(ph)).sum(1))
    got = vis.reshape(-1)[idx].to(torch.complex128)
    err = float((got - ref).abs().max()) / float(ref.abs().max())
    print(f"[large] dirty2vis ({'f32' if single else 'f64'}) 1e5 visibilities: max error / peak {err:.2e}",
          flush=True)
    assert err < 10 * EPS


def test_adjointness(large):
    from rang import device as D

    c = large
    n = c["npix"]
    g = torch.Generator(device="cuda:0").manual_seed(11)
    x = torch.randn((n, n), dtype=torch.float32, device="cuda:0", generator=g)
    v = torch.complex(torch.randn(c["vis"].shape, device="cuda:0", generator=g),
                      torch.randn(c["vis"].shape, device="cuda:0", generator=g))
    ax, _ = D.dirty2vis(opts(c), c["uvw"], c["freq"], x)
    ahv, _, _ = D.vis2dirty(opts(c), c["uvw"], c["freq"], v, single=True)
    lhs = float((ax.to(torch.complex128) * v.to(torch.complex128).conj()).real.sum())
    rhs = float((x.to(torch.float64) * ahv.to(torch.float64)).sum())
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"[large] adjointness {rel:.2e}", flush=True)
    assert rel < 1e-4
