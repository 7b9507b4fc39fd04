"""Seeded synthetic interferometer data for the gridder workloads.

An array of antennas at uniform-random positions in a disk (local east,
north), Earth-rotation synthesis of the baselines' uvw tracks (the rotation
of examples/bench_df2.py:98-121 in rang-core), and Stokes-I visibilities of
point sources by a direct sum, plus complex Gaussian noise. No reference code
runs here; the recipe is the workload's description.
"""

import numpy as np

C_LIGHT = 299792458.0
OMEGA_EARTH = 7.2921159e-5  # rad/s
ARCSEC = np.pi / (180.0 * 3600.0)


def antennas(nant, radius_m, seed):
    """(nant, 3) east, north, up positions uniform in a disk."""
    rng = np.random.default_rng(seed)
    r = radius_m * np.sqrt(rng.uniform(0.0, 1.0, nant))
    phi = rng.uniform(0.0, 2.0 * np.pi, nant)
    return np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(nant)], axis=1)


def baselines_xyz(enu, lat_deg):
    """Baseline vectors (i < j) in the equatorial frame (X to the meridian at
    the equator, Z to the pole) from local east-north-up positions."""
    lat = np.deg2rad(lat_deg)
    i1, i2 = np.triu_indices(enu.shape[0], 1)
    b = enu[i2] - enu[i1]
    e, n, u = b[:, 0], b[:, 1], b[:, 2]
    x = -np.sin(lat) * n + np.cos(lat) * u
    y = e
    z = np.cos(lat) * n + np.sin(lat) * u
    return np.stack([x, y, z], axis=1)


def uvw_tracks(bxyz, ntime, dt_s, dec_deg):
    """(ntime * nbl, 3) uvw in metres, rows time-major, hour angles centred
    on transit."""
    t = (np.arange(ntime) - 0.5 * (ntime - 1)) * dt_s
    ha = OMEGA_EARTH * t
    dec = np.deg2rad(dec_deg)
    sh, ch = np.sin(ha)[:, None], np.cos(ha)[:, None]
    x, y, z = bxyz[None, :, 0], bxyz[None, :, 1], bxyz[None, :, 2]
    u = sh * x + ch * y
    v = -np.sin(dec) * ch * x + np.sin(dec) * sh * y + np.cos(dec) * z
    w = np.cos(dec) * ch * x - np.cos(dec) * sh * y + np.sin(dec) * z
    return np.stack([u, v, w], axis=-1).reshape(-1, 3)


def point_sky(nsrc, radius_deg, flux_lo, flux_hi, seed):
    """Fluxes log-uniform in [flux_lo, flux_hi], positions (l, m) uniform in a
    disk of radius_deg."""
    rng = np.random.default_rng(seed)
    flux = np.exp(rng.uniform(np.log(flux_lo), np.log(flux_hi), nsrc))
    r = np.deg2rad(radius_deg) * np.sqrt(rng.uniform(0.0, 1.0, nsrc))
    phi = rng.uniform(0.0, 2.0 * np.pi, nsrc)
    return flux, r * np.cos(phi), r * np.sin(phi)


def visibilities(uvw, freq, sky, noise_jy, seed, dtype=np.complex64):
    """V = sum_s F_s exp(-2 pi i (u l + v m + w (n-1))) f/c plus complex
    Gaussian noise of noise_jy per component; (nrow, nchan)."""
    flux, l, m = sky
    f = np.asarray(freq) / C_LIGHT
    vis = np.zeros((uvw.shape[0], f.size), np.complex128)
    for s in range(flux.size):
        n1 = np.sqrt(1.0 - l[s] ** 2 - m[s] ** 2) - 1.0
        path = uvw[:, 0] * l[s] + uvw[:, 1] * m[s] + uvw[:, 2] * n1
        vis += flux[s] * np.exp(-2j * np.pi * path[:, None] * f[None, :])
    rng = np.random.default_rng(seed + 1)
    vis += noise_jy * (rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape))
    return vis.astype(dtype)


def imaging_case(ntime, nchan, f_lo, f_hi, npix, pix_arcsec, *, nant=64, radius_m=8000.0,
                 lat_deg=-30.7, dec_deg=-45.0, dt_s=8.0, nsrc=20, sky_radius_deg=0.5,
                 flux=(1e-3, 1.0), noise_jy=1e-3, seed=1234, dtype=np.complex64):
    enu = antennas(nant, radius_m, seed)
    uvw = uvw_tracks(baselines_xyz(enu, lat_deg), ntime, dt_s, dec_deg)
    freq = np.linspace(f_lo, f_hi, nchan)
    sky = point_sky(nsrc, sky_radius_deg, flux[0], flux[1], seed + 7)
    vis = visibilities(uvw, freq, sky, noise_jy, seed, dtype)
    return {
        "uvw": uvw,
        "freq": freq,
        "vis": vis,
        "wgt": np.ones(vis.shape, np.float32 if dtype == np.complex64 else np.float64),
        "npix": npix,
        "pixsize": pix_arcsec * ARCSEC,
        "sky": sky,
    }


def config1(seed=1234, ntime=100):
    """64 antennas within 8 km at latitude -30.7 deg, 2016 baselines x 100
    times of 8 s at dec -45 deg x 16 channels over 1.30-1.45 GHz (3,225,600
    complex64 visibilities), 20 point sources, 1 mJy noise, unit weights,
    1024^2 image at 2 arcsec, epsilon 1e-5."""
    case = imaging_case(ntime, 16, 1.30e9, 1.45e9, 1024, 2.0, seed=seed)
    case["epsilon"] = 1e-5
    return case


def imaging_case_device(ntime, nchan, f_lo, f_hi, npix, pix_arcsec, device, *, nant=64,
                        radius_m=8000.0, lat_deg=-30.7, dec_deg=-45.0, dt_s=8.0, nsrc=20,
                        sky_radius_deg=0.5, flux=(1e-3, 1.0), noise_jy=1e-3, seed=1234,
                        rows_per_chunk=32768, freq=None):
    """The recipe of imaging_case with the visibilities formed on a GPU (for
    workloads whose direct sum is too large for the host): the same array,
    tracks and sky; the sum in float64 on the device, stored complex64; the
    noise from torch's generator seeded with seed + 1 (so its values differ
    from the host recipe's). Returns torch tensors on `device` and the sky."""
    import torch

    enu = antennas(nant, radius_m, seed)
    uvw = uvw_tracks(baselines_xyz(enu, lat_deg), ntime, dt_s, dec_deg)
    freq = np.linspace(f_lo, f_hi, nchan) if freq is None else np.asarray(freq, np.float64)
    nchan = freq.size
    flux_s, l_s, m_s = point_sky(nsrc, sky_radius_deg, flux[0], flux[1], seed + 7)
    d_uvw = torch.from_numpy(uvw).to(device)
    f = torch.from_numpy(freq / C_LIGHT).to(device)
    vis = torch.empty((uvw.shape[0], nchan), dtype=torch.complex64, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed + 1)
    for r0 in range(0, uvw.shape[0], rows_per_chunk):
        r1 = min(uvw.shape[0], r0 + rows_per_chunk)
        u = d_uvw[r0:r1]
        acc = torch.zeros((r1 - r0, nchan), dtype=torch.complex128, device=device)
        for s in range(flux_s.size):
            n1 = np.sqrt(1.0 - l_s[s] ** 2 - m_s[s] ** 2) - 1.0
            path = u[:, 0] * l_s[s] + u[:, 1] * m_s[s] + u[:, 2] * n1
            ph = (-2.0 * np.pi) * path[:, None] * f[None, :]
            acc += flux_s[s] * torch.polar(torch.ones_like(ph), ph)
        noise = torch.randn((r1 - r0, nchan, 2), dtype=torch.float64, device=device, generator=gen)
        acc += noise_jy * torch.complex(noise[..., 0], noise[..., 1])
        vis[r0:r1] = acc.to(torch.complex64)
    return {
        "uvw": d_uvw,
        "freq": torch.from_numpy(freq).to(device),
        "vis": vis,
        "wgt": torch.ones((uvw.shape[0], nchan), dtype=torch.float32, device=device),
        "npix": npix,
        "pixsize": pix_arcsec * ARCSEC,
        "sky": (flux_s, l_s, m_s),
    }


def config2_device(device, seed=1234, ntime=900, nchan=256):
    """The scaled-up workload: the same array and sky recipe, 2016 baselines x
    900 times of 8 s x 256 channels over 0.9-1.7 GHz (464,486,400 complex64
    visibilities), 8192^2 image at 1 arcsec, epsilon 1e-5; formed on the GPU."""
    case = imaging_case_device(ntime, nchan, 0.9e9, 1.7e9, 8192, 1.0, device, seed=seed)
    case["epsilon"] = 1e-5
    return case


def fullpol_case_device(ntime, nchan, f_lo, f_hi, npix, pix_arcsec, device, *, nant=64, radius_m=8000.0,
                        lat_deg=-30.7, dec_deg=-45.0, dt_s=8.0, nsrc=20, sky_radius_deg=1.4,
                        flux=(1e-3, 1.0), noise_jy=1e-3, seed=1234, rows_per_chunk=16384):
    """Four correlations XX, XY, YX, YY per visibility (complex64, (nrow,
    nchan, 4)) of an unpolarised point-source sky spread over the field: XX =
    YY = I + noise, XY = YX = noise (independent complex Gaussian noise of
    sigma `noise_jy` in each), the direct sum in float64 on the device; the
    same array and tracks as imaging_case_device. Also unit weights per
    correlation (float32, (nrow, nchan, 4)) and no flags (uint8). Returns
    torch tensors on `device`, the sky, and the largest |w| in wavelengths."""
    import torch

    enu = antennas(nant, radius_m, seed)
    uvw = uvw_tracks(baselines_xyz(enu, lat_deg), ntime, dt_s, dec_deg)
    freq = np.linspace(f_lo, f_hi, nchan)
    flux_s, l_s, m_s = point_sky(nsrc, sky_radius_deg, flux[0], flux[1], seed + 7)
    d_uvw = torch.from_numpy(uvw).to(device)
    f = torch.from_numpy(freq / C_LIGHT).to(device)
    nrow = uvw.shape[0]
    data = torch.empty((nrow, nchan, 4), dtype=torch.complex64, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed + 1)
    for r0 in range(0, nrow, rows_per_chunk):
        r1 = min(nrow, r0 + rows_per_chunk)
        u = d_uvw[r0:r1]
        acc = torch.zeros((r1 - r0, nchan), dtype=torch.complex128, device=device)
        for s in range(flux_s.size):
            n1 = np.sqrt(1.0 - l_s[s] ** 2 - m_s[s] ** 2) - 1.0
            path = u[:, 0] * l_s[s] + u[:, 1] * m_s[s] + u[:, 2] * n1
            ph = (-2.0 * np.pi) * path[:, None] * f[None, :]
            acc += flux_s[s] * torch.polar(torch.ones_like(ph), ph)
        noise = noise_jy * torch.randn((r1 - r0, nchan, 4, 2), dtype=torch.float32, device=device, generator=gen)
        block = torch.complex(noise[..., 0], noise[..., 1])
        sky = acc.to(torch.complex64)
        block[:, :, 0] += sky
        block[:, :, 3] += sky
        data[r0:r1] = block
    return {
        "uvw": d_uvw,
        "freq": torch.from_numpy(freq).to(device),
        "data": data,
        "weight": torch.ones((nrow, nchan, 4), dtype=torch.float32, device=device),
        "flag": torch.zeros((nrow, nchan, 4), dtype=torch.uint8, device=device),
        "npix": npix,
        "pixsize": pix_arcsec * ARCSEC,
        "sky": (flux_s, l_s, m_s),
        "wmax_lambda": float(np.abs(uvw[:, 2]).max() * freq.max() / C_LIGHT),
    }


def flag_runs(ntime, nbl, nchan, fraction, rng, max_time=8):
    """(ntime * nbl, nchan) bool flags: rectangles of contiguous time steps
    (1..max_time, all baselines of those steps) by contiguous channels
    (1..nchan/4) added until at least `fraction` of the cells is flagged."""
    f = np.zeros((ntime, nchan), bool)
    target = fraction * ntime * nchan
    while f.sum() < target:
        t0 = int(rng.integers(ntime))
        c0 = int(rng.integers(nchan))
        f[t0:t0 + int(rng.integers(1, max_time + 1)), c0:c0 + int(rng.integers(1, max(nchan // 4, 1) + 1))] = True
    return np.repeat(f, nbl, axis=0)


def config3_device(device, seed=1234, nimages=48, ntime=32, pix_arcsec=1.5, robust=0.0):
    """Batched uneven sub-band imaging: `nimages` images with sides drawn from
    512, 1024, 2048 and 4096, each with its own 8-128 channels (a contiguous
    sub-band of 0.9-1.7 GHz at 0.78125 MHz spacing) over the same 2016
    baselines x `ntime` times of 8 s, a flagged fraction uniform in 0-60 %
    in contiguous time/channel runs, and float32 weights: the Briggs
    (robust 0) imaging weight of each sample on the image's padded grid
    (rang.core_weights, on the GPU) times its visibility weight (0 where
    flagged, else 1). Returns a list of per-image cases (device tensors)."""
    import torch

    from . import core_weights

    rng = np.random.default_rng(seed)
    sides = rng.choice([512, 1024, 2048, 4096], nimages)
    df = 0.8e9 / 1024
    cases = []
    for k in range(nimages):
        nchan = int(rng.integers(8, 129))
        c0 = int(rng.integers(0, 1024 - nchan))
        freq = 0.9e9 + (c0 + np.arange(nchan)) * df
        c = imaging_case_device(ntime, nchan, 0.0, 0.0, int(sides[k]), pix_arcsec, device, seed=seed + 101 * k,
                                freq=freq)
        nbl = c["uvw"].shape[0] // ntime
        frac = float(rng.uniform(0.0, 0.6))
        flags = flag_runs(ntime, nbl, nchan, frac, rng)
        vis_w = (~flags).astype(np.float32)
        npad = core_weights.padded_size(int(sides[k]))
        uvw_h = c["uvw"].cpu().numpy()
        res = core_weights.imaging_weights(uvw_h, core_weights.wsclean_wavelengths(freq), vis_w,
                                           np.ones(uvw_h.shape[0], bool), mode="briggs", robust=robust, width=npad,
                                           height=npad, pixel_scale=c["pixsize"], return_grid=False)
        c["wgt"] = torch.from_numpy(vis_w * res["weights"]).to(device)
        c["flag_fraction"] = float(flags.mean())
        c["side"] = int(sides[k])
        cases.append(c)
    return cases
