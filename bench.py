#!/usr/bin/env python3
"""Throughput of the exact w-gridder on one MI355X: vis2dirty + dirty2vis.

Workload (the scaled-up configuration): 64 antennas within 8 km (2016
baselines) x 900 times of 8 s x 256 channels over 0.9-1.7 GHz = 464,486,400
complex64 visibilities from 20 point sources plus 1 mJy noise, unit float32
weights, an 8192 x 8192 image at 1 arcsec, epsilon 1e-5, w-gridding on,
single precision. The visibilities are formed on the GPU (a float64 direct
sum). One step is one vis2dirty of the visibilities followed by one dirty2vis
of the resulting image, each a full call (scan, plan, sort, grid/degrid,
transforms) on inputs resident in HBM. value = visibilities per step x steps
/ wall time of the timed steps, in 10^6 visibilities per second.

The JSON line also carries: the roofline of the step's dominant kernel
(algorithmic bytes of one launch over its HIP-event duration on the call's
stream, against 8 TB/s); the oracle (the CPU restatement of the reference
gridder, OpenMP) timed on this host's cores on a bounded sample of the same
workload and scaled to it; the same measurements on the small workload
(3,225,600 visibilities, 1024 x 1024 at 2 arcsec) under "config_small";
under "sub_band_batch", 48 uneven sub-band images in batch calls (one
gridding launch per call) on tuned plans at epsilon 1e-5 and 1e-4
(--no-batch skips it); under "gram_solve", the exact Gram applications and
seconds that plain CG, the inverse-filter PCG and the inner-outer solve (with
and without the learned warm start) need to reach a relative residual of
1e-4 on a 2048^2 image through 19.4 M visibilities (--no-gram skips it);
under "ingest", blosc-1 zarr chunks expanded on the GPU (the MSv4 ingest
path; --no-ingest skips it); under "published", the
exact gridder and the w-stacking gridder on a 23.8 M-visibility 1024^2 case
beside the reference's published timings for that size (--no-published
skips it).
"""

import argparse
import gc
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "rang-core_amd"))

METRIC = "Mvis/s vis2dirty+dirty2vis at eps=1e-5; Gram applies to tol, learned vs CG"
HBM_PEAK_GBS = 8000.0


def stage_bytes(name, direction, info, nvis, nrow, nchan, npix):
    """Algorithmic HBM bytes of one launch of a stage's kernel (single
    precision): what the kernel must read and write at least once. Gridding:
    the visibilities and weights it consumes (12 bytes each, through the
    sorted records) and the planes written once; the transforms: one read
    and one write of the data they transform, pruned to the image columns;
    degridding: the planes read once, the weights read and the visibilities
    written."""
    nu, nv, P = int(info[2]), int(info[3]), max(int(info[4]), 1)
    nx = ny = npix
    cx, real = 8, 4
    grid = nu * nv * cx
    if direction == "v2d":
        table = {
            "grid": nvis * 12 + P * grid,
            "fft_v": P * grid + P * nu * ny * cx,
            "fft_u": P * nu * ny * cx + P * nx * ny * real,
            "sum": P * nx * ny * real + nx * ny * real,
        }
    else:
        table = {
            "fft_u": nx * ny * real + P * nu * ny * cx,
            "fft_v": P * nu * ny * cx + P * grid,
            "degrid": P * grid + nvis * (4 + cx),
        }
    return table.get(name)


def call_bytes(info, nvis, nrow, nchan, npix):
    """Algorithmic bytes of one vis2dirty call as the reference moves data
    (one plane written, one v pass read and written, the u pass over the
    image columns, the image updated per plane)."""
    nu, nv, P = int(info[2]), int(info[3]), max(int(info[4]), 1)
    G = nu * nv * 8
    img = npix * npix * 4
    return nvis * 12 + nrow * 24 + nchan * 8 + P * (G * (3 + npix / nv) + 2 * img) + 2 * img


def run_gpu(case, steps, warmup, verbose=False):
    import numpy as np
    import torch

    from rang import device as D

    uvw, freq, vis, wgt = case["uvw"], case["freq"], case["vis"], case["wgt"]
    npix, pix, eps = case["npix"], case["pixsize"], case["epsilon"]
    opts = [npix, npix, pix, pix, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    dev = vis.device
    dirty = torch.empty((npix, npix), dtype=torch.float32, device=dev)
    out_vis = torch.empty_like(vis)

    def step(timed_stages=None):
        _, _, info = D.vis2dirty(opts, uvw, freq, vis, wgt, single=True, out=(dirty, None))
        if timed_stages is not None:
            timed_stages.append(("v2d", D.stage_times()))
        D.dirty2vis(opts, uvw, freq, dirty, wgt, out=out_vis)
        if timed_stages is not None:
            timed_stages.append(("d2v", D.stage_times()))
        return info

    for _ in range(warmup):
        info = step()
    D.stage_timing(True)
    stages = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        info = step(stages)
        if verbose:
            progress(f"step {k + 1}/{steps}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    D.stage_timing(False)
    acc = {}
    for direction, st in stages:
        for name, ms in st:
            acc.setdefault((direction, name), []).append(ms)
    means = {k: statistics.fmean(v) for k, v in acc.items()}
    return elapsed, np.asarray(info), means


def run_prepared(case, steps, warmup, verbose=False):
    """vis2dirty + dirty2vis of the workload on a prepared plan (the plan and
    sorted layouts made once, outside the timed steps; rang-core's
    wgridder::Prepared): each step one Prepared::vis2dirty of the
    visibilities and one weighted Prepared::dirty2vis of the resulting image,
    replayed as graphs after the first two. Returns (elapsed, seconds to
    prepare, per-stage means of one more step run eagerly with stage timing,
    the plan, the image and the visibilities of the last step)."""
    import torch

    from rang import device as D
    from rang.prepared import Prepared

    uvw, freq, vis, wgt = case["uvw"], case["freq"], case["vis"], case["wgt"]
    npix, pix, eps = case["npix"], case["pixsize"], case["epsilon"]
    opts = [npix, npix, pix, pix, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = Prepared(opts, uvw, freq, wgt, single=True)
    torch.cuda.synchronize()
    t_prepare = time.perf_counter() - t0
    dirty = torch.empty((npix, npix), dtype=torch.float32, device=vis.device)
    out_vis = torch.empty_like(vis)

    def step():
        plan.vis2dirty(vis, out=dirty)
        plan.dirty2vis(dirty, weighted=True, out=out_vis)

    for _ in range(max(warmup, 2)):  # the first call of each kind eager, the second captured
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step()
        if verbose:
            progress(f"prepared step {k + 1}/{steps}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    D.stage_timing(True)
    means = {}
    plan.vis2dirty(vis, out=dirty)
    for name, ms in D.stage_times():
        means[("v2d", name)] = ms
    plan.dirty2vis(dirty, weighted=True, out=out_vis)
    for name, ms in D.stage_times():
        means[("d2v", name)] = ms
    D.stage_timing(False)
    torch.cuda.synchronize()
    return elapsed, t_prepare, means, plan, dirty, out_vis


def dft_errors(case, dirty, vis_of, n_pix=256, n_vis=100_000, seed=5):
    """(L2, Linf/peak) of vis2dirty on `n_pix` random pixels (the brightest
    source's pixel first) against the FP64 direct sum over every visibility,
    and of dirty2vis of a 20-point image (`vis_of(image)` the gridder's
    visibilities) on `n_vis` random visibilities against its direct sum."""
    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "tests"))
    from direct import dirty_pixels, errors, vis_of_points

    n, pix = case["npix"], case["pixsize"]
    dev = case["vis"].device
    g = torch.Generator(device=dev).manual_seed(seed)
    i = torch.randint(0, n, (n_pix,), device=dev, generator=g)
    j = torch.randint(0, n, (n_pix,), device=dev, generator=g)
    flux, sl, sm = case["sky"]
    b = int(np.argmax(flux))
    i[0], j[0] = int(round(n // 2 + sl[b] / pix)), int(round(n // 2 + sm[b] / pix))
    ref = dirty_pixels(case["uvw"], case["freq"], case["vis"], case["wgt"], i, j, n, n, pix, pix)
    e_v2d = errors(dirty[i, j].double().cpu().numpy(), ref.cpu().numpy())
    rng = np.random.default_rng(seed)
    pi, pj = rng.integers(n // 4, 3 * n // 4, 20), rng.integers(n // 4, 3 * n // 4, 20)
    pf = rng.uniform(0.1, 1.0, 20)
    img = torch.zeros((n, n), dtype=dirty.dtype, device=dev)
    img[torch.from_numpy(pi).to(dev), torch.from_numpy(pj).to(dev)] = torch.from_numpy(pf).to(dev).to(dirty.dtype)
    vis = vis_of(img)
    idx = torch.randint(0, vis.numel(), (n_vis,), device=dev, generator=g)
    nchan = case["freq"].numel()
    refv = vis_of_points(case["uvw"], case["freq"], idx // nchan, idx % nchan, torch.from_numpy(pi).to(dev),
                         torch.from_numpy(pj).to(dev), torch.from_numpy(pf).to(dev), n, n, pix, pix)
    refv = refv * case["wgt"].reshape(-1)[idx].double()
    e_d2v = errors(vis.reshape(-1)[idx].to(torch.complex128).cpu().numpy(), refv.cpu().numpy())
    return {
        "vis2dirty_vs_fp64_dft": {"l2": e_v2d[0], "linf_over_peak": e_v2d[1], "sample": f"{n_pix} pixels"},
        "dirty2vis_vs_fp64_dft": {"l2": e_d2v[0], "linf_over_peak": e_d2v[1],
                                  "sample": f"{n_vis} visibilities of a 20-point image"},
    }


def run_f64(case, steps):
    """The scaled-up workload in double precision: one-shot vis2dirty +
    dirty2vis calls (complex128 visibilities, float64 weights), timed over
    `steps` after one warm-up, with the direct-transform errors."""
    import torch

    from rang import device as D

    npix, pix, eps = case["npix"], case["pixsize"], case["epsilon"]
    opts = [npix, npix, pix, pix, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    c64 = dict(case)
    c64["vis"] = case["vis"].to(torch.complex128)
    c64["wgt"] = case["wgt"].to(torch.float64)
    dirty = torch.empty((npix, npix), dtype=torch.float64, device=c64["vis"].device)
    out_vis = torch.empty_like(c64["vis"])

    def step():
        _, _, info = D.vis2dirty(opts, c64["uvw"], c64["freq"], c64["vis"], c64["wgt"], single=False, out=(dirty, None))
        D.dirty2vis(opts, c64["uvw"], c64["freq"], dirty, c64["wgt"], out=out_vis)
        return info

    info = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        info = step()
        progress(f"f64 step {k + 1}/{steps}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    errs = dft_errors(c64, dirty, lambda img: D.dirty2vis(opts, c64["uvw"], c64["freq"], img, c64["wgt"])[0])
    nvis = c64["vis"].numel()
    del c64, out_vis
    torch.cuda.empty_cache()
    return {
        "value": round(nvis * steps / elapsed / 1e6, 2),
        "unit": "Mvis/s",
        "ms_per_step": round(elapsed / steps * 1e3, 1),
        "grid": [int(info[2]), int(info[3])],
        "support": int(info[1]),
        "w_planes": int(info[4]),
        "calls": "one-shot vis2dirty + dirty2vis (scan, plan, sort, grid, transforms each call)",
        "errors": errs,
    }


def run_host_arrays(case, steps=1):
    """vis2dirty + dirty2vis through the host-array entry points of rang._core
    (the reference's boundary): the arrays cross PCIe both ways inside the
    timed region. Returns Mvis/s and seconds per step."""
    import numpy as np

    from rang import _core

    npix, pix, eps = case["npix"], case["pixsize"], case["epsilon"]
    opts = [npix, npix, pix, pix, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    uvw, freq = case["uvw"].cpu().numpy(), case["freq"].cpu().numpy()
    vis, wgt = case["vis"].cpu().numpy(), case["wgt"].cpu().numpy()
    dirty = np.zeros((npix, npix), np.float32)
    out = np.zeros(vis.shape, np.complex64)
    info = np.zeros(_core.N_INFO)
    times = []
    for _ in range(steps + 1):
        t0 = time.perf_counter()
        st = _core.wgridder_vis2dirty_f32(opts, uvw, freq, vis, wgt, None, None, dirty, None, info)
        st |= _core.wgridder_dirty2vis_f32(opts, uvw, freq, dirty, wgt, None, out.view(np.float32), info)
        times.append(time.perf_counter() - t0)
        if st:
            raise RuntimeError(_core.wgridder_last_error())
    t = statistics.fmean(times[1:])
    return {"value": round(vis.size / t / 1e6, 2), "unit": "Mvis/s", "seconds_per_step": round(t, 3),
            "boundary": "rang_core_wgridder_vis2dirty_ex_f32 + dirty2vis_ex_f32 on host arrays (PCIe both ways)"}


def profiled_traffic(workload, kernel, path=None):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE scaled by a plain
    copy's reading), or None."""
    path = path or ROOT / "profiles" / f"wgridder_{workload}_traffic.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 4 and parts[0] == "traffic" and parts[1] == workload and parts[2] == kernel:
            return int(parts[3])
    return None


def summarise(case, nvis_total, nrow, nchan, steps, elapsed, info, means, workload=None):
    value = nvis_total * steps / elapsed / 1e6
    nvis = int(info[5])
    candidates = []
    for (direction, name), ms in means.items():
        b = stage_bytes(name, direction, info, nvis, nrow, nchan, case["npix"])
        if b is not None:
            candidates.append((ms, direction, name, b))
    ms, direction, name, b = max(candidates)
    achieved = b / (ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": f"{direction}:{name}",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": profiled_traffic(workload, f"{direction}:{name}") if workload else None,
        "algorithmic_bytes": int(b),
        "kernel_ms": round(ms, 4),
    }
    v2d_ms = sum(v for (d, _), v in means.items() if d == "v2d")
    call = call_bytes(info, nvis, nrow, nchan, case["npix"])
    per_call = {
        "vis2dirty_bytes": int(call),
        "vis2dirty_ms": round(v2d_ms, 3),
        "frac": round(call / (v2d_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    return value, roofline, per_call


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_small(case, repeats, gpu_dirty=None, gpu_vis=None):
    """The oracle on the whole small workload (median of `repeats`); with the
    GPU's image and visibilities (the dirty2vis of that image), also their
    errors against the oracle's: vis2dirty of the same data, dirty2vis of the
    same image. Returns (baseline, errors or None)."""
    import numpy as np

    from oracle import oracle as O

    npix, pix, eps = case["npix"], case["pixsize"], case["epsilon"]
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        imgs, _ = O.vis2dirty(case["uvw"], case["freq"], case["vis"], case["wgt"], npix_x=npix,
                              npix_y=npix, pixsize_x=pix, pixsize_y=pix, epsilon=eps, single=True)
        O.dirty2vis(case["uvw"], case["freq"], imgs["dirty"], case["wgt"], pixsize_x=pix,
                    pixsize_y=pix, epsilon=eps)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    errs = None
    if gpu_dirty is not None:
        sys.path.insert(0, str(ROOT / "tests"))
        from direct import errors

        e1 = errors(gpu_dirty, imgs["dirty"])
        vref, _ = O.dirty2vis(case["uvw"], case["freq"], np.ascontiguousarray(gpu_dirty), case["wgt"],
                              pixsize_x=pix, pixsize_y=pix, epsilon=eps)
        e2 = errors(gpu_vis, vref)
        errs = {"vis2dirty_vs_oracle": {"l2": e1[0], "linf_over_peak": e1[1], "sample": "every pixel"},
                "dirty2vis_vs_oracle": {"l2": e2[0], "linf_over_peak": e2[1], "sample": "every visibility"}}
    return {
        "value": round(case["vis"].size / t / 1e6, 3),
        "unit": "Mvis/s",
        "cores": _threads(),
        "kind": "port",
        "sample": f"the whole small workload ({case['vis'].size} visibilities), vis2dirty + "
                  f"dirty2vis, median of {repeats} runs of the oracle on {_threads()} threads",
    }, errs


def cpu_large(info, nvis_total):
    """The oracle's time for the large workload, from a bounded sample: its
    cost is a per-visibility part (scan, sort, gridding and degridding over S
    planes) plus a per-plane part (the plane's transforms both ways). The
    first is timed on rows of the workload (8 time steps spread over the
    track, all 2016 baselines and 256 channels) imaged at 1024^2 with the
    same kernel and oversampling, where the transforms are small; the second
    on one plane of the full 14784^2 grid (w-gridding off, one time step of
    rows), less that run's own per-visibility part. Both use the plan of
    the GPU run (kernel row, grid size); the total is N t_vis + P t_plane."""
    import numpy as np

    from oracle import oracle as O
    from rang import synth

    kidx, nu, nv, P = int(info[0]), int(info[2]), int(info[3]), int(info[4])
    row = O.kernel_row(kidx)
    sigma = row[1]
    enu = synth.antennas(64, 8000.0, 1234)
    tracks = synth.uvw_tracks(synth.baselines_xyz(enu, -30.7), 900, 8.0, -45.0).reshape(900, -1, 3)
    freq = np.linspace(0.9e9, 1.7e9, 256)
    sky = synth.point_sky(20, 0.5, 1e-3, 1.0, 1234 + 7)
    pix = synth.ARCSEC

    def sample(nt):
        uvw = np.ascontiguousarray(tracks[:: 900 // nt][:nt].reshape(-1, 3))
        vis = synth.visibilities(uvw, freq, sky, 1e-3, 1234)
        return uvw, vis, np.ones(vis.shape, np.float32)

    def both(uvw, vis, wgt, npix, kernel, wgrid):
        t0 = time.perf_counter()
        imgs, _ = O.vis2dirty(uvw, freq, vis, wgt, npix_x=npix, npix_y=npix, pixsize_x=pix,
                              pixsize_y=pix, epsilon=1e-5, single=True, kernel=kernel,
                              do_wgridding=wgrid)
        O.dirty2vis(uvw, freq, imgs["dirty"], wgt, pixsize_x=pix, pixsize_y=pix, epsilon=1e-5,
                    kernel=kernel, do_wgridding=wgrid)
        return time.perf_counter() - t0

    small_n = max(2 * O.good_size_complex(int(1024 * sigma * 0.5) + 1), 16)
    uvw8, vis8, wgt8 = sample(8)
    t_small = both(uvw8, vis8, wgt8, 1024, (kidx, small_n, small_n), True)
    t_vis = t_small / vis8.size
    uvw1, vis1, wgt1 = sample(1)
    t_flat = both(uvw1, vis1, wgt1, 8192, (kidx, nu, nv), False)
    t_small_flat = both(uvw1, vis1, wgt1, 1024, (kidx, small_n, small_n), False)
    t_plane = max(t_flat - t_small_flat, 0.0)
    total = nvis_total * t_vis + P * t_plane
    return {
        "value": round(nvis_total / total / 1e6, 3),
        "unit": "Mvis/s",
        "cores": _threads(),
        "kind": "port",
        "sample": f"oracle on {_threads()} threads, scaled from a sample of the workload: "
                  f"{vis8.size} visibilities (8 of 900 time steps) gridded and degridded at "
                  f"1024^2 with kernel row {kidx} ({t_vis * 1e9:.1f} ns per visibility), and one "
                  f"plane of the {nu}x{nv} grid transformed both ways ({t_plane:.2f} s per plane); "
                  f"estimate = N t_vis + {P} planes t_plane = {total:.1f} s per step",
    }


def published(dev, cpu_baseline=True):
    """The reference's own published timings at one size, beside this GPU's
    2016 baselines x 738 times x 16 channels =
    23,804,928 visibilities (the reference's whole-array cases have
    23,811,985), 1024^2 at 1.5 arcsec.

    * exact gridder, dirty image and PSF in one call, float64, epsilon 1e-6:
      1.93 s per image on 16 threads (ducc0 1.81 s), research/df2-wgridder-
      speed.md:290-291, 333-335; here the device call on HBM-resident inputs
      and the host-array C entry point (PCIe included);
    * w-stacking, 64 layers, kernel 7 x 1023, padding 1.2, dirty + PSF:
      2.55 s (f64) and 1.96 s (f32) per image on the reference's GPU path,
      wall including host preparation and transfers (research/x2-wgpu.md:
      421-424); here rang_core_grid_invert_gpu from host arrays, the same
      boundary."""
    import numpy as np
    import torch

    from rang import _core, synth
    from rang import device as D

    c = synth.imaging_case_device(738, 16, 1.30e9, 1.45e9, 1024, 1.5, dev)
    nrow, nchan = int(c["uvw"].shape[0]), int(c["freq"].shape[0])
    pix = c["pixsize"]
    opts = [1024, 1024, pix, pix, 1e-6, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    wgt64 = c["wgt"].to(torch.float64)
    vis128 = c["vis"].to(torch.complex128)
    times = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, info = D.vis2dirty(opts, c["uvw"], c["freq"], vis128, wgt64, single=False, psf=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t_dev = min(times[1:])
    # the host-array entry point (rang._core, as the reference's)
    uvw_h = c["uvw"].cpu().numpy()
    freq_h = c["freq"].cpu().numpy()
    vis_h = vis128.cpu().numpy()
    wgt_h = wgt64.cpu().numpy()
    dirty_h = np.zeros((1024, 1024))
    psf_h = np.zeros((1024, 1024))
    info_h = np.zeros(_core.N_INFO)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        st = _core.wgridder_vis2dirty_f64(opts, uvw_h, freq_h, vis_h, wgt_h, None, None, dirty_h, psf_h,
                                          info_h)
        times.append(time.perf_counter() - t0)
        if st != 0:
            raise RuntimeError(_core.wgridder_last_error())
    t_host = times[-1]
    progress(f"exact gridder f64 dirty+PSF: device {t_dev:.3f} s, host arrays {t_host:.3f} s")
    # the oracle on this host beside the reference's published 1- and
    # 16-thread times: all threads on the whole case, one thread on the
    # first quarter of the rows (its time scaled by 4: the gridding's share
    # dominates; the transforms of 7 planes of 2376^2 take well under a
    # second)
    oracle_cpu = None
    if cpu_baseline:
        from oracle import oracle as O

        def oracle_run(rows):
            t0 = time.perf_counter()
            O.vis2dirty(uvw_h[:rows], freq_h, vis_h[:rows], wgt_h[:rows], npix_x=1024, npix_y=1024,
                        pixsize_x=pix, pixsize_y=pix, epsilon=1e-6, single=False, outputs=("dirty", "psf"))
            return time.perf_counter() - t0

        threads = _threads()
        O.set_threads(threads)
        t_all = oracle_run(nrow)
        O.set_threads(1)
        quarter = nrow // 4
        t_one = oracle_run(quarter) * nrow / quarter
        O.set_threads(threads)
        oracle_cpu = {"threads": threads, "seconds_all_threads": round(t_all, 3),
                      "seconds_one_thread": round(t_one, 2),
                      "one_thread_sample": f"the first {quarter} of {nrow} rows, time scaled by {nrow / quarter:.2f}",
                      "reference_seconds_16_threads": 1.93, "reference_seconds_1_thread": 11.35}
        progress(f"oracle f64 dirty+PSF: {t_all:.2f} s on {threads} threads, {t_one:.1f} s on one (scaled)")
    data = np.ascontiguousarray(c["vis"].cpu().numpy().view(np.float32).reshape(nrow, nchan, 2))
    w32 = np.ascontiguousarray(c["wgt"].cpu().numpy().astype(np.float32))
    ws = {}
    for prec in ("f64", "f32"):
        d = np.zeros(1024 * 1024)
        p = np.zeros(1024 * 1024)
        stats = np.zeros(_core.GRID_N_STATS)
        timings = np.zeros(_core.GRID_N_TIMINGS)
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            st = _core.grid_invert_gpu(nrow, nchan, 1024, 1024, 7, 1023, 64, pix, pix, 1.2, uvw_h, freq_h, data,
                                       w32, w32, prec, d, p, stats, timings)
            t = time.perf_counter() - t0
            if st != 0:
                raise RuntimeError(_core.grid_last_error())
            best = t if best is None else min(best, t)
        ws[prec] = round(best, 3)
        progress(f"w-stacking {prec}: {best:.3f} s")
    del c, vis128, wgt64
    torch.cuda.empty_cache()
    return {
        "workload": "2016 baselines x 738 times x 16 channels (23,804,928 vis, 1.30-1.45 GHz), "
                    "1024x1024 at 1.5 arcsec, dirty + PSF",
        "wgridder_f64_eps1e-6": {
            "grid": [int(info[2]), int(info[3])], "support": int(info[1]), "w_planes": int(info[4]),
            "seconds_device": round(t_dev, 3), "seconds_host_arrays": round(t_host, 3),
            "reference_seconds_16_threads": 1.93, "ducc0_seconds_16_threads": 1.81,
            "speedup_host_arrays": round(1.93 / t_host, 1),
            "oracle_on_this_host": oracle_cpu,
        },
        "wstack_64_layers": {
            "seconds_f64": ws["f64"], "seconds_f32": ws["f32"],
            "reference_gpu_seconds_f64": 2.55, "reference_gpu_seconds_f32": 1.96,
            "speedup_f64": round(2.55 / ws["f64"], 1), "speedup_f32": round(1.96 / ws["f32"], 1),
        },
    }


def run_batch(dev, steps=3, ntime=100, eps_list=(1e-5, 1e-4), parity_images=3, dft_images=3):
    """Batched uneven sub-band imaging: 48 images of 512^2-4096^2 at 1.5
    arcsec, each with its own 8-128 channels of 0.9-1.7 GHz over 2016
    baselines x `ntime` times, 0-60 % flagged in contiguous runs, float32
    weights = Briggs (robust 0) imaging weights x visibility weights. Per
    epsilon: one batch call on the cost model's plans, one call per image,
    the tuner over every image (uncached: a fresh cache file), and the batch
    call on the tuned plans; value = input samples / wall time of a batch
    call. Errors: the `parity_images` images with the fewest samples against
    the oracle at the tuned plan, and `dft_images` images on 256 pixels
    against the FP64 direct sum."""
    import tempfile

    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "tests"))
    from direct import dirty_pixels, errors

    from oracle import oracle as O
    from rang import batch, synth
    from rang import device as D

    cases = synth.config3_device(dev, ntime=ntime)
    nsamp = sum(int(c["vis"].numel()) for c in cases)
    progress(f"sub-band batch formed: {len(cases)} images, {nsamp} samples")

    def opts(c, eps):
        return [c["side"], c["side"], c["pixsize"], c["pixsize"], eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    out = {
        "workload": f"{len(cases)} images, sides {{512, 1024, 2048, 4096}} at 1.5 arcsec, 8-128 channels each "
                    f"(0.78125 MHz spacing in 0.9-1.7 GHz), 2016 baselines x {ntime} times, 0-60 % flagged in "
                    "contiguous runs, Briggs robust 0 x visibility weights (float32), single precision",
        "images": len(cases),
        "sides": sorted(int(c["side"]) for c in cases),
        "samples": nsamp,
        "unit": "Mvis/s (input samples, flagged included)",
        "by_epsilon": {},
    }
    order = sorted(range(len(cases)), key=lambda k: cases[k]["vis"].numel())
    for eps in eps_list:
        images = [batch.Image(opts(c, eps), c["uvw"], c["freq"], c["vis"], c["wgt"]) for c in cases]
        outs = [torch.empty((c["side"], c["side"]), dtype=torch.float32, device=dev) for c in cases]
        t_def = timed(lambda: batch.vis2dirty(images, single=True, out=outs))
        active = int(sum(im.info[5] for im in images))

        def one_by_one():
            for im, o, c in zip(images, outs, cases):
                D.vis2dirty(im.opts, im.uvw, c["freq"], im.vis, im.wgt, single=True, out=(o, None))

        t_one = timed(one_by_one)
        with tempfile.TemporaryDirectory() as tmp:
            t0 = time.perf_counter()
            reps = [batch.tune(im, single=True, cache_path=os.path.join(tmp, "tuning.txt"), kernels=3)
                    for im in images]
            t_tune = time.perf_counter() - t0
        t_tuned = timed(lambda: batch.vis2dirty(images, single=True, out=outs))
        torch.cuda.synchronize()
        errs = []
        for k in order[:parity_images]:
            c, im = cases[k], images[k]
            h = {key: c[key].cpu().numpy() for key in ("uvw", "freq", "vis", "wgt")}
            ref, _ = O.vis2dirty(h["uvw"], h["freq"], h["vis"], h["wgt"], npix_x=c["side"], npix_y=c["side"],
                                 pixsize_x=c["pixsize"], pixsize_y=c["pixsize"], epsilon=eps, single=True,
                                 kernel=(int(im.info[0]), int(im.info[2]), int(im.info[3])))
            e = errors(outs[k].double().cpu().numpy(), ref["dirty"])
            errs.append({"image": k, "side": int(c["side"]), "samples": int(c["vis"].numel()),
                         "vs_oracle": {"l2": e[0], "linf_over_peak": e[1], "sample": "every pixel"}})
        for k in order[-dft_images:]:
            c = cases[k]
            n, pix = int(c["side"]), c["pixsize"]
            g = torch.Generator(device=dev).manual_seed(k)
            i = torch.randint(0, n, (256,), device=dev, generator=g)
            j = torch.randint(0, n, (256,), device=dev, generator=g)
            ref = dirty_pixels(c["uvw"], c["freq"], c["vis"], c["wgt"], i, j, n, n, pix, pix)
            e = errors(outs[k][i, j].double().cpu().numpy(), ref.cpu().numpy())
            errs.append({"image": k, "side": n, "samples": int(c["vis"].numel()),
                         "vs_fp64_dft": {"l2": e[0], "linf_over_peak": e[1], "sample": "256 pixels"}})
        out["by_epsilon"][f"{eps:g}"] = {
            "value": round(nsamp / t_tuned / 1e6, 2),
            "ms_per_batch": round(t_tuned * 1e3, 3),
            "active_visibilities": active,
            "untuned_batch": {"value": round(nsamp / t_def / 1e6, 2), "ms": round(t_def * 1e3, 3)},
            "one_call_per_image": {"value": round(nsamp / t_one / 1e6, 2), "ms": round(t_one * 1e3, 3)},
            "tuning_seconds": round(t_tune, 2),
            "tuned_choices": [{"kernel_index": int(im.opts[15]), "tile": int(im.tile),
                               "model_ms": round(r["default_ms"], 3), "tuned_ms": round(r["ms"], 3)}
                              for im, r in zip(images, reps)],
            "errors": errs,
        }
        progress(f"sub-band batch eps {eps:g}: {nsamp / t_tuned / 1e6:.1f} Mvis/s tuned, "
                 f"{nsamp / t_def / 1e6:.1f} untuned, {nsamp / t_one / 1e6:.1f} one call per image")
    return out


def gram_solve(dev, tol=1e-4, budget=4000, train_steps=60, tau=0.1):
    """The Gram-solve half of the metric at the learned-preconditioner
    workload: a 2048^2 image at 1.5 arcsec through 2016 baselines x 300
    times x 32 channels (19,353,600 visibilities, 1.30-1.45 GHz, unit
    weights), eps 1e-5, single precision; A^H W A x = b, b = A^H W A x_sky
    of a 300-point sky (log-uniform 0.1 mJy - 1 Jy), solved from zero to
    relative residual `tol` judged by the exact operator. Exact-operator
    applications and seconds for: plain CG; CG preconditioned by the
    circulant inverse filter; the inner-outer solve (flexible CG on the
    exact operator, each step a PCG solve to `tau` on the w-aware Toeplitz
    surrogate, which costs no gridding); and the same with the learned
    warm start (U-Net 4 levels x 32 base channels, BF16 on MFMA, trained
    `train_steps` steps of 4 skies on the surrogate's images and on residuals
    of the filter-preconditioned solve)."""
    import math

    import numpy as np
    import torch

    from rang import gram, synth
    from rang import precond as P
    from rang.measurement import WgridMeasurement

    npix, ntime, nchan = 2048, 300, 32
    enu = synth.antennas(64, 8000.0, 1234)
    uvw = synth.uvw_tracks(synth.baselines_xyz(enu, -30.7), ntime, 8.0, -45.0)
    freq = np.linspace(1.30e9, 1.45e9, nchan)
    tu, tf = torch.from_numpy(uvw).to(dev), torch.from_numpy(freq).to(dev)
    vis = torch.zeros((uvw.shape[0], nchan), dtype=torch.complex64, device=dev)
    wgt = torch.ones((uvw.shape[0], nchan), dtype=torch.float32, device=dev)
    pix, eps = 1.5 * synth.ARCSEC, 1e-5
    t0 = time.perf_counter()
    m = WgridMeasurement(tu, tf, vis, wgt, npix, pix, eps, single=True)
    o = [npix, npix, pix, pix, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    psf = gram.GramOperator.no_w(o, tu, tf, m.wgt, eps, single=True)
    filt = psf.inverse_filter(1e-3 * psf.stats()["spectral_peak"])
    G = gram.GramOperator.new(o, tu, tf, m.wgt, gram.WExpansion(8, 1.0, 1e-13), eps, single=True)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    rng = np.random.default_rng(1)
    xs = torch.zeros((npix, npix), dtype=torch.float32, device=dev)
    k = 300
    ii = torch.from_numpy(rng.integers(npix // 8, 7 * npix // 8, k)).to(dev)
    jj = torch.from_numpy(rng.integers(npix // 8, 7 * npix // 8, k)).to(dev)
    xs[ii, jj] = torch.from_numpy(np.exp(rng.uniform(np.log(1e-4), 0, k)).astype(np.float32)).to(dev)
    b = m.plan.gram(xs).clone()

    def judged(x):
        r = b.double() - m.plan.gram(x.contiguous()).double()
        return math.sqrt(float((r * r).sum()) / float((b.double() * b.double()).sum()))

    def clock(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = fn()
        torch.cuda.synchronize()
        return res, round(time.perf_counter() - t, 3)

    (x_cg, n_cg, _), t_cg = clock(lambda: m.plan.conjugate_gradients(b, budget, tol))
    progress(f"Gram solve: CG {n_cg} applications, {t_cg} s")
    (x_p, n_p, _), t_p = clock(lambda: m.plan.preconditioned_cg(filt, b, budget, tol))
    (x_io, c_io, _), t_io = clock(lambda: P.inner_outer_solve(m, G, filt, b, 200, tol, tau=tau))
    net = P.UNet(npix, npix, batch=4, levels=4, base=32, seed=1)
    tr = P.FilteredTrainer(m, net, filt, G)
    _, t_train = clock(lambda: (tr.collect(rhs_count=2, iterations=40, keep=(0, 2, 5, 10, 20, 39)),
                                tr.train(train_steps, lr=1e-3, lr_end=1e-4)))
    learned = P.FilteredPreconditioner(m, net, filt)
    (x_l, c_l, _), t_l = clock(lambda: P.inner_outer_solve(m, G, filt, b, 200, tol, tau=tau, learned=learned))
    return {
        "workload": "2048x2048 at 1.5 arcsec, 2016 baselines x 300 times x 32 channels (19,353,600 vis, "
                    "1.30-1.45 GHz), eps 1e-5, single precision; A^H W A x = A^H W A x_sky (300 points) from zero",
        "tolerance": tol,
        "judge": "relative residual |b - A^H W A x| / |b| with the exact operator, recomputed after each solve",
        "setup_seconds": round(t_setup, 2),
        "cg": {"exact_applies": n_cg, "seconds": t_cg, "relative_residual": judged(x_cg)},
        "pcg_inverse_filter": {"exact_applies": n_p, "seconds": t_p, "relative_residual": judged(x_p)},
        "inner_outer": {"exact_applies": c_io["exact"], "surrogate_applies": c_io["surrogate"],
                        "seconds": t_io, "relative_residual": judged(x_io), "tau": tau,
                        "fallback_at": c_io["switched_at"]},
        "inner_outer_learned": {"exact_applies": c_l["exact"], "surrogate_applies": c_l["surrogate"],
                                "network_applies": c_l["learned"], "seconds": t_l,
                                "relative_residual": judged(x_l), "fallback_at": c_l["switched_at"],
                                "training_seconds": t_train, "training_exact_applies": tr.exact_applies,
                                "training_surrogate_applies": tr.gram_applies,
                                "training_skies": train_steps * net.batch // 2 + 2},
        "network": f"U-Net 4 levels x 32 base channels, BF16 on MFMA, {net.parameters()} weights",
        "exact_over_surrogate_cost": "one surrogate application costs about 0.3 of an exact one",
    }


def wide_field(dev, ntime=1800, nchan=512, f_lo=0.200e9, f_hi=0.252e9, npix=8192, pix_arcsec=1.3, eps=1e-5,
               cycles=10, inner_iterations=20, plain_iterations=30, lam_frac=0.02, kkt_rel=1e-4):
    """Wide-field major-cycle imaging: four-correlation complex64 visibilities
    of 2016 baselines x `ntime` times x `nchan` channels (a 20-source sky over
    the 3 degree field; the band chosen so that |w| reaches about 6000
    wavelengths), Stokes I formed on the GPU by WSClean's reader rule, an
    8192^2 image at 1.3 arcsec, eps 1e-5, single precision. The L1-regularised
    problem (lambda = lam_frac of the dirty peak) solved two ways to the same
    exact optimality residual (kkt_rel of the dirty peak): plain FISTA on the
    exact normal operator (one vis2dirty of dirty2vis per iteration) for at
    most `plain_iterations`, and `cycles` major cycles of defect correction
    (prox_solver::defect_correction), each judged by one exact application and
    solved by `inner_iterations` FISTA steps on the visibility-free w-aware
    surrogate. Exact and surrogate applications, seconds, and the residual
    each reaches."""
    import numpy as np
    import torch

    from rang import core_weights, gram, synth
    from rang import device as D
    from rang.measurement import WgridMeasurement
    from rang.prox_solver import L1Prior, OuterSchedule, Problem, Settings, defect_correction, solve

    def clock(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = fn()
        torch.cuda.synchronize()
        return res, round(time.perf_counter() - t, 3)

    free0, _ = torch.cuda.mem_get_info(dev)
    progress(f"wide field: {free0 / 2**30:.1f} GiB free at the start")
    c, t_gen = clock(lambda: synth.fullpol_case_device(ntime, nchan, f_lo, f_hi, npix, pix_arcsec, dev))
    ncell = c["data"].shape[0] * c["data"].shape[1]
    progress(f"wide field: {ncell} x 4 correlations formed in {t_gen} s, |w| up to {c['wmax_lambda']:.0f}")
    (vis, wgt), t_stokes = clock(lambda: core_weights.stokes_i_device(c["data"], c["weight"], c["flag"]))
    corr_bytes = c["data"].numel() * 8 + c["weight"].numel() * 4 + c["flag"].numel()
    for k in ("data", "weight", "flag"):
        del c[k]
    torch.cuda.empty_cache()
    uvw, freq, pix = c["uvw"], c["freq"], c["pixsize"]
    # the surrogate first, on the weights normalised as the measurement
    # normalises them; its set-up grids through the library's shared work
    # buffers, released before the plan is made
    o = [npix, npix, pix, pix, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
    wn = (wgt.to(torch.float64) / float(wgt.to(torch.float64).sum())).to(torch.float32).contiguous()
    G, t_sur = clock(lambda: gram.GramOperator.new(o, uvw, freq, wn, gram.WExpansion(8, 1.0, 1e-13), eps,
                                                  single=True))
    del wn
    D.release()
    torch.cuda.empty_cache()
    m, t_prep = clock(lambda: WgridMeasurement(uvw, freq, vis, wgt, npix, pix, eps, single=True))
    del wgt
    torch.cuda.empty_cache()
    progress(f"wide field: Stokes I {t_stokes} s, surrogate built in {t_sur} s, plan prepared in {t_prep} s")
    zeros = torch.zeros((npix, npix), dtype=torch.float32, device=dev)
    (dirty, q0), t_dirty = clock(lambda: m.residual(zeros))
    free_bytes, total_bytes = torch.cuda.mem_get_info(dev)
    from rang import core_wgridder as CW

    # the plan of a unit-weight pass (every visibility active, as here)
    plan_info = CW.plan(uvw=uvw.cpu().numpy(), freq=freq.cpu().numpy(), npix_x=npix, npix_y=npix, pixsize_x=pix,
                        pixsize_y=pix, epsilon=eps, single=True)
    m.plan.gram(dirty)
    _, t_exact = clock(lambda: m.plan.gram(dirty))
    G.apply(dirty)  # the first application sizes the operator's scratch
    _, t_cheap = clock(lambda: G.apply(dirty))
    progress(f"wide field: one exact application {t_exact} s, one surrogate application {t_cheap} s")
    case = {"npix": npix, "pixsize": pix, "uvw": uvw, "freq": freq, "vis": m.vis, "wgt": m.wgt, "sky": c["sky"]}
    # the data are not read again once the dirty image and the direct sum's
    # pixels are formed (dft_errors forms those first): the test image's
    # visibilities go into their 14 GiB instead of a new buffer
    torch.cuda.empty_cache()
    errs = dft_errors(case, dirty, lambda img: m.plan.dirty2vis(img, weighted=True, out=m.vis), n_pix=64)
    del case
    m.vis = None
    torch.cuda.empty_cache()
    # the step from the largest eigenvalue, by power iteration on the surrogate
    g = torch.Generator(device=dev).manual_seed(3)
    v = torch.randn((npix, npix), dtype=torch.float32, device=dev, generator=g)
    v /= v.norm()
    lmax = 0.0
    for _ in range(100):
        w_ = G.apply(v)
        nxt = float(w_.double().norm())
        v = (w_ / nxt).contiguous()
        if abs(nxt - lmax) <= 1e-4 * nxt:
            lmax = nxt
            break
        lmax = nxt
    peak = float(dirty.max())
    lam, kkt_tol = lam_frac * peak, kkt_rel * peak
    step = 1.0 / (1.02 * lmax)
    pr = Problem(dirty, q0)
    plain, t_plain = clock(lambda: solve(m, pr, L1Prior(lam, 1.0),
                                         Settings(step=step, max_iterations=plain_iterations, kkt_tolerance=kkt_tol),
                                         zeros))
    progress(f"wide field: plain FISTA {plain.iterations + 1} exact applications, {t_plain} s, "
             f"kkt/peak {plain.last.kkt / peak:.2e}")
    dc, t_dc = clock(lambda: defect_correction(
        m, G, pr, L1Prior(lam, 1.0), Settings(step=step, max_iterations=inner_iterations, kkt_tolerance=0.1 * kkt_tol),
        OuterSchedule(cycles=cycles, kkt=kkt_tol, lam=lam), zeros))
    progress(f"wide field: {len(dc.records) - 1} major cycles, {dc.exact_applies} exact + {dc.cheap_applies} "
             f"surrogate applications, {t_dc} s, kkt/peak {dc.last.kkt / peak:.2e}")
    reach = next((r.iteration + 1 for r in plain.records if r.kkt <= dc.last.kkt), None)
    # exact applications the cycles needed for the residual plain FISTA ended at
    back = next((r.outer + 1 for r in dc.records if r.kkt <= plain.last.kkt), None)
    out = {
        "workload": f"2016 baselines x {ntime} times x {nchan} channels ({f_lo / 1e9:.3f}-{f_hi / 1e9:.3f} GHz), "
                    f"4 correlations (XX, XY, YX, YY) complex64, Stokes I on the GPU; {npix}x{npix} at "
                    f"{pix_arcsec} arcsec (3 degree field), eps {eps:g}, single precision",
        "visibilities": ncell,
        "correlation_bytes": corr_bytes,
        "wmax_lambda": round(c["wmax_lambda"], 1),
        "grid": [int(plan_info["nu"]), int(plan_info["nv"])],
        "support": int(plan_info["supp"]),
        "w_planes": int(plan_info["nplanes"]),
        "plan_device_bytes": int(m.plan.memory()),
        "device_bytes_free_after_setup": int(free_bytes),
        "device_bytes_total": int(total_bytes),
        "seconds": {"generate": t_gen, "stokes_i": t_stokes, "prepare": t_prep, "surrogate": t_sur,
                    "dirty": t_dirty, "exact_application": t_exact, "surrogate_application": t_cheap},
        "problem": {"lambda_over_peak": lam_frac, "kkt_over_peak_target": kkt_rel, "step": step,
                    "largest_eigenvalue": lmax},
        "surrogate": {"w_nodes": int(G.stats()["planes"]), "stencil": 8, "theta": 1.0,
                      "sup_error": G.stats()["sup_error"]},
        "plain_fista": {"exact_applies": plain.iterations + 1, "seconds": t_plain,
                        "kkt_over_peak": plain.last.kkt / peak, "objective": plain.last.objective,
                        "stop": plain.stop},
        "major_cycles": {"cycles": len(dc.records) - 1, "exact_applies": dc.exact_applies,
                         "surrogate_applies": dc.cheap_applies, "inner_iterations": dc.iterations,
                         "seconds": t_dc, "kkt_over_peak": dc.last.kkt / peak, "objective": dc.last.objective,
                         "stop": dc.stop,
                         "kkt_over_peak_by_cycle": [r.kkt / peak for r in dc.records]},
        "plain_applies_to_reach_the_cycles_residual": reach,
        "cycles_exact_applies_to_reach_the_plain_residual": back,
        "errors": errs,
        "note": "the cheap operator of the cycles is the visibility-free surrogate; the learned network is not "
                "used here (at the Gram-solve workload its warm start changed no count)",
    }
    m.plan.close()
    G.close()
    return out


def ingest_chunk(blocks=16, blocksize=1 << 20, seed=3):
    """A synthetic zarr chunk of a visibility variable: `blocks` blocks of
    1 MiB of complex64 values (unit normal, 20 % zeros), byte-shuffled and
    split into 8 LZ4 streams per block (c-blosc 1.21's layout at typesize 8
    for 1 MiB blocks); one block is encoded (tests/blosc_cases.py's writer)
    and stored `blocks` times. Returns (decoded bytes of one block, chunk)."""
    import struct

    sys.path.insert(0, str(ROOT / "tests"))
    import blosc_cases as BC

    data = BC.contents("vis", blocksize, seed)
    one = BC.blosc_chunk(data, 8, blocksize, shuffled=True)
    body = one[20:]
    head = 16 + 4 * blocks
    cbytes = head + blocks * len(body)
    chunk = (struct.pack("<BBBBIII", 2, 1, one[2], 8, blocks * blocksize, blocksize, cbytes)
             + b"".join(struct.pack("<I", head + k * len(body)) for k in range(blocks)) + body * blocks)
    return data, chunk


def run_ingest(dev, nchunks=64, steps=5, warmup=1, cpu_baseline=True):
    """Blosc-1 chunks expanded on the GPU (the MSv4 ingest path): `nchunks`
    chunks of 16 MiB decoded in one call, from device-resident compressed
    bytes and from host bytes (PCIe-inclusive), beside the oracle's decoder
    on one host thread."""
    import numpy as np
    import torch

    from rang import msv4

    block, chunk = ingest_chunk()
    nbytes = 16 * len(block)
    chunks = [chunk] * nchunks
    outs = [torch.empty(nbytes, dtype=torch.uint8, device=dev) for _ in range(nchunks)]
    res = torch.from_numpy(np.frombuffer(chunk, np.uint8).copy()).to(dev)
    resident = [res] * nchunks

    def timed(fn, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    for _ in range(warmup):
        msv4.blosc_decode_device(chunks, outs, resident)
    t_res = timed(lambda: msv4.blosc_decode_device(chunks, outs, resident), steps)
    want = np.frombuffer(block, np.uint8)
    ok = all(np.array_equal(outs[i][k * len(block):(k + 1) * len(block)].cpu().numpy(), want)
             for i in (0, nchunks - 1) for k in (0, 15))
    t_host = timed(lambda: msv4.blosc_decode_device(chunks, outs), max(2, steps // 2))
    total_out = nchunks * nbytes
    total_in = nchunks * len(chunk)
    alg = total_in + total_out
    out = {
        "workload": f"{nchunks} blosc-1 chunks of 16 MiB (16 blocks of 1 MiB complex64, byte shuffle, 8 LZ4 "
                    "streams per block), decoded in one call",
        "decoded_bytes": total_out,
        "compressed_bytes": total_in,
        "ratio": round(total_out / total_in, 4),
        "resident": {"ms": round(t_res * 1e3, 3), "decoded_GBps": round(total_out / t_res / 1e9, 2)},
        "host_bytes": {"ms": round(t_host * 1e3, 3), "decoded_GBps": round(total_out / t_host / 1e9, 2),
                       "note": "compressed bytes uploaded from pageable host memory inside the call"},
        "roofline": {"bound": "hbm", "achieved": round(alg / t_res / 1e9, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / t_res / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic("ingest", "k_lz4_streams",
                                                 ROOT / "profiles" / "blosc_ingest_decode.md"),
                     "note": "algorithmic bytes: compressed read once + decoded written once, over the "
                             "whole call (host plan, table upload, LZ4 and unshuffle kernels); traffic: "
                             "k_lz4_streams per launch"},
        "parity": "equal" if ok else "MISMATCH",
        "cpu_baseline": None,
    }
    if cpu_baseline:
        from oracle import oracle as O

        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            got = O.blosc_decode(chunk)
        t_cpu = (time.perf_counter() - t0) / reps
        assert got[: len(block)] == block
        out["cpu_baseline"] = {"value": round(nbytes / t_cpu / 1e9, 3), "unit": "decoded GB/s", "cores": 1,
                               "kind": "port", "sample": "one 16 MiB chunk, three times, the oracle's scalar "
                               "decoder (oracle/blosc.c)"}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the small workload")
    ap.add_argument("--no-gram", action="store_true", help="skip the Gram-solve comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the sub-band batch")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the cases the reference published timings for")
    ap.add_argument("--only-small", action="store_true",
                    help="measure the small workload alone (no JSON contract line)")
    ap.add_argument("--no-ingest", action="store_true", help="skip the blosc ingest leg")
    ap.add_argument("--no-wide", action="store_true", help="skip the wide-field major-cycle leg")
    ap.add_argument("--only-wide", action="store_true",
                    help="run the wide-field major-cycle leg alone (no JSON contract line)")
    ap.add_argument("--wide-ntime", type=int, default=1800,
                    help="time steps of the wide-field leg (1800 is the workload)")
    ap.add_argument("--profile", choices=("large", "small", "unet", "wstack", "ingest", "gram"),
                    help="only run --warmup + --steps full vis2dirty + dirty2vis calls of one workload and print "
                         "their stage times (for profiler runs; no JSON contract line)")
    ap.add_argument("--copy-kernel", action="store_true",
                    help="only run a plain copy of 1 GiB (8 launches), the reference for HBM "
                         "byte counters")
    args = ap.parse_args()
    if args.copy_kernel:
        import torch

        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(8):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the gridder runs on "
              "one MI355X; a multi-GPU split of the visibilities is a design, not a measured path",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from rang import synth

    dev = torch.device("cuda:0")
    from rang import device as D

    if args.profile == "unet":
        # the preconditioner's network at the Gram-solve size: forward and
        # backward passes of 4 images of 2048^2 on random inputs
        from rang import precond as P

        net = P.UNet(2048, 2048, batch=4, levels=4, base=32, seed=1)
        net.buf[0]["in"].copy_(torch.randn(net.buf[0]["in"].shape, device=dev).to(torch.bfloat16)
                               .view(torch.int16))
        net.g_y.normal_()

        def one():
            net.forward(4)
            net.zero_grad()
            net.backward(4)

        for _ in range(args.warmup):
            one()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            net.forward(4)
        torch.cuda.synchronize()
        t_f = (time.perf_counter() - t0) / max(args.steps, 1)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            one()
        torch.cuda.synchronize()
        t_fb = (time.perf_counter() - t0) / max(args.steps, 1)
        fl = net.forward_flops(4)
        print(json.dumps({"profile": "unet", "images": 4, "side": 2048, "forward_ms": round(t_f * 1e3, 3),
                          "forward_tflops": round(fl / t_f / 1e12, 2),
                          "forward_backward_ms": round(t_fb * 1e3, 3),
                          "forward_backward_tflops": round(3 * fl / t_fb / 1e12, 2),
                          "bf16_dense_peak_tflops": 2500.0}))
        return

    if args.only_wide:
        print(json.dumps({"wide_field": wide_field(dev, ntime=args.wide_ntime)}))
        return

    if args.profile == "gram":
        # the visibility-free operator at the wide-field image (8192^2 at 1.3
        # arcsec, 0.200-0.252 GHz, eight w-nodes) built from 100 time steps
        # (its application does not depend on the visibilities)
        from rang import gram

        c = synth.imaging_case_device(100, 64, 0.200e9, 0.252e9, 8192, 1.3, dev)
        o = [8192, 8192, c["pixsize"], c["pixsize"], 1e-5, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
        wn = (c["wgt"] / c["wgt"].sum()).contiguous()
        G = gram.GramOperator.new(o, c["uvw"], c["freq"], wn, gram.WExpansion(8, 1.0, 1e-13), 1e-5, single=True)
        x = torch.randn((8192, 8192), dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(args.warmup):
            G.apply(x, out=y)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            G.apply(x, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"profile": "gram", "side": 8192, "nodes": int(G.stats()["planes"]),
                          "apply_ms": round((time.perf_counter() - t0) / max(args.steps, 1) * 1e3, 3)}))
        return

    if args.profile == "ingest":
        print(json.dumps({"profile": "ingest", **run_ingest(dev, steps=args.steps, warmup=args.warmup,
                                                            cpu_baseline=False)}))
        return

    if args.profile == "wstack":
        # the w-stacking gridder on the published-comparison case (23.8 M
        # visibilities, 1024^2 at 1.5 arcsec, 64 layers, kernel 7 x 1023,
        # padding 1.2, dirty + PSF) from host arrays, f64 then f32
        import numpy as np

        from rang import _core

        c = synth.imaging_case_device(738, 16, 1.30e9, 1.45e9, 1024, 1.5, dev)
        nrow, nchan = int(c["uvw"].shape[0]), int(c["freq"].shape[0])
        uvw_h, freq_h = c["uvw"].cpu().numpy(), c["freq"].cpu().numpy()
        data = np.ascontiguousarray(c["vis"].cpu().numpy().view(np.float32).reshape(nrow, nchan, 2))
        w32 = np.ascontiguousarray(c["wgt"].cpu().numpy().astype(np.float32))
        out = {}
        for prec in ("f64", "f32"):
            d, p = np.zeros(1024 * 1024), np.zeros(1024 * 1024)
            stats, timings = np.zeros(_core.GRID_N_STATS), np.zeros(_core.GRID_N_TIMINGS)
            for k in range(args.warmup + args.steps):
                t0 = time.perf_counter()
                st = _core.grid_invert_gpu(nrow, nchan, 1024, 1024, 7, 1023, 64, c["pixsize"], c["pixsize"], 1.2,
                                           uvw_h, freq_h, data, w32, w32, prec, d, p, stats, timings)
                if st != 0:
                    raise RuntimeError(_core.grid_last_error())
                progress(f"w-stacking {prec} call {k + 1}: {time.perf_counter() - t0:.3f} s")
            out[prec] = dict(zip(("context_s", "compile_s", "prepare_s", "upload_s", "compute_s", "download_s",
                                  "finish_s", "gpu_bytes", "batches", "layers_per_batch", "grid_s", "fft_s",
                                  "project_s", "bucket_s"), (float(v) for v in timings)))
        print(json.dumps({"profile": "wstack", "timings": out}))
        return

    if args.profile:
        if args.profile == "large":
            case = synth.config2_device(dev)
        else:
            small = synth.config1()
            case = dict(small)
            for k in ("uvw", "freq", "vis", "wgt"):
                case[k] = torch.from_numpy(small[k]).to(dev)
        progress("visibilities formed")
        el, info, means = run_gpu(case, args.steps, args.warmup, verbose=True)
        print(json.dumps({"profile": args.profile, "seconds": round(el, 4), "steps": args.steps,
                          "stages_ms": {f"{d}:{n}": round(v, 4) for (d, n), v in sorted(means.items())}}))
        return

    if not args.only_small:
        big = synth.config2_device(dev)
        progress("visibilities formed")
        nrow, nchan = int(big["uvw"].shape[0]), int(big["freq"].shape[0])
        nvis_total = nrow * nchan
        # the plan cached (the metric's definition): prepared plan, graph replays
        elapsed, t_prepare, means, plan, dirty, out_vis = run_prepared(big, args.steps, args.warmup, verbose=True)
        _, _, info = D.vis2dirty([big["npix"], big["npix"], big["pixsize"], big["pixsize"], big["epsilon"], 0, 0,
                                  1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0], big["uvw"], big["freq"], big["vis"],
                                 big["wgt"], single=True, out=(dirty, None))
        value, roofline, per_call = summarise(big, nvis_total, nrow, nchan, args.steps, elapsed,
                                              info, means, workload="large")
        progress(f"large workload, prepared: {value:.1f} Mvis/s")
        errs_large = dft_errors(big, dirty, lambda img: plan.dirty2vis(img, weighted=True))
        plan.close()
        del plan, out_vis
        torch.cuda.empty_cache()
        # every call in full (scan, plan, sort each time)
        f_el, f_info, f_means = run_gpu(big, max(2, args.steps // 2), 1)
        full_value = nvis_total * max(2, args.steps // 2) / f_el / 1e6
        progress(f"large workload, full calls: {full_value:.1f} Mvis/s")
        host_large = run_host_arrays(big, 1)
        progress(f"large workload, host arrays: {host_large['value']} Mvis/s")
        f64 = run_f64(big, 2)
        progress(f"large workload, f64: {f64['value']} Mvis/s")
        del big
        D.release()
        torch.cuda.empty_cache()

    wide = None
    if not args.no_wide and not args.only_small:
        # the wide-field plan needs most of the GPU: it runs while nothing
        # else is held, every earlier plan and kept buffer released first
        gc.collect()
        D.release()
        torch.cuda.empty_cache()
        try:
            wide = wide_field(dev, ntime=args.wide_ntime)
        except (torch.OutOfMemoryError, RuntimeError, ValueError) as e:
            # the leg needs about 270 GB of the GPU's 288 GiB; a box that
            # cannot hold it reports why instead of losing the other legs
            wide = {"error": f"{type(e).__name__}: {str(e)[:400]}"}
            progress(f"wide field: not measured ({wide['error'][:120]})")
        gc.collect()
        D.release()
        torch.cuda.empty_cache()

    small_line = None
    if not args.no_small:
        small = synth.config1()
        dcase = dict(small)
        for k in ("uvw", "freq", "vis", "wgt"):
            dcase[k] = torch.from_numpy(small[k]).to(dev)
        s_el, s_tp, s_means, s_plan, s_dirty, s_vis = run_prepared(dcase, 50, 3)
        _, _, s_info = D.vis2dirty([1024, 1024, small["pixsize"], small["pixsize"], 1e-5, 0, 0, 1.1, 2.6, 1, 0, 0, 0,
                                    0, 1, -1, 0, 0], dcase["uvw"], dcase["freq"], dcase["vis"], dcase["wgt"],
                                   single=True)
        s_value, s_roof, s_call = summarise(dcase, small["vis"].size, small["uvw"].shape[0],
                                            small["freq"].size, 50, s_el, s_info, s_means)
        s_errs = dft_errors(dcase, s_dirty, lambda img: s_plan.dirty2vis(img, weighted=True))
        sf_el, _, _ = run_gpu(dcase, 20, 3)
        s_host = run_host_arrays(dcase, 3)
        small_line = {
            "workload": "2016 baselines x 100 times x 16 channels (3,225,600 vis), 1024x1024 at "
                        "2 arcsec, eps 1e-5, w-gridding, single precision",
            "value": round(s_value, 2),
            "ms_per_step": round(s_el / 50 * 1e3, 4),
            "prepare_seconds": round(s_tp, 4),
            "full_calls": {"value": round(small["vis"].size * 20 / sf_el / 1e6, 2),
                           "ms_per_step": round(sf_el / 20 * 1e3, 4)},
            "host_arrays": s_host,
            "grid": [int(s_info[2]), int(s_info[3])],
            "w_planes": int(s_info[4]),
            "roofline": s_roof,
            "per_call": s_call,
            "stages_ms": {f"{d}:{n}": round(v, 4) for (d, n), v in sorted(s_means.items())},
            "errors": s_errs,
            "cpu_baseline": None,
        }
        if not args.no_cpu_baseline:
            base, oerr = cpu_small(small, args.cpu_repeats, s_dirty.cpu().numpy(),
                                   s_plan.dirty2vis(s_dirty, weighted=True).cpu().numpy())
            small_line["cpu_baseline"] = base
            small_line["errors"].update(oerr)
        s_plan.close()
        progress(f"small workload: {s_value:.1f} Mvis/s")
        if args.only_small:
            print(json.dumps({"config_small": small_line}))
            return
    cpu = None if args.no_cpu_baseline else cpu_large(info, nvis_total)
    pub = None if args.no_published else published(dev, not args.no_cpu_baseline)
    sub = None
    if not args.no_batch:
        sub = run_batch(dev)
        D.release()
        torch.cuda.empty_cache()
    ingest = None
    if not args.no_ingest:
        ingest = run_ingest(dev, cpu_baseline=not args.no_cpu_baseline)
        progress(f"ingest: {ingest['resident']['decoded_GBps']} GB/s decoded")
        torch.cuda.empty_cache()
    gram = None
    if not args.no_gram:
        gram = gram_solve(dev)
        progress("Gram solve: " + json.dumps({k: gram[k].get("exact_applies") for k in
                                              ("cg", "pcg_inverse_filter", "inner_outer", "inner_outer_learned")}))
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "wgridder vis2dirty+dirty2vis, 2016 baselines x 900 times x 256 channels "
                        "(464,486,400 vis, 0.9-1.7 GHz), 8192x8192 at 1 arcsec, eps 1e-5, "
                        "w-gridding, single precision, plan cached (prepared once, outside the timed steps)",
            "nvis": nvis_total,
            "image": [8192, 8192],
            "grid": [int(info[2]), int(info[3])],
            "support": int(info[1]),
            "w_planes": int(info[4]),
            "parallelism": "one GPU",
        },
        "prepare_seconds": round(t_prepare, 3),
        "full_calls": {"value": round(full_value, 2), "ms_per_step": round(f_el / max(2, args.steps // 2) * 1e3, 3),
                       "stages_ms": {f"{d}:{n}": round(v, 3) for (d, n), v in sorted(f_means.items())}},
        "host_arrays": host_large,
        "f64": f64,
        "errors": errs_large,
        "stages_ms": {f"{d}:{n}": round(v, 3) for (d, n), v in sorted(means.items())},
        "roofline": roofline,
        "per_call": per_call,
        "cpu_baseline": cpu,
        "config_small": small_line,
        "sub_band_batch": sub,
        "ingest": ingest,
        "gram_solve": gram,
        "wide_field": wide,
        "published": pub,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
