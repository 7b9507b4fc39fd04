#!/usr/bin/env python3
"""Throughput of the exact w-gridder on one MI355X: vis2dirty + dirty2vis.

Workload: 64 antennas within 8 km (2016 baselines) x 100 times of 8 s x 16
channels over 1.30-1.45 GHz = 3,225,600 complex64 visibilities from 20
point sources plus noise, unit float32 weights, a 1024 x 1024 image at
2 arcsec, epsilon 1e-5, w-gridding on, single precision. One step is one
vis2dirty of the visibilities followed by one dirty2vis of the resulting
image, each a full call (scan, plan, sort, grid/degrid, transforms) on
inputs already resident in HBM. value = visibilities per step x steps /
wall time of the timed steps, in 10^6 visibilities per second.

The JSON line also carries the roofline of the step's dominant kernel
(algorithmic bytes of one launch over its HIP-event duration on the call's
stream, against 8 TB/s) and the oracle (the CPU restatement of the
reference gridder) timed on this host's cores on the same workload.
"""

import argparse
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
    precision): the data the kernel must read and write at least once."""
    nu, nv, P = int(info[2]), int(info[3]), max(int(info[4]), 1)
    nx = ny = npix
    cx, real, rec = 8, 4, 32
    grid = nu * nv * cx
    if direction == "v2d":
        table = {
            "grid": nvis * rec + P * grid,
            "fft_v": P * grid + P * nu * ny * cx,
            "fft_u": P * nu * ny * cx + P * nx * ny * real,
            "sum": P * nx * ny * real + nx * ny * real,
        }
    else:
        table = {
            "fft_u": nx * ny * real + P * nu * ny * cx,
            "fft_v": P * nu * ny * cx + P * grid,
            "degrid": nvis * rec + P * grid + nrow * nchan * cx,
        }
    return table.get(name)


def run_gpu(case, steps, warmup):
    import numpy as np
    import torch

    from rang import device as D

    dev = torch.device("cuda:0")
    uvw = torch.from_numpy(case["uvw"]).to(dev)
    freq = torch.from_numpy(case["freq"]).to(dev)
    vis = torch.from_numpy(case["vis"]).to(dev)
    wgt = torch.from_numpy(case["wgt"]).to(dev)
    npix, pix, eps = case["npix"], case["pixsize"], case["epsilon"]
    opts = [npix, npix, pix, pix, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0]
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
    for _ in range(steps):
        info = step(stages)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    D.stage_timing(False)
    # per-stage mean milliseconds over the timed steps
    acc = {}
    for direction, st in stages:
        for name, ms in st:
            acc.setdefault((direction, name), []).append(ms)
    means = {k: statistics.fmean(v) for k, v in acc.items()}
    return elapsed, info, means, (dirty.cpu().numpy(), np.asarray(info))


def run_cpu_baseline(case, repeats):
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
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
    return {
        "value": case["vis"].size / t / 1e6,
        "unit": "Mvis/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the full workload ({case['vis'].size} visibilities), vis2dirty + dirty2vis, "
                  f"median of {repeats} runs of the oracle (C restatement of core/src/wgridder.rs, "
                  f"OpenMP) on {threads} host threads",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the gridder runs on "
              "one MI355X; a multi-GPU split of the visibilities is a design, not a measured path",
              file=sys.stderr)
        sys.exit(2)
    from rang import synth

    case = synth.config1()
    nvis_total = case["vis"].size
    elapsed, info, means, _ = run_gpu(case, args.steps, args.warmup)
    nvis = int(info[5])
    value = nvis_total * args.steps / elapsed / 1e6
    # the dominant kernel: the stage with the longest mean duration
    candidates = []
    for (direction, name), ms in means.items():
        b = stage_bytes(name, direction, info, nvis, case["uvw"].shape[0], case["freq"].size, case["npix"])
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
        "traffic": None,
        "algorithmic_bytes": int(b),
        "kernel_ms": round(ms, 4),
    }
    cpu = None if args.no_cpu_baseline else run_cpu_baseline(case, args.cpu_repeats)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "wgridder vis2dirty+dirty2vis, 2016 baselines x 100 times x 16 channels "
                        "(3,225,600 vis), 1024x1024 at 2 arcsec, eps 1e-5, w-gridding, single precision",
            "nvis": nvis_total,
            "image": [case["npix"], case["npix"]],
            "grid": [int(info[2]), int(info[3])],
            "support": int(info[1]),
            "w_planes": int(info[4]),
            "parallelism": "one GPU",
        },
        "stages_ms": {f"{d}:{n}": round(v, 4) for (d, n), v in sorted(means.items())},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
