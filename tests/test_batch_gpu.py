"""Sub-band batches (include/rang_core_batch.h, rang.batch) on the GPU.

Every image of a batch, its gridding sharing one launch with the others',
equals the oracle at that image's plan (same kernel row, grid and planes),
as the single-image entry point does: 2e-5 of the peak in single precision
(the single-image parity tolerance) and 1e-9 in double, where at epsilon
1e-4 and 1e-5 the gridding correction at the image edge (the inverse of a
short kernel's transform) magnifies summation-order differences of order
1e-16 by up to about 1e6. Groups split by
the memory budget give the same images; an image whose samples are all
flagged gives a zero image; the tuner's choice is a plan the oracle
reproduces, never slower than the cost model's own choice, and is read
back from its cache file by a later lookup."""

import shutil

import numpy as np
import pytest
import torch
from conftest import ref_case, rel_linf

from oracle import oracle as O
from rang import batch, device

pytestmark = pytest.mark.gpu

TOL = {True: 2e-5, False: 1e-9}
# side, rows, channels, epsilon, flagged fraction (flags in contiguous runs
# of rows), seed
SHAPES = [(64, 60, 3, 1e-5, 0.0, 1), (48, 30, 8, 1e-4, 0.3, 2), (96, 90, 5, 1e-5, 0.6, 3),
          (32, 20, 2, 1e-4, 0.1, 4), (128, 200, 4, 1e-5, 0.2, 5), (64, 10, 1, 1e-5, 1.0, 6)]


def make(shape, single):
    nx, nrow, nchan, eps, frac, seed = shape
    uvw, freq, vis, wgt, _, _, dx = ref_case(nx=nx, ny=nx, nrow=nrow, seed=seed, nchan=nchan)
    rng = np.random.default_rng(seed)
    nflag = int(round(frac * nrow))
    start = int(rng.integers(0, nrow - nflag + 1))
    wgt[start:start + nflag] = 0.0
    cvis, real = (np.complex64, np.float32) if single else (np.complex128, np.float64)
    vis, wgt = vis.astype(cvis), wgt.astype(real)
    dev = torch.device("cuda:0")
    im = batch.Image([nx, nx, dx, dx, eps, 0, 0, 1.1, 2.6, 1, 0, 0, 0, 0, 1, -1, 0, 0],
                     torch.from_numpy(uvw).to(dev), freq, torch.from_numpy(vis).to(dev),
                     torch.from_numpy(wgt).to(dev))
    host = dict(uvw=uvw, freq=freq, vis=vis, wgt=wgt, nx=nx, dx=dx, eps=eps)
    return im, host


def oracle_at(h, info, single):
    kernel = None if info[5] == 0 else (int(info[0]), int(info[2]), int(info[3]))
    ref, rinfo = O.vis2dirty(h["uvw"], h["freq"], h["vis"], h["wgt"], npix_x=h["nx"], npix_y=h["nx"],
                             pixsize_x=h["dx"], pixsize_y=h["dx"], epsilon=h["eps"], single=single,
                             kernel=kernel)
    return ref["dirty"], rinfo


@pytest.mark.parametrize("single", [True, False])
def test_every_image_of_a_batch_matches_the_oracle(single):
    made = [make(s, single) for s in SHAPES]
    images = [m[0] for m in made]
    out = batch.vis2dirty(images, single=single)
    torch.cuda.synchronize()
    for (im, h), d in zip(made, out):
        got = d.cpu().numpy()
        ref, rinfo = oracle_at(h, im.info, single)
        if rinfo[5] == 0:
            assert not got.any() and im.info[5] == 0
            continue
        assert list(im.info[:6]) == list(rinfo[:6])
        assert rel_linf(got, ref) < TOL[single], (h["nx"], rel_linf(got, ref))
        # the image alone through the single-image entry point
        alone, _, _ = device.vis2dirty(im.opts, im.uvw, torch.from_numpy(im.freq).to(im.uvw.device), im.vis,
                                       im.wgt, single=single)
        assert rel_linf(got, alone.cpu().numpy()) < TOL[single]


def test_groups_split_by_the_budget_give_the_same_images():
    made = [make(s, True) for s in SHAPES]
    images = [m[0] for m in made]
    one = [d.cpu().numpy() for d in batch.vis2dirty(images, single=True)]
    each = [d.cpu().numpy() for d in batch.vis2dirty(images, single=True, budget=1)]
    for a, b in zip(one, each):
        if b.any():
            assert rel_linf(a, b) < 1e-6
        else:
            assert not a.any()


def test_tuner_choice_is_a_plan_the_oracle_reproduces(tmp_path):
    path = tmp_path / "tuning.txt"
    im, h = make(SHAPES[4], True)
    rep = batch.tune(im, single=True, cache_path=path, kernels=3)
    assert not rep["cached"] and 0 < rep["ms"] <= rep["default_ms"]
    assert im.opts[15] >= 0 and im.opts[16] > 0 and im.tile >= 0
    out = batch.vis2dirty([im], single=True)[0].cpu().numpy()
    assert int(im.info[0]) == im.opts[15] and int(im.info[2]) == im.opts[16]
    ref, rinfo = oracle_at(h, im.info, True)
    assert list(im.info[:6]) == list(rinfo[:6])
    assert rel_linf(out, ref) < TOL[True]
    # the same signature again: from the process's table, then from the file
    again, _ = make(SHAPES[4], True)
    rep2 = batch.tune(again, single=True, cache_path=path)
    assert rep2["cached"] and again.opts[15:] == im.opts[15:] and again.tile == im.tile
    copy = tmp_path / "copy.txt"
    shutil.copy(path, copy)
    third, _ = make(SHAPES[4], True)
    rep3 = batch.tune(third, single=True, cache_path=copy)
    assert rep3["cached"] and third.opts[15:] == im.opts[15:] and third.tile == im.tile


def test_a_refused_image_is_named():
    made = [make(s, True) for s in SHAPES[:3]]
    images = [m[0] for m in made]
    images[1].opts[0] = 47
    with pytest.raises(ValueError, match="image 1"):
        batch.vis2dirty(images, single=True)


def test_released_workspaces_are_allocated_again():
    made = [make(s, True) for s in SHAPES[:3]]
    images = [m[0] for m in made]
    first = [d.cpu().numpy() for d in batch.vis2dirty(images, single=True)]
    device.release()
    again = [d.cpu().numpy() for d in batch.vis2dirty(images, single=True)]
    for a, b in zip(first, again):
        assert rel_linf(a, b) < 1e-6


def test_both_gridding_kernels_in_one_batch():
    # a tile side asked for sends an image to the LDS gridder; the others go
    # to the matrix cores: one launch each, every image still the oracle's
    made = [make(s, True) for s in SHAPES[:5]]
    images = [m[0] for m in made]
    images[1].tile = 24
    images[3].tile = 16
    out = batch.vis2dirty(images, single=True)
    torch.cuda.synchronize()
    for (im, h), d in zip(made, out):
        ref, rinfo = oracle_at(h, im.info, True)
        assert list(im.info[:6]) == list(rinfo[:6])
        assert rel_linf(d.cpu().numpy(), ref) < TOL[True]
