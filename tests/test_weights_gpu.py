"""Imaging weights and Stokes I on the GPU (rang.core_weights over the C ABI
of include/rang_core_weights.h) against rang-core's golden vectors and the
oracle (the serial restatement of core/src/weights.rs).

Stokes I is elementwise float arithmetic in the reference's order: bit for
bit. The weight grid's cells are summed with double atomics, so a cell of
three or more samples, the in-grid total and the Briggs mean may differ from
the serial sums in their last bits: which cells are non-zero and the counts
must match exactly, the grids to 1e-12 relative, the float weights to 2
units in the last place (2.4e-7 relative) and the sums to 1e-12."""

import numpy as np
import pytest
from conftest import ROOT
from weights_cases import PIXEL_SCALE, golden_weights_case

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = ROOT / "tests" / "golden" / "weights_golden.npz"


def close(a, b, rel):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b).max(initial=0.0) <= rel * max(np.abs(b).max(initial=0.0), 1e-300)


def test_golden_case():
    from rang import core_weights as W

    z = np.load(GOLD, allow_pickle=False)
    c = golden_weights_case()
    vis, w = W.stokes_i(c["data"], c["weights"], c["flags"])
    assert np.array_equal(vis.view(np.uint32), z["stokes_vis"].view(np.uint32))
    assert np.array_equal(w.view(np.uint32), z["stokes_weight"].view(np.uint32))
    assert [W.padded_size(512), *W.grid_shape(616, 616)] == [int(v) for v in z["sizes"]]
    for mode in ("briggs", "uniform", "natural"):
        r = W.imaging_weights(c["uvw"], c["wavelength"], c["vis_weights"], c["selected"], mode=mode, robust=0.5)
        gold_w, gold_g, gold_s = z[f"{mode}_weights"], z[f"{mode}_grid"], z[f"{mode}_sums"]
        assert np.array_equal(r["weights"] != 0, gold_w != 0) and np.array_equal(r["grid"] != 0, gold_g != 0)
        assert close(r["weights"], gold_w, 2.4e-7) and close(r["grid"], gold_g, 1e-12)
        s = r["sums"]
        got = [s["visibility_weight_sum"], s["total_weight"], s["max_gridded_weight"], s["count"]]
        assert got[3] == gold_s[3] and close(got[:3], gold_s[:3], 1e-12)
        print(f"[weights] golden {mode}: weights max rel diff "
              f"{np.abs(r['weights'] - gold_w).max() / np.abs(gold_w).max():.1e}", flush=True)


@pytest.mark.parametrize("mode", ["briggs", "uniform", "natural"])
def test_many_samples_against_the_oracle(mode):
    # 2000 rows x 64 channels of an 8 km array at 1024^2 / 2 arcsec (padded
    # 1232): many samples per cell near the centre, 10 % of rows deselected,
    # 20 % zero weights; Briggs robust 0
    from rang import core_weights as W
    from rang import synth

    rng = np.random.default_rng(31)
    enu = synth.antennas(64, 8000.0, 99)
    uvw = synth.uvw_tracks(synth.baselines_xyz(enu, -30.7), 1, 8.0, -45.0)[:2000]
    freq = np.linspace(1.30e9, 1.45e9, 64)
    wl = W.wsclean_wavelengths(freq)
    wts = (rng.uniform(0.5, 2.0, (uvw.shape[0], 64)) * (rng.uniform(size=(uvw.shape[0], 64)) > 0.2)).astype(
        np.float32)
    sel = rng.uniform(size=uvw.shape[0]) > 0.1
    kw = dict(mode=mode, robust=0.0, width=1232, height=1232, pixel_scale=W.pixel_scale_radians(2.0))
    r = W.imaging_weights(uvw, wl, wts, sel, **kw)
    ow, og, osum = O.imaging_weights(uvw, wl, wts, sel, **kw)
    assert np.array_equal(r["grid"] != 0, og != 0) and np.array_equal(r["weights"] != 0, ow != 0)
    assert close(r["grid"], og, 1e-12) and close(r["weights"], ow, 2.4e-7)
    s = r["sums"]
    assert s["count"] == osum[3]
    assert close([s["visibility_weight_sum"], s["total_weight"], s["max_gridded_weight"]], osum[:3], 1e-12)
    print(f"[weights] {mode}: {int((og != 0).sum())} cells, max rel diff "
          f"{np.abs(r['grid'] - og).max() / np.abs(og).max():.1e}", flush=True)


def test_device_entry_points():
    import ctypes

    import torch

    from rang import _core

    c = golden_weights_case()
    d = torch.device("cuda:0")
    uvw = torch.from_numpy(c["uvw"]).to(d)
    wl = torch.from_numpy(c["wavelength"]).to(d)
    wts = torch.from_numpy(c["vis_weights"]).to(d)
    sel = torch.from_numpy(c["selected"].astype(np.uint8)).to(d)
    out = torch.zeros((5, 3), dtype=torch.float32, device=d)
    sums = torch.zeros(4, dtype=torch.float64, device=d)
    inp = _core.ImagingWeightInputs(5, 3, uvw.data_ptr(), wl.data_ptr(), wts.data_ptr(), sel.data_ptr(), 2, 0.5,
                                    616, 616, PIXEL_SCALE, PIXEL_SCALE, 1.0, 3.0, 16)
    o = _core.ImagingWeightOutputs(out.data_ptr(), None, sums.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert _core.library().rang_amd_imaging_weights_dev(ctypes.byref(inp), ctypes.byref(o), st) == 0
    z = np.load(GOLD, allow_pickle=False)
    assert close(out.cpu().numpy(), z["briggs_weights"], 2.4e-7)
    # refusals: class out of range, non-positive super weight
    inp.class_ = 3
    assert _core.library().rang_amd_imaging_weights_dev(ctypes.byref(inp), ctypes.byref(o), st) == 1


def test_stokes_i_on_device_tensors_matches_the_golden_case():
    # stokes_i_device (the wide-field ingest's form, device tensors in and
    # out) equals the reference's vectors bit for bit, flags as bool or uint8
    import torch

    from rang import core_weights as W

    z = np.load(GOLD, allow_pickle=False)
    c = golden_weights_case()
    d = torch.device("cuda:0")
    data = torch.from_numpy(np.ascontiguousarray(c["data"], np.complex64)).to(d)
    wts = torch.from_numpy(np.ascontiguousarray(c["weights"], np.float32)).to(d)
    for flags in (torch.from_numpy(np.ascontiguousarray(c["flags"], bool)).to(d),
                  torch.from_numpy(np.ascontiguousarray(c["flags"], np.uint8)).to(d)):
        vis, w = W.stokes_i_device(data, wts, flags)
        assert np.array_equal(vis.cpu().numpy().view(np.uint32), z["stokes_vis"].view(np.uint32))
        assert np.array_equal(w.cpu().numpy().view(np.uint32), z["stokes_weight"].view(np.uint32))
    with pytest.raises(ValueError, match="four correlations"):
        W.stokes_i_device(data[..., :3].contiguous(), wts[..., :3].contiguous(), flags[..., :3].contiguous())
