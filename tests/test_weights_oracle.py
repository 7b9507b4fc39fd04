"""The imaging-weights oracle (serial C restatement of core/src/weights.rs)
pinned to rang-core's golden vectors (tests/data/core-backends-golden, keys
weights.0-12, copied by tests/golden/make_weights_golden.py; their SHA256s
are checked) bit for bit, and to the reference's own known-answer tests
(core/src/weights.rs:574-700): padded sizes, grid shapes, the Briggs /
uniform / natural formulas on one cell, the rank filter's clip, the Stokes
I rules and the header sums. CPU only."""

import hashlib
import math

import numpy as np
from conftest import ROOT
from weights_cases import PIXEL_SCALE, golden_weights_case

from oracle import oracle as O

GOLD = ROOT / "tests" / "golden" / "weights_golden.npz"


def test_fixture_is_the_references_data():
    z = np.load(GOLD, allow_pickle=False)
    keys = ["stokes_vis", "stokes_weight"]
    for cls in ("briggs", "uniform", "natural"):
        keys += [f"{cls}_weights", f"{cls}_grid"]
    for key, sha in zip(keys, z["sha256"]):
        assert hashlib.sha256(np.ascontiguousarray(z[key]).tobytes()).hexdigest() == sha


def test_golden_vectors_bit_for_bit():
    z = np.load(GOLD, allow_pickle=False)
    c = golden_weights_case()
    vis, w = O.stokes_i(c["data"], c["weights"], c["flags"])
    assert np.array_equal(vis.view(np.uint32), z["stokes_vis"].view(np.uint32))
    assert np.array_equal(w.view(np.uint32), z["stokes_weight"].view(np.uint32))
    assert [O.padded_size(512), *O.grid_shape(616, 616)] == [int(v) for v in z["sizes"]]
    for mode in ("briggs", "uniform", "natural"):
        out, grid, sums = O.imaging_weights(c["uvw"], c["wavelength"], c["vis_weights"], c["selected"], mode=mode,
                                            robust=0.5, pixel_scale=PIXEL_SCALE)
        assert np.array_equal(out.view(np.uint32), z[f"{mode}_weights"].view(np.uint32))
        assert np.array_equal(grid.view(np.uint64), z[f"{mode}_grid"].view(np.uint64))
        assert np.array_equal(sums, z[f"{mode}_sums"])


def test_padded_size_and_shape():
    # core/src/weights.rs:580-604
    assert O.padded_size(512, 1.2) == 616 and O.padded_size(1024, 1.2) == 1232
    for trimmed in range(2, 4097, 2):
        e = math.ceil(trimmed * 1.2)
        while e % 4:
            e += 1
        assert O.padded_size(trimmed, 1.2) == e
    assert O.grid_shape(616, 616, 3.0) == (206, 206)
    assert O.grid_shape(610, 610, 4.0) == (154, 154)
    assert O.grid_shape(614, 614, 4.0) == (154, 154)


def test_one_cell_formulas():
    # core/src/weights.rs:606-633: samples of weight 2 and 3 in one cell and
    # one off the grid
    u, v = 120.5, 33.25
    uvw = np.array([[u, v, 0.0], [u, v, 0.0], [u + 1e6, v, 0.0], [-u, -v, 0.0]])
    wl = np.array([1.0])
    wts = np.array([[2.0], [3.0], [1.0], [0.0]], np.float32)
    sel = np.ones(4, bool)
    f = 5.0 * math.exp(-math.log(10.0) * 0.5)
    expected = 1.0 / (1.0 + 5.0 * (f * f / 5.0))
    out, _, _ = O.imaging_weights(uvw, wl, wts, sel, mode="briggs", robust=0.5, pixel_scale=PIXEL_SCALE,
                                  rank_filter=None)
    assert out[0, 0] == np.float32(expected) and out[3, 0] == np.float32(expected) and out[2, 0] == 0.0
    out, _, _ = O.imaging_weights(uvw, wl, wts, sel, mode="uniform", pixel_scale=PIXEL_SCALE, rank_filter=None)
    assert out[0, 0] == np.float32(0.2)
    out, _, sums = O.imaging_weights(uvw, wl, wts, sel, mode="natural", pixel_scale=PIXEL_SCALE, rank_filter=None)
    assert out[0, 0] == 1.0 and sums[3] == 2


def test_rank_filter_clips_to_the_window_mean():
    # core/src/weights.rs:635-648: a cell of 100 among five of 1 (one far
    # away), through uniform weighting of summed weights 1/100 and 1
    scale = PIXEL_SCALE
    w = h = 32

    def uv_of(x, y):
        return ((x - w // 2) / (w * scale) + 1e-9, y / (h * scale) + 1e-9)

    cells = [((5, 5), 0.01), ((6, 5), 1.0), ((5, 6), 1.0), ((4, 5), 1.0), ((5, 4), 1.0), ((9, 9), 1.0)]
    uvw = np.array([[*uv_of(x, y), 0.0] for (x, y), _ in cells])
    wts = np.array([[val] for _, val in cells], np.float32)
    _, grid, _ = O.imaging_weights(uvw, np.array([1.0]), wts, np.ones(len(cells), bool), mode="uniform",
                                   width=w, height=h, pixel_scale=scale, rank_filter=(3.0, 16))
    heavy = 1.0 / float(np.float32(0.01))
    mean = (heavy + 5.0) / 6.0
    assert grid[5, 5] == mean * 3.0
    assert grid[5, 6] == 1.0 and grid[9, 9] == 1.0


def test_stokes_i_rules():
    # core/src/weights.rs:650-676
    nan = np.float32("nan")
    d = np.array([[1, 2, 9, 9, 9, 9, 3, -2], [1, 0, 0, 0, 0, 0, 1, 0], [nan, 0, 0, 0, 0, 0, 1, 0],
                  [2, 0, 0, 0, 0, 0, 4, 0]], np.float32).view(np.complex64).reshape(4, 4)
    wt = np.array([[2, 7, 7, 3], [1, 1, 1, 1], [1, 1, 1, 1], [5, 1, 1, 0.5]], np.float32)
    fl = np.array([[0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 0], [0, 0, 0, 0]], bool)
    vis, w = O.stokes_i(d, wt, fl)
    assert np.array_equal(vis, np.array([2, 1, 0, 3], np.complex64))
    assert np.array_equal(w, np.array([8, 0, 0, 2], np.float32))
