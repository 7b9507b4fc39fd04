"""Copy the w-stacking gridder's golden vectors out of rang-core's own test
data (tests/data/core-backends-golden.{npz,json}, keys grid.0-grid.3, written
by tests/test_core_backends.py:307-338 in rang-core) into ws_golden.npz here:
the dirty image, the PSF, the oversampled kernel and the counters of the
48 x 48, 4-layer case. Data only; run once with the reference mounted."""

import json
import sys
from pathlib import Path

import numpy as np

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "tests" / "data"
HERE = Path(__file__).resolve().parent

z = np.load(SRC / "core-backends-golden.npz", allow_pickle=False)
meta = json.loads((SRC / "core-backends-golden.json").read_text())
names = ("image_weight", "gridded_count", "visibility_weight_sum", "max_gridded_weight", "min_w",
         "max_w", "psf_normalization", "n_w_layers")
stats = []
for n in names:
    v = meta[f"grid.2.{n}"]
    stats.append(float.fromhex(v["float"]) if isinstance(v, dict) else float(v))
np.savez(HERE / "ws_golden.npz", dirty=z["grid.0"], psf=z["grid.1"], kernel=z["grid.3"],
         stats=np.array(stats), sha256=np.array([meta["grid.0"]["sha256"], meta["grid.1"]["sha256"],
                                                 meta["grid.3"]["sha256"]]))
print("wrote", HERE / "ws_golden.npz")
