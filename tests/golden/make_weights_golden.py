"""Copy the imaging weights' golden vectors out of rang-core's own test data
(tests/data/core-backends-golden.{npz,json}, keys weights.0-weights.12,
written by tests/test_core_backends.py:281-304 in rang-core) into
weights_golden.npz here: the Stokes I visibilities and weights of the 5 x 3
four-correlation case, the padded size of 512 and the grid shape of 616,
and for Briggs (robust 0.5), uniform and natural weighting the imaging
weights, the 308 x 616 finished half-plane grid and the four sums. Data
only; run once with the reference mounted."""

import json
import sys
from pathlib import Path

import numpy as np

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "tests" / "data"
HERE = Path(__file__).resolve().parent

z = np.load(SRC / "core-backends-golden.npz", allow_pickle=False)
meta = json.loads((SRC / "core-backends-golden.json").read_text())


def scalar(key):
    v = meta[key]
    return float.fromhex(v["float"]) if isinstance(v, dict) else float(v)


out = dict(stokes_vis=z["weights.0"], stokes_weight=z["weights.1"],
           sizes=np.array([scalar("weights.2"), scalar("weights.3.0"), scalar("weights.3.1")]))
sums = ("visibility_weight_sum", "total_weight", "max_gridded_weight", "count")
sha = [meta["weights.0"]["sha256"], meta["weights.1"]["sha256"]]
for cls, base in (("briggs", 4), ("uniform", 7), ("natural", 10)):
    out[f"{cls}_weights"] = z[f"weights.{base}"]
    out[f"{cls}_grid"] = z[f"weights.{base + 1}"]
    out[f"{cls}_sums"] = np.array([scalar(f"weights.{base + 2}.{s}") for s in sums])
    sha += [meta[f"weights.{base}"]["sha256"], meta[f"weights.{base + 1}"]["sha256"]]
out["sha256"] = np.array(sha)
np.savez_compressed(HERE / "weights_golden.npz", **out)
print("wrote", HERE / "weights_golden.npz")
