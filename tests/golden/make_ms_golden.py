"""Copy rang-core's Measurement Set fixtures (its own test data) here:
tests/data/msv2-edge (three casacore tables written by python-casacore 3.8.1
with arcae 0.5.9's cells in oracle.json; tests/test_core_ms_edge.py),
tests/data/strarray (a casacore table with python-casacore's cells in
oracle.json; tests/test_core_ms_strarray.py) and tests/data/msv4/m05.zip (an
xradio 1.3.1 MSv4 zarr store with the MSv2 it was converted from;
tests/test_msv4.py). Data only (table files, JSON, a zip); run once with the
reference mounted."""

import hashlib
import shutil
import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "tests" / "data"
DST = Path(__file__).resolve().parent / "ms"

CASES = ["msv2-edge", "strarray", "msv4"]

if DST.exists():
    shutil.rmtree(DST)
lines = []
for case in CASES:
    for f in sorted((SRC / case).rglob("*")):
        if f.is_dir():
            continue
        rel = f.relative_to(SRC)
        out = DST / rel
        out.parent.mkdir(parents=True, exist_ok=True)
        shutil.copyfile(f, out)
        lines.append(f"{hashlib.sha256(f.read_bytes()).hexdigest()}  {rel}")
(DST / "SHA256SUMS").write_text("\n".join(lines) + "\n")
print("copied", len(lines), "files to", DST)
