"""The drop-in surface: the vlasolve_hip executable with the reference's own
command line for the 65 x 65 x 32 UDF case (statistics every 25 steps,
snapshots at steps 0 and 200), against what the reference executable writes
for the same line (fixture tests/golden/ref_cli_c1.json, written by
make_golden_long.py): the statistics.txt rows to the printed precision, the
snapshot file names and the exact bytes of their NDfield header blocks
(vlasovSolver21::snapshot, src/vlasovSolver21.hxx:717-773; Save_NDfield,
src/NDfield.hxx:308-362), and the snapshot data against the oracle run in
the reference's order (whose bytes are pinned to the reference) to 1e-11.
"""
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.configs import UDF_C1

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")
from oracle import oracle as O  # noqa: E402
from vlasolve_amd import ndfield  # noqa: E402

FIX = json.loads((Path(__file__).resolve().parent / "golden" / "ref_cli_c1.json").read_text())


def test_executable_matches_reference_outputs(tmp_path):
    assert V.EXE_PATH.exists(), "vlasolve_hip not built"
    print("case vlasolve_hip udf 65x65x32, 200 steps", flush=True)
    res = subprocess.run([str(V.EXE_PATH)] + FIX["argv"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    rows = [[float(x) for x in line.split()] for line in open(tmp_path / "statistics.txt")
            if not line.startswith("#")]
    assert len(rows) == len(FIX["stats"])
    for got, ref in zip(rows, FIX["stats"]):
        assert got[:2] == ref[:2]  # time, step
        for a, b in zip(got[2:], ref[2:]):
            assert a == pytest.approx(b, rel=2e-9, abs=1e-300)
    snaps = sorted(p for p in os.listdir(tmp_path) if p.startswith("snapshot_"))
    assert snaps == sorted(FIX["snapshots"])
    for name, hexhead in FIX["snapshots"].items():
        raw = (tmp_path / name).read_bytes()
        head = bytes.fromhex(hexhead)
        assert raw[:len(head)] == head, f"{name}: header bytes differ"
    f_end = ndfield.read(tmp_path / snaps[-1])[0]
    o = O.Oracle(UDF_C1, nthreads=1)
    f0 = ndfield.read(tmp_path / snaps[0])[0]
    assert np.array_equal(f0, o.f())
    o.step(200)
    fr = o.f()
    rel = np.abs(f_end - fr).max() / np.abs(fr).max()
    print(f"  snapshot at step 200: max rel diff vs the reference order {rel:.3e}", flush=True)
    assert rel <= 1e-11
