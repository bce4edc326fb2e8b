"""The reference-side binding on the GPU: vlasolve built from the reference's
own sources with the vlasovSolver21_hip policy (integration/, output
oracle/_ref/vlasolve_hip_policy) runs the 65 x 65 x 32 UDF case with the
reference's command line.  Its statistics.txt rows and snapshot header bytes
match what the reference executable writes (tests/golden/ref_cli_c1.json);
its final snapshot and final .vls restart hold f bit for bit as the oracle's
device order computes it (the policy feeds the library the geometry, feet and
delay lines of the reference's own set-up, so its step is the library's);
and the restart's delay-line values are the library's own after 200 steps.
Skipped when the policy executable was not built (no reference sources where
the build ran)."""
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
from vlasolve_amd import ndfield, vlsfile  # noqa: E402

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "oracle" / "_ref" / "vlasolve_hip_policy"
FIX = json.loads((ROOT / "tests" / "golden" / "ref_cli_c1.json").read_text())


@pytest.mark.skipif(not EXE.exists(), reason="policy executable not built")
def test_reference_with_hip_policy(tmp_path):
    print("case reference + vlasovSolver21_hip, udf 65x65x32, 200 steps", flush=True)
    res = subprocess.run([str(EXE)] + FIX["argv"] + ["-threads_per_node", "1"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    rows = [[float(x) for x in line.split()] for line in open(tmp_path / "statistics.txt")
            if not line.startswith("#")]
    assert len(rows) == len(FIX["stats"])
    for got, ref in zip(rows, FIX["stats"]):
        assert got[:2] == ref[:2]
        for a, b in zip(got[2:], ref[2:]):
            assert a == pytest.approx(b, rel=2e-9, abs=1e-300)
    snaps = sorted(p for p in os.listdir(tmp_path) if p.startswith("snapshot_"))
    assert snaps == sorted(FIX["snapshots"])
    for name, hexhead in FIX["snapshots"].items():
        head = bytes.fromhex(hexhead)
        assert (tmp_path / name).read_bytes()[:len(head)] == head, name
    o = O.Oracle(UDF_C1, nthreads=0)
    o.step(200)
    f_dev = o.f()
    f_snap = ndfield.read(tmp_path / snaps[-1])[0]
    assert np.array_equal(f_snap, f_dev)
    finals = sorted(p for p in tmp_path.iterdir() if p.name.startswith("final_"))
    assert finals, "no final restart written"
    state = vlsfile.read(finals[-1])
    assert np.array_equal(state["f"], f_dev)
    # the delay-line values the restart holds, front first, against the
    # library's own run
    s = V.Solver(UDF_C1)
    s.step(200)
    head, slots = s.context.get_chains()
    ch = O.Oracle(UDF_C1, nthreads=1).chains()
    front = np.concatenate([slots[o:o + L][(h + np.arange(L)) % L]
                            for h, o, L in zip(head, ch["off"], ch["len"])])
    keys = sorted(state["chains"])
    assert np.array_equal(np.array(keys), ch["point"])
    got = np.concatenate([state["chains"][k][2] for k in keys])
    assert np.array_equal(got, front)
