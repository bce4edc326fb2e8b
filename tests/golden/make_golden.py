#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the reference executable.

The reference (oracle/_ref/vlasolve, built by `make -C oracle ref` from the
reference sources, unmodified) is run at one OpenMP thread on the cases of
tests/configs.py.  Stored per case: the SHA-256 of the exact bytes of f at the
end of the run and at step 0, statistics.txt, and for the smallest case the
full f arrays and the delay-line chains of the final restart file.  Also the
reference's own reproducibility floor: the largest relative difference of f
between one and eight threads on the 65x65x32 case after 200 steps.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from tests.configs import SMALL_CASES  # noqa: E402
from vlasolve_amd import ndfield, vlsfile  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


def run(cfg, nsteps, threads=1):
    cfg = dict(cfg)
    dt = float(cfg.get("solver.dt", 0.005))
    cfg.update({"solver.Tmax": dt * nsteps, "solver.statsEvery": max(1, nsteps // 8),
                "solver.snapshotEvery": nsteps, "solver.restartEvery": -1})
    wd = tempfile.mkdtemp(prefix="golden_", dir="/tmp")
    O.run_reference(cfg, wd, threads=threads)
    snaps = sorted(f for f in os.listdir(wd) if f.startswith("snapshot_"))
    f0 = ndfield.read(os.path.join(wd, snaps[0]))[0]
    f1 = ndfield.read(os.path.join(wd, snaps[-1]))[0]
    stats = [[float(x) for x in line.split()] for line in open(os.path.join(wd, "statistics.txt"))
             if not line.startswith("#")]
    vls = [f for f in os.listdir(wd) if f.startswith("final_")][0]
    state = vlsfile.read(os.path.join(wd, vls))
    return f0, f1, stats, state


def main():
    if not O.REF_EXE.exists():
        sys.exit("build the reference first: make -C oracle ref")
    out = {"generator": "oracle/_ref/vlasolve at OMP_NUM_THREADS=1", "cases": {}}
    for name, (cfg, nsteps) in SMALL_CASES.items():
        f0, f1, stats, state = run(cfg, nsteps)
        assert np.array_equal(state["f"], f1), "restart f differs from the last snapshot"
        out["cases"][name] = {"steps": nsteps, "sha_f0": sha(f0), "sha_f": sha(f1),
                              "stats": stats, "n_chains": len(state["chains"])}
        if name == "udf_tiny":
            keys = np.array(sorted(state["chains"]), np.int64)
            R = np.concatenate([state["chains"][k][0] for k in keys])
            U = np.concatenate([state["chains"][k][1] for k in keys])
            Vv = np.concatenate([state["chains"][k][2] for k in keys])
            L = np.array([len(state["chains"][k][0]) for k in keys], np.int32)
            np.savez_compressed(HERE / "udf_tiny.npz", f0=f0, f=f1, chain_point=keys,
                                chain_len=L, chain_R=R, chain_U=U, chain_val=Vv)
        print(name, "done", flush=True)
    # reproducibility floor of the reference itself
    cfg, nsteps = SMALL_CASES["udf_c1"]
    _, f1, _, _ = run(cfg, nsteps, threads=1)
    _, f8, _, _ = run(cfg, nsteps, threads=8)
    out["reference_floor_1_vs_8_threads_udf_c1"] = float(np.abs(f8 - f1).max() / np.abs(f1).max())
    (HERE / "ref_fixtures.json").write_text(json.dumps(out, indent=1))
    print("floor", out["reference_floor_1_vs_8_threads_udf_c1"])


if __name__ == "__main__":
    main()
