#!/usr/bin/env python3
"""Fixtures of the long and command-line runs, from the reference executable.

The reference (oracle/_ref/vlasolve, built by `make -C oracle ref` from the
reference sources, unmodified) is run on

* the 201 x 201 x 200 UDF grid (Pres 200, Ures 200 200, dt 0.005) for 1000
  steps at one OpenMP thread, statistics every 25 steps.  Stored in
  udf_c2_1000.npz: the statistics.txt rows, two whole j-slices of the final
  f (j = 0, which holds the delay-line points, and j = 199), 4096 seeded
  sample points of it, the SHA-256 of all of its bytes, and the moment,
  enclosed-mass profile, density and potential of the final f evaluated by
  the oracle's restatement of the reference quadrature (one thread).  The
  same run at 8 threads gives the reference's own reproducibility floor:
  its moment sums in a per-thread order, and the two final f differ by
  floor_slice0 / floor_slice199 / floor_samples (max |difference| over
  max |f|), also stored;
* the 65 x 65 x 32 UDF case with the reference's own command line
  (statistics every 25 steps, a snapshot at step 200): the snapshot's file
  name and the exact bytes of its header blocks, stored in
  ref_cli_c1.json, for the test of the vlasolve_hip executable.

    python tests/golden/make_golden_long.py
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
from tests.configs import UDF_C1, UDF_C2  # noqa: E402
from vlasolve_amd import ndfield  # noqa: E402

C2_STEPS, C2_THREADS, C2_FLOOR_THREADS, C2_SAMPLES, C2_SEED = 1000, 1, 8, 4096, 1234


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


def stats_rows(wd):
    return [[float(x) for x in line.split()] for line in open(os.path.join(wd, "statistics.txt"))
            if not line.startswith("#")]


def c2_final(threads):
    cfg = dict(UDF_C2)
    cfg.update({"solver.Tmax": 0.005 * C2_STEPS, "solver.statsEvery": 25,
                "solver.snapshotEvery": C2_STEPS, "solver.restartEvery": -1})
    wd = tempfile.mkdtemp(prefix="golden_c2_", dir="/tmp")
    O.run_reference(cfg, wd, threads=threads, timeout=7200)
    snaps = sorted(f for f in os.listdir(wd) if f.startswith("snapshot_"))
    f = ndfield.read(os.path.join(wd, snaps[-1]))[0]
    assert f.shape == (200, 201, 201)
    return f, stats_rows(wd)


def c2_long():
    f, rows = c2_final(C2_THREADS)
    f8, _ = c2_final(C2_FLOOR_THREADS)
    o = O.Oracle(UDF_C2, nthreads=1)
    R, _, _, _ = o.axes()
    o.set_f(f)
    dM, Mx, mtot = o.moments()
    rho, phi = O.density_potential(R, dM, Mx, G=1.0)
    rng = np.random.default_rng(C2_SEED)
    idx = np.sort(rng.choice(f.size, C2_SAMPLES, replace=False))
    fmax = np.abs(f).max()
    floor = {f"floor_slice{j}": np.abs(f8[j] - f[j]).max() / fmax for j in (0, 199)}
    floor["floor_samples"] = np.abs(f8.reshape(-1)[idx] - f.reshape(-1)[idx]).max() / fmax
    np.savez_compressed(HERE / "udf_c2_1000.npz", steps=C2_STEPS, threads=C2_THREADS,
                        stats=np.array(rows), slice0=f[0], slice199=f[199], sample_idx=idx,
                        sample_f=f.reshape(-1)[idx], sha_f=np.array(sha(f)), fmax=fmax,
                        dM=dM, M=Mx, mtot=mtot, rho=rho, phi=phi,
                        floor_threads=C2_FLOOR_THREADS, **floor)
    print("c2 floor (reference at 8 threads vs 1):", floor, flush=True)
    print("c2 done:", len(rows), "statistics rows", flush=True)


def c1_cli():
    cfg = dict(UDF_C1)
    cfg.update({"solver.Tmax": 1.0, "solver.statsEvery": 25, "solver.snapshotEvery": 200,
                "solver.restartEvery": -1})
    wd = tempfile.mkdtemp(prefix="golden_c1_", dir="/tmp")
    O.run_reference(cfg, wd, threads=1)
    snaps = sorted(f for f in os.listdir(wd) if f.startswith("snapshot_"))
    heads = {}
    for s in snaps:
        raw = open(os.path.join(wd, s), "rb").read()
        # tag block (4+16+4) and header block (4+652+4), then the data block's size
        heads[s] = raw[:24 + 660 + 4].hex()
    out = {"generator": "oracle/_ref/vlasolve at OMP_NUM_THREADS=1",
           "argv": O.to_argv(cfg), "snapshots": heads, "stats": stats_rows(wd)}
    (HERE / "ref_cli_c1.json").write_text(json.dumps(out, indent=1))
    print("c1 cli done:", sorted(heads), flush=True)


if __name__ == "__main__":
    if not O.REF_EXE.exists():
        sys.exit("build the reference first: make -C oracle ref")
    what = sys.argv[1:] or ["c1", "c2"]
    if "c1" in what:
        c1_cli()
    if "c2" in what:
        c2_long()
