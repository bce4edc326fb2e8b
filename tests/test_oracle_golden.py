"""The oracle against the reference executable's own outputs.

Fixtures (tests/golden/, written by make_golden.py from the reference built
from its sources, one thread): SHA-256 of the exact bytes of f at step 0 and
at the end of each case, statistics.txt (10 significant digits), and for the
smallest case the full arrays and the delay-line chains of its final restart
file.  The oracle at one thread must reproduce every byte.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from tests.configs import SMALL_CASES

GOLD = Path(__file__).resolve().parent / "golden"
FIX = json.loads((GOLD / "ref_fixtures.json").read_text())


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("name", list(SMALL_CASES))
def test_trajectory_bytes(name):
    cfg, nsteps = SMALL_CASES[name]
    fx = FIX["cases"][name]
    assert fx["steps"] == nsteps
    o = O.Oracle(cfg, nthreads=1)
    assert sha(o.f()) == fx["sha_f0"], "initial conditions differ"
    assert len(o.chains()["len"]) == fx["n_chains"]
    # statistics are written before the step of their row (vlasovSolver.hxx:170-174);
    # statistics.txt holds the line of all species
    rows = {int(r[1]): r for r in fx["stats"]}
    for k in range(nsteps + 1):
        if k in rows:
            s = o.stats_total()
            ref = rows[k]
            for i, key in enumerate(("M", "E", "Ek", "Ep", "S", "L1", "L2")):
                assert s[key] == pytest.approx(ref[2 + i], rel=2e-9, abs=1e-300), (k, key)
        if k < nsteps:
            o.step(1)
    assert sha(o.f()) == fx["sha_f"], "final f differs"


def test_tiny_arrays_and_chains():
    cfg, nsteps = SMALL_CASES["udf_tiny"]
    z = np.load(GOLD / "udf_tiny.npz")
    o = O.Oracle(cfg, nthreads=1)
    assert np.array_equal(o.f(), z["f0"])
    o.step(nsteps)
    assert np.array_equal(o.f(), z["f"])
    ch = o.chains()
    assert np.array_equal(ch["point"], z["chain_point"])
    assert np.array_equal(ch["len"], z["chain_len"])
    # cell positions are static, values have rolled nsteps times
    assert np.array_equal(ch["R"], z["chain_R"])
    assert np.array_equal(ch["U"], z["chain_U"])
    assert np.array_equal(ch["val"], z["chain_val"])


def test_reference_reproducibility_floor_recorded():
    # the reference itself differs between 1 and 8 threads by summation order
    floor = FIX["reference_floor_1_vs_8_threads_udf_c1"]
    assert 0 < floor < 1e-12


def test_long_fixture_initial_row():
    """The 1000-step reference fixture of the 201 x 201 x 200 grid starts
    from the oracle's initial conditions (its first statistics row), and its
    stored moment, density and potential follow from its stored final f."""
    from tests.configs import UDF_C2

    z = np.load(GOLD / "udf_c2_1000.npz")
    o = O.Oracle(UDF_C2, nthreads=1)
    s = o.stats_total()
    row = z["stats"][0]
    assert row[1] == 0
    for i, key in enumerate(("M", "E", "Ek", "Ep", "S", "L1", "L2")):
        assert s[key] == pytest.approx(row[2 + i], rel=2e-9, abs=1e-300), key
    R = o.axes()[0]
    rho, phi = O.density_potential(R, z["dM"], z["M"], G=1.0)
    assert np.array_equal(rho, z["rho"]) and np.array_equal(phi, z["phi"])
    assert z["M"][-1] + z["dM"][-1] == pytest.approx(float(z["mtot"]), rel=1e-14)


def test_cli_fixture_agrees_with_case_fixture():
    """The command-line fixture and the udf_c1 case fixture are the same run
    (the same statistics rows)."""
    cli = json.loads((GOLD / "ref_cli_c1.json").read_text())
    case = FIX["cases"]["udf_c1"]["stats"]
    assert cli["stats"] == case
