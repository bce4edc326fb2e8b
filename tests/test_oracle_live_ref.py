"""The oracle against a live run of the reference executable (skipped when
oracle/_ref was not built): configurations outside the fixtures."""
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from tests.configs import UDF_TINY
from vlasolve_amd import ndfield

pytestmark = pytest.mark.skipif(not O.REF_EXE.exists(), reason="reference executable not built")

CASES = {
    "udf_alpha": dict(UDF_TINY, **{"init.alpha": 0.8, "init.R0": 1.5, "grid.Pres": 24,
                                   "grid.Ures": [20, 12]}),
    "pldf_n": {"init.type": "PLDF", "init.n": 0.5, "init.eta": 0.2, "init.alpha": 1.5,
               "grid.Pres": 24, "grid.Ures": [24, 10], "grid.Pmax": 20,
               "grid.Umin": [-2, 0], "grid.Umax": [2, 1.6]},
    "udf_nolink_reflect": dict(UDF_TINY, **{"solver.noLinking": 1,
                                            "solver.kernelType": "reflective"}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_live(name):
    cfg = dict(CASES[name])
    n = 15
    cfg.update({"solver.Tmax": 0.005 * n, "solver.statsEvery": -1, "solver.snapshotEvery": n,
                "solver.restartEvery": -1})
    with tempfile.TemporaryDirectory() as wd:
        O.run_reference(cfg, wd, threads=1)
        snaps = sorted(f for f in os.listdir(wd) if f.startswith("snapshot_"))
        ref = ndfield.read(os.path.join(wd, snaps[-1]))[0]
    o = O.Oracle(CASES[name], nthreads=1)
    o.step(n)
    assert np.array_equal(o.f(), ref)
