"""The schedule tuner (vls_autotune): runs of shapes with and without
register line solvers, tuned, follow bit for bit the trajectories of the
same runs with the default schedules (every schedule runs the same
operations per line in the same order), inside the ensemble's graph; the
choices are cached per shape."""
import numpy as np
import pytest

from tests.configs import HERNQUIST_SMALL, UDF_C1, UDF_ODD, UDF_TINY

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")

CFGS = [UDF_TINY, UDF_ODD, dict(UDF_C1, **{"grid.Pres": 96, "grid.Ures": [128, 24]}),
        dict(HERNQUIST_SMALL, **{"grid.Pres": 64, "grid.Ures": [64, 32]})]


def test_tuned_schedules_give_the_same_bits():
    n = 12
    print(f"case autotune of {len(CFGS)} runs, {n} graph steps", flush=True)
    ref = V.Ensemble(CFGS)
    ref.step(n)
    e = V.Ensemble(CFGS)
    desc = e.autotune()
    print("  " + desc, flush=True)
    assert desc.count(";") == len(CFGS) - 1
    e.step(n)
    for k in range(len(CFGS)):
        assert np.array_equal(e.f(k), ref.f(k)), k
    # a second tuning of the same shapes comes from the cache
    again = V.Ensemble(CFGS[:1])
    assert "cached" in again.autotune()
