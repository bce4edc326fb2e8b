"""Ensembles: several runs of different grids in one context, stepped
concurrently through a HIP graph.  Every run must follow exactly the
trajectory it follows alone (bit for bit), across several graph replays
with downloads (synchronisations) between them, and with the graph off."""
import numpy as np
import pytest

from tests.configs import UDF_TINY, UDF_ODD, HERNQUIST_SMALL, PLDF_SMALL

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")

CFGS = [
    UDF_TINY,
    UDF_ODD,
    dict(HERNQUIST_SMALL, **{"grid.Pres": 32, "grid.Ures": [32, 16]}),
    dict(PLDF_SMALL, **{"grid.Pres": 24, "grid.Ures": [40, 6]}),
]


def alone(cfg, n):
    s = V.Solver(cfg)
    s.step(n)
    return s.f()


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_runs_match_single_runs(graph):
    print(f"case ensemble of {len(CFGS)} graph={graph}", flush=True)
    e = V.Ensemble(CFGS)
    e.set_graph(graph)
    done = 0
    for k in (3, 4, 1):  # replays separated by downloads
        e.step(k)
        done += k
        for i, cfg in enumerate(CFGS):
            fe = e.f(i)
            fa = alone(cfg, done)
            bad = np.count_nonzero(fe != fa)
            assert bad == 0, f"run {i} after {done} steps: {bad} points differ"
