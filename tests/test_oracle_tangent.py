"""The oracle's tangent-linear model against central finite differences of
the oracle's own step: the check that makes it a valid reference for the
adjoint of the HIP path.  The step is linear in (f, delay-line values) except
through the kick's acceleration, so the differences are accurate to the
rounding of the step itself."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.configs import UDF_TINY, HERNQUIST_SMALL

CASES = {
    "udf_tiny": (UDF_TINY, 6),
    "hernquist": (dict(HERNQUIST_SMALL, **{"grid.Pres": 24, "grid.Ures": [24, 8]}), 5),
}


def mass_after(cfg, n, f0, vals0):
    o = O.Oracle(cfg, nthreads=1)
    o.set_f(f0)
    if vals0 is not None and vals0.size:
        o.set_chain_vals(vals0)
    o.step(n)
    return o.moments()[1]  # exclusive-scan mass profile


@pytest.mark.parametrize("name", list(CASES))
def test_tangent_matches_central_differences(name):
    cfg, n = CASES[name]
    base = O.Oracle(cfg, nthreads=1)
    f0 = base.f()
    vals0 = base.chains()["val"]
    rng = np.random.default_rng(7)
    df = rng.standard_normal(f0.shape) * np.abs(f0).max() * 1e-1
    dv = rng.standard_normal(vals0.shape) * (np.abs(vals0).max() if vals0.size else 1) * 1e-1
    # tangent
    o = O.Oracle(cfg, nthreads=1)
    o.tl_init(df, dv if dv.size else None)
    o.tl_step(n)
    M, dM = o.tl_mass()
    # central differences; the kick zeroes targets outside (U_min, U_max), a
    # jump that a step of 1e-6 already crosses for some points of the
    # Hernquist grid, while 1e-8 stays on the smooth side and is still far
    # above the rounding of the step
    h = 1e-8
    Mp = mass_after(cfg, n, f0 + h * df, vals0 + h * dv)
    Mm = mass_after(cfg, n, f0 - h * df, vals0 - h * dv)
    fd = (Mp - Mm) / (2 * h)
    scale = np.abs(dM).max()
    assert scale > 0
    assert np.abs(fd - dM).max() <= 1e-6 * scale, np.abs(fd - dM).max() / scale
    # and the state itself is the plain step's
    ref = O.Oracle(cfg, nthreads=1)
    ref.step(n)
    assert np.array_equal(ref.f(), o.f())
