"""Mathematical pins of the oracle's building blocks, independent of the
reference run: each test compares against something the mathematics fixes
(a closed form, a textbook routine, an invariant), chosen so that a dropped
term, a wrong sign or index, or a transposed operand fails."""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle as O
from tests.configs import UDF_TINY

CFG = dict(UDF_TINY, **{"grid.Pres": 20, "grid.Ures": [14, 6]})


@pytest.fixture(scope="module")
def orc():
    return O.Oracle(CFG, nthreads=1)


def knots(orc):
    """spline knots in rescaled coordinates: log r (uniform), v (uniform)"""
    R, U, _, _ = orc.axes()
    return np.log(R), U


def padded_eval(orc, coef):
    c = np.zeros(coef.size + 4 * (orc.nR + 2))
    c[: coef.size] = coef.ravel()
    return c


# ---------------------------------------------------------------- spline
def test_interpolates_at_nodes(orc):
    rng = np.random.default_rng(1)
    f = rng.uniform(-1, 1, (orc.nU, orc.nR))
    c = padded_eval(orc, orc.spline_build(f))
    for u in range(orc.nU):
        for r in range(orc.nR):
            assert orc.spline_evalC(c, r, 0.0, u, 0.0) == pytest.approx(f[u, r], abs=1e-13)


def test_reproduces_linear_data(orc):
    x, y = knots(orc)
    X, Y = np.meshgrid(x, y)
    f = 0.7 - 1.3 * X + 2.1 * Y
    c = padded_eval(orc, orc.spline_build(f))
    rng = np.random.default_rng(2)
    hx, hy = x[1] - x[0], y[1] - y[0]
    for _ in range(200):
        i, j = rng.integers(0, orc.nR - 1), rng.integers(0, orc.nU - 1)
        a, b = rng.uniform(0, 1, 2)
        exact = 0.7 - 1.3 * (x[i] + a * hx) + 2.1 * (y[j] + b * hy)
        assert orc.spline_evalC(c, i, a, j, b) == pytest.approx(exact, abs=1e-12)


def test_natural_end_conditions(orc):
    # natural rows: c0 - 2 c1 + c2 = 0 and c_{M-3} - 2 c_{M-2} + c_{M-1} = 0 on
    # every line of both directions
    rng = np.random.default_rng(3)
    c = orc.spline_build(rng.uniform(0, 1, (orc.nU, orc.nR)))
    scale = np.abs(c).max()
    assert np.abs(c[:, 0] - 2 * c[:, 1] + c[:, 2]).max() < 1e-13 * scale
    assert np.abs(c[:, -3] - 2 * c[:, -2] + c[:, -1]).max() < 1e-13 * scale
    assert np.abs(c[0, :] - 2 * c[1, :] + c[2, :]).max() < 1e-13 * scale
    assert np.abs(c[-3, :] - 2 * c[-2, :] + c[-1, :]).max() < 1e-13 * scale


def collocation(n):
    """dense (n+2)x(n+2) B-spline system: natural rows at both ends, interior
    rows (1,4,1)/6 against the data"""
    A = np.zeros((n + 2, n + 2))
    A[0, :3] = [1, -2, 1]
    A[-1, -3:] = [1, -2, 1]
    for i in range(1, n + 1):
        A[i, i - 1:i + 2] = [1 / 6, 4 / 6, 1 / 6]
    return A


def test_matches_dense_solve(orc):
    rng = np.random.default_rng(4)
    f = rng.uniform(-1, 1, (orc.nU, orc.nR))
    c = orc.spline_build(f)
    Ar, Av = collocation(orc.nR), collocation(orc.nU)
    rhs = np.zeros((orc.nU + 2, orc.nR + 2))
    rhs[1:-1, 1:-1] = f
    # first along r on the data rows, then along v on every column
    rows = np.zeros_like(rhs)
    for u in range(1, orc.nU + 1):
        b = np.zeros(orc.nR + 2)
        b[1:-1] = f[u - 1]
        rows[u] = np.linalg.solve(Ar, b)
    dense = np.zeros_like(rhs)
    for k in range(orc.nR + 2):
        b = np.zeros(orc.nU + 2)
        b[1:-1] = rows[1:-1, k]
        dense[:, k] = np.linalg.solve(Av, b)
    assert np.abs(c - dense).max() < 1e-12 * np.abs(dense).max()


def test_matches_scipy_natural_spline_on_node_rows(orc):
    # on a node row (y fraction 0) the tensor spline is the natural cubic
    # spline of that row's data in the rescaled r coordinate
    x, y = knots(orc)
    rng = np.random.default_rng(5)
    f = rng.uniform(-1, 1, (orc.nU, orc.nR))
    c = padded_eval(orc, orc.spline_build(f))
    for u in (0, 3, orc.nU - 1):
        cs = CubicSpline(x, f[u], bc_type="natural")
        for _ in range(30):
            i = rng.integers(0, orc.nR - 1)
            a = rng.uniform(0, 1)
            got = orc.spline_evalC(c, i, a, u, 0.0)
            assert got == pytest.approx(cs(x[i] + a * (x[1] - x[0])), abs=1e-12)


def test_kick_evaluation_is_v_spline_of_line(orc):
    # evaluateNxPy at node r equals the natural v-spline of column r
    x, y = knots(orc)
    rng = np.random.default_rng(6)
    f = rng.uniform(0, 1, (orc.nU, orc.nR))
    c = padded_eval(orc, orc.spline_build(f))
    for r in (0, 5, orc.nR - 1):
        cs = CubicSpline(y, f[:, r], bc_type="natural")
        for v in rng.uniform(y[0] + 1e-9, y[-1] - 1e-9, 20):
            assert orc.spline_evalNxPy(c, r, v) == pytest.approx(cs(v), abs=1e-12)


# ---------------------------------------------------------------- feet
def energy(R, U, J):
    return 0.5 * (U * U + J * J / (R * R))


@pytest.mark.parametrize("R,U,J,dt", [(1.0, 0.3, 0.5, 0.05), (2.0, -0.4, 0.2, 0.1),
                                      (0.2, 0.5, 0.05, 0.3), (0.5, 0.01, 0.3, 0.5),
                                      (3.0, 0.0, 1.0, 0.2)])
def test_foot_conserves_energy_and_matches_rk4(R, U, J, dt):
    nR, nU = O.advance_drift(R, U, J, dt)
    assert energy(nR, nU, J) == pytest.approx(energy(R, U, J), rel=1e-12)

    # integrate dr/dt = u, du/dt = J^2/r^3 backwards over dt
    def rhs(s):
        return np.array([s[1], J * J / s[0] ** 3])

    s = np.array([R, U])
    n = 20000
    h = -dt / n
    for _ in range(n):
        k1 = rhs(s)
        k2 = rhs(s + 0.5 * h * k1)
        k3 = rhs(s + 0.5 * h * k2)
        k4 = rhs(s + h * k3)
        s = s + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert nR == pytest.approx(s[0], rel=1e-8)
    assert nU == pytest.approx(s[1], rel=1e-7, abs=1e-9)


def test_foot_closed_form_without_pericentre():
    # inward motion (u<0) backwards in time only moves outwards:
    # r(-t)^2 = R^2 - 2 R U t + (U^2 + J^2/R^2) t^2
    for R, U, J, t in [(1.0, -0.3, 0.4, 0.2), (0.3, -1.0, 0.1, 0.05)]:
        nR, _ = O.advance_drift(R, U, J, t)
        assert nR == pytest.approx(math.sqrt(R * R - 2 * R * U * t + (U * U + J * J / R ** 2) * t * t),
                                   rel=1e-13)


# ---------------------------------------------------------------- scales
def test_scales_closed_forms():
    lin = O.gen_scale(-2.0, 2.0, 8, O.SCALES["linear"])
    assert np.allclose(lin, np.linspace(-2, 2, 9), rtol=0, atol=1e-15)
    lg = O.gen_scale(0.01, 25.0, 10, O.SCALES["logarithmic"])
    assert lg[0] == 0.01 and lg[-1] == 25.0
    assert np.allclose(np.diff(np.log(lg)), math.log(2500) / 10, rtol=1e-12)
    q = O.gen_scale(0.0, 1.6, 4, O.SCALES["quadratic"])
    assert np.allclose(q, (np.linspace(0, math.sqrt(1.6), 5)) ** 2, rtol=1e-15)
    qc = O.gen_scale(0.0, 1.6, 4, O.SCALES["quadratic"], cell=True)
    assert np.allclose(qc, 0.5 * (q[1:] + q[:-1]), rtol=1e-15)


# ---------------------------------------------------------------- quadrature
def test_quadrature_closed_forms_for_constant_f(orc):
    R, U, J, JV = orc.axes()
    orc2 = O.Oracle(CFG, nthreads=1)
    orc2.set_f(np.full(orc2.shape, 2.0))
    dJ = np.diff(JV)
    dR, dU = np.diff(R), np.diff(U)
    vol = 8 * math.pi ** 2 * (J * dJ).sum() * dR.sum() * dU.sum()
    s = orc2.stats()
    assert s["M"] == pytest.approx(2.0 * vol, rel=1e-12)
    assert s["L1"] == pytest.approx(2.0 * vol, rel=1e-12)
    assert s["L2"] == pytest.approx(4.0 * vol, rel=1e-12)
    assert s["S"] == pytest.approx(-2.0 * math.log(2.0) * vol, rel=1e-12)
    rc = R[:-1] + 0.5 * dR
    uc = U[:-1] + 0.5 * dU
    w = 8 * math.pi ** 2 * (J * dJ)[:, None, None] * dU[None, :, None] * dR[None, None, :]
    ek = (w * 0.5 * 2.0 * (uc[None, :, None] ** 2 + (J ** 2)[:, None, None] / rc[None, None, :] ** 2)).sum()
    assert s["Ek"] == pytest.approx(ek, rel=1e-12)
    # dM sums to the mass; the exclusive prefix steps by dM
    dM, Mx, mtot = orc2.moments()
    assert mtot == pytest.approx(dM.sum(), rel=1e-14)
    assert Mx[0] == 0 and np.allclose(np.diff(Mx), dM[:-1], rtol=1e-12)
    assert dM[-1] == 0
    # potential energy: shell k meets M(<R_k) at its midpoint
    mid = 0.5 * (R[1:] + R[:-1])
    ep = -(dM[:-1] * Mx[:-1] / mid).sum()
    assert s["Ep"] == pytest.approx(ep, rel=1e-12)


def test_renormalised_mass_is_exact(orc):
    assert orc.stats()["M"] == pytest.approx(1.0, rel=1e-14)


def test_device_order_moment_matches_reference_order():
    """The device order (nthreads=0: per-term ((Ct1 dU 0.25) dR) times the
    corner sum, summed in groups of 16 rows and chunks of 20 slices) sums the
    same terms as the reference's quadrature (nthreads=1), so dM agrees to
    rounding; for constant f both give the closed form.  A dropped 0.25, a
    wrong corner or a wrong weight would be off by far more."""
    dev = O.Oracle(CFG, nthreads=0)
    ref = O.Oracle(CFG, nthreads=1)
    rng = np.random.default_rng(7)
    f = rng.uniform(0.1, 2.0, dev.shape)
    dev.set_f(f)
    ref.set_f(f)
    dMd, Mxd, td = dev.moments()
    dMr, Mxr, tr = ref.moments()
    scale = np.abs(dMr).max()
    assert np.abs(dMd - dMr).max() <= 1e-13 * scale
    assert td == pytest.approx(tr, rel=1e-13)
    # constant f: dM[r] = 2 * 8 pi^2 (sum_j J dJ) (sum_u dU) dR[r]
    R, U, J, JV = dev.axes()
    dev.set_f(np.full(dev.shape, 2.0))
    dMc, _, _ = dev.moments()
    closed = 2.0 * 8 * math.pi ** 2 * (J * np.diff(JV)).sum() * np.diff(U).sum() * np.diff(R)
    assert np.allclose(dMc[:-1], closed, rtol=1e-12)
    assert dMc[-1] == 0
