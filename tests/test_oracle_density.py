"""The oracle's shell density and potential (oracle.density_potential)
against what the mathematics fixes: the mass they integrate to, a direct sum
over shells, the uniform sphere and the discrete force identity.  Each check
fails for a plausible slip (a dropped shell, the midpoint on the wrong side,
a sign, the enclosed mass taken inclusive)."""
import math

import numpy as np

from oracle import oracle as O


def _grid(n=60, rmin=0.01, rmax=25.0):
    return np.exp(np.linspace(math.log(rmin), math.log(rmax), n))


def test_density_integrates_to_the_mass():
    R = _grid()
    rng = np.random.default_rng(3)
    dM = rng.uniform(0.0, 1.0, R.size)
    dM[-1] = 0.0
    rho, _ = O.density_potential(R, dM)
    vol = 4.0 * math.pi / 3.0 * (R[1:] ** 3 - R[:-1] ** 3)
    assert rho[-1] == 0.0
    assert abs((rho[:-1] * vol).sum() - dM.sum()) <= 1e-13 * dM.sum()


def test_potential_is_the_shell_sum():
    """phi_i = -G sum_k dM_k / max-radius: shells below R_i act from the
    centre (1/R_i), the others from their midpoints."""
    R = _grid(40)
    rng = np.random.default_rng(4)
    dM = rng.uniform(0.0, 1.0, R.size)
    dM[-1] = 0.0
    G = 0.7
    _, phi = O.density_potential(R, dM, G=G)
    rbar = 0.5 * (R[:-1] + R[1:])
    for i in range(R.size):
        direct = -G * sum(dM[k] / (R[i] if k < i else rbar[k]) for k in range(R.size - 1))
        assert abs(phi[i] - direct) <= 1e-14 * abs(direct)


def test_uniform_sphere():
    """Shell masses of a uniform sphere of radius R[m]: the density is the
    constant, and outside the sphere the potential is -G M / R."""
    R = _grid(50, 0.1, 10.0)
    m, rho0 = 30, 2.5
    dM = np.zeros(R.size)
    dM[:m] = rho0 * 4.0 * math.pi / 3.0 * (R[1:m + 1] ** 3 - R[:m] ** 3)
    rho, phi = O.density_potential(R, dM)
    assert np.allclose(rho[:m], rho0, rtol=1e-13, atol=0)
    assert np.all(rho[m:] == 0.0)
    Mtot = dM.sum()
    for i in range(m, R.size):
        assert abs(phi[i] + Mtot / R[i]) <= 1e-14 * Mtot / R[i]


def test_discrete_force_identity():
    """phi_{i+1} - phi_i = G M_i (1/R_i - 1/R_{i+1}) + G dM_i (1/rbar_i - 1/R_{i+1}):
    the potential's difference is the force of the enclosed mass over the
    shell plus the shell's own term."""
    R = _grid(45)
    rng = np.random.default_rng(5)
    dM = rng.uniform(0.0, 1.0, R.size)
    dM[-1] = 0.0
    G = 1.3
    _, phi = O.density_potential(R, dM, G=G)
    Mx = np.concatenate([[0.0], np.cumsum(dM)[:-1]])
    for i in range(R.size - 1):
        rbar = 0.5 * (R[i] + R[i + 1])
        want = G * Mx[i] * (1 / R[i] - 1 / R[i + 1]) + G * dM[i] * (1 / rbar - 1 / R[i + 1])
        assert abs((phi[i + 1] - phi[i]) - want) <= 1e-12 * abs(phi).max()


def test_explicit_enclosed_mass_matches_the_scan():
    R = _grid(30)
    dM = np.linspace(0.1, 1.0, R.size)
    dM[-1] = 0.0
    Mx = np.concatenate([[0.0], np.cumsum(dM)[:-1]])
    a = O.density_potential(R, dM)
    b = O.density_potential(R, dM, Mx)
    assert np.array_equal(a[0], b[0]) and np.allclose(a[1], b[1], rtol=1e-15, atol=0)
