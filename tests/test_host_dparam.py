"""The hand-written derivative of the initial state in the initial-condition
parameters (vlsd_host_dparam) against central differences of the host
set-up itself, on the CPU: f0 after the renormalisation and the delay-line
values, for every parameter of the UDF and PLDF generators.  Central
differences with a relative step of 1e-6 carry a truncation error of about
1e-12 and a rounding error of about 1e-10 of the derivative's scale."""
import numpy as np
import pytest

import vlasolve_amd as V
from tests.configs import PLDF_SMALL, UDF_TINY

UDF = dict(UDF_TINY, **{"grid.Pres": 24, "grid.Ures": [20, 10]})
# explicit ranges: the grid does not move with the parameters
PLDF = dict(PLDF_SMALL, **{"grid.Pres": 24, "grid.Ures": [20, 10]})
PLDF_R0 = dict(PLDF, **{"init.R0": 8.0})

CASES = [(UDF, p) for p in ("M", "R0", "alpha", "smoothingRadius")] + \
        [(PLDF, p) for p in ("M", "eta", "alpha", "R0", "smoothingRadius")] + \
        [(PLDF_R0, "eta")]
DEFAULTS = {"UDF": {"M": 1.0, "R0": 2.0, "alpha": 0.5, "smoothingRadius": 0.5},
            "PLDF": {"M": 1.0, "eta": 0.1, "alpha": 2.0, "smoothingRadius": 0.5}}


def _value(cfg, p):
    key = "init." + p
    if key in cfg:
        return float(cfg[key])
    if p == "R0" and cfg["init.type"] == "PLDF":
        return 1.0 / _value(cfg, "eta")
    return DEFAULTS[cfg["init.type"]][p]


@pytest.mark.parametrize("cfg,param", CASES,
                         ids=[f"{c['init.type']}{'-R0set' if 'init.R0' in c else ''}-{p}"
                              for c, p in CASES])
def test_dparam_against_central_differences(cfg, param):
    x = _value(cfg, param)
    h = 1e-6 * abs(x)
    base = V.HostSetup(cfg)
    df, dv = base.dparam(param)
    plus = V.HostSetup(dict(cfg, **{"init." + param: x + h}))
    minus = V.HostSetup(dict(cfg, **{"init." + param: x - h}))
    fd = (plus.f() - minus.f()) / (2 * h)
    scale = np.abs(fd).max()
    assert scale > 0
    assert np.abs(df - fd).max() <= 1e-6 * scale, np.abs(df - fd).max() / scale
    if base.n_cells:
        fdv = (plus.chains()["val"] - minus.chains()["val"]) / (2 * h)
        assert np.abs(dv - fdv).max() <= 1e-6 * max(np.abs(fdv).max(), 1e-300)


def test_dparam_rejects_unknown_parameters():
    with pytest.raises(V.VlsError, match="no derivative"):
        V.HostSetup(UDF).dparam("eta")
    with pytest.raises(V.VlsError, match="no derivative"):
        V.HostSetup({"init.type": "Hernquist", "grid.Pres": 16, "grid.Ures": [16, 8]}).dparam("M")


def test_initial_state_autograd_function():
    """vlasolve_amd.autograd.InitialState: its backward pass (the set-up's
    derivative) against central differences of a linear functional of the
    initial state, for two parameters at once."""
    torch = pytest.importorskip("torch")
    from vlasolve_amd.autograd import InitialState

    cfg, names = PLDF, ["eta", "alpha"]
    theta0 = np.array([0.1, 1.0])
    base = V.HostSetup(cfg)
    rng = np.random.default_rng(3)
    w = rng.standard_normal(base.shape)
    wv = rng.standard_normal(base.n_cells)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    f0, v0 = InitialState().apply(theta, cfg, names)
    L = (f0 * torch.from_numpy(w)).sum() + (v0 * torch.from_numpy(wv)).sum()
    L.backward()

    def lin(th):
        h = V.HostSetup(dict(cfg, **{"init." + k: float(x) for k, x in zip(names, th)}))
        return float((h.f() * w).sum() + (h.chains()["val"] * wv).sum())

    for k in range(2):
        e = np.zeros(2)
        e[k] = 1e-6 * theta0[k]
        fd = (lin(theta0 + e) - lin(theta0 - e)) / (2 * e[k])
        assert abs(float(theta.grad[k]) - fd) <= 1e-6 * abs(fd), (names[k], float(theta.grad[k]), fd)
    with pytest.raises(TypeError, match="float64"):
        InitialState().apply(torch.tensor([0.1], requires_grad=True), cfg, ["eta"])
