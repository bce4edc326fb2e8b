"""ctypes access to the CPU oracle (liboracle.so) and the reference executable.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.

A run configuration is a dict of reference command-line keys ("init.type",
"grid.Pres", "solver.dt", ...) to a value or a list of values, exactly as the
reference's paramsParser reads them (paramsParser.hxx:176-193).  ``resolve``
applies the reference's defaults (vlasovSolver.hxx:61-91, the generators'
defaultGridParams/defaultSolverParams/setupParameters) and returns the oracle's
parameter struct.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_EXE = HERE / "_ref" / "vlasolve"

SCALES = {"linear": 0, "logarithmic": 1, "quadratic": 2, "cubic": 3, "quartic": 4}
INITS = {"UDF": 0, "PLDF": 1, "Hernquist": 2, "Plummer": 3}
KERNELS = {"delayed": 0, "reflective": 1}
MAX_SPECIES = 5  # initGen::MAX_SPECIES, initGen.hxx:21


class Params(C.Structure):
    _fields_ = [
        ("Pmin", C.c_double), ("Pmax", C.c_double), ("Pres", C.c_int), ("Pscale", C.c_int),
        ("Umin", C.c_double * 2), ("Umax", C.c_double * 2), ("Ures", C.c_int * 2),
        ("Uscale", C.c_int * 2),
        ("G", C.c_double), ("dt", C.c_double), ("noLinking", C.c_int), ("kernelType", C.c_int),
        ("kernelBufferSize", C.c_int),
        ("initType", C.c_int),
        ("M", C.c_double), ("R0", C.c_double), ("alpha", C.c_double),
        ("smoothingRadius", C.c_double), ("gap0", C.c_double), ("gap1", C.c_double),
        ("eta", C.c_double), ("n", C.c_double), ("Ra", C.c_double),
        ("nSpecies", C.c_int),
        ("spM", C.c_double * MAX_SPECIES), ("spR0", C.c_double * MAX_SPECIES),
        ("spAlpha", C.c_double * MAX_SPECIES), ("spSR", C.c_double * MAX_SPECIES),
        ("spGap0", C.c_double * MAX_SPECIES), ("spGap1", C.c_double * MAX_SPECIES),
        ("spEta", C.c_double * MAX_SPECIES), ("spN", C.c_double * MAX_SPECIES),
        ("nthreads", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        d = C.c_double
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [P(Params)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_dims.argtypes = [C.c_void_p, P(C.c_long)]
        L.orc_nval.restype = C.c_long
        L.orc_nval.argtypes = [C.c_void_p]
        L.orc_nspecies.restype = C.c_int
        L.orc_nspecies.argtypes = [C.c_void_p]
        for name in ("orc_get_f", "orc_set_f"):
            getattr(L, name).argtypes = [C.c_void_p, P(d)]
        L.orc_get_axes.argtypes = [C.c_void_p, P(d), P(d), P(d), P(d)]
        L.orc_kernel_radius.restype = d
        L.orc_kernel_radius.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_last_moments.argtypes = [C.c_void_p, P(d), P(d), P(d)]
        L.orc_moments.argtypes = [C.c_void_p, P(d), P(d), P(d)]
        L.orc_stats.argtypes = [C.c_void_p, P(d)]
        L.orc_stats_total.argtypes = [C.c_void_p, P(d)]
        L.orc_get_feet.argtypes = [C.c_void_p, P(d), P(d), P(C.c_uint8)]
        L.orc_num_chains.restype = C.c_long
        L.orc_num_chains.argtypes = [C.c_void_p]
        L.orc_chain_cells.restype = C.c_long
        L.orc_chain_cells.argtypes = [C.c_void_p]
        L.orc_get_chains.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64), P(C.c_int32), P(d),
                                     P(d), P(d)]
        L.orc_spline_build.argtypes = [C.c_void_p, P(d), P(d)]
        L.orc_spline_evalC.restype = d
        L.orc_spline_evalC.argtypes = [C.c_void_p, P(d), C.c_int, d, C.c_int, d]
        L.orc_spline_evalNxPy.restype = d
        L.orc_spline_evalNxPy.argtypes = [C.c_void_p, P(d), C.c_int, d]
        L.orc_spline_eval.restype = d
        L.orc_spline_eval.argtypes = [C.c_void_p, P(d), d, d]
        L.orc_advance_drift.argtypes = [d, d, d, d, d, P(d), P(d)]
        L.orc_gen_scale.restype = C.c_long
        L.orc_gen_scale.argtypes = [d, d, C.c_long, C.c_int, C.c_int, P(d)]
        L.orc_set_chain_vals.argtypes = [C.c_void_p, P(d)]
        L.orc_tl_init.restype = C.c_int
        L.orc_tl_init.argtypes = [C.c_void_p, P(d), P(d)]
        L.orc_tl_step.restype = C.c_int
        L.orc_tl_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_tl_get.argtypes = [C.c_void_p, P(d)]
        L.orc_tl_mass.argtypes = [C.c_void_p, P(d), P(d)]
        _lib = L
    return _lib


def _ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


# ---------------------------------------------------------------------------
# configuration handling

def _vals(cfg, key):
    v = cfg.get(key)
    if v is None:
        return None
    return list(v) if isinstance(v, (list, tuple)) else [v]


def _get(cfg, key, default, which=0, conv=float):
    v = _vals(cfg, key)
    if v is None or len(v) <= which:
        return default
    return conv(v[which])


def _bool(x):
    return bool(int(x)) if str(x).lstrip("-").isdigit() else str(x).lower() == "true"


def resolve(cfg: dict, nthreads: int = 1) -> Params:
    """Resolve a reference-style configuration into oracle parameters."""
    p = Params()
    itype = _get(cfg, "init.type", "UDF", conv=str)
    if itype not in INITS:
        raise ValueError(f"unknown initial conditions type {itype}")
    p.initType = INITS[itype]
    # defaultSolverParams (UDF :104-112, PLDF :127-135, Hernquist :77-85, Plummer :72-80)
    p.G = _get(cfg, "solver.G", 1.0)
    p.dt = _get(cfg, "solver.dt", {"Hernquist": 0.025, "Plummer": 0.1}.get(itype, 0.005))
    p.noLinking = int(_get(cfg, "solver.noLinking", False, conv=_bool))
    p.kernelType = KERNELS[_get(cfg, "solver.kernelType", "delayed", conv=str)]
    p.kernelBufferSize = _get(cfg, "solver.kernelBufferSize", 0, conv=int)
    p.nSpecies = _get(cfg, "solver.nSpecies", 1, conv=int)  # vlasovSolver.hxx:72
    p.nthreads = nthreads
    # generator parameters
    if itype == "UDF":  # initGen_UniformDensitySphere.hxx:49-62, value i for species i
        for i in range(min(p.nSpecies, MAX_SPECIES)):
            p.spR0[i] = _get(cfg, "init.R0", 2.0, i)
            p.spM[i] = _get(cfg, "init.M", 1.0, i)
            p.spAlpha[i] = _get(cfg, "init.alpha", 0.5, i)
            p.spSR[i] = _get(cfg, "init.smoothingRadius", 0.5, i)
            p.spGap0[i] = _get(cfg, "init.gap", 0.0, 2 * i)
            p.spGap1[i] = _get(cfg, "init.gap", 0.0, 2 * i + 1)
        p.R0, p.M, p.alpha = p.spR0[0], p.spM[0], p.spAlpha[0]
        p.smoothingRadius, p.gap0, p.gap1 = p.spSR[0], p.spGap0[0], p.spGap1[0]
        grid = dict(Pmin=0.01, Pmax=25.0, Umin=[-2.0, 0.0], Umax=[2.0, 1.6], Pres=200,
                    Ures=[200, 200], Pscale=1, Uscale=[0, 2])
    elif itype == "PLDF":  # initGen_PowerLawDensitySphere.hxx:50-80, value i for species i
        sigma2 = 0.0
        for i in range(min(p.nSpecies, MAX_SPECIES)):
            p.spM[i] = _get(cfg, "init.M", 1.0, i)
            p.spAlpha[i] = _get(cfg, "init.alpha", 2.0, i)
            p.spEta[i] = _get(cfg, "init.eta", 0.1, i)
            p.spN[i] = _get(cfg, "init.n", 1.0, i)
            p.spSR[i] = _get(cfg, "init.smoothingRadius", 0.5, i)
            p.spR0[i] = _get(cfg, "init.R0", 1.0 / p.spEta[i], i)
            n = p.spN[i]
            sigO2 = ((3.0 - n) / (5.0 - 2 * n) * (p.spM[i] * p.spEta[i]) /
                     (p.spR0[i] * (p.spAlpha[i] * 0.5 + 1)))
            sigR2 = sigO2 * 0.5 * p.spAlpha[i]
            sigma2 = max(sigma2, sigO2 + 2 * sigR2)  # defaultGridParams :95-100
        p.M, p.alpha, p.eta, p.n = p.spM[0], p.spAlpha[0], p.spEta[0], p.spN[0]
        p.smoothingRadius, p.R0 = p.spSR[0], p.spR0[0]
        grid = dict(Pmin=0.01, Pmax=200.0, Umin=[-16 * math.sqrt(sigma2), 0.0],
                    Umax=[16 * math.sqrt(sigma2), 1.6], Pres=200, Ures=[200, 200], Pscale=1,
                    Uscale=[0, 2])
    elif itype == "Plummer":  # initGen_Plummer.hxx:37-70
        p.R0 = _get(cfg, "init.R0", 1.0)
        p.M = _get(cfg, "init.M", 1.0)
        grid = dict(Pmin=0.01, Pmax=100.0, Umin=[-1.45, 0.0], Umax=[1.45, 3.24], Pres=180,
                    Ures=[80, 80], Pscale=1, Uscale=[0, 2])
    else:  # initGen_Hernquist.hxx:38-75
        p.R0 = _get(cfg, "init.R0", 1.0)
        p.M = _get(cfg, "init.M", 1.0)
        p.Ra = _get(cfg, "init.Ra", -1.0)
        grid = dict(Pmin=0.01, Pmax=250.0, Umin=[-1.45, 1.0e-4], Umax=[1.45, 3.24 * 4],
                    Pres=90, Ures=[80, 80], Pscale=1, Uscale=[0, 1])
    # gridHandler overrides (grid_handler.hxx:221-234)
    p.Pmin = _get(cfg, "grid.Pmin", grid["Pmin"])
    p.Pmax = _get(cfg, "grid.Pmax", grid["Pmax"])
    p.Pres = _get(cfg, "grid.Pres", grid["Pres"], conv=int)
    p.Pscale = SCALES[_get(cfg, "grid.Pscale", None, conv=str)] if _vals(cfg, "grid.Pscale") \
        else grid["Pscale"]
    for i in range(2):
        p.Umin[i] = _get(cfg, "grid.Umin", grid["Umin"][i], i)
        p.Umax[i] = _get(cfg, "grid.Umax", grid["Umax"][i], i)
        p.Ures[i] = _get(cfg, "grid.Ures", grid["Ures"][i], i, conv=int)
        v = _vals(cfg, "grid.Uscale")
        p.Uscale[i] = SCALES[str(v[i])] if v and len(v) > i else grid["Uscale"][i]
    return p


def to_argv(cfg: dict) -> list[str]:
    """Reference-style command line for a configuration."""
    out = []
    for k, v in cfg.items():
        out.append("-" + k)
        vals = list(v) if isinstance(v, (list, tuple)) else [v]
        out += [repr(x) if isinstance(x, float) else str(x) for x in vals]
    return out


# ---------------------------------------------------------------------------
class Oracle:
    """One oracle run (initial conditions, feet, delay lines, renormalized f)."""

    def __init__(self, cfg: dict, nthreads: int = 1):
        self.params = resolve(cfg, nthreads)
        L = lib()
        self._h = L.orc_create(C.byref(self.params))
        if not self._h:
            raise RuntimeError("oracle: " + L.orc_last_error().decode())
        dims = (C.c_long * 3)()
        L.orc_dims(self._h, dims)
        self.nR, self.nU, self.nJ = dims[0], dims[1], dims[2]
        self.nS = L.orc_nspecies(self._h)
        # the reference's layout: species fastest when there are several
        self.shape = (self.nJ, self.nU, self.nR) + ((self.nS,) if self.nS > 1 else ())

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        lib().orc_step(self._h, int(n))

    def f(self):
        out = np.empty(self.shape, np.float64)
        lib().orc_get_f(self._h, _ptr(out))
        return out

    def set_f(self, f):
        a = np.ascontiguousarray(f, np.float64)
        assert a.size == self.nR * self.nU * self.nJ * self.nS
        lib().orc_set_f(self._h, _ptr(a))

    def axes(self):
        R = np.empty(self.nR)
        U = np.empty(self.nU)
        J = np.empty(self.nJ)
        JV = np.empty(self.nJ + 1)
        lib().orc_get_axes(self._h, _ptr(R), _ptr(U), _ptr(J), _ptr(JV))
        return R, U, J, JV

    def kernel_radius(self):
        return lib().orc_kernel_radius(self._h)

    def moments(self):
        dM = np.empty(self.nR)
        M = np.empty(self.nR)
        t = C.c_double()
        lib().orc_moments(self._h, _ptr(dM), _ptr(M), C.byref(t))
        return dM, M, t.value

    def last_moments(self):
        dM = np.empty(self.nR)
        a = np.empty(self.nR)
        t = C.c_double()
        lib().orc_last_moments(self._h, _ptr(dM), _ptr(a), C.byref(t))
        return dM, a, t.value

    def stats(self):
        """the statistics line of the species files (statistics.txt with one species)"""
        out = np.empty(7)
        lib().orc_stats(self._h, _ptr(out))
        return dict(zip(("M", "E", "Ek", "Ep", "S", "L1", "L2"), out))

    def stats_total(self):
        """the statistics.txt line"""
        out = np.empty(7)
        lib().orc_stats_total(self._h, _ptr(out))
        return dict(zip(("M", "E", "Ek", "Ep", "S", "L1", "L2"), out))

    def feet(self):
        n = self.nR * self.nU * self.nJ
        R = np.empty(n)
        U = np.empty(n)
        fl = np.empty(n, np.uint8)
        lib().orc_get_feet(self._h, _ptr(R), _ptr(U), _ptr(fl, C.c_uint8))
        pts = (self.nJ, self.nU, self.nR)  # one foot per point, whatever the species
        return R.reshape(pts), U.reshape(pts), fl.reshape(pts)

    def chains(self):
        L = lib()
        nc = L.orc_num_chains(self._h)
        ncell = L.orc_chain_cells(self._h)
        point = np.empty(nc, np.int64)
        off = np.empty(nc, np.int64)
        ln = np.empty(nc, np.int32)
        R = np.empty(ncell)
        U = np.empty(ncell)
        V = np.empty(ncell * self.nS)
        L.orc_get_chains(self._h, _ptr(point, C.c_int64), _ptr(off, C.c_int64),
                         _ptr(ln, C.c_int32), _ptr(R), _ptr(U), _ptr(V))
        if self.nS > 1:
            V = V.reshape(ncell, self.nS)
        return dict(point=point, off=off, len=ln, R=R, U=U, val=V)

    # ---- tangent-linear model (checks the adjoint of the HIP path)
    def set_chain_vals(self, vals):
        v = np.ascontiguousarray(vals, np.float64)
        lib().orc_set_chain_vals(self._h, _ptr(v))

    def tl_init(self, df, dvals=None):
        d = np.ascontiguousarray(df, np.float64).reshape(-1)
        assert d.size == self.nR * self.nU * self.nJ
        dv = None if dvals is None else np.ascontiguousarray(dvals, np.float64)
        if lib().orc_tl_init(self._h, _ptr(d), _ptr(dv) if dv is not None else None) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def tl_step(self, n=1):
        if lib().orc_tl_step(self._h, int(n)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def tl_get(self):
        out = np.empty(self.shape)
        lib().orc_tl_get(self._h, _ptr(out))
        return out

    def tl_mass(self):
        """exclusive-scan mass profile M(<R_r) of f and its tangent"""
        M = np.empty(self.nR)
        dM = np.empty(self.nR)
        lib().orc_tl_mass(self._h, _ptr(M), _ptr(dM))
        return M, dM

    def spline_build(self, slice2d):
        a = np.ascontiguousarray(slice2d, np.float64)
        assert a.shape == (self.nU, self.nR)
        coef = np.zeros((self.nU + 2) * (self.nR + 2))
        lib().orc_spline_build(self._h, _ptr(a), _ptr(coef))
        return coef.reshape(self.nU + 2, self.nR + 2)

    def spline_evalC(self, coef, Nx, x, Ny, y):
        c = np.ascontiguousarray(coef, np.float64)
        return lib().orc_spline_evalC(self._h, _ptr(c), int(Nx), float(x), int(Ny), float(y))

    def spline_evalNxPy(self, coef, Nx, yy):
        c = np.ascontiguousarray(coef, np.float64)
        return lib().orc_spline_evalNxPy(self._h, _ptr(c), int(Nx), float(yy))

    def spline_eval(self, coef, xx, yy):
        c = np.ascontiguousarray(coef, np.float64)
        return lib().orc_spline_eval(self._h, _ptr(c), float(xx), float(yy))


def advance_drift(R, U, J, dt, R0=-1.0):
    a, b = C.c_double(), C.c_double()
    lib().orc_advance_drift(R, U, J, dt, R0, C.byref(a), C.byref(b))
    return a.value, b.value


def gen_scale(start, stop, N, sct, cell=False):
    out = np.empty(N + 2)
    n = lib().orc_gen_scale(start, stop, N, sct, int(bool(cell)), _ptr(out))
    return out[:n].copy()


# ---------------------------------------------------------------------------
def run_reference(cfg: dict, workdir, threads: int = 1, timeout: float = 600.0):
    """Run the reference executable (oracle/_ref/vlasolve) in ``workdir``.

    Returns its stdout.  Raises if the executable was not built."""
    if not REF_EXE.exists():
        raise FileNotFoundError(str(REF_EXE))
    workdir = Path(workdir)
    workdir.mkdir(parents=True, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    argv = [str(REF_EXE)] + to_argv(cfg) + ["-threads_per_node", str(threads)]
    res = subprocess.run(argv, cwd=workdir, env=env, capture_output=True, text=True,
                         timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError(f"reference failed ({res.returncode}): {res.stderr[-2000:]}")
    return res.stdout


# ---------------------------------------------------------------------------
def density_potential(R, dM, Mx=None, G=1.0):
    """Density and gravitational potential of the radial shells, from the
    moment dM[r] of integrate1D (quadrature21.hxx:88-142) and the exclusive
    enclosed mass Mx[r] = M(<R_r) (gridHandler::accumulate with addSpare=1,
    grid_handler.hxx:350-407; computed from dM when not given).

    The reference has no rho or potential output of its own: its yorick
    density() integrates snapshots over (u, j) (yorick/vlasolve.i:39-64) and
    potentialEnergy pairs each shell with the mass enclosed below it
    (vlasovSolver21.hxx:627-642).  Defined here from the step's own moment:

      rho[k] = dM[k] / ((4 pi / 3) (R[k+1]^3 - R[k]^3)),  rho[nR-1] = 0
      phi[i] = -G (Mx[i] / R[i] + S[i]),  S[i] = sum_{i <= k < nR-1} dM[k] / rbar[k]

    with rbar[k] = (R[k] + R[k+1]) / 2 the shell's midpoint and S summed from
    the outermost shell inwards.  Evaluated with Python floats (IEEE double,
    one rounding per operation) in exactly this order, so that the device's
    kernel, which runs the same operations, agrees bit for bit."""
    R = [float(x) for x in R]
    dM = [float(x) for x in dM]
    n = len(R)
    if Mx is None:
        Mx, m = [], 0.0
        for k in range(n):
            Mx.append(m)
            m = m + dM[k]
    Mx = [float(x) for x in Mx]
    vol = (4.0 * math.pi) / 3.0
    rho = [0.0] * n
    for k in range(n - 1):
        rho[k] = dM[k] / (vol * (R[k + 1] * R[k + 1] * R[k + 1] - R[k] * R[k] * R[k]))
    phi = [0.0] * n
    S = 0.0
    for i in range(n - 1, -1, -1):
        if i < n - 1:
            S = dM[i] / (0.5 * (R[i] + R[i + 1])) + S
        phi[i] = -(G * (Mx[i] / R[i] + S))
    return np.array(rho), np.array(phi)
