"""vlasolve_amd: MI355X (gfx950) spherical Vlasov-Poisson spline step.

The time step runs in hand-written HIP kernels behind the C ABI declared in
include/vlasolve_hip.h; this module binds that ABI with ctypes.  There is no
CPU fallback: if the library is not built, or no GPU is present, every call
raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libvlasolve_hip.so"
EXE_PATH = PKG / "bin" / "vlasolve_hip"

VLS_FOOT_EVAL, VLS_FOOT_OUT, VLS_FOOT_CHAIN = 0, 1, 2
STAT_NAMES = ("M", "E", "Ek", "Ep", "S", "L1", "L2")

# every exported symbol of include/vlasolve_hip.h
EXPORTS = (
    "vls_create", "vls_destroy", "vls_last_error", "vls_upload_f", "vls_download_f",
    "vls_set_drift", "vls_set_chains", "vls_get_chains", "vls_step", "vls_synchronize",
    "vls_moments", "vls_stats", "vls_download_coef", "vls_set_timing", "vls_kernel_times",
    "vlsd_create", "vlsd_destroy", "vlsd_info", "vlsd_step", "vlsd_get_f", "vlsd_stats",
    "vlsd_moments", "vlsd_run", "vlsd_write_restart", "vlsd_snapshot", "vlsd_context",
    "vlsd_last_error", "vls_thomas_table", "vls_check_division", "vls_device_info",
    "vls_copy_probe", "vlsd_host_create", "vlsd_host_destroy", "vlsd_host_info", "vlsd_host_f",
    "vlsd_host_feet", "vlsd_host_chains", "vls_run_info", "vls_put_chains", "vls_mass_profile",
    "vls_record", "vls_adjoint", "vls_set_graph", "vlsd_ensemble_create", "vlsd_ensemble_destroy",
    "vlsd_ensemble_context", "vls_stats_total", "vlsd_species", "vlsd_stats_total",
    "vlsd_host_species", "vls_memory", "vls_density_potential", "vls_density_potential_noalloc", "vls_set_timing_kernel", "vlsd_host_dparam",
    "vls_record_batch", "vls_adjoint_batch", "vls_autotune",
)


class VlsError(RuntimeError):
    pass


class Geometry(C.Structure):
    _fields_ = [
        ("nR", C.c_int32), ("nU", C.c_int32), ("nJ", C.c_int32),
        ("R_Val", C.POINTER(C.c_double)), ("U_Val", C.POINTER(C.c_double)),
        ("J_Val", C.POINTER(C.c_double)), ("dR", C.POINTER(C.c_double)),
        ("dU", C.POINTER(C.c_double)), ("dJ", C.POINTER(C.c_double)),
        ("spline_start", C.c_double * 2), ("spline_hinv", C.c_double * 2),
    ]


class SolverParams(C.Structure):
    _fields_ = [("G", C.c_double), ("dt", C.c_double), ("no_linking", C.c_int32),
                ("kernel_type", C.c_int32), ("n_species", C.c_int32)]


_lib = None


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise VlsError(f"HIP library not built: {LIB_PATH} (run `python -m vlasolve_amd.build`)")
        L = C.CDLL(str(LIB_PATH))
        P, d, i32, i64, vp = C.POINTER, C.c_double, C.c_int32, C.c_int64, C.c_void_p
        L.vls_last_error.restype = C.c_char_p
        L.vlsd_last_error.restype = C.c_char_p
        L.vls_create.argtypes = [P(Geometry), P(SolverParams), C.c_int, P(vp)]
        L.vls_destroy.argtypes = [vp]
        L.vls_destroy.restype = None
        L.vls_upload_f.argtypes = [vp, C.c_int, P(d)]
        L.vls_download_f.argtypes = [vp, C.c_int, P(d)]
        L.vls_set_drift.argtypes = [vp, C.c_int, P(d), P(d), P(C.c_uint8)]
        L.vls_set_chains.argtypes = [vp, C.c_int, i64, P(i64), P(i32), P(i32), P(d), P(d),
                                     P(C.c_uint8), P(d)]
        L.vls_get_chains.argtypes = [vp, C.c_int, P(i32), P(d)]
        L.vls_step.argtypes = [vp, C.c_int]
        L.vls_synchronize.argtypes = [vp]
        L.vls_moments.argtypes = [vp, C.c_int, P(d), P(d), P(d)]
        L.vls_stats.argtypes = [vp, C.c_int, P(d)]
        L.vls_stats_total.argtypes = [vp, C.c_int, P(d), P(d)]
        L.vls_download_coef.argtypes = [vp, C.c_int, C.c_int, P(d)]
        L.vls_set_timing.argtypes = [vp, C.c_int]
        L.vls_set_timing_kernel.argtypes = [vp, C.c_char_p]
        L.vls_kernel_times.argtypes = [vp, P(C.c_char_p), P(d), P(C.c_int)]
        L.vlsd_create.argtypes = [C.c_int, P(C.c_char_p), P(vp)]
        L.vlsd_destroy.argtypes = [vp]
        L.vlsd_destroy.restype = None
        L.vlsd_info.argtypes = [vp, P(i64), P(d), P(i64)]
        L.vlsd_step.argtypes = [vp, C.c_int]
        L.vlsd_get_f.argtypes = [vp, P(d)]
        L.vlsd_stats.argtypes = [vp, P(d)]
        L.vlsd_stats_total.argtypes = [vp, P(d)]
        L.vlsd_species.argtypes = [vp]
        L.vlsd_host_species.argtypes = [vp]
        L.vlsd_moments.argtypes = [vp, P(d), P(d), P(d)]
        L.vlsd_run.argtypes = [vp]
        L.vlsd_write_restart.argtypes = [vp, C.c_char_p]
        L.vlsd_snapshot.argtypes = [vp]
        L.vlsd_context.argtypes = [vp]
        L.vlsd_context.restype = vp
        L.vls_run_info.argtypes = [vp, C.c_int, P(i64)]
        L.vls_put_chains.argtypes = [vp, C.c_int, P(i32), P(d)]
        L.vls_mass_profile.argtypes = [vp, C.c_int, P(d)]
        L.vls_record.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.vls_adjoint.argtypes = [vp, C.c_int, P(d), P(d), P(d), P(d)]
        L.vls_record_batch.argtypes = [vp, C.c_int, C.c_int]
        L.vls_autotune.argtypes = [vp, C.c_int, C.c_char_p, C.c_int]
        L.vls_adjoint_batch.argtypes = [vp, vp, vp, vp, vp]
        L.vls_set_graph.argtypes = [vp, C.c_int]
        L.vls_memory.argtypes = [vp, C.c_int, P(i64), P(i64)]
        L.vls_density_potential.argtypes = [vp, C.c_int, P(d), P(d)]
        L.vls_density_potential_noalloc.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
        L.vlsd_ensemble_create.argtypes = [C.c_int, P(C.c_int), P(P(C.c_char_p)), P(vp)]
        L.vlsd_ensemble_destroy.argtypes = [vp]
        L.vlsd_ensemble_destroy.restype = None
        L.vlsd_ensemble_context.argtypes = [vp]
        L.vlsd_ensemble_context.restype = vp
        _lib = L
    return _lib


def _ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def _check(rc, driver=False):
    if rc:
        L = lib()
        msg = (L.vlsd_last_error() if driver else L.vls_last_error()) or b""
        raise VlsError(f"status {rc}: {msg.decode(errors='replace')}")


def to_argv(cfg) -> list[str]:
    """Reference-style command line ("-cat.key v ...") from a dict or a list."""
    if isinstance(cfg, (list, tuple)):
        return [str(x) for x in cfg]
    out = []
    for k, v in cfg.items():
        out.append("-" + k)
        vals = list(v) if isinstance(v, (list, tuple)) else [v]
        out += [repr(x) if isinstance(x, float) else str(x) for x in vals]
    return out


def device_info() -> str:
    buf = C.create_string_buffer(512)
    _check(lib().vls_device_info(buf, 512))
    return buf.value.decode()


def copy_probe(nbytes: int) -> None:
    """One device-to-device copy of nbytes (counter calibration)."""
    lib().vls_copy_probe.argtypes = [C.c_size_t]
    _check(lib().vls_copy_probe(int(nbytes)))


class Solver:
    """The drop-in solver: reference parameters in, every step on the GPU."""

    def __init__(self, cfg, verbose: int = 0):
        argv = ["vlasolve_hip"] + to_argv(cfg)
        if not any(a == "-verbose" for a in argv):
            argv += ["-verbose", str(verbose)]
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        h = C.c_void_p()
        _check(lib().vlsd_create(len(argv), arr, C.byref(h)), driver=True)
        self._h = h
        dims = (C.c_int64 * 3)()
        _check(lib().vlsd_info(self._h, dims, None, None), driver=True)
        self.nR, self.nU, self.nJ = dims[0], dims[1], dims[2]
        self.n_species = lib().vlsd_species(self._h)
        # the reference's layout: species fastest when there are several
        self.shape = (self.nJ, self.nU, self.nR) + ((self.n_species,) if self.n_species > 1 else ())

    def close(self):
        if getattr(self, "_h", None):
            lib().vlsd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def time(self):
        t = C.c_double()
        _check(lib().vlsd_info(self._h, None, C.byref(t), None), driver=True)
        return t.value

    @property
    def step_count(self):
        n = C.c_int64()
        _check(lib().vlsd_info(self._h, None, None, C.byref(n)), driver=True)
        return n.value

    def step(self, n=1):
        _check(lib().vlsd_step(self._h, int(n)), driver=True)

    def f(self):
        out = np.empty(self.shape, np.float64)
        _check(lib().vlsd_get_f(self._h, _ptr(out)), driver=True)
        return out

    def stats(self):
        """the statistics line (of the species files with several species)"""
        out = np.empty(7)
        _check(lib().vlsd_stats(self._h, _ptr(out)), driver=True)
        return dict(zip(STAT_NAMES, out))

    def stats_total(self):
        """the statistics.txt line"""
        out = np.empty(7)
        _check(lib().vlsd_stats_total(self._h, _ptr(out)), driver=True)
        return dict(zip(STAT_NAMES, out))

    def moments(self):
        dM = np.empty(self.nR)
        a = np.empty(self.nR)
        t = C.c_double()
        _check(lib().vlsd_moments(self._h, _ptr(dM), _ptr(a), C.byref(t)), driver=True)
        return dM, a, t.value

    def run(self):
        _check(lib().vlsd_run(self._h), driver=True)

    def write_restart(self, fname_no_ext):
        _check(lib().vlsd_write_restart(self._h, str(fname_no_ext).encode()), driver=True)

    def snapshot(self):
        _check(lib().vlsd_snapshot(self._h), driver=True)

    @property
    def context(self):
        return Context(lib().vlsd_context(self._h), owner=self)


class Context:
    """Layer-1 device context (vls_*); created directly or borrowed from a Solver."""

    def __init__(self, handle=None, owner=None, geometries=None, params=None):
        self._owner = owner
        self._own = False
        self.n_runs = getattr(owner, "n", 1) if handle is not None else len(geometries)
        if handle is not None:
            self._h = C.c_void_p(handle)
            return
        n = len(geometries)
        G = (Geometry * n)(*geometries)
        S = (SolverParams * n)(*params)
        h = C.c_void_p()
        _check(lib().vls_create(G, S, n, C.byref(h)))
        self._h = h
        self._own = True

    def close(self):
        if self._own and self._h:
            lib().vls_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        _check(lib().vls_step(self._h, int(n)))

    def synchronize(self):
        _check(lib().vls_synchronize(self._h))

    def upload_f(self, run, f):
        a = np.ascontiguousarray(f, np.float64)
        _check(lib().vls_upload_f(self._h, run, _ptr(a)))

    def download_f(self, run, shape):
        out = np.empty(shape, np.float64)
        _check(lib().vls_download_f(self._h, run, _ptr(out)))
        return out

    def coef(self, run, j, nU, nR):
        out = np.empty((nU + 2, nR + 2))
        _check(lib().vls_download_coef(self._h, run, j, _ptr(out)))
        return out

    def stats(self, run=0):
        out = np.empty(7)
        _check(lib().vls_stats(self._h, run, _ptr(out)))
        return dict(zip(STAT_NAMES, out))

    def set_timing(self, on=True, kernel=None):
        """HIP-event timing of the next vls_step calls: every kernel, or only
        the launches of `kernel` (an event before and after each)"""
        _check(lib().vls_set_timing_kernel(self._h, kernel.encode() if kernel else None))
        _check(lib().vls_set_timing(self._h, int(bool(on))))

    def info(self, run=0):
        """{nR, nU, nJ, chains, cells, species} of a run"""
        out = (C.c_int64 * 6)()
        _check(lib().vls_run_info(self._h, run, out))
        return dict(zip(("nR", "nU", "nJ", "chains", "cells", "species"), list(out)))

    def get_chains(self, run=0):
        """(ring heads, ring values: species fastest) of the delay lines"""
        inf = self.info(run)
        head = np.zeros(inf["chains"], np.int32)
        vals = np.zeros(inf["cells"] * inf["species"])
        _check(lib().vls_get_chains(self._h, run, _ptr(head, C.c_int32), _ptr(vals)))
        return head, vals

    def put_chains(self, head=None, vals=None, run=0):
        h = None if head is None else np.ascontiguousarray(head, np.int32)
        v = None if vals is None else np.ascontiguousarray(vals, np.float64)
        _check(lib().vls_put_chains(self._h, run, _ptr(h, C.c_int32) if h is not None else None,
                                    _ptr(v) if v is not None else None))

    def memory(self, run=0):
        """(allocations, bytes) of device memory the run holds"""
        n, b = C.c_int64(), C.c_int64()
        _check(lib().vls_memory(self._h, run, C.byref(n), C.byref(b)))
        return n.value, b.value

    def clear_chains(self, run=0):
        """Detach every delay line (vls_set_chains with no chains)."""
        _check(lib().vls_set_chains(self._h, run, 0, None, None, None, None, None, None, None))

    def mass_profile(self, run=0):
        """M(<R_r) of the current f"""
        out = np.empty(self.info(run)["nR"])
        _check(lib().vls_mass_profile(self._h, run, _ptr(out)))
        return out

    def density_potential(self, run=0):
        """(rho[nR], phi[nR]) of the current f: shell densities and potential"""
        n = self.info(run)["nR"]
        rho, phi = np.empty(n), np.empty(n)
        _check(lib().vls_density_potential(self._h, run, _ptr(rho), _ptr(phi)))
        return rho, phi

    def density_potential_noalloc(self, f_dev, rho_dev=None, phi_dev=None, M_dev=None, run=0,
                                  stream=None):
        """vls_density_potential_noalloc on device addresses (ints, e.g. a
        torch tensor's data_ptr()): f_dev [nS][nJ][nU][nR] in, rho, phi, M
        [nR] out (None: not wanted).  Returns without waiting."""
        _check(lib().vls_density_potential_noalloc(self._h, run, f_dev, rho_dev, phi_dev, M_dev,
                                                   stream))

    def record(self, n_steps, ck_every, run=0):
        _check(lib().vls_record(self._h, run, int(n_steps), int(ck_every)))

    def adjoint(self, dL_dM=None, dL_df=None, run=0):
        """(grad_f0, grad_vals0) of a loss of the recorded run's final state"""
        inf = self.info(run)
        shape = (inf["nJ"], inf["nU"], inf["nR"])
        gm = None if dL_dM is None else np.ascontiguousarray(dL_dM, np.float64)
        gf = None if dL_df is None else np.ascontiguousarray(dL_df, np.float64).reshape(shape)
        g0 = np.empty(shape)
        gv = np.empty(inf["cells"])
        _check(lib().vls_adjoint(self._h, run, _ptr(gm) if gm is not None else None,
                                 _ptr(gf) if gf is not None else None, _ptr(g0), _ptr(gv)))
        return g0, gv

    def autotune(self, run=-1):
        """vls_autotune: time and keep the fastest line-solve and kick
        schedules of a run (-1: every run); returns their description"""
        buf = C.create_string_buffer(8192)
        _check(lib().vls_autotune(self._h, int(run), buf, len(buf)))
        return buf.value.decode()

    def record_batch(self, n_steps, ck_every):
        """vls_record_batch: every run recorded together"""
        _check(lib().vls_record_batch(self._h, int(n_steps), int(ck_every)))

    def adjoint_batch(self, dL_dM=None, dL_df=None):
        """[(grad_f0, grad_vals0) per run] of vls_adjoint_batch; dL_dM and
        dL_df are lists with one entry (or None) per run"""
        n = self.n_runs
        infos = [self.info(k) for k in range(n)]
        keep = []

        def arr(items, shape_of):
            if items is None:
                return None
            ptrs = (C.c_void_p * n)()
            for k in range(n):
                if items[k] is not None:
                    a = np.ascontiguousarray(items[k], np.float64).reshape(shape_of(k))
                    keep.append(a)
                    ptrs[k] = a.ctypes.data
            return ptrs

        gm = arr(dL_dM, lambda k: (infos[k]["nR"],))
        gf = arr(dL_df, lambda k: (infos[k]["nJ"], infos[k]["nU"], infos[k]["nR"]))
        outs = [(np.empty((i["nJ"], i["nU"], i["nR"])), np.empty(i["cells"])) for i in infos]
        g0 = (C.c_void_p * n)(*[o[0].ctypes.data for o in outs])
        gv = (C.c_void_p * n)(*[o[1].ctypes.data for o in outs])
        _check(lib().vls_adjoint_batch(self._h, gm, gf, g0, gv))
        return outs

    def kernel_times(self, cap=4096):
        names = (C.c_char_p * cap)()
        ms = np.zeros(cap)
        n = C.c_int(cap)
        _check(lib().vls_kernel_times(self._h, names, _ptr(ms), C.byref(n)))
        return [(names[k].decode(), float(ms[k])) for k in range(min(n.value, cap))]


class Ensemble:
    """Several independent runs (grids may differ) in one device context.

    ``Ensemble([cfg0, cfg1, ...]).context`` is a Context whose run k is cfg k;
    ``step`` advances all runs (concurrently, through a HIP graph)."""

    def __init__(self, cfgs, verbose: int = 0):
        argvs = []
        for cfg in cfgs:
            a = ["vlasolve_hip"] + to_argv(cfg) + ["-verbose", str(verbose)]
            argvs.append((C.c_char_p * len(a))(*[x.encode() for x in a]))
        n = len(argvs)
        argc = (C.c_int * n)(*[len(a) for a in argvs])
        argv = (C.POINTER(C.c_char_p) * n)(*[C.cast(a, C.POINTER(C.c_char_p)) for a in argvs])
        h = C.c_void_p()
        _check(lib().vlsd_ensemble_create(n, argc, argv, C.byref(h)), driver=True)
        self._h = h
        self.n = n

    @property
    def context(self):
        """the ensemble's device context (borrowed: it lives as long as the
        ensemble; made on access, so that no reference cycle delays the
        release of the runs' device memory)"""
        if not self._h:
            raise VlsError("ensemble closed")
        return Context(lib().vlsd_ensemble_context(self._h), owner=self)

    def step(self, n=1):
        self.context.step(n)

    def f(self, run):
        inf = self.context.info(run)
        return self.context.download_f(run, (inf["nJ"], inf["nU"], inf["nR"]))

    def set_graph(self, on=True):
        _check(lib().vls_set_graph(self.context._h, int(bool(on))))

    def autotune(self):
        """the fastest schedule of every run (vls_autotune); its description"""
        return self.context.autotune(-1)

    def close(self):
        if getattr(self, "_h", None):
            lib().vlsd_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostSetup:
    """Host-only set-up of a run (initial f, feet, delay lines), no device.

    Runs the same C++ set-up as Solver up to the upload; used to check the
    host logic on machines without a GPU."""

    def __init__(self, cfg, verbose: int = 0):
        L = lib()
        P, d, i64 = C.POINTER, C.c_double, C.c_int64
        L.vlsd_host_create.argtypes = [C.c_int, P(C.c_char_p), P(C.c_void_p)]
        L.vlsd_host_destroy.argtypes = [C.c_void_p]
        L.vlsd_host_destroy.restype = None
        L.vlsd_host_info.argtypes = [C.c_void_p, P(i64), P(i64), P(i64), P(d)]
        L.vlsd_host_f.argtypes = [C.c_void_p, P(d)]
        L.vlsd_host_feet.argtypes = [C.c_void_p, P(d), P(d), P(C.c_uint8)]
        L.vlsd_host_chains.argtypes = [C.c_void_p, P(i64), P(C.c_int32), P(d), P(d), P(d)]
        argv = ["vlasolve_hip"] + to_argv(cfg)
        if not any(a == "-verbose" for a in argv):
            argv += ["-verbose", str(verbose)]
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        h = C.c_void_p()
        _check(L.vlsd_host_create(len(argv), arr, C.byref(h)), driver=True)
        self._h = h
        dims = (i64 * 3)()
        nc, ncell, kr = i64(), i64(), d()
        _check(L.vlsd_host_info(h, dims, C.byref(nc), C.byref(ncell), C.byref(kr)), driver=True)
        self.nR, self.nU, self.nJ = dims[0], dims[1], dims[2]
        self.n_species = L.vlsd_host_species(h)
        self.points = (self.nJ, self.nU, self.nR)
        self.shape = self.points + ((self.n_species,) if self.n_species > 1 else ())
        self.n_chains, self.n_cells, self.kernel_radius = nc.value, ncell.value, kr.value

    def close(self):
        if getattr(self, "_h", None):
            lib().vlsd_host_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def f(self):
        out = np.empty(self.shape)
        _check(lib().vlsd_host_f(self._h, _ptr(out)), driver=True)
        return out

    def dparam(self, param, species=0):
        """(df0, dvals0): derivative of the initial f (layout of f()) and of
        the delay-line values (layout of chains()["val"]) in an
        initial-condition parameter, the grid held fixed"""
        L = lib()
        L.vlsd_host_dparam.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        df = np.empty(self.shape)
        dv = np.empty(self.n_cells * self.n_species)
        _check(L.vlsd_host_dparam(self._h, param.encode(), int(species), _ptr(df), _ptr(dv)),
               driver=True)
        return df, dv

    def feet(self):
        R = np.empty(self.points)
        U = np.empty(self.points)
        fl = np.empty(self.points, np.uint8)
        _check(lib().vlsd_host_feet(self._h, _ptr(R), _ptr(U), _ptr(fl, C.c_uint8)), driver=True)
        return R, U, fl

    def chains(self):
        point = np.empty(self.n_chains, np.int64)
        ln = np.empty(self.n_chains, np.int32)
        R = np.empty(self.n_cells)
        U = np.empty(self.n_cells)
        V = np.empty(self.n_cells * self.n_species)
        _check(lib().vlsd_host_chains(self._h, _ptr(point, C.c_int64), _ptr(ln, C.c_int32),
                                      _ptr(R), _ptr(U), _ptr(V)), driver=True)
        if self.n_species > 1:
            V = V.reshape(self.n_cells, self.n_species)
        return dict(point=point, len=ln, R=R, U=U, val=V)
