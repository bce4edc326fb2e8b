"""GPU step against the oracle, through the C ABI.

The device sums the moment and the statistics in a fixed order; the oracle
run in that same order (nthreads=0) must then agree with the GPU bit for bit
on f, dM and the acceleration after every step, because every per-point
operation is the reference's, in the reference's order, without contraction.
Against the reference's own order (nthreads=1, the order pinned to the
reference executable) the trajectories differ only by summation rounding:
the tolerance 1e-11 relative is the one the hot path is specified to.
"""
import numpy as np
import pytest

from tests.configs import SMALL_CASES

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")
from oracle import oracle as O  # noqa: E402


def same(a, b):
    """bitwise equality up to the sign of zero"""
    return a.shape == b.shape and bool(np.all((a == b)))


@pytest.mark.parametrize("name", list(SMALL_CASES))
def test_trajectory_bit_exact(name):
    cfg, nsteps = SMALL_CASES[name]
    print(f"case {name} steps {nsteps}", flush=True)
    gpu = V.Solver(cfg)
    orc = O.Oracle(cfg, nthreads=0)
    assert same(gpu.f(), orc.f()), "initial f differs"
    checkpoints = sorted({1, 2, 5, nsteps // 2, nsteps})
    done = 0
    for k in checkpoints:
        gpu.step(k - done)
        orc.step(k - done)
        done = k
        fg, fo = gpu.f(), orc.f()
        bad = np.count_nonzero(fg != fo)
        assert bad == 0, f"step {k}: {bad} points differ, max |d| {np.abs(fg - fo).max()}"
        dMg, ag, mg = gpu.moments()
        dMo, ao, mo = orc.last_moments()
        assert same(dMg, dMo) and same(ag, ao) and mg == mo, f"step {k}: moments differ"
    _stats_match(gpu.stats(), orc.stats())
    _stats_match(gpu.stats_total(), orc.stats_total())


def _stats_match(sg, so):
    for key in ("M", "E", "Ek", "Ep", "L1", "L2"):
        assert sg[key] == so[key], (key, sg[key], so[key])
    # the entropy uses the device log, which may differ from glibc by an ulp
    assert abs(sg["S"] - so["S"]) <= 1e-14 * abs(so["S"])


@pytest.mark.parametrize("name", ["udf_c1", "hernquist"])
def test_against_reference_order(name):
    cfg, nsteps = SMALL_CASES[name]
    gpu = V.Solver(cfg)
    ref = O.Oracle(cfg, nthreads=1)
    gpu.step(nsteps)
    ref.step(nsteps)
    fg, fr = gpu.f(), ref.f()
    rel = np.abs(fg - fr).max() / np.abs(fr).max()
    print(f"{name}: max rel diff vs reference order after {nsteps} steps: {rel:.3e}")
    assert rel <= 1e-11
    sg, sr = gpu.stats(), ref.stats()
    for key in sg:
        assert abs(sg[key] - sr[key]) <= 1e-11 * max(abs(sr[key]), 1e-300), key


def test_c2_shape_bit_exact():
    """The 201 x 201 x 200 grid of the throughput benchmark, a few steps."""
    from tests.configs import UDF_C2

    gpu = V.Solver(UDF_C2)
    orc = O.Oracle(UDF_C2, nthreads=0)
    for k in (1, 3):
        gpu.step(k)
        orc.step(k)
        fg, fo = gpu.f(), orc.f()
        bad = np.count_nonzero(fg != fo)
        assert bad == 0, f"{bad} points differ"
    dMg, ag, mg = gpu.moments()
    dMo, ao, mo = orc.last_moments()
    assert same(dMg, dMo) and same(ag, ao) and mg == mo


def _bit_exact_run(cfg, nsteps, label):
    gpu = V.Solver(cfg)
    orc = O.Oracle(cfg, nthreads=0)
    for k in (1, nsteps - 1):
        gpu.step(k)
        orc.step(k)
        fg, fo = gpu.f(), orc.f()
        bad = np.count_nonzero(fg != fo)
        assert bad == 0, f"{label}: {bad} points differ"
    dMg, ag, mg = gpu.moments()
    dMo, ao, mo = orc.last_moments()
    assert same(dMg, dMo) and same(ag, ao) and mg == mo


def test_fused_kick_small_bit_exact(monkeypatch):
    """The fused kick (k_kick_line), which small grids do not take by
    default, forced on the 65 x 65 x 32 grid."""
    from tests.configs import UDF_C1

    print("case udf_c1 with VLS_FUSED_KICK", flush=True)
    monkeypatch.setenv("VLS_FUSED_KICK", "1")
    _bit_exact_run(UDF_C1, 12, "fused kick")


def test_kick_scratch_path_bit_exact(monkeypatch):
    """The fused kick's global-scratch path (taken by waves whose kick targets
    spread wider than its LDS ring), forced for every wave."""
    from tests.configs import UDF_C1

    print("case udf_c1 with VLS_FUSED_KICK, VLS_KICK_SCRATCH", flush=True)
    monkeypatch.setenv("VLS_FUSED_KICK", "1")
    monkeypatch.setenv("VLS_KICK_SCRATCH", "1")
    _bit_exact_run(UDF_C1, 12, "scratch path")


@pytest.mark.parametrize("fused", [False, True])
def test_strong_kick_bit_exact(monkeypatch, fused):
    """Kicks of tens of v cells per step: the column kernels, and the fused
    kick's ring and scratch paths side by side."""
    from tests.configs import UDF_STRONG_KICK

    print(f"case udf_strong_kick fused={fused}", flush=True)
    if fused:
        monkeypatch.setenv("VLS_FUSED_KICK", "1")
    _bit_exact_run(UDF_STRONG_KICK, 10, "strong kick")


def test_unfused_kick_bit_exact(monkeypatch):
    """The column v-solve and evaluation kernels after a separate moment on
    the 201 x 201 x 200 grid, which takes the moment fused into the column
    splines by default."""
    from tests.configs import UDF_C2

    print("case udf_c2 with VLS_NO_KICK_MOMENT, VLS_NO_FUSED_KICK", flush=True)
    monkeypatch.setenv("VLS_NO_KICK_MOMENT", "1")
    monkeypatch.setenv("VLS_NO_FUSED_KICK", "1")
    _bit_exact_run(UDF_C2, 3, "unfused kick")


@pytest.mark.parametrize("env", [
    {"VLS_NO_KICK_MOMENT": "1"},                       # one-launch moment, fused kick
    {"VLS_NO_KICK_MOMENT": "1", "VLS_MOMENT_3K": "1"},  # three-launch moment, fused kick
])
def test_separate_moment_bit_exact(monkeypatch, env):
    """The moment before the kick, in one launch (partials, chunk sums and
    scan joined by completion counters) and in three, on the 201 x 201 x 200
    grid."""
    from tests.configs import UDF_C2

    print(f"case udf_c2 with {env}", flush=True)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    _bit_exact_run(UDF_C2, 3, "separate moment")


@pytest.mark.parametrize("cfg_name", ["UDF_C1", "UDF_C2", "UDF_STRONG_KICK"])
def test_kick_in_rsolve_bit_exact(monkeypatch, cfg_name):
    """The kick folded into the drift's r-solve (its evaluation from the
    column splines staged per chunk in LDS; kicks that spread wider than the
    window read the splines directly, as on the strong-kick grid)."""
    import tests.configs as C

    print(f"case {cfg_name.lower()} with VLS_KICK_RSOLVE", flush=True)
    monkeypatch.setenv("VLS_KICK_RSOLVE", "1")
    _bit_exact_run(getattr(C, cfg_name), 4, "kick in r-solve")


@pytest.mark.parametrize("name", ["udf_c1", "hernquist"])
def test_streaming_kick_moment_bit_exact(monkeypatch, name):
    """The moment fused into the streaming column splines of the kick
    (k_kick_vsolve_moment), forced on grids that have register lines."""
    cfg, _ = SMALL_CASES[name]
    print(f"case {name} with VLS_NO_REGISTER_LINES", flush=True)
    monkeypatch.setenv("VLS_NO_REGISTER_LINES", "1")
    _bit_exact_run(cfg, 6, "streaming kick and moment")


@pytest.mark.parametrize("pres,ures", [(96, [192, 8]), (192, [64, 6]), (128, [96, 8])])
def test_register_lines_uneven_bit_exact(pres, ures):
    """Register lines on grids whose axes differ (the r-direction kernels are
    compiled per nR, the v-direction and kick kernels per nU): 97 x 193,
    193 x 65 and 129 x 97 nodes, the kick, r- and v-solves all in VGPRs."""
    from tests.configs import UDF_C1

    cfg = dict(UDF_C1, **{"grid.Pres": pres, "grid.Ures": ures})
    print(f"case udf {pres + 1} x {ures[0] + 1} x {ures[1]}", flush=True)
    _bit_exact_run(cfg, 5, "uneven register lines")


@pytest.mark.parametrize("name", ["udf_c1", "pldf"])
def test_moment_counters_across_runs(name):
    """The one-launch moment's counters return to zero after every launch:
    two runs of one grid stepped in turn, then one of them alone, against
    the oracle."""
    cfg, nsteps = SMALL_CASES[name]
    print(f"case {name} two runs in turn", flush=True)
    a, b = V.Solver(cfg), V.Solver(cfg)
    orc = O.Oracle(cfg, nthreads=0)
    for _ in range(3):
        a.step(1)
        b.step(1)
        orc.step(1)
    a.step(2)
    b.step(2)
    orc.step(2)
    assert np.count_nonzero(a.f() != orc.f()) == 0
    assert np.count_nonzero(b.f() != orc.f()) == 0
    dMg, ag, mg = b.moments()
    dMo, ao, mo = orc.last_moments()
    assert same(dMg, dMo) and same(ag, ao) and mg == mo


@pytest.mark.parametrize("pres,ures", [(20, [18, 5]), (31, [33, 6])])
def test_streaming_solvers_bit_exact(monkeypatch, pres, ures):
    """The streaming line solvers (taken by lines too long for an LDS tile of
    64 lines, such as the 515 padded nodes of the 513 x 513 grids), forced on
    shapes without register solvers.  With nR = 32 the forward sweep ends on
    a chunk boundary and the last two padded columns form a chunk of their own."""
    from tests.configs import UDF_C1

    cfg = dict(UDF_C1, **{"grid.Pres": pres, "grid.Ures": ures})
    print(f"case udf {pres}x{ures} with VLS_SPLINE_STREAM", flush=True)
    monkeypatch.setenv("VLS_SPLINE_STREAM", "1")
    _bit_exact_run(cfg, 8, "streaming solvers")


def test_two_species_delayed_bit_exact():
    """Two species sharing the delay lines' rings (each cell holds both
    species, the rings turn once per step): f of both species, the moment of
    their sum and both statistics lines, bit for bit."""
    from tests.configs import UDF_2SPECIES_DELAYED

    print("case udf 2 species delayed", flush=True)
    gpu = V.Solver(UDF_2SPECIES_DELAYED)
    orc = O.Oracle(UDF_2SPECIES_DELAYED, nthreads=0)
    assert gpu.f().shape[-1] == 2 and same(gpu.f(), orc.f())
    for k in (1, 12):
        gpu.step(k)
        orc.step(k)
        fg, fo = gpu.f(), orc.f()
        assert np.count_nonzero(fg != fo) == 0
    dMg, ag, mg = gpu.moments()
    dMo, ao, mo = orc.last_moments()
    assert same(dMg, dMo) and same(ag, ao) and mg == mo
    _stats_match(gpu.stats(), orc.stats())
    _stats_match(gpu.stats_total(), orc.stats_total())
