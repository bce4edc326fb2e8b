"""The benchmark configurations on the GPU, at their full sizes.

* PLDF 129 x 129 x 64 (the adjoint workload) and Hernquist 513 x 513 x 256
  (the large workload): the HIP step against the oracle in the device order,
  bit for bit, which exercises the register solvers of 129-node lines and
  the streaming solvers of 515-node lines at their real shapes.
* The seeded 32-run uneven UDF ensemble: every member, stepped inside the
  ensemble's graph, against its own oracle run, bit for bit.
* UDF 201 x 201 x 200 for 1000 steps against the reference executable at
  one OpenMP thread (fixture tests/golden/udf_c2_1000.npz, written by
  make_golden_long.py): statistics every 25 steps to the reference's printed
  precision, and f, the mass profile, density and potential at the end to
  1e-11.  The reference's result depends on its thread count (it sums the
  moment in per-thread chunks): at 8 threads its final f differs from its
  own one-thread f by about 1.5e-11 of max |f| in the slices near j = 0;
  the fixture records that floor and the test prints it beside the GPU's
  difference.
* Density and potential (vls_density_potential) bit for bit against the
  oracle's restatement applied to the oracle's moment.
"""
from pathlib import Path

import numpy as np
import pytest

import bench
from tests.configs import UDF_C1, UDF_C2

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")
from oracle import oracle as O  # noqa: E402

GOLD = Path(__file__).resolve().parent / "golden"


def _same(a, b):
    return a.shape == b.shape and bool(np.all(a == b))


def _compare(gpu, orc, label):
    fg, fo = gpu.f(), orc.f()
    bad = int(np.count_nonzero(fg != fo))
    assert bad == 0, f"{label}: {bad} points differ, max |d| {np.abs(fg - fo).max():.3e}"
    dMg, ag, mg = gpu.moments()
    dMo, ao, mo = orc.last_moments()
    assert _same(dMg, dMo) and _same(ag, ao) and mg == mo, f"{label}: moments differ"


def test_pldf_129x129x64_bit_exact():
    cfg = bench.ADJ_WORKLOAD
    print("case pldf 129x129x64, 20 steps", flush=True)
    gpu = V.Solver(cfg)
    orc = O.Oracle(cfg, nthreads=0)
    assert (gpu.nR, gpu.nU, gpu.nJ) == (129, 129, 64)
    assert _same(gpu.f(), orc.f())
    done = 0
    for k in (1, 10, 20):
        gpu.step(k - done)
        orc.step(k - done)
        done = k
        _compare(gpu, orc, f"step {k}")
        print(f"  step {k} bit-exact", flush=True)
    sg, so = gpu.stats(), orc.stats()
    for key in ("M", "E", "Ek", "Ep", "L1", "L2"):
        assert sg[key] == so[key], key


def test_hernquist_513x513x256_bit_exact():
    cfg = bench.HQ_WORKLOAD
    print("case hernquist 513x513x256, 2 steps", flush=True)
    gpu = V.Solver(cfg)
    print("  gpu set up", flush=True)
    orc = O.Oracle(cfg, nthreads=0)
    print("  oracle set up", flush=True)
    assert (gpu.nR, gpu.nU, gpu.nJ) == (513, 513, 256)
    assert _same(gpu.f(), orc.f())
    for k in (1, 2):
        gpu.step(1)
        orc.step(1)
        _compare(gpu, orc, f"step {k}")
        print(f"  step {k} bit-exact", flush=True)


def test_ensemble_members_bit_exact():
    cfgs = bench.ensemble_cfgs()
    n = 10
    print(f"case {len(cfgs)}-run ensemble, {n} graph steps", flush=True)
    e = V.Ensemble(cfgs)
    e.set_graph(True)
    e.step(n)
    for i, cfg in enumerate(cfgs):
        orc = O.Oracle(cfg, nthreads=0)
        orc.step(n)
        fe, fo = e.f(i), orc.f()
        bad = int(np.count_nonzero(fe != fo))
        assert bad == 0, f"run {i} ({fe.shape}): {bad} points differ"
    print("  every member bit-exact", flush=True)


def test_udf_201_1000_steps_against_reference():
    z = np.load(GOLD / "udf_c2_1000.npz")
    nsteps = int(z["steps"])
    print(f"case udf 201x201x200, {nsteps} steps vs the reference", flush=True)
    gpu = V.Solver(UDF_C2)
    ctx = gpu.context
    # the reference writes statistics before steps 0, 25, ..., 975 (none
    # after the last step)
    rows = {int(round(r[1])): r for r in z["stats"]}
    assert sorted(rows) == list(range(0, nsteps, 25))
    worst = 0.0
    for k in range(0, nsteps + 1, 25):
        if k:
            gpu.step(25)
        if k not in rows:
            continue
        s = gpu.stats_total()
        ref = rows[k]
        for i, key in enumerate(("M", "E", "Ek", "Ep", "S", "L1", "L2")):
            # statistics.txt holds 10 significant digits
            assert s[key] == pytest.approx(ref[2 + i], rel=2e-9, abs=1e-300), (k, key)
    f = gpu.f()
    fmax = float(z["fmax"])
    for j, key in ((0, "slice0"), (199, "slice199")):
        d = np.abs(f[j] - z[key]).max() / fmax
        worst = max(worst, d)
    d = np.abs(f.reshape(-1)[z["sample_idx"]] - z["sample_f"]).max() / fmax
    worst = max(worst, d)
    floor = max(float(z[k]) for k in ("floor_slice0", "floor_slice199", "floor_samples"))
    print(f"  f: max |gpu - reference| / max|f| = {worst:.3e} (the reference at "
          f"{int(z['floor_threads'])} threads against itself at 1: {floor:.3e})", flush=True)
    assert worst <= 1e-11
    M = ctx.mass_profile()
    rho, phi = ctx.density_potential()
    for name, a, b in (("M", M, z["M"]), ("rho", rho, z["rho"]), ("phi", phi, z["phi"])):
        rel = np.abs(a - b).max() / np.abs(b).max()
        print(f"  {name}: max rel diff {rel:.3e}", flush=True)
        assert rel <= 1e-11, name


def test_density_potential_bit_exact():
    print("case udf 65x65x32 density and potential", flush=True)
    gpu = V.Solver(UDF_C1)
    orc = O.Oracle(UDF_C1, nthreads=0)
    gpu.step(20)
    orc.step(20)
    ctx = gpu.context
    rho, phi = ctx.density_potential()
    dM, Mx, _ = orc.moments()
    R = orc.axes()[0]
    rho_o, phi_o = O.density_potential(R, dM, Mx, G=1.0)
    assert _same(rho, rho_o) and _same(phi, phi_o)
    assert _same(ctx.mass_profile(), Mx)
