"""The discrete adjoint at the benchmark configurations.

* PLDF 129 x 129 x 64 (the adjoint workload): the adjoint identity
  <w, J v> = <J^T w, v> against the oracle's tangent-linear model over two
  checkpoint segments of 50 steps, with the segments kept by the recorded run
  and with both recomputed in the reverse sweep; this runs the register
  adjoints of 131-node lines.
* Hernquist 513 x 513 x 256: the same identity over two steps (two
  segments), which runs the long-column adjoints (the column walk of the
  kick's transpose and the streaming line transposes) at 515 nodes.
* UDF 201 x 201 x 200 (the step benchmark's grid): the same identity over
  two steps, which runs the long-column adjoints at 203 nodes.
* PLDF 129 x 129 x 64 for the full 2000 steps with checkpoints every 50:
  the gradient of L = sum_r (M_final(r) - M_target(r))^2, M_target the final
  mass profile of a run from f0 perturbed by a seeded multiplicative 1e-3,
  against central finite differences with h = 1e-6, in a direction of f0
  (second order) and in the virial ratio eta (through the set-up's
  derivative; fourth order, the second-order difference being limited by its
  own h^2 truncation there), to 1e-6.
"""
import numpy as np
import pytest

import bench
from tests.test_gpu_adjoint import chain_layout, ring_to_front

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")
from oracle import oracle as O  # noqa: E402


def _identity(cfg, n, ck, seed):
    solver = V.Solver(cfg)
    ctx = solver.context
    f0 = solver.f()
    head0, slots0 = ctx.get_chains()
    off, ln = chain_layout(cfg)
    ctx.record(n, ck)
    M = ctx.mass_profile()
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(M.shape)
    g0, gv = ctx.adjoint(dL_dM=w)
    gv_front = ring_to_front(gv, head0, off, ln)
    solver.close()
    orc = O.Oracle(cfg, nthreads=0)
    assert np.array_equal(orc.f(), f0)
    df = rng.standard_normal(f0.shape) * np.abs(f0).max()
    dv = rng.standard_normal(gv.shape) * (np.abs(slots0).max() if slots0.size else 1.0)
    orc.tl_init(df, dv if dv.size else None)
    orc.tl_step(n)
    Mo, dMo = orc.tl_mass()
    assert np.allclose(Mo, M, rtol=1e-12, atol=1e-15)
    lhs = float(w @ dMo)
    rhs = float((g0 * df).sum() + (gv_front * dv).sum())
    print(f"  <w, J v> = {lhs:.15e}  <J^T w, v> = {rhs:.15e}  rel {abs(lhs - rhs) / abs(lhs):.2e}",
          flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("keep", [None, "0"])
def test_identity_pldf_129x129x64(monkeypatch, keep):
    print(f"case pldf 129x129x64 adjoint identity, 100 steps, ck 50, keep={keep}", flush=True)
    if keep is not None:
        monkeypatch.setenv("VLS_ADJ_KEEP_STEPS", keep)
    _identity(bench.ADJ_WORKLOAD, 100, 50, 3)


@pytest.mark.parametrize("keep", [None, "0"])
def test_identity_hernquist_513x513x256(monkeypatch, keep):
    print(f"case hernquist 513x513x256 adjoint identity, 2 steps, ck 1, keep={keep}", flush=True)
    if keep is not None:
        monkeypatch.setenv("VLS_ADJ_KEEP_STEPS", keep)
    _identity(bench.HQ_WORKLOAD, 2, 1, 4)


@pytest.mark.parametrize("keep", [None, "0"])
def test_identity_udf_201x201x200(monkeypatch, keep):
    print(f"case udf 201x201x200 adjoint identity, 2 steps, ck 1, keep={keep}", flush=True)
    if keep is not None:
        monkeypatch.setenv("VLS_ADJ_KEEP_STEPS", keep)
    _identity(bench.WORKLOAD, 2, 1, 5)


def _final_mass(ctx, f0, head0, slots0, n):
    ctx.upload_f(0, f0)
    ctx.put_chains(head0, slots0)
    ctx.step(n)
    return ctx.mass_profile()


def test_pldf_129x129x64_2000_steps_against_finite_differences():
    cfg = bench.ADJ_WORKLOAD
    n, ck, h, seed = bench.ADJ_STEPS, bench.ADJ_CK, 1e-6, 2024
    print(f"case pldf 129x129x64 finite differences, {n} steps, ck {ck}, h {h}", flush=True)
    solver = V.Solver(cfg)
    ctx = solver.context
    f0 = solver.f()
    head0, slots0 = ctx.get_chains()
    off, ln = chain_layout(cfg)
    rng = np.random.default_rng(seed)
    Mt = _final_mass(ctx, f0 * (1 + 1e-3 * rng.standard_normal(f0.shape)), head0, slots0, n)

    def loss(M):
        return float(((M - Mt) ** 2).sum())

    ctx.upload_f(0, f0)
    ctx.put_chains(head0, slots0)
    ctx.record(n, ck)
    M = ctx.mass_profile()
    g0, gv = ctx.adjoint(dL_dM=2 * (M - Mt))
    assert np.isfinite(g0).all() and np.isfinite(gv).all()

    # a direction of f0: a seeded relative perturbation of every point
    v = f0 * rng.standard_normal(f0.shape)
    Lp = loss(_final_mass(ctx, f0 + h * v, head0, slots0, n))
    Lm = loss(_final_mass(ctx, f0 - h * v, head0, slots0, n))
    fd = (Lp - Lm) / (2 * h)
    ad = float((g0 * v).sum())
    print(f"  f0 direction: finite differences {fd:.12e}  adjoint {ad:.12e}  "
          f"rel {abs(fd - ad) / abs(ad):.2e}", flush=True)
    assert abs(fd - ad) <= 1e-6 * abs(ad)

    # the virial ratio eta: the run's loss at eta +- h (the grid is fixed by
    # the explicit ranges; R0 = 1/eta follows), against the adjoint through
    # the set-up's derivative
    eta = float(cfg["init.eta"])

    def loss_at(e):
        s = V.Solver(dict(cfg, **{"init.eta": e}))
        c = s.context
        c.step(n)
        val = loss(c.mass_profile())
        s.close()
        return val

    # The loss is strongly curved in eta: the second-order central difference
    # converges as h^2 (relative error 2.8e-4, 7.1e-5, 1.8e-5, 4.4e-6, 1.1e-6
    # at h = 8, 4, 2, 1, 0.5e-6), so at h = 1e-6 it carries a truncation error
    # of 4.4e-6; the fourth-order central difference (h and 2h combined)
    # removes it and agrees with the adjoint to about 1e-9.
    Lp, Lm = loss_at(eta + h), loss_at(eta - h)
    fd2 = (Lp - Lm) / (2 * h)
    fd4 = (8 * (Lp - Lm) - (loss_at(eta + 2 * h) - loss_at(eta - 2 * h))) / (12 * h)
    df, dvals = V.HostSetup(cfg).dparam("eta")
    gv_front = ring_to_front(gv, head0, off, ln)
    ad_eta = float((g0 * df).sum() + (gv_front * dvals).sum())
    print(f"  eta: finite differences {fd2:.12e} (2nd order, rel "
          f"{abs(fd2 - ad_eta) / abs(ad_eta):.2e}) {fd4:.12e} (4th order, rel "
          f"{abs(fd4 - ad_eta) / abs(ad_eta):.2e})  adjoint {ad_eta:.12e}", flush=True)
    assert abs(fd4 - ad_eta) <= 1e-6 * abs(ad_eta)
    assert abs(fd2 - ad_eta) <= 1e-5 * abs(ad_eta)
