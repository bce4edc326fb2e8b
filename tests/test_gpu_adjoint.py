"""The discrete adjoint of the HIP step, through the C ABI.

Checks: (1) against the oracle's tangent-linear model (an independent,
forward-mode derivative of the same discrete map, itself pinned against
finite differences on the CPU): the adjoint identity <w, J v> = <J^T w, v>
for random directions, over several checkpoint segments; (2) against central
finite differences of the GPU forward run, to the 1e-6 the adjoint is
specified to; (3) the torch.autograd harness; (4) the gradient with respect
to an initial-condition parameter (the PLDF virial ratio eta), through the
host set-up's derivative."""
import numpy as np
import pytest

from tests.configs import UDF_TINY, PLDF_SMALL

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")
from oracle import oracle as O  # noqa: E402


def ring_to_front(slots, head, off, ln):
    out = np.empty_like(slots)
    for k in range(len(head)):
        L, o = ln[k], off[k]
        idx = (head[k] + np.arange(L)) % L
        out[o:o + L] = slots[o + idx]
    return out


def chain_layout(cfg):
    ch = O.Oracle(cfg, nthreads=1).chains()
    return ch["off"], ch["len"]


@pytest.mark.parametrize("n,ck,fused,keep", [(6, 4, False, None), (5, 1, False, "0"),
                                             (6, 4, True, "2"), (6, 4, True, "0")])
def test_adjoint_matches_tangent_linear(monkeypatch, n, ck, fused, keep):
    """fused: the recorded steps keep the kick's column splines from the
    fused kick (small grids run the column kernels by default).  keep caps
    the steps the recorded run keeps itself (None: all of them here; "0":
    every segment recomputed in the reverse sweep; "2": the last segment
    kept, the first recomputed)."""
    print(f"case udf_tiny adjoint n={n} ck={ck} fused={fused} keep={keep}", flush=True)
    if fused:
        monkeypatch.setenv("VLS_FUSED_KICK", "1")
    if keep is not None:
        monkeypatch.setenv("VLS_ADJ_KEEP_STEPS", keep)
    cfg = UDF_TINY
    solver = V.Solver(cfg)
    ctx = solver.context
    f0 = solver.f()
    head0, slots0 = ctx.get_chains()
    off, ln = chain_layout(cfg)
    assert len(head0) > 0, "the case must exercise the delay lines"
    ctx.record(n, ck)
    M = ctx.mass_profile()
    rng = np.random.default_rng(11)
    w = rng.standard_normal(M.shape)
    g0, gv = ctx.adjoint(dL_dM=w)
    gv_front = ring_to_front(gv, head0, off, ln)

    orc = O.Oracle(cfg, nthreads=0)
    assert np.array_equal(orc.f(), f0)
    df = rng.standard_normal(f0.shape) * np.abs(f0).max()
    dv = rng.standard_normal(gv.shape) * np.abs(slots0).max()
    orc.tl_init(df, dv)
    orc.tl_step(n)
    Mo, dMo = orc.tl_mass()
    assert np.allclose(Mo, M, rtol=1e-12, atol=1e-15)
    lhs = float(w @ dMo)
    rhs = float((g0 * df).sum() + (gv_front * dv).sum())
    print(f"<w, J v> = {lhs:.15e}  <J^T w, v> = {rhs:.15e}")
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # the context is back in the state after the recorded run
    assert np.array_equal(ctx.download_f(0, f0.shape), _after(cfg, n))


@pytest.mark.parametrize("pres,ures,keep", [(15, [15, 6], None), (20, [18, 5], None),
                                            (20, [18, 5], "0")])
def test_adjoint_generic_lines(monkeypatch, pres, ures, keep):
    """Shapes without register line solvers: the streaming line adjoints.
    With 18 padded nodes per line the last row's two coefficients fall on
    both sides of a 16-element chunk boundary."""
    from tests.configs import UDF_C1

    cfg = dict(UDF_C1, **{"grid.Pres": pres, "grid.Ures": ures})
    n, ck = 5, 2
    print(f"case udf {pres}x{ures} adjoint n={n} ck={ck} keep={keep}", flush=True)
    if keep is not None:
        monkeypatch.setenv("VLS_ADJ_KEEP_STEPS", keep)
    solver = V.Solver(cfg)
    ctx = solver.context
    f0 = solver.f()
    head0, slots0 = ctx.get_chains()
    off, ln = chain_layout(cfg)
    ctx.record(n, ck)
    M = ctx.mass_profile()
    rng = np.random.default_rng(5)
    w = rng.standard_normal(M.shape)
    g0, gv = ctx.adjoint(dL_dM=w)
    gv_front = ring_to_front(gv, head0, off, ln)
    orc = O.Oracle(cfg, nthreads=0)
    df = rng.standard_normal(f0.shape) * np.abs(f0).max()
    dv = rng.standard_normal(gv.shape) * (np.abs(slots0).max() if slots0.size else 1.0)
    orc.tl_init(df, dv)
    orc.tl_step(n)
    Mo, dMo = orc.tl_mass()
    assert np.allclose(Mo, M, rtol=1e-12, atol=1e-15)
    lhs = float(w @ dMo)
    rhs = float((g0 * df).sum() + (gv_front * dv).sum())
    print(f"<w, J v> = {lhs:.15e}  <J^T w, v> = {rhs:.15e}")
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def _after(cfg, n):
    o = O.Oracle(cfg, nthreads=0)
    o.step(n)
    return o.f()


def _loss_run(ctx, f0, head0, slots0, n, Mt):
    ctx.upload_f(0, f0)
    ctx.put_chains(head0, slots0)
    ctx.step(n)
    M = ctx.mass_profile()
    return float(((M - Mt) ** 2).sum()), M


def test_adjoint_against_gpu_finite_differences():
    print("case udf_tiny finite differences", flush=True)
    cfg = UDF_TINY
    n = 8
    solver = V.Solver(cfg)
    ctx = solver.context
    f0 = solver.f()
    head0, slots0 = ctx.get_chains()
    rng = np.random.default_rng(5)
    _, Mref = _loss_run(ctx, f0, head0, slots0, n, 0.0)
    Mt = Mref * (1 + 0.01 * rng.standard_normal(Mref.shape))
    # gradient
    ctx.upload_f(0, f0)
    ctx.put_chains(head0, slots0)
    ctx.record(n, 3)
    M = ctx.mass_profile()
    g0, gv = ctx.adjoint(dL_dM=2 * (M - Mt))
    v = rng.standard_normal(f0.shape) * np.abs(f0).max()
    vv = rng.standard_normal(slots0.shape) * np.abs(slots0).max()
    h = 1e-7
    Lp, _ = _loss_run(ctx, f0 + h * v, head0, slots0 + h * vv, n, Mt)
    Lm, _ = _loss_run(ctx, f0 - h * v, head0, slots0 - h * vv, n, Mt)
    fd = (Lp - Lm) / (2 * h)
    ad = float((g0 * v).sum() + (gv * vv).sum())
    print(f"finite differences {fd:.12e}  adjoint {ad:.12e}")
    assert abs(fd - ad) <= 1e-6 * abs(ad)


def test_autograd_harness():
    torch = pytest.importorskip("torch")
    from vlasolve_amd.autograd import MassProfileRun

    print("case udf_tiny torch.autograd", flush=True)
    cfg = UDF_TINY
    n = 4
    solver = V.Solver(cfg)
    ctx = solver.context
    f0 = solver.f()
    head0, slots0 = ctx.get_chains()
    Fn = MassProfileRun()
    f0_t = torch.tensor(f0, requires_grad=True)
    v0_t = torch.tensor(slots0, requires_grad=True)
    M = Fn.apply(f0_t, v0_t, ctx, n, 2, head0)
    Mt = M.detach() * 1.01
    loss = ((M - Mt) ** 2).sum()
    loss.backward()
    # the same gradient straight from the C ABI
    ctx.upload_f(0, f0)
    ctx.put_chains(head0, slots0)
    ctx.record(n, 2)
    M2 = ctx.mass_profile()
    g0, gv = ctx.adjoint(dL_dM=2 * (M2 - Mt.numpy()))
    # the drift's transpose sums with atomics, so two calls agree to rounding
    # relative to the gradient's scale, not bit for bit
    assert np.abs(f0_t.grad.numpy() - g0).max() <= 1e-12 * np.abs(g0).max()
    assert np.abs(v0_t.grad.numpy() - gv).max() <= 1e-12 * np.abs(gv).max()


def test_gradient_wrt_pldf_eta():
    """dL/deta = <dL/df0, df0/deta> + <dL/dvals0, dvals0/deta>, the
    set-up's derivative by central differences of the host initial
    conditions, against central differences of the whole run in eta."""
    print("case pldf eta gradient", flush=True)
    base = dict(PLDF_SMALL, **{"grid.Pres": 32, "grid.Ures": [32, 12]})
    eta, n, h = 0.1, 6, 1e-6

    def cfg_at(e):
        return dict(base, **{"init.eta": e})

    solver = V.Solver(cfg_at(eta))
    ctx = solver.context
    f0 = solver.f()
    head0, slots0 = ctx.get_chains()
    off, ln = chain_layout(cfg_at(eta))
    _, Mref = _loss_run(ctx, f0, head0, slots0, n, 0.0)
    Mt = Mref * 1.02

    def run_loss(e):
        s = V.Solver(cfg_at(e))
        c = s.context
        c.step(n)
        M = c.mass_profile()
        return float(((M - Mt) ** 2).sum())

    fd = (run_loss(eta + h) - run_loss(eta - h)) / (2 * h)
    ctx.upload_f(0, f0)
    ctx.put_chains(head0, slots0)
    ctx.record(n, n)
    M = ctx.mass_profile()
    g0, gv = ctx.adjoint(dL_dM=2 * (M - Mt))
    hp, hm = V.HostSetup(cfg_at(eta + h)), V.HostSetup(cfg_at(eta - h))
    df = (hp.f() - hm.f()) / (2 * h)
    dv_front = (hp.chains()["val"] - hm.chains()["val"]) / (2 * h)
    gv_front = ring_to_front(gv, head0, off, ln)
    ad = float((g0 * df).sum() + (gv_front * dv_front).sum())
    print(f"dL/deta finite differences {fd:.10e}  adjoint {ad:.10e}")
    assert abs(fd - ad) <= 1e-6 * abs(fd)


def test_autograd_parameter_gradient():
    """The initial-condition parameter through the torch harness: the PLDF
    virial ratio eta -> initial state (InitialState, the set-up's
    derivative) -> run (MassProfileRun, the adjoint) -> mass-profile loss,
    against central differences of the whole run in eta."""
    torch = pytest.importorskip("torch")
    from vlasolve_amd.autograd import InitialState, MassProfileRun

    print("case pldf eta through the autograd harness", flush=True)
    cfg = dict(PLDF_SMALL, **{"grid.Pres": 32, "grid.Ures": [32, 12]})
    eta, n, ck, h = 0.1, 6, 3, 1e-6
    solver = V.Solver(cfg)
    ctx = solver.context
    head0, _ = ctx.get_chains()
    assert not head0.any()  # a new run: the ring layout is front first
    ctx.step(n)
    Mt = ctx.mass_profile() * 1.02

    theta = torch.tensor([eta], dtype=torch.float64, requires_grad=True)
    f0, v0 = InitialState().apply(theta, cfg, ["eta"])
    M = MassProfileRun().apply(f0, v0, ctx, n, ck, head0)
    loss = ((M - torch.from_numpy(Mt)) ** 2).sum()
    loss.backward()

    def run_loss(e):
        s = V.Solver(dict(cfg, **{"init.eta": e}))
        c = s.context
        c.step(n)
        val = float(((c.mass_profile() - Mt) ** 2).sum())
        s.close()
        return val

    # the loss is strongly curved in eta: the second-order central difference
    # at h carries a truncation error of about 1e-6 here, the fourth-order one
    # (h and 2h combined) none that matters
    Lp, Lm = run_loss(eta + h), run_loss(eta - h)
    fd2 = (Lp - Lm) / (2 * h)
    fd4 = (8 * (Lp - Lm) - (run_loss(eta + 2 * h) - run_loss(eta - 2 * h))) / (12 * h)
    g = float(theta.grad[0])
    print(f"dL/deta finite differences {fd2:.10e} (2nd order) {fd4:.10e} (4th order)  "
          f"harness {g:.10e}")
    assert abs(fd4 - g) <= 1e-6 * abs(fd4)
    assert abs(fd2 - g) <= 1e-5 * abs(fd2)
