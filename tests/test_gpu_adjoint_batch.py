"""The batched adjoint (vls_record_batch, vls_adjoint_batch): several runs of
different grids recorded and reversed together, each on a stream of its own.

* Every member's gradient equals the gradient of the same run recorded and
  reversed alone, to the rounding of the floating-point atomics of the
  drift's transpose (1e-12 of the gradient's scale).
* The adjoint identity against the oracle's tangent-linear model for every
  member of a small batch.
"""
import numpy as np
import pytest

from tests.configs import UDF_C1, UDF_TINY
from tests.test_gpu_adjoint import chain_layout, ring_to_front

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")
from oracle import oracle as O  # noqa: E402

BATCH = [dict(UDF_TINY, **{"grid.Pres": 16, "grid.Ures": [16, 8]}),
         dict(UDF_C1, **{"grid.Pres": 24, "grid.Ures": [20, 10], "init.R0": 1.5}),
         dict(UDF_C1, **{"grid.Pres": 40, "grid.Ures": [36, 12], "init.alpha": 0.8})]


def _weights(ens, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(ens.context.info(k)["nR"]) for k in range(ens.n)]


def test_batch_equals_single_runs():
    n, ck = 7, 3
    print(f"case adjoint batch of {len(BATCH)} udf runs, {n} steps, ck {ck}", flush=True)
    e = V.Ensemble(BATCH)
    w = _weights(e, 1)
    f0 = [e.f(k) for k in range(e.n)]
    ch0 = [e.context.get_chains(k) for k in range(e.n)]
    e.context.record_batch(n, ck)
    grads = e.context.adjoint_batch(dL_dM=w)
    for k, cfg in enumerate(BATCH):
        s = V.Solver(cfg)
        ctx = s.context
        assert np.array_equal(s.f(), f0[k])
        ctx.record(n, ck)
        g0, gv = ctx.adjoint(dL_dM=w[k])
        bg0, bgv = grads[k]
        assert np.abs(bg0 - g0).max() <= 1e-12 * np.abs(g0).max(), k
        if gv.size:
            assert np.abs(bgv - gv).max() <= 1e-12 * np.abs(gv).max(), k
        # the member is back in its state after the recorded run
        assert np.array_equal(e.f(k), ctx.download_f(0, f0[k].shape)), k
        assert ch0[k][0].shape == e.context.get_chains(k)[0].shape


def test_batch_adjoint_identity():
    n, ck = 6, 2
    print(f"case adjoint batch identity, {len(BATCH)} runs", flush=True)
    e = V.Ensemble(BATCH)
    w = _weights(e, 2)
    f0 = [e.f(k) for k in range(e.n)]
    head0 = [e.context.get_chains(k)[0] for k in range(e.n)]
    slots0 = [e.context.get_chains(k)[1] for k in range(e.n)]
    e.context.record_batch(n, ck)
    grads = e.context.adjoint_batch(dL_dM=w)
    rng = np.random.default_rng(9)
    for k, cfg in enumerate(BATCH):
        g0, gv = grads[k]
        off, ln = chain_layout(cfg)
        gv_front = ring_to_front(gv, head0[k], off, ln)
        orc = O.Oracle(cfg, nthreads=0)
        assert np.array_equal(orc.f(), f0[k])
        df = rng.standard_normal(f0[k].shape) * np.abs(f0[k]).max()
        dv = rng.standard_normal(gv.shape) * (np.abs(slots0[k]).max() if slots0[k].size else 1.0)
        orc.tl_init(df, dv if dv.size else None)
        orc.tl_step(n)
        _, dMo = orc.tl_mass()
        lhs = float(w[k] @ dMo)
        rhs = float((g0 * df).sum() + (gv_front * dv).sum())
        print(f"  run {k}: <w, J v> = {lhs:.15e}  <J^T w, v> = {rhs:.15e}", flush=True)
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs), k
