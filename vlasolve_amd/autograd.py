"""torch.autograd wrappers of the discrete adjoint (a thin harness over the C ABI).

MassProfileRun.apply(f0, vals0, ctx, n_steps, ck_every, head0) runs n_steps
steps of the device context `ctx` (vlasolve_amd.Context) from the initial
distribution f0 ([nJ, nU, nR], float64) and delay-line values vals0 (ring
layout of vls_set_chains, with ring heads head0), and returns the enclosed
mass profile M(<R_r) of the final state.  Its backward pass is vls_adjoint:
the gradients with respect to f0 and vals0 of any loss of that profile.

InitialState.apply(theta, cfg, names) returns the initial state (f0, vals0)
of the run configuration cfg with its initial-condition parameters `names`
("eta", "M", ...) set to the values of the tensor theta, as the host set-up
builds it (renormalisation included; the delay-line values front first, the
ring layout of a new run, whose heads are 0).  Its backward pass is the
set-up's hand-written derivative (vlsd_host_dparam): chained with
MassProfileRun it gives the gradient of a loss with respect to the
initial-condition parameters.

Tensors live on the CPU; the run itself stays on the GPU.
"""
from __future__ import annotations

import numpy as np


def _torch():
    import torch

    return torch


def _make():
    torch = _torch()

    class MassProfileRun(torch.autograd.Function):
        @staticmethod
        def forward(fctx, f0, vals0, ctx, n_steps, ck_every, head0):
            ctx.upload_f(0, f0.detach().cpu().numpy())
            ctx.put_chains(head0, vals0.detach().cpu().numpy())
            ctx.record(n_steps, ck_every)
            M = ctx.mass_profile()
            fctx.vls = ctx
            return torch.from_numpy(M)

        @staticmethod
        def backward(fctx, gM):
            g0, gv = fctx.vls.adjoint(dL_dM=gM.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g0), torch.from_numpy(gv), None, None, None, None

    return MassProfileRun


def _make_initial():
    torch = _torch()

    class InitialState(torch.autograd.Function):
        @staticmethod
        def forward(fctx, theta, cfg, names):
            from . import HostSetup

            if theta.dtype != torch.float64:
                raise TypeError("InitialState: theta must be float64 (a float32 parameter "
                                "already differs from the value meant)")
            vals = [float(x) for x in theta.detach().cpu().numpy().reshape(-1)]
            if len(vals) != len(names):
                raise ValueError("one value per parameter name")
            run = dict(cfg, **{"init." + k: v for k, v in zip(names, vals)})
            h = HostSetup(run)
            f0 = h.f()
            v0 = h.chains()["val"]
            fctx.host, fctx.names = h, list(names)
            return torch.from_numpy(f0), torch.from_numpy(v0)

        @staticmethod
        def backward(fctx, g_f0, g_vals):
            g = param_gradient(fctx.host, g_f0.detach().cpu().numpy(),
                               g_vals.detach().cpu().numpy(), fctx.names)
            return torch.tensor(g, dtype=torch.float64), None, None

    return InitialState


def param_gradient(host, grad_f0, grad_vals0, names, species=0):
    """<grad_f0, df0/dp> + <grad_vals0, dvals0/dp> for each parameter p in
    names, with the set-up's derivative (vlasolve_amd.HostSetup.dparam);
    grad_vals0 in the front-first layout of a new run's delay lines."""
    out = []
    for p in names:
        df, dv = host.dparam(p, species)
        g = float((np.asarray(grad_f0) * df).sum())
        if dv.size:
            g += float((np.asarray(grad_vals0) * dv).sum())
        out.append(g)
    return out


_cache = {}


def MassProfileRun():
    """The autograd Function class (created on first use, torch imported lazily)."""
    if "fn" not in _cache:
        _cache["fn"] = _make()
    return _cache["fn"]


def InitialState():
    """The autograd Function class of the initial state (torch imported lazily)."""
    if "init" not in _cache:
        _cache["init"] = _make_initial()
    return _cache["init"]
