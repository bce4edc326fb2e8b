"""torch.autograd wrapper of the discrete adjoint (a thin harness over the C ABI).

MassProfileRun.apply(f0, vals0, ctx, n_steps, ck_every, head0) runs n_steps
steps of the device context `ctx` (vlasolve_amd.Context) from the initial
distribution f0 ([nJ, nU, nR], float64) and delay-line values vals0 (ring
layout of vls_set_chains, with ring heads head0), and returns the enclosed
mass profile M(<R_r) of the final state.  Its backward pass is vls_adjoint:
the gradients with respect to f0 and vals0 of any loss of that profile.
Tensors live on the CPU; the run itself stays on the GPU.
"""
from __future__ import annotations

import numpy as np


def _torch():
    import torch

    return torch


class _Fn:
    pass


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


_cache = {}


def MassProfileRun():
    """The autograd Function class (created on first use, torch imported lazily)."""
    if "fn" not in _cache:
        _cache["fn"] = _make()
    return _cache["fn"]
