"""The allocation-free entry point vls_density_potential_noalloc against
vls_density_potential and vls_mass_profile, through the C ABI: the same
bits for f held in a caller's device array, on two calls in a row (the
second on a smaller grid, so that the shared scratch holds the first call's
larger contents), the context's own moments left as they were, and a grid
whose partial sums exceed the static scratch refused with VLS_ERR_ARG."""
import numpy as np
import pytest

from tests.configs import UDF_C1, UDF_TINY

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")
torch = pytest.importorskip("torch")


def _check_run(cfg, steps):
    s = V.Solver(cfg)
    s.step(steps)
    ctx = s.context
    n = s.nR
    rho_ref, phi_ref = ctx.density_potential()
    M_ref = ctx.mass_profile()
    before = s.moments()
    f = torch.from_numpy(s.f()).to("cuda")
    out = torch.full((3, n), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.density_potential_noalloc(f.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                                  out[2].data_ptr())
    ctx.synchronize()
    rho, phi, M = out.cpu().numpy()
    assert np.array_equal(rho, rho_ref) and np.array_equal(phi, phi_ref)
    assert np.array_equal(M, M_ref)
    after = s.moments()
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
    # outputs may be omitted
    out.fill_(float("nan"))
    torch.cuda.synchronize()
    ctx.density_potential_noalloc(f.data_ptr(), None, out[1].data_ptr(), None)
    ctx.synchronize()
    assert np.array_equal(out[1].cpu().numpy(), phi_ref)
    assert torch.isnan(out[0]).all() and torch.isnan(out[2]).all()
    return s


def test_noalloc_matches_allocating_entry_points():
    print("case noalloc udf 65x65x32 then udf 17x17x8", flush=True)
    _check_run(UDF_C1, 7)
    _check_run(UDF_TINY, 3)


def test_noalloc_rejects_oversized_runs():
    print("case noalloc oversized grid (1101 r nodes x 240 j slices)", flush=True)
    cfg = dict(UDF_TINY, **{"grid.Pres": 1100, "grid.Ures": [8, 240]})
    s = V.Solver(cfg)
    f = torch.zeros(s.nJ * s.nU * s.nR, dtype=torch.float64, device="cuda")
    M = torch.zeros(s.nR, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(V.VlsError, match="status 1"):
        s.context.density_potential_noalloc(f.data_ptr(), None, None, M.data_ptr())
    with pytest.raises(V.VlsError, match="status 1"):
        s.context.density_potential_noalloc(f.data_ptr(), None, None, None)
