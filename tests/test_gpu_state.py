"""State and memory rules of the C ABI around the adjoint (vls_record,
vls_adjoint, vls_set_chains): a recording made before the delay lines change
is refused, and repeated recordings and delay-line uploads release the device
memory they replace."""
import numpy as np
import pytest

from tests.configs import UDF_TINY

pytestmark = pytest.mark.gpu

V = pytest.importorskip("vlasolve_amd")


def test_repeated_record_keeps_memory_flat():
    print("case udf_tiny repeated vls_record", flush=True)
    s = V.Solver(UDF_TINY)
    ctx = s.context
    ctx.record(4, 2)
    n1, b1 = ctx.memory()
    ctx.record(4, 2)
    assert ctx.memory() == (n1, b1)
    ctx.record(8, 2)  # more checkpoints: the checkpoint arrays are replaced
    n2, b2 = ctx.memory()
    assert n2 == n1 and b2 > b1
    ctx.record(8, 4)  # longer segments: the segment buffer is replaced
    n3, b3 = ctx.memory()
    assert n3 == n2 and b3 > b2
    # smaller recordings reuse what is there
    ctx.record(4, 2)
    ctx.record(8, 4)
    ctx.record(2, 1)
    assert ctx.memory() == (n3, b3)


def test_adjoint_after_chain_change_is_refused():
    print("case udf_tiny set_chains then adjoint", flush=True)
    s = V.Solver(UDF_TINY)
    ctx = s.context
    assert ctx.info()["chains"] > 0
    ctx.record(3, 2)
    n0, b0 = ctx.memory()
    ctx.clear_chains()
    assert ctx.info()["chains"] == 0
    n1, _ = ctx.memory()
    assert n1 == n0  # the eight delay-line arrays were replaced, not added
    with pytest.raises(V.VlsError, match="status 3"):
        ctx.adjoint(dL_dM=[1.0] * ctx.info()["nR"])
    # a new recording with the current delay lines is accepted, and its
    # transposed foot table (rebuilt without the delay lines' back cells)
    # gives the gradient of a run whose first table was built without them
    # (the same state: three steps with the delay lines, which the recording
    # above leaves behind, then none)
    ctx.record(3, 2)
    g, _ = ctx.adjoint(dL_dM=[1.0] * ctx.info()["nR"])
    fs = V.Solver(UDF_TINY)
    fs.step(3)
    fresh = fs.context
    fresh.clear_chains()
    fresh.record(3, 2)
    g0, _ = fresh.adjoint(dL_dM=[1.0] * fresh.info()["nR"])
    assert np.allclose(g, g0, rtol=1e-14, atol=1e-300)
