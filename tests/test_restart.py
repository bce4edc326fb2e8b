""".vls restarts: the host reader against files the reference writes, the
GPU solver's write -> reload round trip, and the reference continuing from a
file this build wrote (vlasovSolver::init with -reload,
src/vlasovSolver.hxx:104-132; layout src/vlasovSolver.hxx:248-314,
src/vlasovSolver21.hxx:550-583, src/sphericalKernel.hxx:88-130)."""
import numpy as np
import pytest

import vlasolve_amd as V
from oracle import oracle as O
from tests.configs import UDF_TINY
from vlasolve_amd import vlsfile

needs_ref = pytest.mark.skipif(not O.REF_EXE.exists(), reason="reference executable not built")
DT = 0.005


def _ref_final(cfg, nsteps, wd, extra=None):
    c = dict(cfg)
    c.update({"solver.Tmax": DT * nsteps, "solver.statsEvery": -1,
              "solver.snapshotEvery": -1, "solver.restartEvery": -1})
    c.update(extra or {})
    O.run_reference(c, wd, threads=1)
    vls = sorted(p for p in wd.iterdir() if p.name.startswith("final_") and p.suffix == ".vls")
    return vls[-1]


def _front_first(state):
    keys = sorted(state["chains"])
    return np.concatenate([state["chains"][k][2] for k in keys]) if keys else np.zeros(0)


@needs_ref
def test_host_reads_reference_restart(tmp_path):
    path = _ref_final(UDF_TINY, 9, tmp_path)
    state = vlsfile.read(path)
    h = V.HostSetup({"reload": str(path.with_suffix(""))})
    assert np.array_equal(h.f(), state["f"])
    ch = h.chains()
    keys = np.array(sorted(state["chains"]), np.int64)
    assert np.array_equal(ch["point"], keys)
    assert np.array_equal(ch["val"], _front_first(state))


@needs_ref
def test_host_reads_reference_two_species_restart(tmp_path):
    """A restart the reference writes with two species (nFields = 2, reflective
    kernel): f interleaved per point."""
    from tests.configs import UDF_2SPECIES

    path = _ref_final(UDF_2SPECIES, 6, tmp_path)
    state = vlsfile.read(path)
    assert state["grid"]["nFields"] == 2 and state["f"].shape[-1] == 2
    h = V.HostSetup({"reload": str(path.with_suffix(""))})
    assert h.n_species == 2 and np.array_equal(h.f(), state["f"])


@pytest.mark.gpu
def test_gpu_two_species_restart_round_trip(tmp_path):
    """Two species with delay lines: f and both species' ring values survive
    write -> reload, and the continuation is bit-identical."""
    from tests.configs import UDF_2SPECIES_DELAYED

    print("case udf 2 species restart round trip", flush=True)
    a = V.Solver(UDF_2SPECIES_DELAYED)
    a.step(6)
    a.write_restart(tmp_path / "mid")
    state = vlsfile.read(tmp_path / "mid.vls")
    assert np.array_equal(state["f"], a.f())
    b = V.Solver({"reload": str(tmp_path / "mid")})
    assert np.array_equal(b.f(), a.f())
    a.step(5)
    b.step(5)
    assert np.array_equal(a.f(), b.f())


@pytest.mark.gpu
def test_gpu_restart_round_trip(tmp_path):
    print("case udf_tiny restart round trip", flush=True)
    a = V.Solver(UDF_TINY)
    a.step(7)
    a.write_restart(tmp_path / "mid")
    state = vlsfile.read(tmp_path / "mid.vls")
    assert np.array_equal(state["f"], a.f())
    assert state["stepCount"] == 7 and abs(state["time"] - 7 * DT) < 1e-12
    b = V.Solver({"reload": str(tmp_path / "mid")})
    assert np.array_equal(b.f(), a.f())
    a.step(5)
    b.step(5)
    assert np.array_equal(a.f(), b.f())


@pytest.mark.gpu
@needs_ref
def test_reference_continues_from_our_restart(tmp_path):
    """The reference executable (built from its sources, a CPU program) reloads
    the file written from the GPU state and continues; the GPU continues too.
    The two differ only by the order of the moment sums."""
    print("case udf_tiny reference reload", flush=True)
    a = V.Solver(UDF_TINY)
    a.step(6)
    a.write_restart(tmp_path / "ours")
    final = _ref_final(UDF_TINY, 10, tmp_path, {"reload": str(tmp_path / "ours")})
    ref = vlsfile.read(final)
    assert ref["stepCount"] == 10
    a.step(4)
    fg = a.f()
    rel = np.abs(fg - ref["f"]).max() / np.abs(ref["f"]).max()
    print(f"reference continuation vs GPU: max rel diff {rel:.2e}")
    assert rel <= 1e-11


@needs_ref
def test_host_refuses_nonzero_kernel_mass(tmp_path):
    """The reference adds a restart's kernelMass to the enclosed mass of the
    kick (vlasovSolver21_spline.hxx:738-742) and always writes 0; the device
    step has no such term, so a file with another value is refused."""
    import struct

    path = _ref_final(UDF_TINY, 3, tmp_path)
    raw = bytearray(path.read_bytes())
    tag = b"VLASOV_SOLVER21 v0.12"
    at = raw.find(struct.pack("<i", len(tag)) + tag)
    assert at > 0
    o = at + 4 + len(tag) + 4  # the tag, then kernelType (int), then kernelMass (double)
    assert struct.unpack_from("<d", raw, o)[0] == 0.0
    struct.pack_into("<d", raw, o, 0.5)
    bad = tmp_path / "patched.vls"
    bad.write_bytes(bytes(raw))
    with pytest.raises(V.VlsError, match="kernelMass"):
        V.HostSetup({"reload": str(bad.with_suffix(""))})
