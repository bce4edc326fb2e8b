"""Host logic of the product on the CPU: the C++ set-up behind the C ABI
(initial conditions, renormalisation, feet, delay lines, parameter parsing)
against the oracle, without a device."""
import os
import re
import tempfile

import numpy as np
import pytest

import vlasolve_amd as V
from oracle import oracle as O
from tests.configs import SMALL_CASES, UDF_TINY


@pytest.mark.parametrize("name", list(SMALL_CASES))
def test_host_setup_bit_exact(name):
    cfg, _ = SMALL_CASES[name]
    h = V.HostSetup(cfg)
    o = O.Oracle(cfg, nthreads=1)
    assert np.array_equal(h.f(), o.f())
    Rh, Uh, fh = h.feet()
    Ro, Uo, fo = o.feet()
    assert np.array_equal(Rh, Ro) and np.array_equal(Uh, Uo) and np.array_equal(fh, fo)
    ch, co = h.chains(), o.chains()
    for k in ("point", "len", "R", "U", "val"):
        assert np.array_equal(ch[k], co[k]), k
    assert h.kernel_radius == o.kernel_radius()


def test_parameter_file_equals_command_line():
    # "key v..." lines with the reference's delimiters and comments
    with tempfile.TemporaryDirectory() as wd:
        p = os.path.join(wd, "p.ini")
        with open(p, "w") as fh:
            fh.write("# comment line\n")
            fh.write("init.type = UDF   # trailing comment\n")
            fh.write("grid.Pres : 16\n")
            fh.write("grid.Ures [16, 8]\n")
            fh.write("grid.Umin {-2 0}\n")
            fh.write("solver.dt\t0.005\n")
        a = V.HostSetup([p])
    b = V.HostSetup({"init.type": "UDF", "grid.Pres": 16, "grid.Ures": [16, 8],
                     "grid.Umin": [-2, 0], "solver.dt": 0.005})
    assert np.array_equal(a.f(), b.f())


def test_command_line_wins_over_file():
    with tempfile.TemporaryDirectory() as wd:
        p = os.path.join(wd, "p.ini")
        with open(p, "w") as fh:
            fh.write("grid.Pres 40\ngrid.Ures 16 8\n")
        a = V.HostSetup([p, "-grid.Pres", "16"])
    assert a.nR == 17 and a.nU == 17 and a.nJ == 8


@pytest.mark.parametrize("argv,msg", [
    (["-init.type", "Plumber"], "unknown initial conditions type"),
    (["-grid.Pscale", "LOGARITHMIC"], "unknown scale type"),
    (["-a.b.c", "1"], "ill formated keyword"),
    (["positional", "extra"], "ERROR parsing command line"),
    (["-solver.kernelType", "sticky"], "unknown kernel type"),
    (["-solver.nSpecies", "6"], "nSpecies"),
    (["-grid.Uscale", "quadratic", "quadratic"], "linear scale"),
])
def test_errors(argv, msg):
    with pytest.raises(V.VlsError, match=re.escape(msg) if "(" in msg else msg):
        V.HostSetup(argv + ["-grid.Pres", "8", "-grid.Ures", "8", "4"])


def test_negative_numbers_are_values():
    h = V.HostSetup({"grid.Umin": [-1.5, 0], "grid.Umax": [1.5, 1.6], "grid.Pres": 8,
                     "grid.Ures": [8, 4]})
    o = O.Oracle({"grid.Umin": [-1.5, 0], "grid.Umax": [1.5, 1.6], "grid.Pres": 8,
                  "grid.Ures": [8, 4]}, nthreads=1)
    assert np.array_equal(h.f(), o.f())


def test_host_two_species_delayed_bit_exact():
    """Two species with delay lines: every cell carries both species' values
    (the oracle's semantics where the reference's back cells hold one)."""
    from tests.configs import UDF_2SPECIES_DELAYED

    h = V.HostSetup(UDF_2SPECIES_DELAYED)
    o = O.Oracle(UDF_2SPECIES_DELAYED, nthreads=1)
    assert h.shape[-1] == 2 and np.array_equal(h.f(), o.f())
    ch, co = h.chains(), o.chains()
    assert len(ch["point"]) > 0
    for k in ("point", "len", "R", "U", "val"):
        assert np.array_equal(ch[k], co[k]), k
    # the back cell of every chain starts at 0 for both species
    back = np.cumsum(ch["len"]) - 1
    assert np.all(ch["val"][back] == 0)


def test_species_limited_to_udf_and_pldf():
    with pytest.raises(V.VlsError, match="nSpecies"):
        V.HostSetup({"init.type": "Hernquist", "solver.nSpecies": 2, "grid.Pres": 16,
                     "grid.Ures": [16, 8]})
