"""The C-ABI library on the CPU: it loads, exports every function the public
header declares, and its host-side helpers behave; device calls fail cleanly
where there is no GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import vlasolve_amd as V

HEADER = Path(__file__).resolve().parents[1] / "include" / "vlasolve_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vlsd?_\w+)\s*\(", text)))


def test_exports_every_declared_function():
    lib = V.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(V.EXPORTS) | set(names)


def test_no_device_is_reported():
    if V.lib().vls_create and _has_gpu():
        pytest.skip("a GPU is present")
    with pytest.raises(V.VlsError, match="no HIP device"):
        V.Solver({"grid.Pres": 8, "grid.Ures": [8, 4]})


def _has_gpu():
    try:
        V.device_info()
        return True
    except V.VlsError:
        return False


def thomas(dim, hinv):
    L = V.lib()
    P = C.POINTER(C.c_double)
    L.vls_thomas_table.argtypes = [C.c_int, C.c_double, P, P, P, P]
    u, d, rd, cs = np.zeros(dim + 2), np.zeros(dim + 2), np.zeros(dim + 2), np.zeros(8)
    assert L.vls_thomas_table(dim, hinv, *(a.ctypes.data_as(P) for a in (u, d, rd, cs))) == 0
    return u, d, rd, cs


def test_pivot_table_fixed_point():
    # the pivots of the natural spline solve settle exactly after a few rows and
    # do not depend on the knot spacing
    u1, d1, rd1, c1 = thomas(201, 31.4)
    u2, d2, rd2, c2 = thomas(65, 3.0)
    assert d1[1] == 1.0 and u1[1] == 0.0  # first interior row after the natural row
    assert np.all(d1[17:202] == d1[201]) and np.all(u1[17:202] == u1[201])
    assert np.array_equal(d1[1:40], d2[1:40])
    assert np.all(rd1[1:202] == 1.0 / d1[1:202])


def test_known_divisor_division_is_correctly_rounded():
    L = V.lib()
    P = C.POINTER(C.c_double)
    L.vls_check_division.argtypes = [P, C.c_int, C.c_int64, C.c_uint64, C.POINTER(C.c_int64)]
    divs = []
    for dim, hinv in [(201, 25.0), (513, 60.0), (17, 2.0)]:
        u, d, rd, cs = thomas(dim, hinv)
        divs += list(d[1:dim + 1]) + [cs[6]]
    divs += list(np.random.default_rng(0).uniform(1e-3, 1e3, 200))
    b = np.array(divs)
    bad = C.c_int64(-1)
    assert L.vls_check_division(b.ctypes.data_as(P), len(b), 20000, 99, C.byref(bad)) == 0
    assert bad.value == 0


def test_ndfield_roundtrip(tmp_path):
    from vlasolve_amd import ndfield

    a = np.random.default_rng(1).uniform(size=(4, 5, 6))
    p = tmp_path / "x.ND"
    ndfield.write(p, a, [0.01, -2, 0], [24.99, 4, 1.6], comment="1 1 0 1 2 0")
    b, meta = ndfield.read(p)
    assert np.array_equal(a, b) and meta["dims"] == [6, 5, 4]
    assert meta["comment"] == "1 1 0 1 2 0"
