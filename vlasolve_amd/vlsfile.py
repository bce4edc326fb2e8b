"""Reader for .vls restart files (the reference's solver-state format).

Layout written by vlasovSolver::write and the classes it calls
(src/vlasovSolver.hxx:272-287, src/grid_handler.hxx:325-347,
src/grid_base.hxx:123-153,492-497, src/vlasovSolver21.hxx:568-583,
src/sphericalKernel.hxx:195-213, src/cellChain.hxx:91-97,
src/dummyCell21_vect.hxx:49-56): tags are an int length followed by the
characters; little-endian, as written on x86-64.
"""
from __future__ import annotations

import struct

import numpy as np


class _R:
    def __init__(self, raw):
        self.raw, self.o = raw, 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.raw, self.o)
        self.o += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def tag(self, expect):
        n = self.take("i")
        s = self.raw[self.o:self.o + n].decode()
        self.o += n
        if s != expect:
            raise ValueError(f"tag {s!r} != {expect!r}")

    def skip(self, n):
        self.o += n


def _grid_params(r):
    d = r.take("6d")
    iv = r.take("19i")
    r.take("q")
    r.skip(2048 - 4 * 6 - 8)
    return dict(Pmin=d[0], Pmax=d[1], Umin0=d[3], Umax=(d[4], d[5]), Pres=iv[0],
                Ures=(iv[1], iv[2]), Pscale=iv[3], Uscale=(iv[4], iv[5]), nFields=iv[12])


def read(path):
    """Return dict(time, dt, stepCount, noLinking, grid, f (nJ, nU, nR), or
    (nJ, nU, nR, nFields) with several fields, kernelType, chains = {point
    index: (R[], U[], val[])} front first; val[] has one value per cell, or
    one row of nFields values per cell (zero-padded) with several fields)."""
    raw = open(path, "rb").read()
    r = _R(raw)
    if r.take("i") != 1:
        raise ValueError("big-endian files are not supported")
    r.tag("VLASOV_SOLVER v0.13")
    G, T0, Tmax, dt = r.take("4d")
    r.take("i")
    r.take("3q")
    time_, dt2, step = r.take("d"), r.take("d"), r.take("q")
    nolink = r.take("i")
    r.tag("GRID_HANDLER v0.11")
    r.take("i")
    r.take("3i")
    r.take("3q")
    _grid_params(r)
    r.tag("GRID_BASE v0.12")
    g = _grid_params(r)
    nR, nU, nJ, nF = g["Pres"] + 1, g["Ures"][0] + 1, g["Ures"][1], max(g["nFields"], 1)
    shape = (nJ, nU, nR) + ((nF,) if nF > 1 else ())
    f = np.frombuffer(raw, "<f8", nR * nU * nJ * nF, r.o).reshape(shape).copy()
    r.o += 8 * nR * nU * nJ * nF
    r.tag("GRID_21 v0.11")
    r.skip(3 * 2048)
    r.tag("VLASOV_SOLVER21 v0.12")
    ktype = r.take("i")
    r.take("d")
    r.take("i")
    r.skip(2048)
    chains = {}
    if ktype == 0:
        r.tag("KERNEL_CELLS v0.10")
        for _ in range(r.take("q")):
            for _ in range(r.take("q")):
                key = r.take("q")
                n = r.take("q")
                R, U, V = [], [], []
                for _ in range(n):
                    R.append(r.take("d"))
                    U.append(r.take("d"))
                    nv = r.take("i")
                    vals = r.take(f"{nv}d") if nv else ()
                    vals = (vals,) if isinstance(vals, float) else tuple(vals)
                    if nF > 1:
                        V.append(list(vals[:nF]) + [0.0] * (nF - len(vals[:nF])))
                    else:
                        V.append(vals[0] if nv else 0.0)
                chains[key] = (np.array(R), np.array(U), np.array(V))
    r.tag("VLASOV_SOLVER21_SPLINE v0.11")
    return dict(time=time_, dt=dt2, stepCount=step, noLinking=nolink, grid=g, f=f,
                kernelType=ktype, chains=chains, G=G)
