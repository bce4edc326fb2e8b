"""Run configurations shared by the tests, as reference command-line keys."""

# The reference's first CPU-runnable case: UDF sphere (M=1, R0=2, alpha=0.5,
# erf smoothing 0.5, G=1) on 65 log-r x 65 v x 32 quadratic-j nodes, dt=0.005.
UDF_C1 = {
    "init.type": "UDF", "init.M": 1, "init.R0": 2, "init.alpha": 0.5,
    "init.smoothingRadius": 0.5, "solver.G": 1,
    "grid.Pmin": 0.01, "grid.Pmax": 25, "grid.Pres": 64, "grid.Pscale": "logarithmic",
    "grid.Umin": [-2, 0], "grid.Umax": [2, 1.6], "grid.Ures": [64, 32],
    "grid.Uscale": ["linear", "quadratic"], "solver.dt": 0.005,
}

UDF_TINY = dict(UDF_C1, **{"grid.Pres": 16, "grid.Ures": [16, 8]})

PLDF_SMALL = {
    "init.type": "PLDF", "init.n": 1, "init.alpha": 1, "init.eta": 0.1, "init.M": 1,
    "init.smoothingRadius": 0.5, "grid.Pres": 48, "grid.Ures": [48, 24], "grid.Pmax": 30,
    "grid.Umin": [-1.5, 0], "grid.Umax": [1.5, 1.6], "solver.dt": 0.005,
}

HERNQUIST_SMALL = {"init.type": "Hernquist", "init.M": 1, "init.R0": 1, "grid.Pres": 48,
                   "grid.Ures": [48, 24], "solver.dt": 0.0025}

UDF_NOLINK = dict(UDF_C1, **{"grid.Pres": 32, "grid.Ures": [32, 16], "solver.noLinking": 1})
UDF_REFLECT = dict(UDF_C1, **{"grid.Pres": 32, "grid.Ures": [32, 16],
                              "solver.kernelType": "reflective"})
UDF_BUFFER = dict(UDF_C1, **{"grid.Pres": 32, "grid.Ures": [32, 16],
                             "solver.kernelBufferSize": 2, "init.gap": [0.5, 1.0]})
# the v axis coarser than the j axis: slices with j >= Ures[0] weigh nothing in
# the reference's quadrature (slice21Iterator.hxx:86)
UDF_DJ_QUIRK = dict(UDF_C1, **{"grid.Pres": 24, "grid.Ures": [12, 20]})

# the default 200^3 grid of the UDF generator
UDF_C2 = dict(UDF_C1, **{"grid.Pres": 200, "grid.Ures": [200, 200]})

SMALL_CASES = {
    "udf_tiny": (UDF_TINY, 30),
    "udf_c1": (UDF_C1, 200),
    "pldf": (PLDF_SMALL, 60),
    "hernquist": (HERNQUIST_SMALL, 60),
    "udf_nolinking": (UDF_NOLINK, 40),
    "udf_reflective": (UDF_REFLECT, 40),
    "udf_buffer_gap": (UDF_BUFFER, 40),
    "udf_dj_quirk": (UDF_DJ_QUIRK, 40),
}

# a shape without a register line-solver instantiation: exercises the
# LDS-tiled spline kernels
UDF_ODD = dict(UDF_C1, **{"grid.Pres": 40, "grid.Ures": [36, 10]})
SMALL_CASES["udf_odd_shape"] = (UDF_ODD, 30)

# Osipkov-Merritt anisotropic Hernquist sphere (the Ra > 0 branch of the DF)
HERNQUIST_OM = dict(HERNQUIST_SMALL, **{"init.Ra": 1.5, "grid.Pres": 32, "grid.Ures": [32, 16]})
SMALL_CASES["hernquist_om"] = (HERNQUIST_OM, 40)

# Plummer sphere (the isotropic Plummer distribution function) on its default
# ranges (r 0.01..100 log, v +-1.45, j 0..3.24 quadratic)
PLUMMER_SMALL = {"init.type": "Plummer", "init.M": 1, "init.R0": 1, "grid.Pres": 48,
                 "grid.Ures": [40, 20], "solver.dt": 0.05}
SMALL_CASES["plummer"] = (PLUMMER_SMALL, 40)

# Two species (solver.nSpecies 2, one value of each UDF parameter per
# species) with the reflective kernel.  With the delayed kernel the
# reference's back cells hold one value (dummyCell21(R,U,1),
# vlasovSolver21.hxx:322) and its species-1 writes and reads fall outside
# them, so the reference is pinned on this case and the delayed one
# (UDF_2SPECIES_DELAYED) is checked against the oracle's well-defined
# semantics (every cell holds each species, back cells start at 0).
UDF_2SPECIES = dict(UDF_C1, **{"grid.Pres": 24, "grid.Ures": [20, 12], "solver.nSpecies": 2,
                               "init.M": [1, 0.5], "init.R0": [2, 1.5],
                               "init.alpha": [0.5, 0.8], "solver.kernelType": "reflective"})
SMALL_CASES["udf_2species_reflective"] = (UDF_2SPECIES, 30)
UDF_2SPECIES_DELAYED = dict(UDF_2SPECIES, **{"solver.kernelType": "delayed"})

# Strong kick: accelerations that move f by tens of v cells in one step, so
# that the target rows of the kick spread widely across neighbouring r nodes
# (physically extreme, numerically valid).
UDF_STRONG_KICK = dict(UDF_C1, **{"solver.G": 1000, "solver.dt": 0.02})
