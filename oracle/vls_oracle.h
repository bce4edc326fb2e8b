/*
 * vls_oracle.h -- CPU restatement of vlasolve's spherically symmetric
 * semi-Lagrangian spline time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product (vlasolve_amd/) never links, loads or calls anything here.
 *
 * Everything is FP64 and follows the reference's operation order so that,
 * built with -ffp-contract=off on x86-64, results are bit-identical to the
 * reference built from its own sources (see vls_oracle.c for the per-function
 * reference citations and pins).
 */
#ifndef VLS_ORACLE_H
#define VLS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* scale types, same codes as the reference (scale.hxx:17-19) */
enum { ORC_LINEAR = 0, ORC_LOG = 1, ORC_QUADRATIC = 2, ORC_CUBIC = 3, ORC_QUARTIC = 4 };
/* initial-condition generators (initGen_all.hxx:19-25) */
enum { ORC_UDF = 0, ORC_PLDF = 1, ORC_HERNQUIST = 2, ORC_PLUMMER = 3 };
/* kernel types (vlasovSolver21.hxx:66-68) */
enum { ORC_DELAYED = 0, ORC_REFLECTIVE = 1 };
/* initGen::MAX_SPECIES (initGen.hxx:21) */
enum { ORC_MAX_SPECIES = 5 };

/* Fully resolved run parameters (defaults already applied by the caller). */
typedef struct orc_params {
  /* grid (grid_base.hxx:59-84; gridHandler overrides grid_handler.hxx:221-234) */
  double Pmin, Pmax;
  int Pres, Pscale;
  double Umin[2], Umax[2];
  int Ures[2], Uscale[2];
  /* solver (vlasovSolver.hxx:67-78, vlasovSolver21.hxx:537-539) */
  double G, dt;
  int noLinking;
  int kernelType;
  int kernelBufferSize;
  /* initial conditions */
  int initType;
  double M, R0, alpha, smoothingRadius, gap0, gap1; /* UDF */
  double eta, n;                                    /* PLDF (R0, alpha, M, smoothingRadius shared) */
  double Ra;                                        /* Hernquist */
  /* species (solver.nSpecies, vlasovSolver.hxx:72): UDF and PLDF take one
   * value of each parameter per species (species 0's equal the scalars
   * above); Hernquist and Plummer have one species */
  int nSpecies;
  double spM[ORC_MAX_SPECIES], spR0[ORC_MAX_SPECIES], spAlpha[ORC_MAX_SPECIES];
  double spSR[ORC_MAX_SPECIES], spGap0[ORC_MAX_SPECIES], spGap1[ORC_MAX_SPECIES];
  double spEta[ORC_MAX_SPECIES], spN[ORC_MAX_SPECIES];
  /* Evaluation order.  nthreads >= 1: the reference's, with that many
   * OpenMP threads under a static schedule (quadrature21.hxx:96-139).
   * nthreads == 0: the device order, which the HIP kernels use: the moment
   * per slice and r as serial sums over groups of 16 rows u, combined four
   * to a block of 64 rows and then over the blocks, then over j in chunks of
   * 20 slices (serial within a chunk, then over the chunks); and the kick in
   * column form (the natural v-spline of each column of f, evaluated with
   * weights B_k/6), equal to the reference's 2D evaluation on r nodes to
   * rounding. */
  int nthreads;
} orc_params;

typedef struct orc_state orc_state;

orc_state *orc_create(const orc_params *p);           /* init + renormalize */
void orc_destroy(orc_state *s);
const char *orc_last_error(void);

void orc_dims(const orc_state *s, long dims[3]);      /* nR, nU, nJ */
long orc_nval(const orc_state *s);                    /* points (per species) */
int orc_nspecies(const orc_state *s);
/* f in the reference's layout [nJ][nU][nR][nSpecies] (species fastest) */
void orc_get_f(const orc_state *s, double *out);
void orc_set_f(orc_state *s, const double *in);
void orc_get_axes(const orc_state *s, double *R_Val, double *U_Val, double *J_Val,
                  double *J_V);
double orc_kernel_radius(const orc_state *s);

/* one makeOneStep each (vlasovSolver21_spline.hxx:680-801) */
void orc_step(orc_state *s, int nsteps);
/* dM[nR] (integrate1D), accel[nR] (M_excl*G/R^2*dt) and mTot of the last step */
void orc_last_moments(const orc_state *s, double *dM, double *accel, double *mTot);
/* moment of the current f without stepping: dM[nR], Mexcl[nR], mTot */
void orc_moments(orc_state *s, double *dM, double *Mexcl, double *mTot);
/* statistics (vlasovSolver21.hxx:644-711): M, E, Ek, Ep, S, L1, L2 as the
 * species files statistics_S%d.txt hold them (with several species M and
 * Ep count every species, Ek, S, L1 and L2 species 0 only, in every file);
 * total: the line of statistics.txt (the same line with one species) */
void orc_stats(orc_state *s, double out[7]);
void orc_stats_total(orc_state *s, double out[7]);

/* feet / chain introspection (for fixtures and device-table checks) */
void orc_get_feet(const orc_state *s, double *driftR, double *driftU, uint8_t *flag);
long orc_num_chains(const orc_state *s);
long orc_chain_cells(const orc_state *s);
/* per chain: point index, slice j, offset into cells, length; per cell R, U, and
 * nSpecies values (list order, front first; cellVal[cell * nSpecies + s]) */
void orc_get_chains(const orc_state *s, int64_t *point, int64_t *off, int32_t *len,
                    double *cellR, double *cellU, double *cellVal);

/* spline unit access: build the natural 2D spline of one slice (nU x nR, r
 * fastest) into padded coefficients (nU+2)x(nR+2), and evaluate */
void orc_spline_build(const orc_state *s, const double *slice, double *coef);
double orc_spline_evalC(const orc_state *s, const double *coef, int Nx, double x, int Ny,
                        double y);
double orc_spline_evalNxPy(const orc_state *s, const double *coef, int Nx, double yy);
double orc_spline_eval(const orc_state *s, const double *coef, double xx, double yy);

/* free-streaming backward foot (vlasovSolver21.hxx:328-408) */
void orc_advance_drift(double R, double U, double J, double dt, double R0, double *newR,
                       double *newU);
/* scale generation (scale.hxx:114-151); returns the number of values */
long orc_gen_scale(double start, double stop, long N, int sct, int cell, double *out);

/* Tangent-linear model (directional derivative of the step; linking steps
 * and a linear v scale).  dvals: tangents of the delay-line cells in the
 * order of orc_get_chains, or NULL for zero.  orc_tl_step advances the state
 * and its tangent together; orc_tl_mass returns the exclusive-scan mass
 * profile M(<R_r) of the current f and its tangent. */
int orc_tl_init(orc_state *s, const double *df, const double *dvals);
/* the delay-line cell values, in the layout of orc_get_chains */
void orc_set_chain_vals(orc_state *s, const double *vals);
int orc_tl_step(orc_state *s, int nsteps);
void orc_tl_get(const orc_state *s, double *df);
void orc_tl_mass(orc_state *s, double *M, double *dM);

#ifdef __cplusplus
}
#endif
#endif
