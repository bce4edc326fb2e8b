/*
 * vlasolve_hip.h -- C ABI of the MI355X (gfx950) spherical Vlasov-Poisson
 * spline time step.
 *
 * Two layers, both plain C (pointers, sizes, status codes; no C++ or torch
 * types):
 *
 *  1. vls_*  : the hot path.  A device context owns the whole (r, v_r, j)
 *              distribution, its spline coefficients, the static foot tables
 *              and the delay-line rings in HBM, and advances them with HIP
 *              kernels.  This replaces the work of
 *              vlasovSolver21_spline::makeOneStep
 *              (src/vlasovSolver21_spline.hxx:680-801) and its helpers.
 *  2. vlsd_* : the solver surface.  A C++ host solver that keeps the
 *              reference's parameter keys, initial conditions, foot and
 *              delay-line construction, statistics.txt, NDfield snapshots and
 *              .vls restarts, and calls layer 1 for every step.  It replaces
 *              vlasovSolver<...>::init/run (src/vlasovSolver.hxx:61-190) as
 *              driven by main (src/vlasolve.cpp:20-55).
 *
 * Layout: f is double[nJ][nU][nR] with r fastest, as the reference's grid21
 * stores it (src/grid_base.hxx:285-310, src/slice21Iterator.hxx:63-64).
 * Every function returns 0 on success and a non-zero status otherwise; the
 * message is in vls_last_error() (thread-local).  The reference prints and
 * calls exit(-1) instead; the vlasolve_hip executable keeps that behaviour.
 */
#ifndef VLASOLVE_HIP_H
#define VLASOLVE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VLS_OK 0
#define VLS_ERR_ARG 1     /* invalid argument or size */
#define VLS_ERR_DEVICE 2  /* HIP runtime error or no device */
#define VLS_ERR_STATE 3   /* call out of order */
#define VLS_ERR_IO 4      /* file error */

/* Point flags of the drift step (vlasovSolver21_spline.hxx:792-795). */
#define VLS_FOOT_EVAL 0   /* f = spline at the foot */
#define VLS_FOOT_OUT 1    /* foot outside the grid: f = 0 (driftOut) */
#define VLS_FOOT_CHAIN 2  /* foot inside the kernel radius: delay line */

typedef struct vls_ctx vls_ctx;

/* Grid geometry of one run: node values and quadrature widths, as the
 * reference's grid21 / slice21Iterator provide them
 * (src/slice21Iterator.hxx:81-86), and the uniform knot spacing of the
 * spline in rescaled coordinates (src/localSpline.hxx:60-93). */
typedef struct vls_geometry {
  int32_t nR;             /* r nodes  = Pres+1   (VERTEX)            */
  int32_t nU;             /* v nodes  = Ures[0]+1 (VERTEX)            */
  int32_t nJ;             /* j slices = Ures[1]   (CELL)              */
  const double *R_Val;    /* [nR] */
  const double *U_Val;    /* [nU] */
  const double *J_Val;    /* [nJ] */
  const double *dR;       /* [nR] R_V[r+1]-R_V[r], 0 at the last node */
  const double *dU;       /* [nU] */
  const double *dJ;       /* [nJ] get_dJ, including its v-size test  */
  double spline_start[2]; /* rescaled first knot (r, v)               */
  double spline_hinv[2];  /* 1/knot spacing (r, v)                    */
} vls_geometry;

typedef struct vls_solver_params {
  double G;               /* solver.G  */
  double dt;              /* solver.dt */
  int32_t no_linking;     /* solver.noLinking: D(dt/2) K(dt) D(dt/2) */
  int32_t kernel_type;    /* 0 delayed, 1 reflective (solver.kernelType) */
  int32_t n_species;      /* solver.nSpecies, 1..5 (0 means 1): the fields of
                             gridBase (nFields, src/grid_base.hxx:78) */
} vls_solver_params;

/* ---------------------------------------------------------------- layer 1 */

/* Create a device context holding nRuns independent runs (an ensemble; the
 * runs may have different grids).  g and sp are arrays of nRuns entries.
 * Replaces the solver's setup (vlasovSolver21_spline::setup,
 * src/vlasovSolver21_spline.hxx:296-374) as far as device state goes. */
int vls_create(const vls_geometry *g, const vls_solver_params *sp, int nRuns, vls_ctx **out);
void vls_destroy(vls_ctx *ctx);
const char *vls_last_error(void);

/* Distribution function of one run, host layout [nJ][nU][nR], with several
 * species [nJ][nU][nR][nSpecies] (species fastest, as the reference
 * interleaves them).  Replaces the grid21 data the reference's step reads
 * and writes (gridBase::getDataPtr, src/grid_base.hxx:383-385). */
int vls_upload_f(vls_ctx *ctx, int run, const double *f);
int vls_download_f(vls_ctx *ctx, int run, double *f);

/* Static drift feet of every point.  xr, xu are the foot coordinates in
 * spline units, (rescale(p) - start) * hinv, whose integer and fractional
 * parts are the reference's driftR_I/driftR_F and driftU_I/driftU_F
 * (decomposeDisplacement, src/vlasovSolver21_spline.hxx:76-100); flag is a
 * VLS_FOOT_* value (driftR < kernelRadius, then driftOut;
 * src/vlasovSolver21.hxx:140-143).  Arrays of nJ*nU*nR entries. */
int vls_set_drift(vls_ctx *ctx, int run, const double *xr, const double *xu,
                  const uint8_t *flag);

/* Delay lines of the delayed central kernel (sphericalKernel,
 * src/sphericalKernel.hxx:48-54, cellChain::readRoll src/cellChain.hxx:49-63).
 * Chain c feeds grid point point[c]; it holds len[c] >= 2 values in a ring
 * stored at vals[off[c] .. off[c]+len[c]), off = exclusive prefix sum of len,
 * whose front is at ring index head[c].  back_xr/back_xu are the spline-unit
 * coordinates of the chain's last cell and back_out[c] is 1 when that cell's
 * v lies outside (U_min, U_max), where the reference stores 0
 * (src/vlasovSolver21_spline.hxx:776-786).  With several species every cell
 * holds nSpecies values, vals[cell * nSpecies + s]; the species share the
 * ring, which turns once per step. */
int vls_set_chains(vls_ctx *ctx, int run, int64_t nChains, const int64_t *point,
                   const int32_t *len, const int32_t *head, const double *back_xr,
                   const double *back_xu, const uint8_t *back_out, const double *vals);
/* Current ring state (for restarts and checkpoints). */
int vls_get_chains(vls_ctx *ctx, int run, int32_t *head, double *vals);

/* Advance every run by nSteps steps (vlasovSolver21_spline::makeOneStep,
 * src/vlasovSolver21_spline.hxx:680-801), stream-ordered on the context's
 * stream.  Does not wait for completion. */
int vls_step(vls_ctx *ctx, int nSteps);
int vls_synchronize(vls_ctx *ctx);
/* Ensembles (nRuns > 1): by default vls_step captures one step of all runs
 * as a HIP graph, each run on a stream of its own so that runs of different
 * sizes execute concurrently, and replays it once per step (the graph is
 * rebuilt after vls_set_chains).  on = 0 steps the runs one after the other
 * on the context stream.  No counterpart in the reference, which runs one
 * model per process. */
int vls_set_graph(vls_ctx *ctx, int on);

/* Moments used by the last step's kick: dM[nR] (quadrature21::integrate1D,
 * src/quadrature21.hxx:88-142), accel[nR] = M(<r) G/r^2 dt
 * (src/vlasovSolver21_spline.hxx:736-746) and the total mass mTot. */
int vls_moments(vls_ctx *ctx, int run, double *dM, double *accel, double *mTot);

/* Statistics of the current f (vlasovSolver21::statistics,
 * src/vlasovSolver21.hxx:627-711): out = {Mass, E, Ek, Ep, S, L1, L2}, the
 * line of statistics.txt.  With several species this is the line of the
 * files statistics_S%d.txt, the same in each as the reference computes it:
 * Mass and Ep of all species, Ek, S, L1 and L2 of species 0. */
int vls_stats(vls_ctx *ctx, int run, double out[7]);
/* Both lines: the species files' (line) and statistics.txt's (total: the
 * species line added once per species, with S and L2 of the summed species;
 * equal to line with one species).  Either pointer may be NULL. */
int vls_stats_total(vls_ctx *ctx, int run, double line[7], double total[7]);

/* Coefficients of the last drift spline of slice j, (nU+2)*(nR+2) doubles,
 * the padded array of localSpline<2> (src/localSpline.hxx:60-150). */
int vls_download_coef(vls_ctx *ctx, int run, int j, double *coef);

/* Per-kernel timing of the last vls_step call on the context's stream
 * (HIP events).  names/ms are arrays of at least *n entries; on return *n is
 * the number of kernels recorded.  Enabled with vls_set_timing(ctx, 1). */
int vls_set_timing(vls_ctx *ctx, int on);
int vls_kernel_times(vls_ctx *ctx, const char **names, double *ms, int *n);
/* Schedule tuner: times the run's line-solve schedules on its current data
 * and keeps the fastest -- register lines (shapes that have them), or the
 * LDS-tiled solvers with the tile size chosen from a set of candidates, or
 * the streaming solvers, for each direction -- then the fused or two-kernel
 * kick on a copy of f.  run = -1 tunes every run of the context.  Every
 * schedule runs the same operations in the same order: the choice changes
 * the time of a step, never its bits.  Choices are cached per grid shape for
 * the life of the process; desc (may be NULL) receives one line per run
 * describing them with the measured times.  No counterpart in the reference,
 * which runs one model per process with a fixed schedule. */
int vls_autotune(vls_ctx *ctx, int run, char *desc, int descLen);

/* Restrict the timing to one kernel ("kick_line", "drift", "spline_r",
 * "spline_v", "moment", "kick"): an event before and after each of its
 * launches only, so that the other kernels run as without timing; NULL
 * times every kernel again.  No counterpart in the reference. */
int vls_set_timing_kernel(vls_ctx *ctx, const char *kernel);

/* Host-side helpers for tests (no device needed).  vls_thomas_table returns
 * the data-independent factors of the natural spline solve of one axis
 * (u, d, rd of dim+2 entries; consts = {c16, A01, A02, A03, hi2, p, q, rq});
 * vls_check_division counts the numerators for which the known-divisor
 * division of the line solves differs from IEEE division. */
int vls_thomas_table(int dim, double hinv, double *u, double *d, double *rd, double consts[8]);
/* One-line description of the current device (CUs, clock, LDS, L2). */
int vls_device_info(char *buf, int len);
/* One device-to-device copy of `bytes` bytes, to calibrate memory counters. */
int vls_copy_probe(size_t bytes);
int vls_check_division(const double *b, int nb, int64_t samples, uint64_t seed,
                       int64_t *mismatches);

/* Sizes of one run: {nR, nU, nJ, number of delay lines, delay-line cells,
 * species}. */
int vls_run_info(vls_ctx *ctx, int run, int64_t out[6]);
/* Device memory the run holds: number of allocations and their bytes (the
 * adjoint's checkpoints included).  No counterpart in the reference. */
int vls_memory(vls_ctx *ctx, int run, int64_t *count, int64_t *bytes);
/* Overwrite the delay-line state (the inverse of vls_get_chains): ring heads
 * and ring values in the layout of vls_set_chains.  Either may be NULL. */
int vls_put_chains(vls_ctx *ctx, int run, const int32_t *head, const double *vals);
/* Enclosed-mass profile of the current f: M[r] = M(<R_r), the exclusive
 * radial sum of the moment (quadrature21::integrate1D and
 * gridHandler::accumulate, src/quadrature21.hxx:88-142,
 * src/grid_handler.hxx:350-407).  Refreshes the moments of vls_moments. */
int vls_mass_profile(vls_ctx *ctx, int run, double *M);

/* Shell density and gravitational potential of the current f, from its
 * moment dM (as vls_mass_profile computes it) and the enclosed mass
 * Mx[r] = M(<R_r): rho[k] = dM[k] / ((4 pi/3)(R[k+1]^3 - R[k]^3)) for
 * k < nR-1, rho[nR-1] = 0; phi[i] = -G (Mx[i]/R[i] + sum_{i<=k<nR-1} dM[k]/rbar[k])
 * with rbar[k] = (R[k]+R[k+1])/2, the sum taken from the outermost shell in.
 * The reference has no such output: its yorick density() integrates
 * snapshots over (u, j) (yorick/vlasolve.i:39-64) and potentialEnergy pairs
 * each shell with the mass below it (src/vlasovSolver21.hxx:627-642); these
 * are the same physics from the step's own quadrature.  Either pointer may
 * be NULL.  Refreshes the moments of vls_moments. */
int vls_density_potential(vls_ctx *ctx, int run, double *rho, double *phi);

/* Allocation-free variant of vls_density_potential.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  f_dev is the caller's device array of the run's shape,
 * [nS][nJ][nU][nR] (one plane per species); rho_dev, phi_dev and M_dev are
 * device arrays of nR entries (any may be NULL, not all) that receive the
 * density and potential of vls_density_potential and the enclosed mass
 * M(<R_r) of vls_mass_profile, bit for bit.  The moment's partial sums go to
 * a static device array of VLS_NOALLOC_SCRATCH_DOUBLES doubles per GPU (2 MiB,
 * held until the process ends): a run with nJ*nR + 5*nR + 1 above it is
 * rejected with VLS_ERR_ARG.  The call allocates nothing, leaves the
 * context's state (f, the moments of vls_moments) untouched, launches on
 * `stream` (a hipStream_t, or NULL for the context's stream) and returns
 * without waiting.  The context's device must be current.  Every call on one
 * GPU shares that array: calls that overlap on one GPU race. */
#define VLS_NOALLOC_SCRATCH_DOUBLES (1 << 18)
int vls_density_potential_noalloc(vls_ctx *ctx, int run, const double *f_dev, double *rho_dev,
                                  double *phi_dev, double *M_dev, void *stream);

/* Discrete adjoint of the step.  No counterpart in the reference (it has no
 * derivative); it differentiates exactly the discrete map that vls_step
 * applies (linking steps).
 *
 * vls_record advances run `run` by nSteps steps like vls_step, keeping f and
 * the delay-line state every ckEvery steps in HBM (checkpoints).
 * vls_adjoint then returns the gradient of a loss L(f_N) with respect to the
 * f and the delay-line values of the first recorded step, given dL_dM
 * ([nR], the loss's derivative in the mass profile of f_N as returned by
 * vls_mass_profile; may be NULL) and dL_df ([nJ][nU][nR], its direct
 * derivative in f_N; may be NULL).  The recorded run also keeps, for every
 * step of the trailing checkpoint segments that fit in seven eighths of the
 * free HBM (all of them for grids like 129x129x64; the environment variable
 * VLS_ADJ_KEEP_STEPS caps the kept steps), the kick's column splines and
 * acceleration; vls_adjoint recomputes the other segments from their
 * checkpoints.  The context ends in the state after the recorded run.  grad_vals0 (may be NULL) has the ring layout of
 * vls_set_chains at the first recorded step.  The drift's transpose adds
 * the points' contributions into tiles of coefficients with LDS atomics and
 * the delay lines' back cells add theirs with global floating-point atomics,
 * so the last bits of the gradient may vary from call to call.  vls_record
 * builds, on first use after vls_set_drift, the transposed foot table on the
 * host (about 1.4 entries of 4 bytes per evaluated point, plus the feet
 * interleaved, 16 bytes per point) and holds, per step of a checkpoint
 * segment, room for the kick's column splines and acceleration in HBM. */
int vls_record(vls_ctx *ctx, int run, int nSteps, int ckEvery);
int vls_adjoint(vls_ctx *ctx, int run, const double *dL_dM, const double *dL_df,
                double *grad_f0, double *grad_vals0);

/* The adjoint of every run of an ensemble at once (runs of different grids
 * allowed): vls_record_batch records all runs for the same nSteps and
 * ckEvery, each run on a stream of its own so that they execute
 * concurrently, with the kept steps sharing seven eighths of the free HBM in
 * proportion to the runs' sizes; vls_adjoint_batch then runs every reverse
 * sweep concurrently the same way.  The arrays hold one pointer per run
 * (dL_dM, dL_df, grad_vals0 and their entries may be NULL, as for
 * vls_adjoint).  Each run's gradient is what vls_record and vls_adjoint give
 * for that run alone, up to the rounding of the floating-point atomics.  No
 * counterpart in the reference. */
int vls_record_batch(vls_ctx *ctx, int nSteps, int ckEvery);
int vls_adjoint_batch(vls_ctx *ctx, const double *const *dL_dM, const double *const *dL_df,
                      double *const *grad_f0, double *const *grad_vals0);

/* ---------------------------------------------------------------- layer 2 */

typedef struct vlsd_solver vlsd_solver;

/* Build a solver from reference-style arguments ("-init.type UDF -grid.Pres
 * 64 ..." and an optional parameter file as first argument; paramsParser,
 * src/paramsParser.hxx:195-218), run init or reload, and upload to the GPU.
 * argv[0] is the program name, as for main (src/vlasolve.cpp:20-48). */
int vlsd_create(int argc, const char *const *argv, vlsd_solver **out);
void vlsd_destroy(vlsd_solver *s);
/* geometry: {nR, nU, nJ}; time, step count */
int vlsd_info(vlsd_solver *s, int64_t dims[3], double *time, int64_t *stepCount);
int vlsd_species(vlsd_solver *s);                   /* solver.nSpecies */
int vlsd_step(vlsd_solver *s, int nSteps);          /* makeOneStep x n, time += dt */
int vlsd_get_f(vlsd_solver *s, double *f);          /* layout of vls_download_f */
int vlsd_stats(vlsd_solver *s, double out[7]);      /* line of vls_stats */
int vlsd_stats_total(vlsd_solver *s, double out[7]); /* statistics.txt line (vls_stats_total) */
int vlsd_moments(vlsd_solver *s, double *dM, double *accel, double *mTot);
int vlsd_run(vlsd_solver *s);                       /* vlasovSolver::run, src/vlasovSolver.hxx:134-190 */
int vlsd_write_restart(vlsd_solver *s, const char *fname_no_ext);   /* .vls */
int vlsd_snapshot(vlsd_solver *s);                  /* snapshot_%4.4f.ND */
vls_ctx *vlsd_context(vlsd_solver *s);
const char *vlsd_last_error(void);

/* An ensemble: nRuns independent runs, each from its own reference-style
 * argument list (argc[k], argv[k]; initial conditions, not restarts), set up
 * on the host like vlsd_create and loaded into one device context whose runs
 * may have different grids.  The context (vlsd_ensemble_context) is stepped
 * with vls_step; it belongs to the ensemble. */
typedef struct vlsd_ensemble vlsd_ensemble;
int vlsd_ensemble_create(int nRuns, const int *argc, const char *const *const *argv,
                         vlsd_ensemble **out);
void vlsd_ensemble_destroy(vlsd_ensemble *e);
vls_ctx *vlsd_ensemble_context(vlsd_ensemble *e);

/* Host set-up only (no device): the same initial conditions, renormalisation,
 * static feet and delay lines as vlsd_create, for inspection and for
 * machines without a GPU.  dims = {nR, nU, nJ}; feet are the free-streaming
 * foot positions and VLS_FOOT_* flags of every point; chains as in
 * vls_set_chains, cells front first with their positions and values
 * (nSpecies values per cell). */
typedef struct vlsd_host vlsd_host;
int vlsd_host_create(int argc, const char *const *argv, vlsd_host **out);
void vlsd_host_destroy(vlsd_host *h);
int vlsd_host_info(vlsd_host *h, int64_t dims[3], int64_t *nChains, int64_t *nCells,
                   double *kernelRadius);
int vlsd_host_species(vlsd_host *h);
int vlsd_host_f(vlsd_host *h, double *f);  /* layout of vls_upload_f */
int vlsd_host_feet(vlsd_host *h, double *driftR, double *driftU, uint8_t *flag);
int vlsd_host_chains(vlsd_host *h, int64_t *point, int32_t *len, double *cellR, double *cellU,
                     double *val);
/* Derivative of the initial state in an initial-condition parameter of
 * species `species`, the grid held fixed: df (layout of vlsd_host_f) and
 * dvals (layout of vlsd_host_chains' val) are d/dparam of f0 and of the
 * delay-line values after the renormalisation to the species mass
 * (vlasovSolver21::renormalize, src/vlasovSolver21.hxx:468-524), written out
 * by hand from the generators' formulas: UDF "M", "R0", "alpha",
 * "smoothingRadius" (initGen_UniformDensitySphere.hxx:43-76,158-173); PLDF
 * "M", "eta" (with R0 = 1/eta following it unless init.R0 is given),
 * "alpha", "R0", "smoothingRadius" (initGen_PowerLawDensitySphere.hxx:44-83,
 * 181-189).  With the gradients of vls_adjoint (grad_f0, and grad_vals0 at a
 * first recorded step whose rings have head 0, as a new run's do) the
 * gradient in the parameter is <grad_f0, df> + <grad_vals0, dvals>.  No
 * counterpart in the reference. */
int vlsd_host_dparam(vlsd_host *h, const char *param, int species, double *df, double *dvals);

#ifdef __cplusplus
}
#endif
#endif
