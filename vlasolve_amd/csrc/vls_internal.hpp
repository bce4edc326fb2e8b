// Internal types shared by the device code and the C ABI of the spherical
// spline step.  Not part of the public interface (include/vlasolve_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vlasolve_hip.h"

namespace vls {

// Data-independent part of the natural cubic B-spline Thomas solve along one
// axis (localSpline::create_1D with BT_NATURAL at both ends,
// src/localSpline.hxx:953-1182).  The reference rebuilds these numbers in
// every call; they depend only on the knot spacing, so they are computed once
// on the host by the very same operations and read from a table.
struct ThomasAxis {
  int M;        // dim + 2 rows (padded line length)
  double c16;   // 1.0/6.0: sub-diagonal of interior rows
  double A01;   // first row after normalisation: coefficient of c1
  double A02;   //                                coefficient of c2
  double A03;   //                                right-hand side
  double hi2;   // last row, coefficient of c_{M-3}
  double p;     // last row, coefficient of c_{M-2} after eliminating c_{M-3}
  double q;     // last row, diagonal after eliminating c_{M-2}
  double rq;    // RN(1/q)
  // The pivot recurrence d_k = 4/6 - u_{k-1}/6, u_k = (1/6)/d_k does not
  // depend on the data or the knot spacing and reaches an exact fixed point:
  // from row K0 on, d_k = dS, RN(1/d_k) = rdS and u_k = uS.
  int K0;
  double dS, rdS, uS;
  const double *u;   // device [M]: super-diagonal after division, rows 1..M-2
  const double *d;   // device [M]: divisor of row k before normalisation
  const double *rd;  // device [M]: RN(1/d[k])
};

// Build the table on the host: u[k], d[k], rd[k] for k=1..M-2 and the constants.
void thomas_table(int dim, double hinv, ThomasAxis &ax, std::vector<double> &u,
                  std::vector<double> &d, std::vector<double> &rd);

// Per-kernel timing with HIP events on the stream of the launches.  Every
// kernel: an event after each launch (consecutive differences).  One kernel
// (`only` set): an event before and after each of its launches, so that the
// other kernels run undisturbed.
struct Timing {
  bool on = false;
  std::string only;
  std::vector<hipEvent_t> ev;
  std::vector<const char *> names;
  std::vector<double> ms;
};

// One run of an ensemble: every array of its state in HBM.
struct Run {
  int nR = 0, nU = 0, nJ = 0;
  int64_t nVal = 0;     // points per species
  int nS = 1;           // species: f and the delay-line values hold nS planes
  bool rollChains = true;  // the drift turns the delay lines (last species of a step)
  int64_t cs = 0;       // padded coefficient slice size (nU+2)*(nR+2)
  int stride = 0;       // padded row stride nR+2
  int TR = 32, LS = 0;  // r-sweep tile: rows per block, odd LDS row stride (>= TR)
  int TC = 32;          // v-sweep tile: columns per block
  bool streamR = false, streamV = false;  // streaming solvers for long lines
  // schedule choices (vls_autotune): -1 from the shape and the environment;
  // lineMode 0 / 1: LDS-tiled or streaming solvers / register lines (shapes
  // that have them); kickMode 0 / 1 / 2 / 3: two-kernel / fused kick / the
  // moment fused into the kick's column splines / that, and the kick's
  // evaluation folded into the drift's r-solve
  int lineMode = -1, kickMode = -1;
  vls_solver_params sp{};
  double start_v = 0, hinv_v = 0, Ulo = 0, Uhi = 0;

  // state
  double *f = nullptr;      // [nS][nJ][nU][nR]
  double *c = nullptr;      // [nJ][nU+2][nR+2] spline coefficients
  double *gsc = nullptr;    // fused kick: per-wave scratch for wide kicks
  bool fRows = false;       // f's buffer holds nJ*cs doubles: the kick folded into the r-solve
                            // leaves its rows there (f is rewritten by the drift)
  bool kickScratch = false; // fused kick: always take the scratch path
  // static feet, interleaved (xr, xu) per point: one 16-byte load each
  // (xr <= -1 marks points outside the grid or on delay lines)
  double *feet = nullptr;  // [nVal][2]
  std::vector<uint8_t> hflag;  // host copy of the VLS_FOOT_* kinds
  // moment
  double *P = nullptr;      // [nJ][nR] per-slice partials
  double *dM = nullptr, *Mx = nullptr, *acc = nullptr, *mTot = nullptr;
  unsigned *mCnt = nullptr;  // counters of the one-launch moment (zero between launches)
  double *rhoPhi = nullptr;  // [2][nR]: shell density and potential (vls_density_potential)
  // tables
  double *R_Val = nullptr, *U_Val = nullptr, *J_Val = nullptr;
  double *dR = nullptr, *dU = nullptr, *Ct1 = nullptr, *gfac = nullptr;
  ThomasAxis tr{}, tv{};
  // delay lines
  int64_t nChains = 0, nCells = 0;
  int64_t *ch_point = nullptr, *ch_off = nullptr;
  int32_t *ch_len = nullptr, *ch_head = nullptr;
  double *ch_xr = nullptr, *ch_xu = nullptr, *ch_vals = nullptr;
  uint8_t *ch_out = nullptr;
  // statistics
  double *st_part = nullptr;  // [7][nJ][nR]
  double *st_q = nullptr;     // [6][nJ] then dM [nR]
  double *st_out = nullptr;   // [14]: the species line, the statistics.txt line

  // adjoint: recorded run and work arrays (allocated by vls_record)
  struct Adjoint {
    int nSteps = 0, ckEvery = 0, nCk = 0, capCk = 0, capSeg = 0;
    int64_t cells = -1, chains = -1;  // delay-line sizes the chain arrays were made for
    double *ckF = nullptr;      // [capCk][nVal] f at steps 0, ckEvery, ...
    double *ckVals = nullptr;   // [capCk][nCells] delay-line values there
    int32_t *ckHead = nullptr;  // [capCk][nChains] ring heads there
    double *endF = nullptr, *endVals = nullptr;  // state after the recorded run
    int32_t *endHead = nullptr;
    double *segC = nullptr;     // [capSeg][coef] the kick coefficients of each state
    double *segAcc = nullptr;   // [capSeg][nR] the kick accelerations
    // the trailing segments kept by the recorded run itself (no recomputation
    // in the reverse sweep): steps keptFrom..nSteps-1
    int keptFrom = 0, capKeep = 0;
    double *keepC = nullptr, *keepAcc = nullptr;
    double *fb = nullptr, *fsb = nullptr, *cb = nullptr;
    double *abp = nullptr, *dMb = nullptr, *Mbar = nullptr, *ringb = nullptr;
    int64_t *tileOff = nullptr;  // transposed foot table of the drift, per coefficient tile
    uint32_t *tilePts = nullptr;
    double *chainXb = nullptr;   // [chains] adjoint of each delay line's back cell
    bool tilesValid = false;
    bool recorded = false;
  } adj;
  std::vector<double> hxr, hxu;  // host copies of the folded foot table
  // host copies of the delay lines' points and back cells
  std::vector<int64_t> hch_point;
  std::vector<double> hch_xr, hch_xu;
  std::vector<uint8_t> hch_out;
  std::vector<double> hU;        // host copy of U_Val

  bool have_f = false, have_drift = false, have_chains = false, stepped = false;
  std::vector<void *> allocs;
  std::vector<size_t> allocBytes;  // size of each entry of allocs
};

struct Context {
  int device = 0;
  int maxLds = 0;  // LDS bytes per workgroup (opt-in limit)
  hipStream_t stream = nullptr;
  std::vector<Run> runs;
  Timing timing;
  // ensembles: one stream per run, the step of all runs captured as a graph
  // (fork from `stream`, each run on its own stream, join back)
  std::vector<hipStream_t> runStreams;
  hipEvent_t forkEv = nullptr;
  std::vector<hipEvent_t> joinEv;
  hipGraphExec_t graph = nullptr;
  bool graphOn = true;
};

// kernel launchers (vls_kernels.hip)
void configure_tiles(Run &r, int maxLdsPerBlock);
// Times the schedules of the run's line solves and kick on its current f
// (f is left as it was) and keeps the fastest; cached per grid shape.
// Returns a one-line description of the choice.
std::string autotune(Run &r, hipStream_t s, int maxLdsPerBlock);
void launch_moment(Run &r, hipStream_t s);       // partials of f, then the scan
void launch_moment_scan(Run &r, hipStream_t s);  // scan of the partials in P
// The same into other buffers: the moment of f (the run's shape and species
// planes) with partials P [nJ][nR] (overwritten), dM, Mx, acc [nR], mTot [1]
struct MomentBufs {
  const double *f;
  double *P, *dM, *Mx, *acc, *mTot;
  unsigned *cnt;  // counters of the one-launch moment (zeroed), or null: three launches
};
// unsigned counters the one-launch moment needs
size_t moment_counters(const Run &r);
void launch_moment_into(Run &r, const MomentBufs &b, hipStream_t s);
void launch_moment_scan_into(Run &r, const MomentBufs &b, hipStream_t s);
void launch_spline(Run &r, const double *src, hipStream_t s, Timing *t);
// the kick of the plane r.f in column form: v-spline of each column of f
// into d (slice stride cs, row stride nR), then the evaluation
void launch_kick(Run &r, hipStream_t s, double *d);
// the kick's evaluation from the column splines d (layout of launch_kick)
void launch_kick_eval(Run &r, hipStream_t s, const double *d);
// the kick folded into the drift's r-solve: the shapes that have it (f's
// buffer is then allocated with nJ*cs doubles), whether this run takes it,
// and the launch (moment and the kick's column splines into c, then the
// kicked rows' r-solve into f's buffer); then launch_spline_v_timed (those
// rows into c) and the drift
bool kick_rsolve_shape(const Run &r);
bool kick_in_rsolve(const Run &r);
void launch_kick_rsolve(Run &r, hipStream_t s, Timing *t);
void launch_spline_v_timed(Run &r, hipStream_t s, Timing *t);
// (moment,) kick of the plane r.f; without the moment the acceleration of
// the last launch_moment is used; with dKeep the kick's column coefficients
// are also left there ([j][m][r], slice stride cs)
void launch_spline_kick(Run &r, hipStream_t s, Timing *t, bool moment = true,
                        double *dKeep = nullptr);
size_t kick_scratch_doubles(const Run &r);
void launch_drift(Run &r, hipStream_t s);
void mark_chains(Run &r, hipStream_t s);  // attach the delay lines to their points
void launch_stats(Run &r, hipStream_t s);
// rho[nR], phi[nR] (device) from the moment and enclosed mass of the last launch_moment
void launch_density_potential(Run &r, double *rho, double *phi, hipStream_t s);
void launch_density_potential_from(Run &r, const double *dM, const double *Mx, double *rho,
                                   double *phi, hipStream_t s);

void time_mark(Timing *t, hipStream_t s, const char *name);   // after a launch
void time_begin(Timing *t, hipStream_t s, const char *name);  // before a launch

// adjoint launchers (vls_adjoint.hip)
void launch_fill(double *p, int64_t n, double v, hipStream_t s);
// ab[2i] = a[i], ab[2i+1] = b[i]
void launch_adj_drift(Run &r, const double *fb, double *cb, double *ringb, const int32_t *head0,
                      int64_t stepOff, hipStream_t s);
// transposed spline solves, cb (destroyed) -> out
void launch_adj_spline(Run &r, double *cb, double *out, hipStream_t s);
void launch_adj_kick(Run &r, const double *fsb, const double *cK, const double *accK, double *cb,
                     double *abp, hipStream_t s);
void launch_adj_scan(Run &r, const double *abp, const double *MbarIn, double *dMb, hipStream_t s);
// the kick's transposed v-solve of g (left in cb by launch_adj_kick, destroyed;
// for long columns launch_adj_kick has already run its ascending sweep) into
// fb, plus the transposed moment quadrature of dMb when given
void launch_adj_kick_cols(Run &r, double *g, double *fb, const double *dMb, hipStream_t s);
void build_drift_tiles(Run &r, const std::vector<double> &xr, const std::vector<double> &xu,
                       std::vector<int64_t> &off, std::vector<uint32_t> &pts);
void launch_adj_mass(Run &r, const double *abp, const double *MbarIn, double *dMb, double *fb,
                     bool accumulate, hipStream_t s);

}  // namespace vls

#define VLS_HIP_CHECK(expr)                                                         \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess)                                                           \
      throw vls::Error(VLS_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

namespace vls {
struct Error {
  int code;
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};
}  // namespace vls
