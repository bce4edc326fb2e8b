// C ABI of the device context (include/vlasolve_hip.h, layer 1).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vls_div.hpp"
#include "vls_internal.hpp"

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

template <class F>
int guarded(F &&fn) {
  try {
    fn();
    return VLS_OK;
  } catch (const vls::Error &e) {
    return fail(e.code, e.msg);
  } catch (const std::exception &e) {
    return fail(VLS_ERR_ARG, e.what());
  }
}

template <class T>
T *dalloc(vls::Run &r, size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  VLS_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
  r.allocs.push_back(p);
  r.allocBytes.push_back(n * sizeof(T));
  return static_cast<T *>(p);
}

// Release an allocation of dalloc before the run ends (and null the pointer).
template <class T>
void dfree(vls::Run &r, T *&p) {
  if (!p) return;
  for (size_t i = 0; i < r.allocs.size(); i++)
    if (r.allocs[i] == (void *)p) {
      r.allocs.erase(r.allocs.begin() + i);
      r.allocBytes.erase(r.allocBytes.begin() + i);
      break;
    }
  VLS_HIP_CHECK(hipFree((void *)p));
  p = nullptr;
}

size_t kCoefPad(const vls::Run &r) { return 2 * (size_t)(r.nR + 2) + 8; }

// The reference interleaves the species of a point (arr[i * nFields + s],
// src/grid_base.hxx:385); the device keeps one plane per species.
std::vector<double> to_planes(const double *in, int64_t n, int nS) {
  std::vector<double> out((size_t)n * nS);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < nS; k++) out[(size_t)k * n + i] = in[i * nS + k];
  return out;
}
void from_planes(const double *in, double *out, int64_t n, int nS) {
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < nS; k++) out[i * nS + k] = in[(size_t)k * n + i];
}

// A foot can be evaluated when its integer parts lie in [0, dim-1] on both
// axes (dim-1 only on the last knot).
bool foot_in_domain(double xr, double xu, const vls::Run &r) {
  double a = std::trunc(xr), b = std::trunc(xu);
  return a >= 0 && a <= r.nR - 1 && b >= 0 && b <= r.nU - 1;
}

template <class T>
void h2d(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) VLS_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}

template <class T>
void d2d(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) VLS_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}

vls::Run &run_of(vls_ctx *ctx, int run);

// One step of a run on stream s (vlasovSolver21_spline::makeOneStep,
// src/vlasovSolver21_spline.hxx:680-801): with linking, kick (moment, spline,
// evaluation) then drift (spline, evaluation with the delay lines); without,
// half drift, kick, half drift.
// Points r.f and r.ch_vals at species k (and lets only the last species turn
// the delay lines) for the life of the object.
struct SpeciesPlane {
  vls::Run &r;
  double *f0, *v0;
  SpeciesPlane(vls::Run &run, int k) : r(run), f0(run.f), v0(run.ch_vals) {
    r.f = f0 + (size_t)k * r.nVal;
    if (v0) r.ch_vals = v0 + (size_t)k * r.nCells;
    r.rollChains = k == r.nS - 1;
  }
  ~SpeciesPlane() {
    r.f = f0;
    r.ch_vals = v0;
    r.rollChains = true;
  }
};

// makeOneStep with several species (src/vlasovSolver21_spline.hxx:680-801):
// the moment of their sum, then the kick of each species with the shared
// acceleration, then the drift of each.
void step_species(vls::Run &r, hipStream_t s, vls::Timing *t) {
  if (r.sp.no_linking)
    for (int k = 0; k < r.nS; k++) {
      SpeciesPlane sp(r, k);
      vls::launch_spline(r, r.f, s, t);
      vls::time_begin(t, s, "drift");
      vls::launch_drift(r, s);
      vls::time_mark(t, s, "drift");
    }
  vls::launch_moment(r, s);
  vls::time_mark(t, s, "moment");
  for (int k = 0; k < r.nS; k++) {
    SpeciesPlane sp(r, k);
    vls::launch_spline_kick(r, s, t, false);
  }
  for (int k = 0; k < r.nS; k++) {
    SpeciesPlane sp(r, k);
    vls::launch_spline(r, r.f, s, t);
    vls::time_begin(t, s, "drift");
    vls::launch_drift(r, s);
    vls::time_mark(t, s, "drift");
  }
  r.stepped = true;
}

void step_once(vls::Run &r, hipStream_t s, vls::Timing *t) {
  if (r.nS > 1) return step_species(r, s, t);
  if (r.sp.no_linking) {
    vls::launch_spline(r, r.f, s, t);
    vls::time_begin(t, s, "drift");
    vls::launch_drift(r, s);
    vls::time_mark(t, s, "drift");
  }
  if (vls::kick_in_rsolve(r)) {
    vls::launch_kick_rsolve(r, s, t);
    vls::launch_spline_v_timed(r, s, t);
  } else {
    vls::launch_spline_kick(r, s, t);
    vls::launch_spline(r, r.f, s, t);
  }
  vls::time_begin(t, s, "drift");
  vls::launch_drift(r, s);
  vls::time_mark(t, s, "drift");
  r.stepped = true;
}

// One step that also keeps what the adjoint of the step needs: the kick's
// acceleration (accOut, [nR]) and the v-spline coefficients of the columns
// of f (cK, [j][m][r] with slice stride cs), which the kick leaves there.
// Without advance the step stops after the kick (the state after it is not
// needed: the last step of a recomputed segment).
void step_keep(vls::Run &r, hipStream_t s, double *cK, double *accOut, bool advance = true) {
  // the moment writes the acceleration straight into accOut (the step's
  // kernels take r.acc when launched); r.acc is left as it was
  struct AccSwap {
    vls::Run &r;
    double *own;
    ~AccSwap() { r.acc = own; }
  } swap{r, r.acc};
  r.acc = accOut;
  vls::launch_spline_kick(r, s, nullptr, true, cK);
  r.acc = swap.own;
  if (!advance) return;
  vls::launch_spline(r, r.f, s, nullptr);
  vls::launch_drift(r, s);
}

// Capture one step of every run of an ensemble: each run's kernels go to a
// stream of their own, forked from and joined back into the context stream,
// so that small runs fill the GPU together; replayed once per step.
// A stream (default priority) and a join event per run.
void ensure_run_streams(vls::Context &C) {
  const int n = (int)C.runs.size();
  for (int k = (int)C.runStreams.size(); k < n; k++) {
    hipStream_t st;
    VLS_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    C.runStreams.push_back(st);
    hipEvent_t ev;
    VLS_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    C.joinEv.push_back(ev);
  }
}

void build_ensemble_graph(vls::Context &C) {
  const int n = (int)C.runs.size();
  ensure_run_streams(C);
  if (!C.forkEv) VLS_HIP_CHECK(hipEventCreateWithFlags(&C.forkEv, hipEventDisableTiming));
  hipGraph_t g = nullptr;
  VLS_HIP_CHECK(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
  try {
    VLS_HIP_CHECK(hipEventRecord(C.forkEv, C.stream));
    for (int k = 0; k < n; k++) {
      VLS_HIP_CHECK(hipStreamWaitEvent(C.runStreams[k], C.forkEv, 0));
      step_once(C.runs[k], C.runStreams[k], nullptr);
      VLS_HIP_CHECK(hipEventRecord(C.joinEv[k], C.runStreams[k]));
      VLS_HIP_CHECK(hipStreamWaitEvent(C.stream, C.joinEv[k], 0));
    }
  } catch (...) {
    (void)hipStreamEndCapture(C.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  VLS_HIP_CHECK(hipStreamEndCapture(C.stream, &g));
  hipError_t e = hipGraphInstantiate(&C.graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    C.graph = nullptr;
    throw vls::Error(VLS_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
}

void drop_graph(vls::Context &C) {
  if (C.graph) (void)hipGraphExecDestroy(C.graph);
  C.graph = nullptr;
}
}  // namespace

struct vls_ctx {
  vls::Context c;
};

// Scratch of vls_density_potential_noalloc: the moment's partials and radial
// sums, one array per GPU, held until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: kernels receive
// the device address that hipGetSymbolAddress returns for the current device
// (scratch_dev, resolved once per device by vls_create).
static __device__ __align__(256) double scratch[VLS_NOALLOC_SCRATCH_DOUBLES];

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_dev;  // index: device ordinal

// With `device` current: its address of `scratch`, resolved on first use.
double *resolve_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize(device + 1, nullptr);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    VLS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = static_cast<double *>(p);
  }
  return g_scratch_dev[device];
}

double *scratch_dev(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device < (int)g_scratch_dev.size() ? g_scratch_dev[device] : nullptr;
}
}  // namespace

namespace {
vls::Run &run_of(vls_ctx *ctx, int run) {
  if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
  if (run < 0 || run >= (int)ctx->c.runs.size())
    throw vls::Error(VLS_ERR_ARG, "run index out of range");
  return ctx->c.runs[run];
}
}  // namespace

namespace vls {

// Runs the elimination of create_1D (src/localSpline.hxx:968-1153) on the
// data-independent coefficients and records what the data path needs.
void thomas_table(int dim, double hinv, ThomasAxis &ax, std::vector<double> &u,
                  std::vector<double> &d, std::vector<double> &rd) {
  const int M = dim + 2;
  std::vector<double> A(4 * (size_t)M);
  const double hi2 = hinv * hinv;
  A[0] = hi2; A[1] = -2.0 * hi2; A[2] = hi2; A[3] = 0;
  for (int i = 1; i < M - 1; i++) {
    A[4 * i + 0] = 1.0 / 6.0;
    A[4 * i + 1] = 4.0 / 6.0;
    A[4 * i + 2] = 1.0 / 6.0;
    A[4 * i + 3] = 0;
  }
  double *L = &A[4 * (M - 1)];
  L[0] = hi2; L[1] = -2.0 * hi2; L[2] = hi2; L[3] = 0;
  u.assign(M, 0.0);
  d.assign(M, 0.0);
  A[1] /= A[0];
  A[2] /= A[0];
  A[3] /= A[0];
  A[0] = 1.0;
  A[4 + 1] -= A[4 + 0] * A[1];
  A[4 + 2] -= A[4 + 0] * A[2];
  d[1] = A[4 + 1];
  A[4 + 2] /= A[4 + 1];
  u[1] = A[4 + 2];
  for (int row = 2; row < M - 1; row++) {
    double *a = &A[4 * (row - 1)];
    a[4 + 1] -= a[4 + 0] * a[2];
    d[row] = a[4 + 1];
    a[4 + 2] /= a[4 + 1];
    u[row] = a[4 + 2];
  }
  ax.M = M;
  ax.c16 = 1.0 / 6.0;
  ax.A01 = A[1];
  ax.A02 = A[2];
  ax.A03 = A[3];
  ax.hi2 = L[0];
  double p = L[1] - L[0] * A[4 * (M - 3) + 2];
  double q = L[2] - p * A[4 * (M - 2) + 2];
  ax.p = p;
  ax.q = q;
  ax.rq = 1.0 / q;
  rd.assign(M, 0.0);
  for (int k = 1; k <= M - 2; k++) rd[k] = 1.0 / d[k];
  // first row of the constant regime (the table is used before it)
  int K0 = M - 2;
  while (K0 > 1 && d[K0 - 1] == d[M - 2] && u[K0 - 1] == u[M - 2] && rd[K0 - 1] == rd[M - 2]) K0--;
  if (K0 - 1 > 16)  // kHead of the kernels
    throw Error(VLS_ERR_ARG, "spline pivots do not settle within 16 rows");
  ax.K0 = K0;
  ax.dS = d[M - 2];
  ax.rdS = rd[M - 2];
  ax.uS = u[M - 2];
}

}  // namespace vls

extern "C" {

const char *vls_last_error(void) { return g_err.c_str(); }

// Short description of device 0 (name, CUs, clock, LDS, L2) for reports.
int vls_device_info(char *buf, int len) {
  return guarded([&] {
    if (!buf || len < 1) throw vls::Error(VLS_ERR_ARG, "bad buffer");
    int dev = 0, cus = 0, clk = 0, ldsBlk = 0, ldsOpt = 0, ldsCU = 0, l2 = 0, mclk = 0, bus = 0;
    VLS_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    VLS_HIP_CHECK(hipGetDeviceProperties(&p, dev));
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, dev);
    hipDeviceGetAttribute(&ldsBlk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    hipDeviceGetAttribute(&ldsOpt, hipDeviceAttributeSharedMemPerBlockOptin, dev);
    hipDeviceGetAttribute(&ldsCU, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
    hipDeviceGetAttribute(&l2, hipDeviceAttributeL2CacheSize, dev);
    hipDeviceGetAttribute(&mclk, hipDeviceAttributeMemoryClockRate, dev);
    hipDeviceGetAttribute(&bus, hipDeviceAttributeMemoryBusWidth, dev);
    snprintf(buf, len,
             "%s %s: %d CUs, %d MHz, LDS %d B/block (opt-in %d, %d per CU), L2 %d B, "
             "memory %d MHz x %d bit",
             p.name, p.gcnArchName, cus, clk / 1000, ldsBlk, ldsOpt, ldsCU, l2, mclk / 1000, bus);
  });
}

// One device-to-device copy of `bytes` bytes (the runtime's copy kernel), to
// calibrate memory counters against a transfer of known size.
int vls_copy_probe(size_t bytes) {
  return guarded([&] {
    void *a = nullptr, *b = nullptr;
    VLS_HIP_CHECK(hipMalloc(&a, bytes));
    VLS_HIP_CHECK(hipMalloc(&b, bytes));
    VLS_HIP_CHECK(hipMemset(a, 1, bytes));
    VLS_HIP_CHECK(hipDeviceSynchronize());
    VLS_HIP_CHECK(hipMemcpy(b, a, bytes, hipMemcpyDeviceToDevice));
    VLS_HIP_CHECK(hipDeviceSynchronize());
    (void)hipFree(a);
    (void)hipFree(b);
  });
}

// The data-independent Thomas table of one axis (dim nodes, knot spacing
// 1/hinv): u, d, rd have dim+2 entries; consts = {c16, A01, A02, A03, hi2, p,
// q, rq}.
int vls_thomas_table(int dim, double hinv, double *u, double *d, double *rd, double consts[8]) {
  return guarded([&] {
    if (dim < 2 || !u || !d || !rd || !consts) throw vls::Error(VLS_ERR_ARG, "bad arguments");
    vls::ThomasAxis ax{};
    std::vector<double> U, D, RD;
    vls::thomas_table(dim, hinv, ax, U, D, RD);
    std::copy(U.begin(), U.end(), u);
    std::copy(D.begin(), D.end(), d);
    std::copy(RD.begin(), RD.end(), rd);
    double c[8] = {ax.c16, ax.A01, ax.A02, ax.A03, ax.hi2, ax.p, ax.q, ax.rq};
    std::copy(c, c + 8, consts);
  });
}

// Host check of the known-divisor division against IEEE division: for each of
// the nb divisors, `samples` numerators spread over many binades (random
// significands and signs, exact multiples, values next to multiples).
int vls_check_division(const double *b, int nb, int64_t samples, uint64_t seed,
                       int64_t *mismatches) {
  if (!b || nb < 0 || samples < 0 || !mismatches) return fail(VLS_ERR_ARG, "bad arguments");
  uint64_t x = seed ? seed : 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  int64_t bad = 0;
  for (int i = 0; i < nb; i++) {
    const double bi = b[i], yb = 1.0 / bi;
    for (int64_t k = 0; k < samples; k++) {
      uint64_t r = next();
      double a;
      switch (k % 4) {
        case 0: {  // random significand, exponent in [-600, 600]
          uint64_t mant = r & ((1ull << 52) - 1);
          int e = (int)((r >> 52) % 1201) - 600;
          a = std::ldexp(1.0 + (double)mant / 4503599627370496.0, e);
          break;
        }
        case 1:  // exact multiple
          a = bi * (double)(int64_t)((r % 2000001) - 1000000);
          break;
        case 2: {  // next to a multiple
          double m = bi * (double)(int64_t)((r % 2000001) - 1000000);
          a = std::nextafter(m, (r >> 40) & 1 ? 1e300 : -1e300);
          break;
        }
        default: {  // random in [-10, 10]
          a = ((double)(r >> 11) / 9007199254740992.0 - 0.5) * 20.0;
          break;
        }
      }
      if ((r >> 63) & 1) a = -a;
      double q = a / bi, p = vls::div_known(a, bi, yb);
      if (!(q == p)) bad++;
    }
  }
  *mismatches = bad;
  return VLS_OK;
}

int vls_create(const vls_geometry *g, const vls_solver_params *sp, int nRuns, vls_ctx **out) {
  if (!out) return fail(VLS_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (!g || !sp || nRuns < 1) return fail(VLS_ERR_ARG, "vls_create: need nRuns >= 1 geometries");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(VLS_ERR_DEVICE, "vls_create: no HIP device");
  vls_ctx *ctx = new vls_ctx();
  int rc = guarded([&] {
    VLS_HIP_CHECK(hipGetDevice(&ctx->c.device));
    (void)resolve_scratch(ctx->c.device);
    int &maxLds = ctx->c.maxLds;
    VLS_HIP_CHECK(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeSharedMemPerBlockOptin,
                                        ctx->c.device));
    if (maxLds <= 0)
      VLS_HIP_CHECK(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                          ctx->c.device));
    VLS_HIP_CHECK(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    hipStream_t s = ctx->c.stream;
    ctx->c.runs.resize(nRuns);
    for (int k = 0; k < nRuns; k++) {
      const vls_geometry &G = g[k];
      vls::Run &r = ctx->c.runs[k];
      if (G.nR < 4 || G.nU < 4 || G.nJ < 1)
        throw vls::Error(VLS_ERR_ARG, "vls_create: grid too small (need nR, nU >= 4, nJ >= 1)");
      if (!G.R_Val || !G.U_Val || !G.J_Val || !G.dR || !G.dU || !G.dJ)
        throw vls::Error(VLS_ERR_ARG, "vls_create: missing geometry table");
      r.nR = G.nR;
      r.nU = G.nU;
      r.nJ = G.nJ;
      r.sp = sp[k];
      r.nVal = (int64_t)r.nR * r.nU * r.nJ;
      r.nS = sp[k].n_species > 1 ? sp[k].n_species : 1;
      if (r.nS > 5) throw vls::Error(VLS_ERR_ARG, "vls_create: at most 5 species");
      r.stride = r.nR + 2;
      r.cs = (int64_t)(r.nR + 2) * (r.nU + 2);
      r.start_v = G.spline_start[1];
      r.hinv_v = G.spline_hinv[1];
      r.hU.assign(G.U_Val, G.U_Val + r.nU);
      r.Ulo = G.U_Val[0];
      r.Uhi = G.U_Val[r.nU - 1];
      vls::configure_tiles(r, maxLds);
      // the kick folded into the r-solve leaves the solved rows (c's layout) in
      // f's buffer between the kick and the drift
      r.fRows = vls::kick_rsolve_shape(r);
      r.f = dalloc<double>(r, r.fRows ? std::max<size_t>(r.nVal, (size_t)r.nJ * r.cs)
                                      : (size_t)(r.nVal * r.nS));
      // A foot on the last knot (integer part dim-1, fraction 0) makes the
      // 4x4 stencil touch one column past its padded row and one row past its
      // slice, with weight zero, as in the reference; the tail pad keeps
      // those reads inside the allocation.
      r.c = dalloc<double>(r, (size_t)r.nJ * r.cs + kCoefPad(r));
      VLS_HIP_CHECK(hipMemsetAsync(r.c, 0, ((size_t)r.nJ * r.cs + kCoefPad(r)) * sizeof(double), s));
      if (size_t n = vls::kick_scratch_doubles(r)) r.gsc = dalloc<double>(r, n);
      r.kickScratch = getenv("VLS_KICK_SCRATCH") != nullptr;
      r.feet = dalloc<double>(r, 2 * r.nVal);
      r.P = dalloc<double>(r, (size_t)r.nJ * r.nR);
      r.dM = dalloc<double>(r, r.nR);
      r.Mx = dalloc<double>(r, r.nR);
      r.acc = dalloc<double>(r, r.nR);
      r.mTot = dalloc<double>(r, 1);
      r.mCnt = dalloc<unsigned>(r, vls::moment_counters(r));
      VLS_HIP_CHECK(hipMemsetAsync(r.mCnt, 0, vls::moment_counters(r) * sizeof(unsigned), s));
      r.rhoPhi = dalloc<double>(r, 2 * (size_t)r.nR);
      r.st_part = dalloc<double>(r, 7 * (size_t)r.nJ * r.nR);
      r.st_q = dalloc<double>(r, 6 * (size_t)r.nJ + r.nR);
      r.st_out = dalloc<double>(r, 14);
      // tables
      std::vector<double> Ct1(r.nJ), gfac(r.nR);
      for (int j = 0; j < r.nJ; j++) Ct1[j] = 8 * M_PI * M_PI * G.J_Val[j] * G.dJ[j];
      for (int i = 0; i < r.nR; i++) gfac[i] = r.sp.G / (G.R_Val[i] * G.R_Val[i]) * r.sp.dt;
      r.R_Val = dalloc<double>(r, r.nR);
      r.U_Val = dalloc<double>(r, r.nU);
      r.J_Val = dalloc<double>(r, r.nJ);
      r.dR = dalloc<double>(r, r.nR);
      r.dU = dalloc<double>(r, r.nU);
      r.Ct1 = dalloc<double>(r, r.nJ);
      r.gfac = dalloc<double>(r, r.nR);
      h2d(r.R_Val, G.R_Val, r.nR, s);
      h2d(r.U_Val, G.U_Val, r.nU, s);
      h2d(r.J_Val, G.J_Val, r.nJ, s);
      h2d(r.dR, G.dR, r.nR, s);
      h2d(r.dU, G.dU, r.nU, s);
      h2d(r.Ct1, Ct1.data(), r.nJ, s);
      h2d(r.gfac, gfac.data(), r.nR, s);
      for (int axis = 0; axis < 2; axis++) {
        std::vector<double> u, d, rd;
        vls::ThomasAxis &ax = axis == 0 ? r.tr : r.tv;
        vls::thomas_table(axis == 0 ? r.nR : r.nU, G.spline_hinv[axis], ax, u, d, rd);
        double *dev = dalloc<double>(r, 3 * (size_t)ax.M);
        h2d(dev, u.data(), ax.M, s);
        h2d(dev + ax.M, d.data(), ax.M, s);
        h2d(dev + 2 * ax.M, rd.data(), ax.M, s);
        ax.u = dev;
        ax.d = dev + ax.M;
        ax.rd = dev + 2 * ax.M;
        VLS_HIP_CHECK(hipStreamSynchronize(s));  // u, d, rd go out of scope
      }
    }
  });
  if (rc) {
    vls_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return VLS_OK;
}

void vls_destroy(vls_ctx *ctx) {
  if (!ctx) return;
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  drop_graph(ctx->c);
  for (auto st : ctx->c.runStreams) (void)hipStreamDestroy(st);
  for (auto e : ctx->c.joinEv) (void)hipEventDestroy(e);
  if (ctx->c.forkEv) (void)hipEventDestroy(ctx->c.forkEv);
  for (auto &r : ctx->c.runs)
    for (void *p : r.allocs) (void)hipFree(p);
  for (auto e : ctx->c.timing.ev) (void)hipEventDestroy(e);
  if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

int vls_set_graph(vls_ctx *ctx, int on) {
  return guarded([&] {
    if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
    VLS_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    drop_graph(ctx->c);
    ctx->c.graphOn = on != 0;
  });
}

int vls_upload_f(vls_ctx *ctx, int run, const double *f) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!f) throw vls::Error(VLS_ERR_ARG, "null f");
    if (r.nS > 1) {
      std::vector<double> planes = to_planes(f, r.nVal, r.nS);
      VLS_HIP_CHECK(hipMemcpyAsync(r.f, planes.data(), planes.size() * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->c.stream));
      VLS_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    } else {
      VLS_HIP_CHECK(hipMemcpyAsync(r.f, f, r.nVal * sizeof(double), hipMemcpyHostToDevice,
                                   ctx->c.stream));
      VLS_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    }
    r.have_f = true;
  });
}

int vls_download_f(vls_ctx *ctx, int run, double *f) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!f) throw vls::Error(VLS_ERR_ARG, "null f");
    if (r.nS > 1) {
      std::vector<double> planes((size_t)r.nVal * r.nS);
      VLS_HIP_CHECK(hipMemcpyAsync(planes.data(), r.f, planes.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->c.stream));
      VLS_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
      from_planes(planes.data(), f, r.nVal, r.nS);
    } else {
      VLS_HIP_CHECK(hipMemcpyAsync(f, r.f, r.nVal * sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->c.stream));
      VLS_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    }
  });
}

int vls_set_drift(vls_ctx *ctx, int run, const double *xr, const double *xu, const uint8_t *flag) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!xr || !xu || !flag) throw vls::Error(VLS_ERR_ARG, "null foot table");
    // every evaluated foot must index inside the padded coefficient array
    for (int64_t i = 0; i < r.nVal; i++) {
      if (flag[i] > VLS_FOOT_CHAIN) throw vls::Error(VLS_ERR_ARG, "invalid foot flag");
      if (flag[i] != VLS_FOOT_EVAL) continue;
      if (!foot_in_domain(xr[i], xu[i], r))
        throw vls::Error(VLS_ERR_ARG, "vls_set_drift: foot outside the spline domain at point " +
                                          std::to_string(i));
    }
    // the drift kernel reads the point's kind from xr (-1 outside, -2 delay
    // line) and a delay-line point's chain index from xu (-1: none yet)
    std::vector<double> xf(xr, xr + r.nVal), uf(xu, xu + r.nVal);
    for (int64_t i = 0; i < r.nVal; i++) {
      if (flag[i] == VLS_FOOT_OUT) xf[i] = -1.0;
      if (flag[i] == VLS_FOOT_CHAIN) xf[i] = -2.0, uf[i] = -1.0;
    }
    r.hflag.assign(flag, flag + r.nVal);
    hipStream_t s = ctx->c.stream;
    std::vector<double> ft(2 * r.nVal);
    for (int64_t i = 0; i < r.nVal; i++) ft[2 * i] = xf[i], ft[2 * i + 1] = uf[i];
    h2d(r.feet, ft.data(), 2 * r.nVal, s);
    VLS_HIP_CHECK(hipStreamSynchronize(s));
    r.hxr.swap(xf);
    r.hxu.swap(uf);
    r.adj.tilesValid = false;  // the adjoint's transposed foot table follows the feet
    r.adj.recorded = false;    // a recording made with other feet is void
    vls::mark_chains(r, s);
    VLS_HIP_CHECK(hipGetLastError());
    VLS_HIP_CHECK(hipStreamSynchronize(s));
    r.have_drift = true;
  });
}

int vls_set_chains(vls_ctx *ctx, int run, int64_t nChains, const int64_t *point,
                   const int32_t *len, const int32_t *head, const double *back_xr,
                   const double *back_xu, const uint8_t *back_out, const double *vals) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (nChains < 0) throw vls::Error(VLS_ERR_ARG, "negative chain count");
    if (nChains > 0 && (!point || !len || !head || !back_xr || !back_xu || !back_out || !vals))
      throw vls::Error(VLS_ERR_ARG, "null chain table");
    std::vector<int64_t> off(nChains);
    std::vector<uint8_t> seen(nChains ? r.nVal : 0, 0);
    int64_t cells = 0;
    for (int64_t k = 0; k < nChains; k++) {
      if (len[k] < 2) throw vls::Error(VLS_ERR_ARG, "chain shorter than 2 cells");
      if (head[k] < 0 || head[k] >= len[k]) throw vls::Error(VLS_ERR_ARG, "chain head out of range");
      if (point[k] < 0 || point[k] >= r.nVal) throw vls::Error(VLS_ERR_ARG, "chain point out of range");
      if (!r.hflag.empty() && r.hflag[point[k]] != VLS_FOOT_CHAIN)
        throw vls::Error(VLS_ERR_ARG, "chain attached to a point whose foot is not a delay line");
      if (seen[point[k]]++) throw vls::Error(VLS_ERR_ARG, "two chains attached to one point");
      if (!back_out[k] && !foot_in_domain(back_xr[k], back_xu[k], r))
        throw vls::Error(VLS_ERR_ARG, "chain back cell outside the spline domain");
      off[k] = cells;
      cells += len[k];
    }
    // replace any previous chain arrays (after the work that may still read
    // them); a recording made with other delay lines is void
    hipStream_t s = ctx->c.stream;
    VLS_HIP_CHECK(hipStreamSynchronize(s));
    drop_graph(ctx->c);
    dfree(r, r.ch_point);
    dfree(r, r.ch_off);
    dfree(r, r.ch_len);
    dfree(r, r.ch_head);
    dfree(r, r.ch_xr);
    dfree(r, r.ch_xu);
    dfree(r, r.ch_out);
    dfree(r, r.ch_vals);
    r.adj.recorded = false;
    r.adj.tilesValid = false;  // the transposed foot table lists the back cells
    r.nChains = nChains;
    r.nCells = cells;
    r.hch_point.assign(point, point + nChains);
    r.hch_xr.assign(back_xr, back_xr + nChains);
    r.hch_xu.assign(back_xu, back_xu + nChains);
    r.hch_out.assign(back_out, back_out + nChains);
    r.ch_point = dalloc<int64_t>(r, nChains);
    r.ch_off = dalloc<int64_t>(r, nChains);
    r.ch_len = dalloc<int32_t>(r, nChains);
    r.ch_head = dalloc<int32_t>(r, nChains);
    r.ch_xr = dalloc<double>(r, nChains);
    r.ch_xu = dalloc<double>(r, nChains);
    r.ch_out = dalloc<uint8_t>(r, nChains);
    r.ch_vals = dalloc<double>(r, cells * r.nS);
    h2d(r.ch_point, point, nChains, s);
    h2d(r.ch_off, off.data(), nChains, s);
    h2d(r.ch_len, len, nChains, s);
    h2d(r.ch_head, head, nChains, s);
    h2d(r.ch_xr, back_xr, nChains, s);
    h2d(r.ch_xu, back_xu, nChains, s);
    h2d(r.ch_out, back_out, nChains, s);
    if (r.nS > 1) {
      std::vector<double> planes = to_planes(vals, cells, r.nS);
      h2d(r.ch_vals, planes.data(), planes.size(), s);
      VLS_HIP_CHECK(hipStreamSynchronize(s));
    } else {
      h2d(r.ch_vals, vals, cells, s);
    }
    if (r.have_drift) {
      // clear the marks of any earlier delay lines, then attach these
      std::vector<double> ft(2 * r.nVal);
      for (int64_t i = 0; i < r.nVal; i++) ft[2 * i] = r.hxr[i], ft[2 * i + 1] = r.hxu[i];
      h2d(r.feet, ft.data(), 2 * r.nVal, s);
      vls::mark_chains(r, s);
      VLS_HIP_CHECK(hipStreamSynchronize(s));
    }
    VLS_HIP_CHECK(hipGetLastError());
    VLS_HIP_CHECK(hipStreamSynchronize(s));
    r.have_chains = true;
  });
}

int vls_get_chains(vls_ctx *ctx, int run, int32_t *head, double *vals) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    hipStream_t s = ctx->c.stream;
    if (head && r.nChains)
      VLS_HIP_CHECK(hipMemcpyAsync(head, r.ch_head, r.nChains * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, s));
    if (vals && r.nCells) {
      std::vector<double> planes((size_t)r.nCells * r.nS);
      VLS_HIP_CHECK(hipMemcpyAsync(planes.data(), r.ch_vals, planes.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, s));
      VLS_HIP_CHECK(hipStreamSynchronize(s));
      from_planes(planes.data(), vals, r.nCells, r.nS);
    }
    VLS_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int vls_step(vls_ctx *ctx, int nSteps) {
  return guarded([&] {
    if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
    hipStream_t s = ctx->c.stream;
    vls::Timing *t = &ctx->c.timing;
    if (t->on) {
      for (auto e : t->ev) (void)hipEventDestroy(e);
      t->ev.clear();
      t->names.clear();
      vls::time_mark(t, s, "start");  // every kernel: the first interval starts here
    }
    for (auto &r : ctx->c.runs)
      if (!r.have_f || !r.have_drift)
        throw vls::Error(VLS_ERR_STATE, "vls_step: f and drift feet must be set first");
    vls::Context &C = ctx->c;
    if (C.runs.size() > 1 && C.graphOn && !t->on) {
      // an ensemble: the runs advance concurrently, one graph launch a step
      if (!C.graph) build_ensemble_graph(C);
      for (int n = 0; n < nSteps; n++) VLS_HIP_CHECK(hipGraphLaunch(C.graph, s));
      for (auto &r : C.runs) r.stepped = true;
    } else {
      for (int n = 0; n < nSteps; n++)
        for (auto &r : C.runs) step_once(r, s, t);
    }
    VLS_HIP_CHECK(hipGetLastError());
  });
}

int vls_synchronize(vls_ctx *ctx) {
  return guarded([&] {
    if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
    VLS_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
  });
}

int vls_moments(vls_ctx *ctx, int run, double *dM, double *accel, double *mTot) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!r.stepped) throw vls::Error(VLS_ERR_STATE, "vls_moments: no step taken yet");
    hipStream_t s = ctx->c.stream;
    if (dM) VLS_HIP_CHECK(hipMemcpyAsync(dM, r.dM, r.nR * sizeof(double), hipMemcpyDeviceToHost, s));
    if (accel)
      VLS_HIP_CHECK(hipMemcpyAsync(accel, r.acc, r.nR * sizeof(double), hipMemcpyDeviceToHost, s));
    if (mTot) VLS_HIP_CHECK(hipMemcpyAsync(mTot, r.mTot, sizeof(double), hipMemcpyDeviceToHost, s));
    VLS_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int vls_stats(vls_ctx *ctx, int run, double out[7]) { return vls_stats_total(ctx, run, out, nullptr); }

int vls_stats_total(vls_ctx *ctx, int run, double line[7], double total[7]) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!line && !total) throw vls::Error(VLS_ERR_ARG, "null output");
    if (!r.have_f) throw vls::Error(VLS_ERR_STATE, "vls_stats: f not set");
    hipStream_t s = ctx->c.stream;
    vls::launch_stats(r, s);
    VLS_HIP_CHECK(hipGetLastError());
    double out[14];
    VLS_HIP_CHECK(hipMemcpyAsync(out, r.st_out, 14 * sizeof(double), hipMemcpyDeviceToHost, s));
    VLS_HIP_CHECK(hipStreamSynchronize(s));
    if (line) std::copy(out, out + 7, line);
    if (total) std::copy(out + 7, out + 14, total);
  });
}

int vls_download_coef(vls_ctx *ctx, int run, int j, double *coef) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (j < 0 || j >= r.nJ || !coef) throw vls::Error(VLS_ERR_ARG, "bad slice or null output");
    VLS_HIP_CHECK(hipMemcpyAsync(coef, r.c + (int64_t)j * r.cs, r.cs * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->c.stream));
    VLS_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
  });
}

int vls_set_timing(vls_ctx *ctx, int on) {
  return guarded([&] {
    if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
    ctx->c.timing.on = on != 0;
  });
}

int vls_autotune(vls_ctx *ctx, int run, char *desc, int descLen) {
  return guarded([&] {
    if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
    vls::Context &C = ctx->c;
    if (run < -1 || run >= (int)C.runs.size()) throw vls::Error(VLS_ERR_ARG, "run index out of range");
    std::string all;
    for (int k = 0; k < (int)C.runs.size(); k++) {
      if (run >= 0 && k != run) continue;
      vls::Run &r = C.runs[k];
      if (!r.have_f) throw vls::Error(VLS_ERR_STATE, "vls_autotune: f not set");
      if (!all.empty()) all += "; ";
      all += vls::autotune(r, C.stream, C.maxLds);
    }
    drop_graph(C);  // the captured step holds the old launches
    VLS_HIP_CHECK(hipStreamSynchronize(C.stream));
    if (desc && descLen > 0) {
      strncpy(desc, all.c_str(), (size_t)descLen - 1);
      desc[descLen - 1] = 0;
    }
  });
}

int vls_set_timing_kernel(vls_ctx *ctx, const char *kernel) {
  return guarded([&] {
    if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
    ctx->c.timing.only = kernel ? kernel : "";
  });
}

int vls_run_info(vls_ctx *ctx, int run, int64_t out[6]) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!out) throw vls::Error(VLS_ERR_ARG, "null output");
    out[0] = r.nR;
    out[1] = r.nU;
    out[2] = r.nJ;
    out[3] = r.nChains;
    out[4] = r.nCells;
    out[5] = r.nS;
  });
}

int vls_memory(vls_ctx *ctx, int run, int64_t *count, int64_t *bytes) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (count) *count = (int64_t)r.allocs.size();
    if (bytes) {
      int64_t b = 0;
      for (size_t x : r.allocBytes) b += (int64_t)x;
      *bytes = b;
    }
  });
}

int vls_put_chains(vls_ctx *ctx, int run, const int32_t *head, const double *vals) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    hipStream_t s = ctx->c.stream;
    if (r.nChains && head) {
      std::vector<int32_t> len(r.nChains);
      VLS_HIP_CHECK(hipMemcpyAsync(len.data(), r.ch_len, r.nChains * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, s));
      VLS_HIP_CHECK(hipStreamSynchronize(s));
      for (int64_t k = 0; k < r.nChains; k++)
        if (head[k] < 0 || head[k] >= len[k]) throw vls::Error(VLS_ERR_ARG, "chain head out of range");
      h2d(r.ch_head, head, r.nChains, s);
    }
    if (r.nCells && vals) {
      std::vector<double> planes = to_planes(vals, r.nCells, r.nS);
      h2d(r.ch_vals, planes.data(), planes.size(), s);
    }
    VLS_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int vls_mass_profile(vls_ctx *ctx, int run, double *M) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!M) throw vls::Error(VLS_ERR_ARG, "null output");
    if (!r.have_f) throw vls::Error(VLS_ERR_STATE, "vls_mass_profile: f not set");
    hipStream_t s = ctx->c.stream;
    vls::launch_moment(r, s);
    VLS_HIP_CHECK(hipGetLastError());
    VLS_HIP_CHECK(hipMemcpyAsync(M, r.Mx, r.nR * sizeof(double), hipMemcpyDeviceToHost, s));
    VLS_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int vls_density_potential(vls_ctx *ctx, int run, double *rho, double *phi) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!rho && !phi) throw vls::Error(VLS_ERR_ARG, "null output");
    if (!r.have_f) throw vls::Error(VLS_ERR_STATE, "vls_density_potential: f not set");
    hipStream_t s = ctx->c.stream;
    vls::launch_moment(r, s);
    vls::launch_density_potential(r, r.rhoPhi, r.rhoPhi + r.nR, s);
    VLS_HIP_CHECK(hipGetLastError());
    if (rho)
      VLS_HIP_CHECK(hipMemcpyAsync(rho, r.rhoPhi, r.nR * sizeof(double), hipMemcpyDeviceToHost, s));
    if (phi)
      VLS_HIP_CHECK(hipMemcpyAsync(phi, r.rhoPhi + r.nR, r.nR * sizeof(double),
                                   hipMemcpyDeviceToHost, s));
    VLS_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int vls_density_potential_noalloc(vls_ctx *ctx, int run, const double *f_dev, double *rho_dev,
                                  double *phi_dev, double *M_dev, void *stream) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    if (!f_dev) throw vls::Error(VLS_ERR_ARG, "vls_density_potential_noalloc: null f");
    if (!rho_dev && !phi_dev && !M_dev)
      throw vls::Error(VLS_ERR_ARG, "vls_density_potential_noalloc: no output");
    // P [nJ][nR], dM, acc, and Mx, rho, phi where the caller passes none, mTot
    const size_t need = (size_t)r.nJ * r.nR + 5 * (size_t)r.nR + 1;
    if (need > (size_t)VLS_NOALLOC_SCRATCH_DOUBLES)
      throw vls::Error(VLS_ERR_ARG,
                       "vls_density_potential_noalloc: run needs " + std::to_string(need) +
                           " doubles of scratch (nJ*nR + 5*nR + 1), the maximum is " +
                           std::to_string(VLS_NOALLOC_SCRATCH_DOUBLES));
    int cur = -1;
    VLS_HIP_CHECK(hipGetDevice(&cur));
    if (cur != ctx->c.device)
      throw vls::Error(VLS_ERR_STATE, "vls_density_potential_noalloc: the context's device is "
                                      "not the current device");
    double *w = scratch_dev(cur);
    if (!w) throw vls::Error(VLS_ERR_STATE, "vls_density_potential_noalloc: no scratch");
    // every entry below is written before it is read: nothing to clear
    double *P = w, *dM = P + (size_t)r.nJ * r.nR, *acc = dM + r.nR, *Mx = acc + r.nR;
    double *rho = Mx + r.nR, *phi = rho + r.nR, *mTot = phi + r.nR;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->c.stream;
    const vls::MomentBufs b{f_dev, P, dM, M_dev ? M_dev : Mx, acc, mTot, nullptr};
    vls::launch_moment_into(r, b, s);
    if (rho_dev || phi_dev)
      vls::launch_density_potential_from(r, dM, b.Mx, rho_dev ? rho_dev : rho,
                                         phi_dev ? phi_dev : phi, s);
    VLS_HIP_CHECK(hipGetLastError());
  });
}

namespace {
// The recording and the reverse sweep of one run, in stages that enqueue on
// a stream of the caller's choice, so that a batch of runs can interleave
// them on streams of their own (vls_record_batch, vls_adjoint_batch).

// Checks and every buffer of a recording except the kept steps.
void record_buffers(vls::Run &r, hipStream_t s, int nSteps, int ckEvery) {
  if (nSteps < 1 || ckEvery < 1) throw vls::Error(VLS_ERR_ARG, "vls_record: need nSteps, ckEvery >= 1");
  if (!r.have_f || !r.have_drift)
    throw vls::Error(VLS_ERR_STATE, "vls_record: f and drift feet must be set first");
  if (r.sp.no_linking)
    throw vls::Error(VLS_ERR_ARG, "vls_record: the adjoint covers linking steps only");
  if (r.nS > 1) throw vls::Error(VLS_ERR_ARG, "vls_record: the adjoint covers one species only");
  // the kick's transpose finds the points of a coefficient row from
  // n ~ u + a hinv, which needs v nodes on the uniform spline knots (to
  // 0.01 of a cell; a linear v scale is there to rounding)
  for (int u = 0; u < r.nU; u++)
    if (std::fabs((r.hU[u] - r.start_v) * r.hinv_v - u) > 0.01)
      throw vls::Error(VLS_ERR_ARG, "vls_record: the adjoint needs a uniform v grid");
  auto &A = r.adj;
  A.recorded = false;
  const int nCk = (nSteps + ckEvery - 1) / ckEvery;
  const int seg = std::min(ckEvery, nSteps);
  const size_t nv = r.nVal, nc = r.nCells, nch = r.nChains;
  const size_t coef = (size_t)r.nJ * r.cs + kCoefPad(r);
  if (!A.fb) {
    A.fb = dalloc<double>(r, nv);
    A.fsb = dalloc<double>(r, nv);
    A.cb = dalloc<double>(r, coef);
    A.abp = dalloc<double>(r, (size_t)r.nJ * r.nR);
    A.dMb = dalloc<double>(r, r.nR);
    A.Mbar = dalloc<double>(r, r.nR);
    A.endF = dalloc<double>(r, nv);
  }
  // the chain-sized arrays follow the current delay lines
  const bool chainsChanged = A.cells != (int64_t)nc || A.chains != (int64_t)nch;
  if (chainsChanged || A.capCk < nCk || A.capSeg < seg)
    VLS_HIP_CHECK(hipStreamSynchronize(s));  // earlier adjoint work may still use them
  if (chainsChanged) {
    dfree(r, A.ringb);
    dfree(r, A.endVals);
    dfree(r, A.endHead);
    dfree(r, A.chainXb);
    A.ringb = dalloc<double>(r, nc);
    A.endVals = dalloc<double>(r, nc);
    A.endHead = dalloc<int32_t>(r, nch);
    A.chainXb = dalloc<double>(r, nch);
    A.cells = nc;
    A.chains = nch;
  }
  if (A.capCk < nCk || chainsChanged) {
    dfree(r, A.ckF);
    dfree(r, A.ckVals);
    dfree(r, A.ckHead);
    A.ckF = dalloc<double>(r, nv * nCk);
    A.ckVals = dalloc<double>(r, nc * nCk);
    A.ckHead = dalloc<int32_t>(r, nch * nCk);
    A.capCk = nCk;
  }
  if (A.capSeg < seg) {
    dfree(r, A.segC);
    dfree(r, A.segAcc);
    A.segC = dalloc<double>(r, coef * seg);
    A.segAcc = dalloc<double>(r, (size_t)r.nR * seg);
    A.capSeg = seg;
  }
  if (!A.tilesValid) {
    std::vector<int64_t> off;
    std::vector<uint32_t> pts;
    vls::build_drift_tiles(r, r.hxr, r.hxu, off, pts);
    VLS_HIP_CHECK(hipStreamSynchronize(s));  // earlier tables may still be in use
    dfree(r, A.tileOff);
    dfree(r, A.tilePts);
    A.tileOff = dalloc<int64_t>(r, off.size());
    A.tilePts = dalloc<uint32_t>(r, pts.size());
    h2d(A.tileOff, off.data(), off.size(), s);
    h2d(A.tilePts, pts.data(), pts.size(), s);
    VLS_HIP_CHECK(hipStreamSynchronize(s));
    A.tilesValid = true;
  }
  A.nSteps = nSteps;
  A.ckEvery = ckEvery;
  A.nCk = nCk;
}

// Device bytes of one kept step: the kick's column splines and acceleration.
size_t kept_step_bytes(const vls::Run &r) {
  return ((size_t)r.nJ * r.cs + kCoefPad(r) + (size_t)r.nR) * sizeof(double);
}

// Bytes the run holds for kept steps.
size_t kept_held_bytes(const vls::Run &r) { return (size_t)r.adj.capKeep * kept_step_bytes(r); }

// The trailing segments whose kept steps fit in `budget` bytes (what the run
// already holds for them included) are kept by the recorded run itself, so
// that the reverse sweep need not recompute them (VLS_ADJ_KEEP_STEPS caps
// the kept steps).
void record_keep(vls::Run &r, hipStream_t s, size_t budget) {
  auto &A = r.adj;
  const size_t coef = (size_t)r.nJ * r.cs + kCoefPad(r);
  const size_t stepB = kept_step_bytes(r);
  int64_t cap = (int64_t)(budget / stepB);
  if (const char *e = getenv("VLS_ADJ_KEEP_STEPS")) cap = std::min<int64_t>(cap, atoll(e));
  int keepSegs = 0;
  while (keepSegs < A.nCk) {
    const int from = (A.nCk - keepSegs - 1) * A.ckEvery;
    if (A.nSteps - from > cap) break;
    keepSegs++;
  }
  A.keptFrom = (A.nCk - keepSegs) * A.ckEvery;
  const int kept = A.nSteps - A.keptFrom;
  if (kept > A.capKeep) {
    VLS_HIP_CHECK(hipStreamSynchronize(s));
    dfree(r, A.keepC);
    dfree(r, A.keepAcc);
    A.keepC = dalloc<double>(r, coef * kept);
    A.keepAcc = dalloc<double>(r, (size_t)r.nR * kept);
    A.capKeep = kept;
  }
}

// Step n of the recorded run: the checkpoint at the start of a segment, then
// the step (keeping its kick coefficients in the kept segments).
void record_step(vls::Run &r, hipStream_t s, int n) {
  auto &A = r.adj;
  const size_t nv = r.nVal, nc = r.nCells, nch = r.nChains;
  const size_t coef = (size_t)r.nJ * r.cs + kCoefPad(r);
  if (n % A.ckEvery == 0) {
    const int k = n / A.ckEvery;
    d2d(A.ckF + nv * k, r.f, nv, s);
    d2d(A.ckVals + nc * k, r.ch_vals, nc, s);
    d2d(A.ckHead + nch * k, r.ch_head, nch, s);
  }
  if (n >= A.keptFrom)
    step_keep(r, s, A.keepC + coef * (n - A.keptFrom), A.keepAcc + (size_t)r.nR * (n - A.keptFrom));
  else
    step_once(r, s, nullptr);
}

void record_end(vls::Run &r, hipStream_t s) {
  auto &A = r.adj;
  // the acceleration of the last step, as after plain steps
  if (A.keptFrom < A.nSteps)
    d2d(r.acc, A.keepAcc + (size_t)r.nR * (A.nSteps - 1 - A.keptFrom), r.nR, s);
  d2d(A.endF, r.f, r.nVal, s);
  d2d(A.endVals, r.ch_vals, r.nCells, s);
  d2d(A.endHead, r.ch_head, r.nChains, s);
  A.recorded = true;
}

void adjoint_check(const vls::Run &r) {
  const auto &A = r.adj;
  if (!A.recorded) throw vls::Error(VLS_ERR_STATE, "vls_adjoint: no recorded run (vls_record)");
  if (A.cells != r.nCells || A.chains != r.nChains || !A.tilesValid)
    throw vls::Error(VLS_ERR_STATE, "vls_adjoint: the delay lines or feet changed since vls_record");
}

// lambda_N: the loss's derivative in f_N, directly and through M(<r).
void adjoint_seed(vls::Run &r, hipStream_t s, const double *dL_dM, const double *dL_df) {
  auto &A = r.adj;
  if (dL_df)
    h2d(A.fb, dL_df, r.nVal, s);
  else
    vls::launch_fill(A.fb, (int64_t)r.nVal, 0.0, s);
  if (dL_dM) {
    h2d(A.Mbar, dL_dM, r.nR, s);
    vls::launch_adj_mass(r, nullptr, A.Mbar, A.dMb, A.fb, true, s);
  }
  vls::launch_fill(A.ringb, (int64_t)r.nCells, 0.0, s);
}

// The reverse sweep through checkpoint segment sg.
void adjoint_segment(vls::Run &r, hipStream_t s, int sg) {
  auto &A = r.adj;
  const size_t nv = r.nVal, nc = r.nCells, nch = r.nChains;
  const size_t coef = (size_t)r.nJ * r.cs + kCoefPad(r);
  const int n0 = sg * A.ckEvery, n1 = std::min(n0 + A.ckEvery, A.nSteps);
  const bool kept = n0 >= A.keptFrom;
  if (!kept) {
    // recompute the segment from its checkpoint, keeping each step's kick
    // coefficients and acceleration; the state after its last step is not
    // needed (the next segment has its own checkpoint, and the recorded end
    // state is restored at the end)
    d2d(r.f, A.ckF + nv * sg, nv, s);
    d2d(r.ch_vals, A.ckVals + nc * sg, nc, s);
    d2d(r.ch_head, A.ckHead + nch * sg, nch, s);
    for (int n = n0; n < n1; n++)
      step_keep(r, s, A.segC + coef * (n - n0), A.segAcc + (size_t)r.nR * (n - n0), n + 1 < n1);
  }
  for (int n = n1 - 1; n >= n0; n--) {
    const double *cK = kept ? A.keepC + coef * (n - A.keptFrom) : A.segC + coef * (n - n0);
    const double *accK =
        kept ? A.keepAcc + (size_t)r.nR * (n - A.keptFrom) : A.segAcc + (size_t)r.nR * (n - n0);
    vls::launch_adj_drift(r, A.fb, A.cb, A.ringb, A.ckHead + nch * sg, n - n0, s);
    vls::launch_adj_spline(r, A.cb, A.fsb, s);
    vls::launch_adj_kick(r, A.fsb, cK, accK, A.cb, A.abp, s);
    vls::launch_adj_scan(r, A.abp, nullptr, A.dMb, s);
    vls::launch_adj_kick_cols(r, A.cb, A.fb, A.dMb, s);  // plus the moment's transpose
  }
}

// The gradients to the host, and the run back in the state after the
// recorded run.
void adjoint_end(vls::Run &r, hipStream_t s, double *grad_f0, double *grad_vals0) {
  auto &A = r.adj;
  const size_t nv = r.nVal, nc = r.nCells, nch = r.nChains;
  VLS_HIP_CHECK(hipMemcpyAsync(grad_f0, A.fb, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  if (grad_vals0 && nc)
    VLS_HIP_CHECK(hipMemcpyAsync(grad_vals0, A.ringb, nc * sizeof(double), hipMemcpyDeviceToHost, s));
  d2d(r.f, A.endF, nv, s);
  d2d(r.ch_vals, A.endVals, nc, s);
  d2d(r.ch_head, A.endHead, nch, s);
}

size_t free_hbm() {
  size_t freeB = 0, totalB = 0;
  VLS_HIP_CHECK(hipMemGetInfo(&freeB, &totalB));
  return freeB;
}

// The ensemble's run streams, forked from the context stream.
void fork_runs(vls::Context &C) {
  ensure_run_streams(C);
  if (!C.forkEv) VLS_HIP_CHECK(hipEventCreateWithFlags(&C.forkEv, hipEventDisableTiming));
  VLS_HIP_CHECK(hipEventRecord(C.forkEv, C.stream));
  for (size_t k = 0; k < C.runs.size(); k++)
    VLS_HIP_CHECK(hipStreamWaitEvent(C.runStreams[k], C.forkEv, 0));
}

void join_runs(vls::Context &C) {
  for (size_t k = 0; k < C.runs.size(); k++) {
    VLS_HIP_CHECK(hipEventRecord(C.joinEv[k], C.runStreams[k]));
    VLS_HIP_CHECK(hipStreamWaitEvent(C.stream, C.joinEv[k], 0));
  }
}
}  // namespace

int vls_record(vls_ctx *ctx, int run, int nSteps, int ckEvery) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    hipStream_t s = ctx->c.stream;
    record_buffers(r, s, nSteps, ckEvery);
    // the kept steps may fill seven eighths of the free HBM and of what the
    // run already holds for them
    record_keep(r, s, (free_hbm() + kept_held_bytes(r)) / 8 * 7);
    for (int n = 0; n < nSteps; n++) record_step(r, s, n);
    record_end(r, s);
    VLS_HIP_CHECK(hipGetLastError());
    VLS_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int vls_adjoint(vls_ctx *ctx, int run, const double *dL_dM, const double *dL_df,
                double *grad_f0, double *grad_vals0) {
  return guarded([&] {
    vls::Run &r = run_of(ctx, run);
    adjoint_check(r);
    if (!grad_f0) throw vls::Error(VLS_ERR_ARG, "null gradient output");
    hipStream_t s = ctx->c.stream;
    adjoint_seed(r, s, dL_dM, dL_df);
    for (int sg = r.adj.nCk - 1; sg >= 0; sg--) adjoint_segment(r, s, sg);
    VLS_HIP_CHECK(hipGetLastError());
    adjoint_end(r, s, grad_f0, grad_vals0);
    VLS_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int vls_record_batch(vls_ctx *ctx, int nSteps, int ckEvery) {
  return guarded([&] {
    if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
    vls::Context &C = ctx->c;
    const int nRuns = (int)C.runs.size();
    ensure_run_streams(C);
    for (int k = 0; k < nRuns; k++) record_buffers(C.runs[k], C.runStreams[k], nSteps, ckEvery);
    // seven eighths of the free HBM and of what the runs hold for kept steps,
    // shared in proportion to the runs' step sizes, so that every run keeps
    // the same trailing steps
    size_t held = 0, total = 0;
    for (auto &r : C.runs) {
      held += kept_held_bytes(r);
      total += kept_step_bytes(r);
    }
    const size_t budget = (free_hbm() + held) / 8 * 7;
    for (int k = 0; k < nRuns; k++)
      record_keep(C.runs[k], C.runStreams[k],
                  (size_t)((double)budget * kept_step_bytes(C.runs[k]) / (double)total));
    VLS_HIP_CHECK(hipStreamSynchronize(C.stream));
    fork_runs(C);
    for (int n = 0; n < nSteps; n++)
      for (int k = 0; k < nRuns; k++) record_step(C.runs[k], C.runStreams[k], n);
    for (int k = 0; k < nRuns; k++) record_end(C.runs[k], C.runStreams[k]);
    join_runs(C);
    VLS_HIP_CHECK(hipGetLastError());
    VLS_HIP_CHECK(hipStreamSynchronize(C.stream));
  });
}

int vls_adjoint_batch(vls_ctx *ctx, const double *const *dL_dM, const double *const *dL_df,
                      double *const *grad_f0, double *const *grad_vals0) {
  return guarded([&] {
    if (!ctx) throw vls::Error(VLS_ERR_ARG, "null context");
    vls::Context &C = ctx->c;
    const int nRuns = (int)C.runs.size();
    if (!grad_f0) throw vls::Error(VLS_ERR_ARG, "null gradient outputs");
    int nCk = -1;
    for (int k = 0; k < nRuns; k++) {
      adjoint_check(C.runs[k]);
      if (!grad_f0[k]) throw vls::Error(VLS_ERR_ARG, "null gradient output");
      if (nCk >= 0 && C.runs[k].adj.nCk != nCk)
        throw vls::Error(VLS_ERR_STATE, "vls_adjoint_batch: the runs were not recorded together "
                                        "(vls_record_batch)");
      nCk = C.runs[k].adj.nCk;
    }
    fork_runs(C);
    for (int k = 0; k < nRuns; k++)
      adjoint_seed(C.runs[k], C.runStreams[k], dL_dM ? dL_dM[k] : nullptr,
                   dL_df ? dL_df[k] : nullptr);
    for (int sg = nCk - 1; sg >= 0; sg--)
      for (int k = 0; k < nRuns; k++) adjoint_segment(C.runs[k], C.runStreams[k], sg);
    for (int k = 0; k < nRuns; k++)
      adjoint_end(C.runs[k], C.runStreams[k], grad_f0[k], grad_vals0 ? grad_vals0[k] : nullptr);
    join_runs(C);
    VLS_HIP_CHECK(hipGetLastError());
    VLS_HIP_CHECK(hipStreamSynchronize(C.stream));
  });
}

int vls_kernel_times(vls_ctx *ctx, const char **names, double *ms, int *n) {
  return guarded([&] {
    if (!ctx || !n) throw vls::Error(VLS_ERR_ARG, "null argument");
    vls::Timing &t = ctx->c.timing;
    VLS_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    // every kernel: consecutive events; one kernel: (before, after) pairs
    const bool pairs = !t.only.empty();
    int cnt = pairs ? (int)t.ev.size() / 2 : (t.ev.empty() ? 0 : (int)t.ev.size() - 1);
    int cap = *n;
    for (int k = 0; k < cnt && k < cap; k++) {
      const int a = pairs ? 2 * k : k, b = a + 1;
      float x = 0;
      VLS_HIP_CHECK(hipEventElapsedTime(&x, t.ev[a], t.ev[b]));
      if (names) names[k] = t.names[b];
      if (ms) ms[k] = x;
    }
    *n = cnt;
  });
}

}  // extern "C"
