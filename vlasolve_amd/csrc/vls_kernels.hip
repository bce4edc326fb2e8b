// HIP kernels of the spherical spline time step (gfx950, FP64).
//
// Every per-point expression keeps the operation order of the reference's
// C++ (cited per kernel) and the file is compiled with -ffp-contract=off, so
// each kernel rounds exactly as the reference does; the only freedom taken is
// the order of the two reductions (moment and statistics), which is fixed
// and deterministic here.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "vls_internal.hpp"

namespace vls {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double BV2(double x) { return x * x * x; }
__device__ __forceinline__ double BV1(double x) {
  return 1.0 + 3.0 * x + 3.0 * x * x - 3.0 * x * x * x;
}

// modf for the finite, non-negative coordinates used here: integer part by
// truncation, exact fractional remainder (localSpline::pos2Coord,
// src/localSpline.hxx:352-358).
__device__ __forceinline__ void split(double x, int &n, double &t) {
  double ip = trunc(x);
  t = x - ip;
  n = (int)ip;
}

// ---------------------------------------------------------------- moment
// Per (j, r) partial of quadrature21::integrate1D (src/quadrature21.hxx:104-133):
// sum over u < nU-1 of ((Ct1_j*dU_u)*dR_r) * avg4, serial in u.
__global__ void __launch_bounds__(kBlock)
k_moment_partial(const double *__restrict__ f, const double *__restrict__ Ct1,
                 const double *__restrict__ dU, const double *__restrict__ dR,
                 double *__restrict__ P, int nR, int nU, int nJ) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (r >= nR) return;
  double acc = 0.0;
  if (r < nR - 1) {
    const double ct = Ct1[j];
    const double dr = dR[r];
    const double *p = f + (int64_t)j * nU * nR + r;
    double a0 = p[0], a1 = p[1];
    for (int u = 0; u < nU - 1; u++) {
      const double *q = p + (int64_t)(u + 1) * nR;
      double b0 = q[0], b1 = q[1];
      double avg = (((a0 + a1) + b0) + b1) * 0.25;
      acc += ((ct * dU[u]) * dr) * avg;
      a0 = b0;
      a1 = b1;
    }
  }
  P[(int64_t)j * nR + r] = acc;
}

// dM[r] = serial sum over j; exclusive radial prefix (gridHandler::accumulate
// with addSpare=1, src/grid_handler.hxx:350-407); acceleration
// M(<r) * (G/R^2 * dt) (src/vlasovSolver21_spline.hxx:746).
__global__ void __launch_bounds__(kBlock)
k_moment_scan(const double *__restrict__ P, const double *__restrict__ gfac, double *__restrict__ dM,
              double *__restrict__ Mx, double *__restrict__ acc, double *__restrict__ mTot, int nR,
              int nJ) {
  for (int r = threadIdx.x; r < nR; r += blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < nJ; j++) s += P[(int64_t)j * nR + r];
    dM[r] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double prev = 0.0;
    for (int r = 0; r < nR; r++) {
      Mx[r] = prev;
      prev += dM[r];
    }
    *mTot = prev;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nR; r += blockDim.x) acc[r] = Mx[r] * gfac[r];
}

// ---------------------------------------------------------------- spline
// Thomas solve of one padded line in place (create_1D with natural rows,
// src/localSpline.hxx:1045-1172).  rhs(k) for k = 1..M-2 is read through
// `in` with stride `is`; the M coefficients are written through `out` with
// stride `os` (out may alias in: each rhs is read before its slot is written).
__device__ __forceinline__ void thomas_line(const double *in, int64_t is, double *out, int64_t os,
                                            const ThomasAxis &ax) {
  const int M = ax.M;
  double y = (in[0] - ax.c16 * ax.A03) / ax.d[1];
  out[os] = y;
  double ym2 = 0.0, ym1 = y;  // y_{k-2}, y_{k-1}
  for (int k = 2; k <= M - 2; k++) {
    double yk = (in[(int64_t)(k - 1) * is] - ax.c16 * ym1) / ax.d[k];
    out[(int64_t)k * os] = yk;
    ym2 = ym1;
    ym1 = yk;
  }
  // last row: ym2 = y_{M-3}, ym1 = y_{M-2}
  double t = 0.0 - ax.hi2 * ym2;
  t = t - ax.p * ym1;
  double cn = t / ax.q;
  out[(int64_t)(M - 1) * os] = cn;
  double c1 = 0.0, c2 = cn;
  for (int k = M - 2; k >= 1; k--) {
    double ck = out[(int64_t)k * os] - ax.u[k] * cn;
    out[(int64_t)k * os] = ck;
    c2 = cn;
    cn = ck;
    c1 = ck;
  }
  out[0] = (ax.A03 - ax.A01 * c1) - ax.A02 * c2;
}

// r-direction solves of the nU interior rows of every slice
// (create_2D first loop, src/localSpline.hxx:1248-1250).
__global__ void __launch_bounds__(kBlock)
k_spline_r(const double *__restrict__ f, double *__restrict__ c, int nR, int nU, int nJ,
           int64_t cs, ThomasAxis ax) {
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (int64_t)nJ * nU) return;
  int64_t j = row / nU, u = row % nU;
  const double *src = f + row * nR;  // rhs(k) = f[k-1]
  double *dst = c + j * cs + (u + 1) * (nR + 2);
  thomas_line(src, 1, dst, 1, ax);
}

// v-direction solves of all nR+2 padded columns (create_2D second loop,
// src/localSpline.hxx:1252-1254); consecutive lanes take consecutive columns.
__global__ void __launch_bounds__(kBlock)
k_spline_v(double *__restrict__ c, int nR, int nU, int nJ, int64_t cs, ThomasAxis ax) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (col >= nR + 2) return;
  const int64_t st = nR + 2;
  double *tab = c + (int64_t)j * cs + col;
  thomas_line(tab + st, st, tab, st, ax);
}

// ---------------------------------------------------------------- kick
// vlasovSolver21_spline.hxx:750-765 with localSpline::evaluateNxPy
// (src/localSpline.hxx:423-441).
__global__ void __launch_bounds__(kBlock)
k_kick(double *__restrict__ f, const double *__restrict__ c, const double *__restrict__ U_Val,
       const double *__restrict__ acc, int nR, int nU, int nJ, int64_t cs, double Ulo, double Uhi,
       double start_v, double hinv_v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t nVal = (int64_t)nJ * nU * nR;
  if (i >= nVal) return;
  int r = (int)(i % nR);
  int64_t ju = i / nR;
  int u = (int)(ju % nU);
  int64_t j = ju / nU;
  double newU = U_Val[u] + acc[r];
  if ((newU <= Ulo) || (newU >= Uhi)) {
    f[i] = 0;
    return;
  }
  double y = (newU - start_v) * hinv_v;
  int n;
  split(y, n, y);
  const int Ny = 1 + n, Nx = r + 1;
  const int64_t st = nR + 2;
  const double *d = c + j * cs;
  const double *r0 = d + Nx + (int64_t)(Ny - 1) * st;
  const double *r1 = r0 + st;
  const double *r2 = r1 + st;
  const double *r3 = r2 + st;
  double result = 1.0 / 36.0;
  result *= BV2(1. - y) * (r0[-1] + r0[0] * 4 + r0[1]) + BV1(1. - y) * (r1[-1] + r1[0] * 4 + r1[1]) +
            BV1(y) * (r2[-1] + r2[0] * 4 + r2[1]) + BV2(y) * (r3[-1] + r3[0] * 4 + r3[1]);
  f[i] = result;
}

// ---------------------------------------------------------------- drift
// localSpline::evaluateC(Nx,x,Ny,y), src/localSpline.hxx:381-398
__device__ __forceinline__ double eval_c(const double *d, int64_t st, double xr, double xu) {
  int Nx, Ny;
  double x, y;
  split(xr, Nx, x);
  split(xu, Ny, y);
  Nx++;
  Ny++;
  const double *p0 = d + Nx + (int64_t)(Ny - 1) * st;
  const double *p1 = p0 + st;
  const double *p2 = p1 + st;
  const double *p3 = p2 + st;
  const double b2xI = BV2(1. - x);
  const double b1xI = BV1(1. - x);
  const double b2x = BV2(x);
  const double b1x = BV1(x);
  double result = 1.0 / 36.0;
  result *= BV2(1. - y) * (p0[-1] * b2xI + p0[0] * b1xI + p0[1] * b1x + p0[2] * b2x) +
            BV1(1. - y) * (p1[-1] * b2xI + p1[0] * b1xI + p1[1] * b1x + p1[2] * b2x) +
            BV1(y) * (p2[-1] * b2xI + p2[0] * b1xI + p2[1] * b1x + p2[2] * b2x) +
            BV2(y) * (p3[-1] * b2xI + p3[0] * b1xI + p3[1] * b1x + p3[2] * b2x);
  return result;
}

// drift loop, src/vlasovSolver21_spline.hxx:788-796 (delay-line points are
// written by k_chains)
__global__ void __launch_bounds__(kBlock)
k_drift(double *__restrict__ f, const double *__restrict__ c, const double *__restrict__ xr,
        const double *__restrict__ xu, const uint8_t *__restrict__ flag, int nR, int nU,
        int64_t nVal, int64_t cs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nVal) return;
  uint8_t fl = flag[i];
  if (fl == VLS_FOOT_CHAIN) return;
  if (fl == VLS_FOOT_OUT) {
    f[i] = 0;
    return;
  }
  int64_t j = i / ((int64_t)nU * nR);
  f[i] = eval_c(c + j * cs, nR + 2, xr[i], xu[i]);
}

// delayed central kernel: the back cell takes the spline value at its
// position (src/vlasovSolver21_spline.hxx:776-786), the grid point reads the
// front (sphericalKernel::readAt, src/sphericalKernel.hxx:155-161) and the
// ring turns by one (cellChain::roll, src/cellChain.hxx:49-56).
__global__ void __launch_bounds__(kBlock)
k_chains(double *__restrict__ f, const double *__restrict__ c, const int64_t *__restrict__ point,
         const int64_t *__restrict__ off, const int32_t *__restrict__ len,
         int32_t *__restrict__ head, const double *__restrict__ bxr, const double *__restrict__ bxu,
         const uint8_t *__restrict__ bout, double *__restrict__ vals, int64_t nChains, int nR,
         int nU, int64_t cs) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nChains) return;
  int64_t pt = point[k];
  int64_t j = pt / ((int64_t)nU * nR);
  int L = len[k];
  int h = head[k];
  double *v = vals + off[k];
  int back = h + L - 1;
  if (back >= L) back -= L;
  v[back] = bout[k] ? 0.0 : eval_c(c + j * cs, nR + 2, bxr[k], bxu[k]);
  f[pt] = v[h];
  head[k] = (h + 1 == L) ? 0 : h + 1;
}

// ---------------------------------------------------------------- statistics
// Per (j, r) partials, serial in u, of the moment and of the integrands of
// quadrature21::integrate (src/quadrature21.hxx:57-85; integrands
// src/integrand21.hxx:82-145).
__global__ void __launch_bounds__(kBlock)
k_stats_partial(const double *__restrict__ f, const double *__restrict__ Ct1,
                const double *__restrict__ dU, const double *__restrict__ dR,
                const double *__restrict__ R_Val, const double *__restrict__ U_Val,
                const double *__restrict__ J_Val, double *__restrict__ part, int nR, int nU,
                int nJ) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (r >= nR) return;
  double m = 0.0, ek = 0.0, s = 0.0, l1 = 0.0, l2 = 0.0;
  if (r < nR - 1) {
    const double ct = Ct1[j];
    const double dr = dR[r];
    const double rc = R_Val[r] + 0.5 * dr;
    const double J = J_Val[j];
    const double *p = f + (int64_t)j * nU * nR + r;
    double a0 = p[0], a1 = p[1];
    for (int u = 0; u < nU - 1; u++) {
      const double *q = p + (int64_t)(u + 1) * nR;
      double b0 = q[0], b1 = q[1];
      double v = (((a0 + a1) + b0) + b1) * 0.25;
      double du = dU[u];
      double w = (ct * du) * dr;
      double uc = U_Val[u] + 0.5 * du;
      m += w * v;
      ek += w * (0.5 * v * (uc * uc + J * J / (rc * rc)));
      s += w * ((v <= 0) ? 0.0 : -v * log(v));
      l1 += w * fabs(v);
      l2 += w * (v * v);
      a0 = b0;
      a1 = b1;
    }
  }
  const int64_t o = (int64_t)j * nR + r, n = (int64_t)nJ * nR;
  part[o] = m;
  part[n + o] = ek;
  part[2 * n + o] = s;
  part[3 * n + o] = l1;
  part[4 * n + o] = l2;
}

// Serial sums over j (moment) and over r (integrands), then over j; the
// potential energy of vlasovSolver21::potentialEnergy
// (src/vlasovSolver21.hxx:627-642).  out = {M, E, Ek, Ep, S, L1, L2}.
__global__ void __launch_bounds__(kBlock)
k_stats_final(const double *__restrict__ part, const double *__restrict__ R_Val, double G,
              double *__restrict__ dM, double *__restrict__ q, double *__restrict__ out, int nR,
              int nJ) {
  const int64_t n = (int64_t)nJ * nR;
  for (int r = threadIdx.x; r < nR; r += blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < nJ; j++) s += part[(int64_t)j * nR + r];
    dM[r] = s;
  }
  for (int t = threadIdx.x; t < 4 * nJ; t += blockDim.x) {
    int k = t / nJ, j = t % nJ;
    const double *p = part + (int64_t)(k + 1) * n + (int64_t)j * nR;
    double s = 0.0;
    for (int r = 0; r < nR - 1; r++) s += p[r];
    q[t] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double M = dM[0] + 0.0, Ep = 0.0;
    for (int i = 1; i < nR; i++) {
      double rr = 0.5 * (R_Val[i - 1] + R_Val[i]);
      Ep -= (dM[i] * G * M) / rr;
      M += dM[i];
    }
    double tot[4];
    for (int k = 0; k < 4; k++) {
      double s = 0.0;
      for (int j = 0; j < nJ; j++) s += q[k * nJ + j];
      tot[k] = s;
    }
    out[0] = M;
    out[1] = tot[0] + Ep;
    out[2] = tot[0];
    out[3] = Ep;
    out[4] = tot[1];
    out[5] = tot[2];
    out[6] = tot[3];
  }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

void time_mark(Timing *t, hipStream_t s, const char *name) {
  if (!t || !t->on) return;
  hipEvent_t e;
  VLS_HIP_CHECK(hipEventCreate(&e));
  VLS_HIP_CHECK(hipEventRecord(e, s));
  t->ev.push_back(e);
  t->names.push_back(name);
}

void launch_moment(Run &r, hipStream_t s) {
  dim3 g(blocks_for(r.nR), r.nJ);
  hipLaunchKernelGGL(k_moment_partial, g, dim3(kBlock), 0, s, r.f, r.Ct1, r.dU, r.dR, r.P, r.nR,
                     r.nU, r.nJ);
  hipLaunchKernelGGL(k_moment_scan, dim3(1), dim3(kBlock), 0, s, r.P, r.gfac, r.dM, r.Mx, r.acc,
                     r.mTot, r.nR, r.nJ);
}

void launch_spline(Run &r, const double *src, hipStream_t s, Timing *t) {
  hipLaunchKernelGGL(k_spline_r, dim3(blocks_for((int64_t)r.nJ * r.nU)), dim3(kBlock), 0, s, src,
                     r.c, r.nR, r.nU, r.nJ, r.cs, r.tr);
  time_mark(t, s, "spline_r");
  hipLaunchKernelGGL(k_spline_v, dim3(blocks_for(r.nR + 2), r.nJ), dim3(kBlock), 0, s, r.c, r.nR,
                     r.nU, r.nJ, r.cs, r.tv);
  time_mark(t, s, "spline_v");
}

void launch_kick(Run &r, hipStream_t s) {
  hipLaunchKernelGGL(k_kick, dim3(blocks_for(r.nVal)), dim3(kBlock), 0, s, r.f, r.c, r.U_Val,
                     r.acc, r.nR, r.nU, r.nJ, r.cs, r.Ulo, r.Uhi, r.start_v, r.hinv_v);
}

void launch_drift(Run &r, hipStream_t s) {
  hipLaunchKernelGGL(k_drift, dim3(blocks_for(r.nVal)), dim3(kBlock), 0, s, r.f, r.c, r.xr, r.xu,
                     r.flag, r.nR, r.nU, r.nVal, r.cs);
}

void launch_chains(Run &r, hipStream_t s) {
  if (r.nChains == 0) return;
  hipLaunchKernelGGL(k_chains, dim3(blocks_for(r.nChains)), dim3(kBlock), 0, s, r.f, r.c,
                     r.ch_point, r.ch_off, r.ch_len, r.ch_head, r.ch_xr, r.ch_xu, r.ch_out,
                     r.ch_vals, r.nChains, r.nR, r.nU, r.cs);
}

void launch_stats(Run &r, hipStream_t s) {
  dim3 g(blocks_for(r.nR), r.nJ);
  hipLaunchKernelGGL(k_stats_partial, g, dim3(kBlock), 0, s, r.f, r.Ct1, r.dU, r.dR, r.R_Val,
                     r.U_Val, r.J_Val, r.st_part, r.nR, r.nU, r.nJ);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kBlock), 0, s, r.st_part, r.R_Val, r.sp.G,
                     r.st_q + 4 * (int64_t)r.nJ, r.st_q, r.st_out, r.nR, r.nJ);
}

}  // namespace vls
