// Discrete adjoint of the spherical spline step (gfx950, FP64).
//
// The step f_{n+1} = Drift(Spline(Kick(Spline(f_n), a(f_n)))) is linear in
// (f, delay-line values) except through the kick's acceleration a, which is
// linear in f (moment, exclusive radial scan).  Its adjoint, applied to
// lambda_{n+1} = (fb, ringb), is, in reverse order of the forward phases:
//   drift^T   each coefficient gathers fb * (B-spline weight) from the points
//             whose 4x4 foot stencil covers it, through a transposed foot
//             table built once (the feet are static), in point order, so the
//             result does not depend on scheduling; a delay-line point moves
//             its fb into the ring slot it read, and the slot written in this
//             step passes its adjoint to the back cell's stencil (atomics,
//             a few hundred cells);
//   spline^T  the transposed natural-spline solves, v lines then r rows, by
//             reverse-mode differentiation of the Thomas elimination;
//   kick^T    per point, the acceleration term fb* hinv/36 sum_k B_k'(t)
//             g_{n+k}, summed over u per (j, r); per coefficient, the points
//             whose four target rows include it (a window of u), through
//             the column combination (1, 4, 1);
//   spline^T  again, giving the direct part of f_n's adjoint;
//   scan^T    a-bar * G/R^2 dt summed over j, then the suffix sums of the
//             exclusive scan, then the transposed moment quadrature, added to
//             f_n's adjoint.
// None of this has a counterpart in the reference, which has no derivative;
// the forward phases it differentiates are those of
// vlasovSolver21_spline::makeOneStep (src/vlasovSolver21_spline.hxx:680-801).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "vls_internal.hpp"

namespace vls {

namespace {

constexpr int kB = 256;

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

__device__ __forceinline__ double BV2(double x) { return x * x * x; }
__device__ __forceinline__ double BV1(double x) {
  return 1.0 + 3.0 * x + 3.0 * x * x - 3.0 * x * x * x;
}
__device__ __forceinline__ double dBV2(double x) { return 3.0 * x * x; }
__device__ __forceinline__ double dBV1(double x) { return 3.0 + 6.0 * x - 9.0 * x * x; }

__device__ __forceinline__ void split(double x, int &n, double &t) {
  double ip = trunc(x);
  t = x - ip;
  n = (int)ip;
}

__global__ void __launch_bounds__(kB) k_fill(double *__restrict__ p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = v;
}

// Adjoint of evaluateC at spline coordinates (xr, xu): scatter lam times the
// 16 weights into the coefficient adjoint of the slice.
__device__ __forceinline__ void eval_c_adj(double *cb, int st, double xr, double xu, double lam) {
  int nx, ny;
  double x, y;
  split(xr, nx, x);
  split(xu, ny, y);
  const double wx[4] = {BV2(1. - x), BV1(1. - x), BV1(x), BV2(x)};
  const double wy[4] = {BV2(1. - y), BV1(1. - y), BV1(y), BV2(y)};
  const double l36 = lam / 36.0;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const double la = l36 * wy[a];
    double *row = cb + (int64_t)(ny + a) * st + nx;
#pragma unroll
    for (int b = 0; b < 4; b++) atomicAdd(row + b, la * wx[b]);
  }
}

// drift^T with the delay lines.  head0/stepOff: the ring heads at step n are
// (head0 + stepOff) mod len.
__global__ void __launch_bounds__(kB)
k_adj_drift(const double *__restrict__ fb, double *__restrict__ cb, const double *__restrict__ xr,
            const double *__restrict__ xu, int64_t nVal, int nSlice, int64_t cs, int st,
            const int64_t *__restrict__ off, const int32_t *__restrict__ len,
            const int32_t *__restrict__ head0, int64_t stepOff, const double *__restrict__ bxr,
            const double *__restrict__ bxu, const uint8_t *__restrict__ bout,
            double *__restrict__ ringb, int64_t nChains) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= nVal) return;
  const double x = xr[i];
  const double lam = fb[i];
  double *cbj = cb + (i / nSlice) * cs;
  if (x > -1.0) {
    if (lam != 0.0) eval_c_adj(cbj, st, x, xu[i], lam);
  } else if (x == -2.0) {
    const double kd = xu[i];
    if (!(kd >= 0) || kd >= (double)nChains) return;
    const int64_t k = (int64_t)kd;
    const int L = len[k];
    const int h = (int)((head0[k] + stepOff) % L);
    const int back = (h + L - 1) % L;
    double *rb = ringb + off[k];
    rb[h] += lam;  // f[point] read the front slot
    const double xb = rb[back];  // the back slot was written from the spline
    rb[back] = 0.0;
    if (!bout[k] && xb != 0.0) eval_c_adj(cbj, st, bxr[k], bxu[k], xb);
  }
}

// Reverse-mode derivative of the natural-spline line solve (the Thomas
// elimination of the forward kernels): in place, c-bar of the M padded
// coefficients in, rhs-bar of rows 1..M-2 out.
template <class ACC>
__device__ __forceinline__ void thomas_adjoint(const ThomasAxis &ax, ACC c) {
  const int M = ax.M;
  const double c0 = c(0);
  c(1) -= ax.A01 * c0;
  c(2) -= ax.A02 * c0;
  for (int k = 1; k <= M - 2; k++) c(k + 1) -= ax.u[k] * c(k);
  const double tt = c(M - 1) / ax.q;
  c(M - 3) -= ax.hi2 * tt;
  c(M - 2) -= ax.p * tt;
  for (int k = M - 2; k >= 1; k--) {
    const double r = c(k) / ax.d[k];
    c(k) = r;
    if (k >= 2) c(k - 1) -= ax.c16 * r;
  }
}

// v lines (lane = padded column), in place
__global__ void __launch_bounds__(64)
k_adj_v(double *__restrict__ cb, int S, int64_t cs, ThomasAxis ax) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= S) return;
  double *col = cb + blockIdx.y * cs + k;
  thomas_adjoint(ax, [&](int m) -> double & { return col[(int64_t)m * S]; });
}

// r rows (lane = padded row 1..nU), rows staged through LDS so that the
// global accesses are row segments; rhs-bar of row u+1 goes to out[j][u][.]
__global__ void __launch_bounds__(64)
k_adj_r(const double *__restrict__ cb, double *__restrict__ out, int nR, int nU, int64_t cs,
        ThomasAxis ax) {
  extern __shared__ double tile[];  // [64][S + 1]
  const int S = nR + 2, P = S + 1;
  const int j = blockIdx.y;
  const int u0 = blockIdx.x * 64;
  const int rows = min(64, nU - u0);
  const double *src = cb + j * cs + (int64_t)(u0 + 1) * S;
  for (int e = threadIdx.x; e < rows * S; e += 64) tile[(e / S) * P + e % S] = src[e];
  __syncthreads();
  if (threadIdx.x < rows) {
    double *row = tile + threadIdx.x * P;
    thomas_adjoint(ax, [&](int m) -> double & { return row[m]; });
  }
  __syncthreads();
  double *dst = out + ((int64_t)j * nU + u0) * nR;
  for (int e = threadIdx.x; e < rows * nR; e += 64) dst[e] = tile[(e / nR) * P + e % nR + 1];
}

// kick^T per (j, r): coefficient adjoint of the column combination g (into
// gb, zeroed before; the thread owns column r) and the acceleration adjoint
__global__ void __launch_bounds__(kB)
k_adj_kick(const double *__restrict__ fsb, const double *__restrict__ c,
           const double *__restrict__ U_Val, const double *__restrict__ acc,
           double *__restrict__ gb, double *__restrict__ abp, int nR, int nU, int64_t cs,
           double Ulo, double Uhi, double start_v, double hinv_v) {
  const int r = blockIdx.x * kB + threadIdx.x;
  const int j = blockIdx.y;
  if (r >= nR) return;
  const int S = nR + 2;
  const double a = acc[r];
  const double *cj = c + j * cs;
  double *gj = gb + j * cs;
  double ab = 0.0;
  for (int u = 0; u < nU; u++) {
    const double lam = fsb[((int64_t)j * nU + u) * nR + r];
    if (lam == 0.0) continue;
    const double newU = U_Val[u] + a;
    if ((newU <= Ulo) || (newU >= Uhi)) continue;
    double t;
    int n;
    split((newU - start_v) * hinv_v, n, t);
    const double w[4] = {BV2(1. - t), BV1(1. - t), BV1(t), BV2(t)};
    const double dw[4] = {-dBV2(1. - t), -dBV1(1. - t), dBV1(t), dBV2(t)};
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (n + k > nU + 1) break;  // past the padded column: weight 0 (target on the last knot)
      const double *rowc = cj + (int64_t)(n + k) * S + r;
      const double g = (rowc[0] + rowc[1] * 4) + rowc[2];
      s += dw[k] * g;
      gj[(int64_t)(n + k) * S + r] += lam * w[k] / 36.0;
    }
    ab += lam * (hinv_v / 36.0) * s;
  }
  abp[(int64_t)j * nR + r] = ab;
}

// c-bar[m][k] = gb[m][k] + 4 gb[m][k-1] + gb[m][k-2] over the columns r < nR
__global__ void __launch_bounds__(kB)
k_adj_kick_gather(const double *__restrict__ gb, double *__restrict__ cb, int nR, int nU,
                  int64_t cs, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= n) return;
  const int S = nR + 2;
  const int64_t j = e / cs, rem = e - j * cs;
  const int k = (int)(rem % S);
  const double *g = gb + e;
  double v = 0.0;
  if (k < nR) v += g[0];
  if (k >= 1 && k - 1 < nR) v += 4.0 * g[-1];
  if (k >= 2 && k - 2 < nR) v += g[-2];
  cb[e] = v;
}

// scan^T: Mbar[r] = (sum_j abp[j][r]) * gfac[r] (or the given Mbar), then
// dMbar[k] = sum_{r > k} Mbar[r].  One block.
__global__ void __launch_bounds__(256)
k_adj_scan(const double *__restrict__ abp, const double *__restrict__ MbarIn,
           const double *__restrict__ gfac, double *__restrict__ dMb, int nR, int nJ) {
  extern __shared__ double mb[];
  for (int r = threadIdx.x; r < nR; r += blockDim.x) {
    double v;
    if (MbarIn) {
      v = MbarIn[r];
    } else {
      constexpr int kJ = 16;
      double s = 0.0;
      int j = 0;
      for (; j + kJ <= nJ; j += kJ) {
        double t[kJ];
#pragma unroll
        for (int q = 0; q < kJ; q++) t[q] = abp[(int64_t)(j + q) * nR + r];
#pragma unroll
        for (int q = 0; q < kJ; q++) s += t[q];
      }
      for (; j < nJ; j++) s += abp[(int64_t)j * nR + r];
      v = s * gfac[r];
    }
    mb[r] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = nR - 1; k >= 0; k--) {
      const double v = mb[k];
      mb[k] = s;
      s += v;
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nR; r += blockDim.x) dMb[r] = mb[r];
}

// moment^T: fb[j][u][r] (+)= sum over the quadrature cells (u', r') that
// contain the point of 0.25 ((Ct1_j dU_u') dR_r') dMbar[r']
__global__ void __launch_bounds__(kB)
k_adj_moment(double *__restrict__ fb, const double *__restrict__ dMb,
             const double *__restrict__ Ct1, const double *__restrict__ dU,
             const double *__restrict__ dR, int nR, int nU, int64_t nVal, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= nVal) return;
  const int64_t nSlice = (int64_t)nU * nR;
  const int j = (int)(i / nSlice);
  const int il = (int)(i - j * nSlice);
  const int u = il / nR, r = il - u * nR;
  const double ct = Ct1[j];
  double s = 0.0;
#pragma unroll
  for (int du = -1; du <= 0; du++) {
    const int uu = u + du;
    if (uu < 0 || uu > nU - 2) continue;
    const double cu = ct * dU[uu];
#pragma unroll
    for (int dr = -1; dr <= 0; dr++) {
      const int rr = r + dr;
      if (rr < 0 || rr > nR - 2) continue;
      s += 0.25 * ((cu * dR[rr]) * dMb[rr]);
    }
  }
  fb[i] = accumulate ? fb[i] + s : s;
}

// drift^T as a gather: coefficient e (global index j*cs + m*S + k) takes
// lam(p) * wy[a] * wx[b] / 36 for each entry (p, a*4 + b) of its list, p the
// point's index inside slice j.  ent[e0..e1) are in increasing point order.
__global__ void __launch_bounds__(kB)
k_adj_drift_gather(const double *__restrict__ fb, double *__restrict__ cb,
                   const double *__restrict__ xr, const double *__restrict__ xu,
                   const int64_t *__restrict__ off, const uint32_t *__restrict__ ent,
                   int64_t nSlice, int64_t cs, int64_t nCoef) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= nCoef) return;
  const int64_t j = e / cs;
  const double *fj = fb + j * nSlice;
  const double *xrj = xr + j * nSlice, *xuj = xu + j * nSlice;
  // entries in batches of kQ, all loads of a batch issued before their use
  constexpr int kQ = 8;
  double acc = 0.0;
  const int64_t q0 = off[e], q1 = off[e + 1];
  for (int64_t q = q0; q < q1; q += kQ) {
    uint32_t v[kQ];
    double lam[kQ], xx[kQ], yy[kQ];
#pragma unroll
    for (int t = 0; t < kQ; t++) v[t] = ent[q + t < q1 ? q + t : q0];
#pragma unroll
    for (int t = 0; t < kQ; t++) {
      const int64_t p = v[t] >> 4;
      lam[t] = fj[p];
      xx[t] = xrj[p];
      yy[t] = xuj[p];
    }
#pragma unroll
    for (int t = 0; t < kQ; t++) {
      if (q + t >= q1 || lam[t] == 0.0) continue;
      int nx, ny;
      double x, y;
      split(xx[t], nx, x);
      split(yy[t], ny, y);
      const int a = (v[t] >> 2) & 3, b = v[t] & 3;
      const double wx = b == 0 ? BV2(1. - x) : b == 1 ? BV1(1. - x) : b == 2 ? BV1(x) : BV2(x);
      const double wy = a == 0 ? BV2(1. - y) : a == 1 ? BV1(1. - y) : a == 2 ? BV1(y) : BV2(y);
      acc += lam[t] / 36.0 * wy * wx;
    }
  }
  cb[e] = acc;
}

// drift^T of the delay lines, one thread per chain (after the gather)
__global__ void __launch_bounds__(kB)
k_adj_chains(const double *__restrict__ fb, double *__restrict__ cb,
             const int64_t *__restrict__ point, const int64_t *__restrict__ off,
             const int32_t *__restrict__ len, const int32_t *__restrict__ head0, int64_t stepOff,
             const double *__restrict__ bxr, const double *__restrict__ bxu,
             const uint8_t *__restrict__ bout, double *__restrict__ ringb, int64_t nChains,
             int64_t nSlice, int64_t cs, int st) {
  const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (k >= nChains) return;
  const int64_t pt = point[k];
  const int L = len[k];
  const int h = (int)((head0[k] + stepOff) % L);
  const int back = (h + L - 1) % L;
  double *rb = ringb + off[k];
  rb[h] += fb[pt];
  const double xb = rb[back];
  rb[back] = 0.0;
  if (!bout[k] && xb != 0.0) eval_c_adj(cb + (pt / nSlice) * cs, st, bxr[k], bxu[k], xb);
}

// kick^T, acceleration part: term[i] = fb*[i] hinv/36 sum_k B_k'(t) g_{n+k}
__global__ void __launch_bounds__(kB)
k_adj_kick_terms(const double *__restrict__ fsb, const double *__restrict__ c,
                 const double *__restrict__ U_Val, const double *__restrict__ acc,
                 double *__restrict__ term, int nR, int nU, int64_t cs, int64_t nVal, double Ulo,
                 double Uhi, double start_v, double hinv_v) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i >= nVal) return;
  const int64_t nSlice = (int64_t)nU * nR;
  const int64_t j = i / nSlice;
  const int il = (int)(i - j * nSlice);
  const int u = il / nR, r = il - u * nR;
  const double lam = fsb[i];
  double out = 0.0;
  const double newU = U_Val[u] + acc[r];
  if (lam != 0.0 && !(newU <= Ulo) && !(newU >= Uhi)) {
    double t;
    int n;
    split((newU - start_v) * hinv_v, n, t);
    const double dw[4] = {-dBV2(1. - t), -dBV1(1. - t), dBV1(t), dBV2(t)};
    const int S = nR + 2;
    const double *cj = c + j * cs;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (n + k > nU + 1) break;  // past the padded column: weight 0
      const double *rowc = cj + (int64_t)(n + k) * S + r;
      s += dw[k] * ((rowc[0] + rowc[1] * 4) + rowc[2]);
    }
    out = lam * (hinv_v / 36.0) * s;
  }
  term[i] = out;
}

// abp[j][r] = sum over u of term[j][u][r]
__global__ void __launch_bounds__(kB)
k_adj_kick_acc(const double *__restrict__ term, double *__restrict__ abp, int nR, int nU) {
  const int r = blockIdx.x * kB + threadIdx.x;
  const int j = blockIdx.y;
  if (r >= nR) return;
  const double *t = term + (int64_t)j * nU * nR + r;
  constexpr int kU = 16;
  double s = 0.0;
  int u = 0;
  for (; u + kU <= nU; u += kU) {
    double v[kU];
#pragma unroll
    for (int q = 0; q < kU; q++) v[q] = t[(int64_t)(u + q) * nR];
#pragma unroll
    for (int q = 0; q < kU; q++) s += v[q];
  }
  for (; u < nU; u++) s += t[(int64_t)u * nR];
  abp[(int64_t)j * nR + r] = s;
}

// kick^T, coefficient part: cb[j][m][k] = g(m, k) + 4 g(m, k-1) + g(m, k-2),
// g(m, r) = sum over the points u of column r whose rows n..n+3 include m of
// fb*[u] B_{m-n}(t) / 36.  The target row of point u is n = u + a_r hinv up to
// rounding, so the points of row m lie in a short window of u.
__global__ void __launch_bounds__(kB)
k_adj_kick_coef(const double *__restrict__ fsb, const double *__restrict__ U_Val,
                const double *__restrict__ acc, double *__restrict__ cb, int nR, int nU,
                int64_t cs, int64_t nCoef, double Ulo, double Uhi, double start_v,
                double hinv_v) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= nCoef) return;
  const int S = nR + 2;
  const int64_t j = e / cs;
  const int rem = (int)(e - j * cs);
  const int m = rem / S, k = rem - m * S;
  const double *fj = fsb + j * (int64_t)nU * nR;
  double out = 0.0;
#pragma unroll
  for (int dr = 0; dr < 3; dr++) {
    const int r = k - dr;
    if (r < 0 || r >= nR) continue;
    const double a = acc[r];
    const double sh = a * hinv_v;
    // n = u + a hinv up to rounding: rows n..n+3 include m for u in
    // [m - 3 - a hinv, m - a hinv], widened by two on each side
    const int ulo = (int)floor((double)m - 3.0 - sh) - 2;
    constexpr int kW = 10;
    double lam[kW], uu[kW];
#pragma unroll
    for (int q = 0; q < kW; q++) {
      const int u = min(max(ulo + q, 0), nU - 1);
      lam[q] = fj[(int64_t)u * nR + r];
      uu[q] = U_Val[u];
    }
    double g = 0.0;
#pragma unroll
    for (int q = 0; q < kW; q++) {
      const int u = ulo + q;
      if (u < 0 || u >= nU || lam[q] == 0.0) continue;
      const double newU = uu[q] + a;
      if ((newU <= Ulo) || (newU >= Uhi)) continue;
      double t;
      int n;
      split((newU - start_v) * hinv_v, n, t);
      const int k4 = m - n;
      if (k4 < 0 || k4 > 3) continue;
      const double w = k4 == 0 ? BV2(1. - t) : k4 == 1 ? BV1(1. - t) : k4 == 2 ? BV1(t) : BV2(t);
      g += lam[q] * w / 36.0;
    }
    out += (dr == 1 ? 4.0 : 1.0) * g;
  }
  cb[e] = out;
}

// Line adjoints in registers for a line length known at compile time
// (the reverse of k_line / k_line_rows); divisions become products with
// the tabulated reciprocals.
template <int M>
__device__ __forceinline__ void thomas_adjoint_reg(const ThomasAxis &ax, double (&c)[M], int lz) {
  c[1] -= ax.A01 * c[0];
  c[2] -= ax.A02 * c[0];
#pragma unroll
  for (int k = 1; k <= M - 2; k++) c[k + 1] -= ax.u[k + lz] * c[k];
  const double tt = c[M - 1] * ax.rq;
  c[M - 3] -= ax.hi2 * tt;
  c[M - 2] -= ax.p * tt;
#pragma unroll
  for (int k = M - 2; k >= 1; k--) {
    c[k] = c[k] * ax.rd[k + lz];
    if (k >= 2) c[k - 1] -= ax.c16 * c[k];
  }
}

template <int M, int S>
__global__ void __launch_bounds__(64)
k_adj_v_reg(double *__restrict__ cb, int64_t cs, int lzmask, ThomasAxis ax) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int lz = threadIdx.x & lzmask;
  if (k >= S) return;
  double *col = cb + blockIdx.y * cs + k;
  double c[M];
#pragma unroll
  for (int m = 0; m < M; m++) c[m] = col[(int64_t)m * S];
  thomas_adjoint_reg<M>(ax, c, lz);
#pragma unroll
  for (int m = 1; m <= M - 2; m++) col[(int64_t)m * S] = c[m];
}

template <int M>
__global__ void __launch_bounds__(64)
k_adj_r_reg(const double *__restrict__ cb, double *__restrict__ out, int nU, int64_t cs,
            int lzmask, ThomasAxis ax) {
  constexpr int nR = M - 2, P = M | 1;  // odd pitch: the lanes' row reads spread over the banks
  extern __shared__ double tile[];  // [64][P]
  const int j = blockIdx.y;
  const int u0 = blockIdx.x * 64;
  const int rows = min(64, nU - u0);
  const int lz = threadIdx.x & lzmask;
  const double *src = cb + j * cs + (int64_t)(u0 + 1) * M;
  for (int e = threadIdx.x; e < rows * M; e += 64) tile[(e / M) * P + e % M] = src[e];
  __syncthreads();
  if (threadIdx.x < rows) {
    double *row = tile + threadIdx.x * P;
    double c[M];
#pragma unroll
    for (int m = 0; m < M; m++) c[m] = row[m];
    thomas_adjoint_reg<M>(ax, c, lz);
#pragma unroll
    for (int m = 1; m <= M - 2; m++) row[m] = c[m];
  }
  __syncthreads();
  double *dst = out + ((int64_t)j * nU + u0) * nR;
  for (int e = threadIdx.x; e < rows * nR; e += 64) dst[e] = tile[(e / nR) * P + e % nR + 1];
}

}  // namespace

void launch_fill(double *p, int64_t n, double v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kB), 0, s, p, n, v);
}

// register line adjoints for the shapes of the forward's register solvers
using AdjV = void (*)(double *, int64_t, int, ThomasAxis);
using AdjR = void (*)(const double *, double *, int, int64_t, int, ThomasAxis);
struct AdjShape {
  int nR, nU;
  AdjV v;
  AdjR r;
};
#define VLS_ADJ_SHAPE(R, U) {R, U, k_adj_v_reg<(U) + 2, (R) + 2>, k_adj_r_reg<(R) + 2>}
static const AdjShape kAdjShapes[] = {
    VLS_ADJ_SHAPE(17, 17),   VLS_ADJ_SHAPE(25, 13),   VLS_ADJ_SHAPE(33, 33),
    VLS_ADJ_SHAPE(49, 49),   VLS_ADJ_SHAPE(65, 65),   VLS_ADJ_SHAPE(129, 129),
    VLS_ADJ_SHAPE(201, 201),
};
#undef VLS_ADJ_SHAPE

static const AdjShape *adj_shape(const Run &r) {
  for (const AdjShape &a : kAdjShapes)
    if (a.nR == r.nR && a.nU == r.nU) return &a;
  return nullptr;
}

void launch_adj_drift(Run &r, const double *fb, double *cb, double *ringb, const int32_t *head0,
                      int64_t stepOff, hipStream_t s) {
  const int64_t nCoef = (int64_t)r.nJ * r.cs;
  hipLaunchKernelGGL(k_adj_drift_gather, dim3(nblocks(nCoef)), dim3(kB), 0, s, fb, cb, r.xr, r.xu,
                     r.adj.csrOff, r.adj.csrEnt, (int64_t)r.nU * r.nR, r.cs, nCoef);
  if (r.nChains)
    hipLaunchKernelGGL(k_adj_chains, dim3(nblocks(r.nChains)), dim3(kB), 0, s, fb, cb, r.ch_point,
                       r.ch_off, r.ch_len, head0, stepOff, r.ch_xr, r.ch_xu, r.ch_out, ringb,
                       r.nChains, (int64_t)r.nU * r.nR, r.cs, r.nR + 2);
}

void launch_adj_spline(Run &r, double *cb, double *out, hipStream_t s) {
  const size_t lds = 64 * (size_t)(r.nR + 3) * sizeof(double);
  if (const AdjShape *a = adj_shape(r)) {
    hipLaunchKernelGGL(a->v, dim3((r.nR + 2 + 63) / 64, r.nJ), dim3(64), 0, s, cb, r.cs, 0, r.tv);
    hipLaunchKernelGGL(a->r, dim3((r.nU + 63) / 64, r.nJ), dim3(64), lds, s, cb, out, r.nU, r.cs, 0,
                       r.tr);
    return;
  }
  hipLaunchKernelGGL(k_adj_v, dim3((r.nR + 2 + 63) / 64, r.nJ), dim3(64), 0, s, cb, r.nR + 2, r.cs,
                     r.tv);
  hipLaunchKernelGGL(k_adj_r, dim3((r.nU + 63) / 64, r.nJ), dim3(64), lds, s, cb, out, r.nR, r.nU,
                     r.cs, r.tr);
}

void launch_adj_kick(Run &r, const double *fsb, const double *cK, const double *accK, double *term,
                     double *cb, double *abp, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_kick_terms, dim3(nblocks(r.nVal)), dim3(kB), 0, s, fsb, cK, r.U_Val,
                     accK, term, r.nR, r.nU, r.cs, r.nVal, r.Ulo, r.Uhi, r.start_v, r.hinv_v);
  hipLaunchKernelGGL(k_adj_kick_acc, dim3(nblocks(r.nR), r.nJ), dim3(kB), 0, s, term, abp, r.nR,
                     r.nU);
  const int64_t nCoef = (int64_t)r.nJ * r.cs;
  hipLaunchKernelGGL(k_adj_kick_coef, dim3(nblocks(nCoef)), dim3(kB), 0, s, fsb, r.U_Val, accK, cb,
                     r.nR, r.nU, r.cs, nCoef, r.Ulo, r.Uhi, r.start_v, r.hinv_v);
}

void launch_adj_mass(Run &r, const double *abp, const double *MbarIn, double *dMb, double *fb,
                     bool accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_scan, dim3(1), dim3(256), r.nR * sizeof(double), s, abp, MbarIn,
                     r.gfac, dMb, r.nR, r.nJ);
  hipLaunchKernelGGL(k_adj_moment, dim3(nblocks(r.nVal)), dim3(kB), 0, s, fb, dMb, r.Ct1, r.dU,
                     r.dR, r.nR, r.nU, r.nVal, accumulate ? 1 : 0);
}

void configure_adjoint(int maxLdsPerBlock, int nR) {
  const size_t lds = 64 * (size_t)(nR + 3) * sizeof(double);
  if (lds > (size_t)maxLdsPerBlock)
    throw Error(VLS_ERR_ARG, "grid too large for the adjoint's row tiles");
  VLS_HIP_CHECK(hipFuncSetAttribute((const void *)k_adj_r,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  for (const AdjShape &a : kAdjShapes)
    if (a.nR == nR)
      VLS_HIP_CHECK(hipFuncSetAttribute((const void *)a.r,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

// The transposed foot table of the drift (built on the host from the folded
// foot table): for every coefficient of the padded slices, the points whose
// 4x4 stencil covers it with a non-zero weight, in point order.
void build_drift_transpose(Run &r, const std::vector<double> &xr, const std::vector<double> &xu,
                           std::vector<int64_t> &off, std::vector<uint32_t> &ent) {
  const int64_t nSlice = (int64_t)r.nU * r.nR, nCoef = (int64_t)r.nJ * r.cs;
  const int S = r.nR + 2, Mv = r.nU + 2;
  if (nSlice >= (int64_t)1 << 28) throw Error(VLS_ERR_ARG, "slice too large for the adjoint's tables");
  off.assign(nCoef + 1, 0);
  auto each = [&](auto &&fn) {
    for (int64_t j = 0; j < r.nJ; j++)
      for (int64_t p = 0; p < nSlice; p++) {
        const int64_t i = j * nSlice + p;
        const double x = xr[i];
        if (x <= -1.0) continue;  // outside the grid or a delay line
        const double ix = std::trunc(x), iy = std::trunc(xu[i]);
        const double fx = x - ix, fy = xu[i] - iy;
        const int nx = (int)ix, ny = (int)iy;
        for (int a = 0; a < 4; a++) {
          if (a == 3 && fy == 0.0) continue;  // weight BV2(0) = 0
          for (int b = 0; b < 4; b++) {
            if (b == 3 && fx == 0.0) continue;
            const int row = ny + a, col = nx + b;
            if (row >= Mv || col >= S) continue;
            fn(j * r.cs + (int64_t)row * S + col, (uint32_t)((p << 4) | (a << 2) | b));
          }
        }
      }
  };
  each([&](int64_t e, uint32_t) { off[e + 1]++; });
  for (int64_t e = 0; e < nCoef; e++) off[e + 1] += off[e];
  ent.assign(off[nCoef], 0);
  std::vector<int64_t> cur(off.begin(), off.end() - 1);
  each([&](int64_t e, uint32_t v) { ent[cur[e]++] = v; });
}

}  // namespace vls
