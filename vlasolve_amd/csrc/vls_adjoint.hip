// Discrete adjoint of the spherical spline step (gfx950, FP64).
//
// The step f_{n+1} = Drift(Spline(Kick(Spline(f_n), a(f_n)))) is linear in
// (f, delay-line values) except through the kick's acceleration a, which is
// linear in f (moment, exclusive radial scan).  Its adjoint, applied to
// lambda_{n+1} = (fb, ringb), is, in reverse order of the forward phases:
//   drift^T   per tile of coefficients, the points whose 4x4 foot stencil
//             touches it (a table built once; the feet are static) add
//             fb * (B-spline weight) into the tile in LDS, and the tile is
//             stored; a delay-line point moves its fb into the ring slot it
//             read, and the slot written in this step passes its adjoint to
//             the back cell's stencil (an entry of the same tile lists);
//   spline^T  the transposed natural-spline solves, v lines then r rows, by
//             reverse-mode differentiation of the Thomas elimination;
//   kick^T    per point, the acceleration term fb* hinv/36 sum_k B_k'(t)
//             g_{n+k}, summed over u per (j, r); the direct part through the
//             kick's equivalent 1D v-spline of each column (see
//             k_adj_kick_coef): per row of the column, the points whose four
//             target rows include it (a window of u), then the transposed
//             v-solve, giving the direct part of f_n's adjoint;
//   scan^T    a-bar * G/R^2 dt summed over j, then the suffix sums of the
//             exclusive scan, then the transposed moment quadrature, added to
//             f_n's adjoint as the column solve stores it.
// None of this has a counterpart in the reference, which has no derivative;
// the forward phases it differentiates are those of
// vlasovSolver21_spline::makeOneStep (src/vlasovSolver21_spline.hxx:680-801).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "vls_internal.hpp"
#include "vls_tile.hpp"

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

// Weight of node b (0..3) of a cubic B-spline segment at fraction x, the
// same value as BV2(1-x), BV1(1-x), BV1(x), BV2(x) for b = 0..3 up to
// rounding: with t = x for b >= 2 and 1 - x otherwise, t^3 for the outer
// nodes and 1 + 3t + 3t^2 - 3t^3 for the inner ones.  Branch-free, for lanes
// whose b differ.
__device__ __forceinline__ double bweight(int b, double x) {
  const double t = b >= 2 ? x : 1.0 - x;
  const double t3 = t * t * t;
  const double inner = fma(t, fma(t, fma(-3.0, t, 3.0), 3.0), 1.0);
  return (b == 0 || b == 3) ? t3 : inner;
}

__device__ __forceinline__ void split(double x, int &n, double &t) {
  double ip = trunc(x);
  t = x - ip;
  n = (int)ip;
}

__global__ void __launch_bounds__(kB) k_fill(double *__restrict__ p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = v;
}

// The transposed moment quadrature factors over the cell pairs around a
// point: 1/4 (cu(u-1) + cu(u)) (w(r-1) + w(r)), with cu(u) = Ct1[j] dU[u] and
// w(r) = dR[r] dMb[r] for the cells that exist (u <= nU-2, r <= nR-2) and 0
// otherwise.  mom_cu is a row's factor, mom_w a column's.
__device__ __forceinline__ double mom_cu(int u, int nU, double ct, const double *__restrict__ dU) {
  const double a = (u >= 1 && u - 1 <= nU - 2) ? ct * dU[u - 1] : 0.0;
  const double b = (u >= 0 && u <= nU - 2) ? ct * dU[u] : 0.0;
  return 0.25 * (a + b);
}
__device__ __forceinline__ double mom_w(int r, int nR, const double *__restrict__ dR,
                                        const double *__restrict__ dMb) {
  const double a = (r >= 1 && r - 1 <= nR - 2) ? dR[r - 1] * dMb[r - 1] : 0.0;
  const double b = (r >= 0 && r <= nR - 2) ? dR[r] * dMb[r] : 0.0;
  return a + b;
}

// Transposed moment quadrature at point (u, r) of a slice (the reverse of
// the cell averages of k_moment_cols): a quarter of each adjacent cell's
// weight Ct1[j] dU dR times that cell's dM adjoint.
__device__ __forceinline__ double moment_adj(int u, int r, int nU, int nR, double ct,
                                             const double *__restrict__ dU,
                                             const double *__restrict__ dR,
                                             const double *__restrict__ dMb) {
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
  return s;
}

// Line adjoints for any line length M (the reverse of the LDS-tiled forward
// solvers): the reverse-mode derivative of the Thomas elimination, run as two
// streaming sweeps over the padded line, a chunk of kLC elements at a time.
// The ascending sweep (c1 -= A01 c0, c2 -= A02 c0, c[t] -= u[t-1] c[t-1])
// writes its values back in place; the descending one applies the last row
// (tt = c[M-1] / q into c[M-3] and c[M-2]), scales by 1/d[k] and carries
// c16 c[k] down to c[k-1].  Only the carry lives across chunks, so nothing
// of the line has to fit on chip.
constexpr int kLC = 16;

// Pivot table rows read from the table: past kHead the pivots sit at their
// fixed point (uS, rdS), which every line adjoint uses from a register.
constexpr int kHead = 16;

// one element of the descending sweep: element k (1..M-2) of the line after
// the ascending sweep, the carry from element k+1 in and out
__device__ __forceinline__ double adj_down(const ThomasAxis &ax, int k, double v, double tt,
                                           double &carry) {
  v -= carry;
  if (k == ax.M - 2) v = fma(-ax.p, tt, v);
  if (k == ax.M - 3) v = fma(-ax.hi2, tt, v);
  v *= k <= kHead ? ax.rd[k] : ax.rdS;  // the table is the fixed point past kHead
  carry = ax.c16 * v;
  return v;
}

// The head rows of the pivot tables (u, rd for rows 0..kHead) in LDS, for the
// streaming line adjoints: a table load inside their loops would wait for
// every load in flight (the vector memory counter retires in order), and
// the chunk after the current one is loaded ahead.  Every thread of the
// block calls it (a barrier).
struct AdjHead {
  double u[kHead + 1], rd[kHead + 1];
};
__device__ __forceinline__ void adj_head_load(AdjHead &h, const ThomasAxis &ax) {
  const int t = threadIdx.x;
  if (t <= kHead && t < ax.M) {
    h.u[t] = ax.u[t];
    h.rd[t] = ax.rd[t];
  }
  __syncthreads();
}
// adj_down with the head rows from LDS
__device__ __forceinline__ double adj_down_h(const AdjHead &h, const ThomasAxis &ax, int k,
                                             double v, double tt, double &carry) {
  v -= carry;
  if (k == ax.M - 2) v = fma(-ax.p, tt, v);
  if (k == ax.M - 3) v = fma(-ax.hi2, tt, v);
  v *= k <= kHead ? h.rd[k] : ax.rdS;
  carry = ax.c16 * v;
  return v;
}

// v lines (lane = padded column), in place: coalesced rows of 64 columns,
// a chunk of kLC rows at a time with the next chunk's loads in flight
__global__ void __launch_bounds__(64)
k_adj_v(double *__restrict__ cb, int S, int64_t cs, ThomasAxis ax) {
  __shared__ AdjHead hd;
  adj_head_load(hd, ax);
  const Tile2 tl = xcd_tile2((S + 63) / 64);
  const int k = tl.x * 64 + threadIdx.x;
  if (k >= S) return;  // no barriers below
  const int M = ax.M;
  double *col = cb + tl.j * cs + k;
  double v[kLC], nv[kLC];
  auto load_up = [&](int m0) {
#pragma unroll
    for (int i = 0; i < kLC; i++)
      if (m0 + i < M) nv[i] = col[(int64_t)(m0 + i) * S];
  };
  auto load_down = [&](int m1) {
#pragma unroll
    for (int i = 0; i < kLC; i++)
      if (m1 - i >= 1) nv[i] = col[(int64_t)(m1 - i) * S];
  };
  double c0 = 0.0, prev = 0.0;
  load_up(0);
  for (int m0 = 0; m0 < M; m0 += kLC) {
#pragma unroll
    for (int i = 0; i < kLC; i++) v[i] = nv[i];
    if (m0 + kLC < M) load_up(m0 + kLC);
#pragma unroll
    for (int i = 0; i < kLC; i++) {
      const int t = m0 + i;
      if (t < M) {
        if (t == 0) c0 = v[i];
        if (t == 1) v[i] = fma(-ax.A01, c0, v[i]);
        if (t == 2) v[i] = fma(-ax.A02, c0, v[i]);
        if (t >= 2) v[i] = fma(-(t - 1 <= kHead ? hd.u[t - 1] : ax.uS), prev, v[i]);
        prev = v[i];
      }
    }
#pragma unroll
    for (int i = 0; i < kLC; i++)
      if (m0 + i < M) col[(int64_t)(m0 + i) * S] = v[i];
  }
  const double tt = prev * ax.rq;  // prev = c[M-1]
  double carry = 0.0;
  load_down(M - 2);
  for (int m1 = M - 2; m1 >= 1; m1 -= kLC) {  // elements m1, m1-1, ... down to max(1, m1-kLC+1)
#pragma unroll
    for (int i = 0; i < kLC; i++) v[i] = nv[i];
    if (m1 - kLC >= 1) load_down(m1 - kLC);
#pragma unroll
    for (int i = 0; i < kLC; i++)
      if (m1 - i >= 1) v[i] = adj_down_h(hd, ax, m1 - i, v[i], tt, carry);
#pragma unroll
    for (int i = 0; i < kLC; i++)
      if (m1 - i >= 1) col[(int64_t)(m1 - i) * S] = v[i];
  }
}

// The kick's transpose for columns too long for registers, in two kernels
// around the scan.  k_adj_kick_walk: one lane per column r walks the points
// u in order and, for each point in the v range, adds its acceleration term
// fb*[u] sum_k B_k'(t) d_{n+k} (d the kept v-spline coefficients of the
// column) to the column's sum, and its B_k(t) / 6 fb*[u] to rows n..n+3 of the
// column's g, kept as a window of four rows.  n never decreases with u, so a
// row is complete once a point targets rows above it: it then takes its
// step of the ascending sweep of the transposed v-solve and is stored over g
// (g[j][m][r], slice stride cs, row stride nR).  k_adj_gcol_down then runs
// the descending sweep and writes fb (plus the transposed moment quadrature).
__global__ void __launch_bounds__(64)
k_adj_kick_walk(const double *__restrict__ fsb, const double *__restrict__ c,
                const double *__restrict__ U_Val, const double *__restrict__ acc,
                double *__restrict__ g, double *__restrict__ abp, int nR, int nU, int64_t cs,
                double Ulo, double Uhi, double start_v, double hinv_v, ThomasAxis ax) {
  extern __shared__ double Us[];  // [nU]
  // the head pivots in LDS: a table load inside the walk's loop would make
  // every row's step wait for all the points' loads in flight (the vector
  // memory counter retires in order)
  __shared__ double uh[kHead + 1];
  for (int u = threadIdx.x; u < nU; u += 64) Us[u] = U_Val[u];
  if (threadIdx.x <= kHead && (int)threadIdx.x < ax.M) uh[threadIdx.x] = ax.u[threadIdx.x];
  __syncthreads();
  const Tile2 tl = xcd_tile2((nR + 63) / 64);
  const int r = tl.x * 64 + threadIdx.x;
  const int j = tl.j;
  if (r >= nR) return;  // no barriers below
  const int Mv = nU + 2;
  const double a = acc[r];
  const double *fj = fsb + (int64_t)j * nU * nR + r;
  const double *dj = c + j * cs + r;
  double *gc = g + j * cs + r;
  double w0 = 0, w1 = 0, w2 = 0, w3 = 0;  // g of rows base..base+3
  int base = 0;
  double c0 = 0, prev = 0, sAcc = 0;
  // row m of g is complete: its step of the ascending sweep
  // (c1 -= A01 c0, c2 -= A02 c0, c[t] -= u[t-1] c[t-1]), stored in place
  auto finish = [&]() {
    const int m = base;
    double v = w0;
    if (m == 0) c0 = v;
    if (m == 1) v = fma(-ax.A01, c0, v);
    if (m == 2) v = fma(-ax.A02, c0, v);
    if (m >= 2) v = fma(-(m - 1 <= kHead ? uh[m - 1] : ax.uS), prev, v);
    prev = v;
    gc[(int64_t)m * nR] = v;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = 0;
    base++;
  };
  // kQ points ahead: the loads of point u+kQ (its value, its target row and
  // the coefficient rows there) fly while points u..u+kQ-1 go through the
  // window; the v nodes come from LDS
  // (the target row comes from the v nodes alone, so that the point's five
  // loads leave together; whether its value is 0 is looked at when it is used)
  auto load = [&](int u, double &lam, double &t, int &n, bool &ok, double (&gk)[4]) {
    const int uc = min(u, nU - 1);
    lam = fj[(int64_t)uc * nR];
    const double newU = Us[uc] + a;
    ok = u < nU && !(newU <= Ulo) && !(newU >= Uhi);
    split(ok ? (newU - start_v) * hinv_v : 0.0, n, t);
#pragma unroll
    for (int k = 0; k < 4; k++) gk[k] = dj[(int64_t)min(n + k, nU + 1) * nR];
  };
  constexpr int kQ = 4;
  double lamQ[kQ], tQ[kQ], gkQ[kQ][4];
  int nQ[kQ];
  bool okQ[kQ];
#pragma unroll
  for (int q = 0; q < kQ; q++) load(q, lamQ[q], tQ[q], nQ[q], okQ[q], gkQ[q]);
  for (int u0 = 0; u0 < nU; u0 += kQ) {
#pragma unroll
    for (int q = 0; q < kQ; q++) {
      if (okQ[q] && lamQ[q] != 0.0) {  // (false past the last point)
        const double lam = lamQ[q], t = tQ[q];
        const int n = nQ[q];
        // acceleration term over the rows n..n+3 inside the padded column
        const double dw[4] = {-dBV2(1. - t), -dBV1(1. - t), dBV1(t), dBV2(t)};
        double sk = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) sk += n + k <= nU + 1 ? dw[k] * gkQ[q][k] : 0.0;
        sAcc += lam * sk;
        while (base < n) finish();
        const double l6 = lam * (1.0 / 6.0);
        w0 += l6 * BV2(1. - t);
        w1 += l6 * BV1(1. - t);
        w2 += l6 * BV1(t);
        w3 += l6 * BV2(t);
      }
      load(u0 + q + kQ, lamQ[q], tQ[q], nQ[q], okQ[q], gkQ[q]);
    }
  }
  while (base < Mv) finish();
  abp[(int64_t)j * nR + r] = sAcc * (hinv_v / 6.0);
}

// The descending sweep of the kick's transposed v-solve over the rows the
// walk left in g (see k_adj_kick_walk), rows 1..nU to fb[j][u][r] plus, with
// dMb, the transposed moment quadrature.
__global__ void __launch_bounds__(64)
k_adj_gcol_down(const double *__restrict__ g, double *__restrict__ fb, int nR, int nU,
                int64_t cs, ThomasAxis ax, const double *__restrict__ dMb,
                const double *__restrict__ Ct1, const double *__restrict__ dU,
                const double *__restrict__ dR) {
  __shared__ AdjHead hd;
  adj_head_load(hd, ax);
  const Tile2 tl = xcd_tile2((nR + 63) / 64);
  const int r = tl.x * 64 + threadIdx.x;
  const int j = tl.j;
  if (r >= nR) return;  // no barriers below
  const int M = ax.M;
  const double *col = g + j * cs + r;
  const double tt = col[(int64_t)(M - 1) * nR] * ax.rq;
  const double ct = dMb ? Ct1[j] : 0.0, w = dMb ? mom_w(r, nR, dR, dMb) : 0.0;
  double *out = fb + (int64_t)j * nU * nR + r;
  double carry = 0.0;
  double v[kLC], nv[kLC];
  auto load_down = [&](int m1) {  // the next chunk's rows, in flight during this one
#pragma unroll
    for (int i = 0; i < kLC; i++)
      if (m1 - i >= 1) nv[i] = col[(int64_t)(m1 - i) * nR];
  };
  load_down(M - 2);
  for (int m1 = M - 2; m1 >= 1; m1 -= kLC) {
#pragma unroll
    for (int i = 0; i < kLC; i++) v[i] = nv[i];
    if (m1 - kLC >= 1) load_down(m1 - kLC);
#pragma unroll
    for (int i = 0; i < kLC; i++) {
      const int m = m1 - i;
      if (m >= 1) {
        v[i] = adj_down_h(hd, ax, m, v[i], tt, carry);
        out[(int64_t)(m - 1) * nR] = dMb ? v[i] + mom_cu(m - 1, nU, ct, dU) * w : v[i];
      }
    }
  }
}

// r rows (lane = padded row u+1 of 64 consecutive rows), rhs-bar of row u+1
// to out[j][u][.]: chunks of kLC columns through an LDS transpose, so that
// the global accesses are row segments; the ascending sweep writes back into
// cb (a work array here), the descending one into out.
__global__ void __launch_bounds__(64)
k_adj_r(double *__restrict__ cb, double *__restrict__ out, int nR, int nU, int64_t cs,
        ThomasAxis ax) {
  constexpr int kP = kLC + 1;
  __shared__ double chunk[64 * kP];
  __shared__ AdjHead hd;
  adj_head_load(hd, ax);
  const int M = nR + 2;
  const int lane = threadIdx.x;
  const int j = blockIdx.y;
  const int u0 = blockIdx.x * 64;
  const int rows = min(64, nU - u0);
  const int rsub = lane / kLC, ksub = lane % kLC;  // staging role: 4 rows x kLC columns per pass
  double *src = cb + (size_t)j * cs + (size_t)(u0 + 1) * M;
  double *dst = out + ((size_t)j * nU + u0) * nR;
  double v[kLC], tn[kLC], tc[kLC];
  // rows x columns [e0, e0 + kLC) of a (rows x width) array: the loads into
  // tn (the chunk after the current one is loaded while this one is
  // solved), then from tc through LDS into the lane-row view v
  auto gload = [&](const double *a, int width, int e0) {
    const int e = e0 + ksub;
#pragma unroll
    for (int q = 0; q < kLC; q++) {
      const int row = min(q * 4 + rsub, rows - 1);
      tn[q] = a[(size_t)row * width + min(e, width - 1)];
    }
  };
  auto gput = [&](int width, int e0) {
    const int e = e0 + ksub;
#pragma unroll
    for (int q = 0; q < kLC; q++) chunk[(q * 4 + rsub) * kP + ksub] = e < width ? tc[q] : 0.0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kLC; i++) v[i] = chunk[lane * kP + i];
    __syncthreads();
  };
  auto scatter = [&](double *a, int width, int e0) {
#pragma unroll
    for (int i = 0; i < kLC; i++) chunk[lane * kP + i] = v[i];
    __syncthreads();
    const int e = e0 + ksub;
#pragma unroll
    for (int q = 0; q < kLC; q++) {
      const int row = q * 4 + rsub;
      if (row < rows && e < width) a[(size_t)row * width + e] = chunk[row * kP + ksub];
    }
    __syncthreads();
  };
  double c0 = 0.0, prev = 0.0;
  gload(src, M, 0);
  for (int m0 = 0; m0 < M; m0 += kLC) {
#pragma unroll
    for (int q = 0; q < kLC; q++) tc[q] = tn[q];
    if (m0 + kLC < M) gload(src, M, m0 + kLC);  // columns this chunk's stores do not touch
    gput(M, m0);
#pragma unroll
    for (int i = 0; i < kLC; i++) {
      const int t = m0 + i;
      if (t < M) {
        if (t == 0) c0 = v[i];
        if (t == 1) v[i] = fma(-ax.A01, c0, v[i]);
        if (t == 2) v[i] = fma(-ax.A02, c0, v[i]);
        if (t >= 2) v[i] = fma(-(t - 1 <= kHead ? hd.u[t - 1] : ax.uS), prev, v[i]);
        prev = v[i];
      }
    }
    scatter(src, M, m0);
  }
  const double tt = prev * ax.rq;
  double carry = 0.0;
  // descending over chunks of the output columns e = k - 1 (k = 1..M-2)
  const int eLast = ((nR - 1) / kLC) * kLC;
  gload(src, M, eLast + 1);
  for (int e0 = eLast; e0 >= 0; e0 -= kLC) {
#pragma unroll
    for (int q = 0; q < kLC; q++) tc[q] = tn[q];
    if (e0 - kLC >= 0) gload(src, M, e0 - kLC + 1);
    gput(M, e0 + 1);
#pragma unroll
    for (int i = kLC - 1; i >= 0; i--)
      if (e0 + i < nR) v[i] = adj_down_h(hd, ax, e0 + i + 1, v[i], tt, carry);
    scatter(dst, nR, e0);
  }
}

// scan^T: Mbar[r] = (sum_j abp[j][r]) * gfac[r] (or the given Mbar), then
// dMbar[k] = sum_{r > k} Mbar[r].  One block of kAdjScanThreads: a thread
// per radius (up to that many), the slices' terms kJ at a time in flight
// (the block is alone on the GPU, so each batch is a round trip to memory),
// each radius still adding them in slice order.
constexpr int kAdjScanThreads = 1024;
__global__ void __launch_bounds__(kAdjScanThreads)
k_adj_scan(const double *__restrict__ abp, const double *__restrict__ MbarIn,
           const double *__restrict__ gfac, double *__restrict__ dMb, int nR, int nJ) {
  extern __shared__ double mb[];
  for (int r = threadIdx.x; r < nR; r += blockDim.x) {
    double v;
    if (MbarIn) {
      v = MbarIn[r];
    } else {
      constexpr int kJ = 64;
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
    // descending, kS entries a batch: the batch's LDS reads leave together
    // (one at a time, each would wait for the one before), then its adds in
    // order
    constexpr int kS = 16;
    int k = nR;
    for (; k >= kS; k -= kS) {
      double v[kS];
#pragma unroll
      for (int q = 0; q < kS; q++) v[q] = mb[k - 1 - q];
#pragma unroll
      for (int q = 0; q < kS; q++) {
        const double d = v[q];
        v[q] = s;
        s += d;
      }
#pragma unroll
      for (int q = 0; q < kS; q++) mb[k - 1 - q] = v[q];
    }
    for (k = k - 1; k >= 0; k--) {
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
  const double s = moment_adj(u, r, nU, nR, Ct1[j], dU, dR, dMb);
  fb[i] = accumulate ? fb[i] + s : s;
}

// drift^T over tiles of kTile x kTile coefficients of a padded slice: the
// block (one wave) takes the points whose 4x4 foot stencil touches its tile
// (a list built once on the host; the feet are static), adds each point's
// lam/36 * wy[a] * wx[b] into the tile in LDS and stores the tile.  Every
// coefficient belongs to one tile, so the global stores need no atomics,
// and the table holds about 1.4 entries per point (16 for a list per
// coefficient).  The points come in increasing order and one wave adds them,
// a chunk of 64 at a time.  After a tile's points come the delay lines whose
// back cell's stencil touches it (entry kTileChain | k): their adjoint xb[k]
// (k_adj_chain_rings) goes in at the back foot (bxr[k], bxu[k]) the same way.
constexpr int kTile = 16;
constexpr uint32_t kTileChain = 0x80000000u;
__global__ void __launch_bounds__(64)
k_adj_drift_tiles(const double *__restrict__ fb, double *__restrict__ cb,
                  const double2 *__restrict__ xy, const int64_t *__restrict__ toff,
                  const uint32_t *__restrict__ tpts, const double *__restrict__ xb,
                  const double *__restrict__ bxr, const double *__restrict__ bxu,
                  int64_t nSlice, int64_t cs, int nR, int nU, int ntk) {
  __shared__ double acc[kTile * kTile];
  const int tile = blockIdx.x, j = blockIdx.y;
  const int tm = tile / ntk, tk = tile - tm * ntk;
  const int m0 = tm * kTile, k0 = tk * kTile;
  const int S = nR + 2, Mv = nU + 2;
  for (int i = threadIdx.x; i < kTile * kTile; i += 64) acc[i] = 0.0;
  __syncthreads();
  const int64_t t = (int64_t)j * gridDim.x + tile;
  const int64_t q1 = toff[t + 1];
  const double *fj = fb + j * nSlice;
  const double2 *xyj = xy + j * nSlice;
  for (int64_t q = toff[t] + threadIdx.x; q < q1; q += 64) {
    const uint32_t p = tpts[q];
    double lam;
    double2 ft;
    if (p & kTileChain) {
      const uint32_t k = p & ~kTileChain;
      lam = xb[k] * (1.0 / 36.0);
      ft = make_double2(bxr[k], bxu[k]);
    } else {
      lam = fj[p] * (1.0 / 36.0);
      ft = xyj[p];
    }
    int nx, ny;
    double x, y;
    split(ft.x, nx, x);
    split(ft.y, ny, y);
    const double wx[4] = {BV2(1. - x), BV1(1. - x), BV1(x), BV2(x)};
    const double wy[4] = {BV2(1. - y), BV1(1. - y), BV1(y), BV2(y)};
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int m = ny + a - m0;
      if (m < 0 || m >= kTile) continue;
      const double la = lam * wy[a];
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int k = nx + b - k0;
        if (k >= 0 && k < kTile) atomicAdd(&acc[m * kTile + k], la * wx[b]);
      }
    }
  }
  __syncthreads();
  double *cj = cb + j * cs;
  for (int i = threadIdx.x; i < kTile * kTile; i += 64) {
    const int m = m0 + i / kTile, k = k0 + i % kTile;
    if (m < Mv && k < S) cj[(int64_t)m * S + k] = acc[i];
  }
}

// drift^T of the delay lines, one thread per chain: the point's adjoint goes
// into the ring slot it read, and the slot written in the step hands its
// adjoint (0 when the back cell is outside the grid) to xb[k], which
// k_adj_drift_tiles adds into the back cell's stencil.
__global__ void __launch_bounds__(kB)
k_adj_chain_rings(const double *__restrict__ fb, const int64_t *__restrict__ point,
                  const int64_t *__restrict__ off, const int32_t *__restrict__ len,
                  const int32_t *__restrict__ head0, int64_t stepOff,
                  const uint8_t *__restrict__ bout, double *__restrict__ ringb,
                  double *__restrict__ xb, int64_t nChains) {
  const int64_t k = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (k >= nChains) return;
  const int L = len[k];
  const int h = (int)((head0[k] + stepOff) % L);
  const int back = (h + L - 1) % L;
  double *rb = ringb + off[k];
  rb[h] += fb[point[k]];
  const double x = rb[back];
  rb[back] = 0.0;
  xb[k] = bout[k] ? 0.0 : x;
}

// kick^T, acceleration part: abp[j][r] = hinv/6 sum_u fb*[j][u][r]
// sum_k B_k'(t) d_{n+k}, d the kept v-spline coefficients of column r.
// A block takes kAccCols columns; its kAccGroups groups of kAccCols lanes
// (two groups to a wave) take every kAccGroups-th row u and add up their
// rows.  The two groups of a wave are combined by a lane swap, the waves'
// sums through LDS.  Narrow column tiles give each lane fewer rows of
// dependent loads and the grid more blocks.
constexpr int kAccCols = 32, kAccGroups = 32;
__global__ void __launch_bounds__(kAccCols * kAccGroups)
k_adj_kick_accel(const double *__restrict__ fsb, const double *__restrict__ c,
                 const double *__restrict__ U_Val, const double *__restrict__ acc,
                 double *__restrict__ abp, int nR, int nU, int64_t cs, double Ulo, double Uhi,
                 double start_v, double hinv_v) {
  __shared__ double part[kAccGroups / 2][kAccCols];
  const int lane = threadIdx.x & (kAccCols - 1), g = threadIdx.x / kAccCols;
  const Tile2 tl = xcd_tile2((nR + kAccCols - 1) / kAccCols);
  const int r = tl.x * kAccCols + lane;
  const int j = tl.j;
  double s = 0.0;
  if (r < nR) {
    const double a = acc[r];
    const double *fj = fsb + (int64_t)j * nU * nR + r;
    const double *dj = c + j * cs + r;
    // straight-line body (points outside the v range read row 0 and count
    // 0), two rows per trip so that their loads overlap
#pragma unroll 2
    for (int u = g; u < nU; u += kAccGroups) {
      const double lam = fj[(int64_t)u * nR];
      const double newU = U_Val[u] + a;
      const bool in = !(newU <= Ulo) && !(newU >= Uhi);
      double t;
      int n;
      split(in ? (newU - start_v) * hinv_v : 0.0, n, t);
      const double dw[4] = {-dBV2(1. - t), -dBV1(1. - t), dBV1(t), dBV2(t)};
      double sk = 0.0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        // past the padded column the weight is 0: read the last row instead
        const double dk = dj[(int64_t)min(n + k, nU + 1) * nR];
        sk += n + k <= nU + 1 ? dw[k] * dk : 0.0;
      }
      s += in ? lam * sk : 0.0;
    }
  }
  s += __shfl_xor(s, kAccCols);  // lanes l and l + 32: groups 2w and 2w + 1
  if ((threadIdx.x & 63) < kAccCols) part[threadIdx.x >> 6][lane] = s;
  __syncthreads();
  if (threadIdx.x < kAccCols && r < nR) {
    double t = part[0][lane];
#pragma unroll
    for (int q = 1; q < kAccGroups / 2; q++) t += part[q][lane];
    abp[(int64_t)j * nR + r] = t * (hinv_v / 6.0);
  }
}

// kick^T, coefficient part.  The forward kick is the 1D natural v-spline of
// each column of f, d = V^-1 f, evaluated with the weights B_k(t) / 6 (the
// column form of the reference's 2D evaluation on r nodes; see
// k_kick_line).  Its transpose therefore needs no r-direction adjoint:
// g(m, r) = sum over the points u of
// column r whose rows n..n+3 include m of fb*[u] B_{m-n}(t) / 6, then the
// transposed v-solve of each column (k_adj_gcol).
//
// g goes to the coefficient buffer as [j][m][r] (slice stride cs, row
// stride nR).  The target row of point u is n = floor(u + a_r hinv + d_u),
// d_u the v node's offset from its knot plus rounding, below 0.01 of a cell
// (checked when the run is recorded), so the points of row m satisfy
// m - 3 - a_r hinv - 0.01 < u < m + 1 - a_r hinv + 0.01: the kW = 5 values of
// u from floor(m - 3 - a_r hinv - 0.01) + 1 on.  One thread per (m, r).
__global__ void __launch_bounds__(kB)
k_adj_kick_coef(const double *__restrict__ fsb, const double *__restrict__ U_Val,
                const double *__restrict__ acc, double *__restrict__ g, int nR, int nU,
                int64_t cs, double Ulo, double Uhi, double start_v, double hinv_v) {
  // one thread per (m, r) of the slice's nU + 2 rows, r fastest
  const int i = blockIdx.x * kB + threadIdx.x, j = blockIdx.y;
  if (i >= (nU + 2) * nR) return;
  const int m = i / nR, r = i - m * nR;
  const double *fj = fsb + (int64_t)j * nU * nR;
  double *gm = g + (int64_t)j * cs + (int64_t)m * nR;
  const double a = acc[r];
  const double sh = a * hinv_v;
  const int ulo = (int)floor((double)m - 3.0 - sh - 0.01) + 1;
  constexpr int kW = 5;
  double lam[kW], uu[kW];
#pragma unroll
  for (int q = 0; q < kW; q++) {
    const int u = min(max(ulo + q, 0), nU - 1);
    lam[q] = fj[(int64_t)u * nR + r];
    uu[q] = U_Val[u];
  }
  double acc6 = 0.0;
#pragma unroll
  for (int q = 0; q < kW; q++) {
    const int u = ulo + q;
    if (u < 0 || u >= nU) continue;
    const double newU = uu[q] + a;
    if ((newU <= Ulo) || (newU >= Uhi)) continue;
    double t;
    int n;
    split((newU - start_v) * hinv_v, n, t);
    const int k4 = m - n;
    if (k4 < 0 || k4 > 3) continue;
    acc6 += lam[q] * bweight(k4, t);
  }
  gm[r] = acc6 * (1.0 / 6.0);
}

// Line adjoints in registers for a line length known at compile time
// (the reverse of k_line / k_line_rows); divisions become products with
// the tabulated reciprocals.  Past row kHead the pivots are at their fixed
// point (the table repeats uS and rdS there; the set-up checks K0 - 1 <=
// kHead), so only the head is read from the table.  The adjoint is not held
// to the forward's operation order, so its updates are fused multiply-adds.
template <int M>
__device__ __forceinline__ void thomas_adjoint_reg(const ThomasAxis &ax, double (&c)[M], int lz) {
  const double uS = ax.uS + (double)lz, rdS = ax.rdS + (double)lz;
  c[1] = fma(-ax.A01, c[0], c[1]);
  c[2] = fma(-ax.A02, c[0], c[2]);
#pragma unroll
  for (int k = 1; k <= M - 2; k++) {
    const double uv = k <= kHead ? ax.u[k + lz] : uS;
    c[k + 1] = fma(-uv, c[k], c[k + 1]);
  }
  const double tt = c[M - 1] * ax.rq;
  c[M - 3] = fma(-ax.hi2, tt, c[M - 3]);
  c[M - 2] = fma(-ax.p, tt, c[M - 2]);
#pragma unroll
  for (int k = M - 2; k >= 1; k--) {
    c[k] *= k <= kHead ? ax.rd[k + lz] : rdS;
    if (k >= 2) c[k - 1] = fma(-ax.c16, c[k], c[k - 1]);
  }
}

template <int M, int S>
__global__ void __launch_bounds__(64)
k_adj_v_reg(double *__restrict__ cb, int64_t cs, int lzmask, ThomasAxis ax) {
  const Tile2 tl = xcd_tile2((S + 63) / 64);
  const int k = tl.x * 64 + threadIdx.x;
  const int lz = threadIdx.x & lzmask;
  if (k >= S) return;
  double *col = cb + tl.j * cs + k;
  double c[M];
#pragma unroll
  for (int m = 0; m < M; m++) c[m] = col[(int64_t)m * S];
  thomas_adjoint_reg<M>(ax, c, lz);
#pragma unroll
  for (int m = 1; m <= M - 2; m++) col[(int64_t)m * S] = c[m];
}

// The kick's transposed v-solve: the columns r of g ([j][m][r], M = nU + 2
// rows) in registers, their rows 1..nU to fb[j][u][r] plus, with dMb, the
// transposed moment quadrature.
template <int M>
__global__ void __launch_bounds__(64)
k_adj_gcol(const double *__restrict__ g, double *__restrict__ fb, int nR, int64_t cs, int lzmask,
           ThomasAxis ax, const double *__restrict__ dMb, const double *__restrict__ Ct1,
           const double *__restrict__ dU, const double *__restrict__ dR) {
  constexpr int nU = M - 2;
  const Tile2 tl = xcd_tile2((nR + 63) / 64);
  const int r = tl.x * 64 + threadIdx.x;
  const int j = tl.j;
  const int lz = threadIdx.x & lzmask;
  if (r >= nR) return;
  const double *col = g + j * cs + r;
  double c[M];
#pragma unroll
  for (int m = 0; m < M; m++) c[m] = col[(int64_t)m * nR];
  thomas_adjoint_reg<M>(ax, c, lz);
  double *out = fb + (int64_t)j * nU * nR + r;
  if (dMb) {
    const double ct = Ct1[j], w = mom_w(r, nR, dR, dMb);
#pragma unroll
    for (int u = 0; u < nU; u++) out[(int64_t)u * nR] = c[u + 1] + mom_cu(u, nU, ct, dU) * w;
  } else {
#pragma unroll
    for (int u = 0; u < nU; u++) out[(int64_t)u * nR] = c[u + 1];
  }
}

// r-row adjoint with coalesced memory access, the transpose of k_line_rows:
// a wave takes 64 consecutive rows u0.. of slice j.  The padded coefficient
// rows of the adjoint arrive in chunks of kC columns through a small LDS
// transpose and feed the ascending sweep of the transposed elimination as
// they come (the loads of the next chunk fly meanwhile); the descending
// sweep releases the adjoint of f a chunk at a time, last chunk first.
// Loads are clamped into the wave's rows; the extra lanes of a partial wave
// solve a copy and store nothing.
template <int M>
__global__ void __launch_bounds__(64)
k_adj_rows(const double *__restrict__ cb, double *__restrict__ out, int nU, int64_t cs,
           int lzmask, ThomasAxis ax) {
  constexpr int kC = 16, kP = kC + 1;
  constexpr int nR = M - 2;
  constexpr int nChIn = (M + kC - 1) / kC, nChOut = (nR + kC - 1) / kC;
  __shared__ double chunk[64 * kP];
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const int j = blockIdx.y;
  const int u0 = blockIdx.x * 64;
  const int rows = min(64, nU - u0);
  const int rsub = lane / kC, ksub = lane % kC;  // staging role: 4 rows x kC columns per pass
  const double *src = cb + (size_t)j * cs + (size_t)(u0 + 1) * M;
  double *dst = out + ((size_t)j * nU + u0) * nR;
  const double uS = ax.uS + (double)lz, rdS = ax.rdS + (double)lz;
  double c[M];
  double tmp[kC];
  auto load_chunk = [&](int ch) {
    const int e = min(ch * kC + ksub, M - 1);
#pragma unroll
    for (int q = 0; q < kC; q++) tmp[q] = src[min(q * 4 + rsub, rows - 1) * M + e];
  };
  load_chunk(0);
#pragma unroll
  for (int ch = 0; ch < nChIn; ch++) {
#pragma unroll
    for (int q = 0; q < kC; q++) chunk[(q * 4 + rsub) * kP + ksub] = tmp[q];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kC; i++)
      if (ch * kC + i < M) c[ch * kC + i] = chunk[lane * kP + i];
    __syncthreads();
    if (ch + 1 < nChIn) load_chunk(ch + 1);
#pragma unroll
    for (int i = 0; i < kC; i++) {
      const int t = ch * kC + i;  // c[t] -= u[t-1] c[t-1], ascending
      if (t == 2) {
        c[1] = fma(-ax.A01, c[0], c[1]);
        c[2] = fma(-ax.A02, c[0], c[2]);
      }
      if (t >= 2 && t <= M - 1) {
        const double uv = t - 1 <= kHead ? ax.u[t - 1 + lz] : uS;
        c[t] = fma(-uv, c[t - 1], c[t]);
      }
    }
  }
  const double tt = c[M - 1] * ax.rq;
  c[M - 3] = fma(-ax.hi2, tt, c[M - 3]);
  c[M - 2] = fma(-ax.p, tt, c[M - 2]);
#pragma unroll
  for (int ch = nChOut - 1; ch >= 0; ch--) {
#pragma unroll
    for (int i = kC - 1; i >= 0; i--) {
      const int k = ch * kC + i + 1;  // produces output column k - 1
      if (k <= M - 2) {
        c[k] *= k <= kHead ? ax.rd[k + lz] : rdS;
        if (k >= 2) c[k - 1] = fma(-ax.c16, c[k], c[k - 1]);
      }
    }
#pragma unroll
    for (int i = 0; i < kC; i++)
      if (ch * kC + i < nR) chunk[lane * kP + i] = c[ch * kC + i + 1];
    __syncthreads();
    // all LDS reads first, so that their latencies overlap, then the stores
    const int e = ch * kC + ksub;
    double sv[kC];
#pragma unroll
    for (int q = 0; q < kC; q++) sv[q] = chunk[(q * 4 + rsub) * kP + ksub];
    __syncthreads();
    if (rows == 64 && (ch + 1) * kC <= nR) {
#pragma unroll
      for (int q = 0; q < kC; q++) dst[(q * 4 + rsub) * nR + e] = sv[q];
    } else {
#pragma unroll
      for (int q = 0; q < kC; q++) {
        const int row = q * 4 + rsub;
        if (row < rows && e < nR) dst[row * nR + e] = sv[q];
      }
    }
  }
}

}  // namespace

void launch_fill(double *p, int64_t n, double v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kB), 0, s, p, n, v);
}

// register line adjoints for the shapes of the forward's register solvers
// (up to 131 nodes: at 203 the line transposes spill, and the streaming
// adjoints reverse a 201x201x200 step in 0.575 ms against 0.626)
using AdjV = void (*)(double *, int64_t, int, ThomasAxis);
using AdjR = void (*)(const double *, double *, int, int64_t, int, ThomasAxis);
using AdjG = void (*)(const double *, double *, int, int64_t, int, ThomasAxis, const double *,
                      const double *, const double *, const double *);
struct AdjShape {
  int nR, nU;
  AdjV v;
  AdjR r;
  AdjG gcol;
};
#define VLS_ADJ_SHAPE(R, U) \
  {R, U, k_adj_v_reg<(U) + 2, (R) + 2>, k_adj_rows<(R) + 2>, k_adj_gcol<(U) + 2>}
static const AdjShape kAdjShapes[] = {
    VLS_ADJ_SHAPE(17, 17),   VLS_ADJ_SHAPE(25, 13),   VLS_ADJ_SHAPE(33, 33),
    VLS_ADJ_SHAPE(49, 49),   VLS_ADJ_SHAPE(65, 65),   VLS_ADJ_SHAPE(129, 129),
};
#undef VLS_ADJ_SHAPE

static const AdjShape *adj_shape(const Run &r) {
  for (const AdjShape &a : kAdjShapes)
    if (a.nR == r.nR && a.nU == r.nU) return &a;
  return nullptr;
}

void launch_adj_drift(Run &r, const double *fb, double *cb, double *ringb, const int32_t *head0,
                      int64_t stepOff, hipStream_t s) {
  const int ntm = (r.nU + 2 + kTile - 1) / kTile, ntk = (r.nR + 2 + kTile - 1) / kTile;
  if (r.nChains)
    hipLaunchKernelGGL(k_adj_chain_rings, dim3(nblocks(r.nChains)), dim3(kB), 0, s, fb,
                       r.ch_point, r.ch_off, r.ch_len, head0, stepOff, r.ch_out, ringb,
                       r.adj.chainXb, r.nChains);
  hipLaunchKernelGGL(k_adj_drift_tiles, dim3(ntm * ntk, r.nJ), dim3(64), 0, s, fb, cb,
                     (const double2 *)r.feet, r.adj.tileOff, r.adj.tilePts, r.adj.chainXb,
                     r.ch_xr, r.ch_xu, (int64_t)r.nU * r.nR, r.cs, r.nR, r.nU, ntk);
}

void launch_adj_spline(Run &r, double *cb, double *out, hipStream_t s) {
  if (const AdjShape *a = adj_shape(r)) {
    hipLaunchKernelGGL(a->v, dim3((r.nR + 2 + 63) / 64 * r.nJ), dim3(64), 0, s, cb, r.cs, 0, r.tv);
    hipLaunchKernelGGL(a->r, dim3((r.nU + 63) / 64, r.nJ), dim3(64), 0, s, cb, out, r.nU, r.cs, 0,
                       r.tr);
    return;
  }
  hipLaunchKernelGGL(k_adj_v, dim3((r.nR + 2 + 63) / 64 * r.nJ), dim3(64), 0, s, cb, r.nR + 2, r.cs,
                     r.tv);
  hipLaunchKernelGGL(k_adj_r, dim3((r.nU + 63) / 64, r.nJ), dim3(64), 0, s, cb, out, r.nR, r.nU,
                     r.cs, r.tr);
}

void launch_adj_kick(Run &r, const double *fsb, const double *cK, const double *accK, double *cb,
                     double *abp, hipStream_t s) {
  if (!adj_shape(r)) {  // long columns: the walk (its ascending sweep included)
    hipLaunchKernelGGL(k_adj_kick_walk, dim3((r.nR + 63) / 64 * r.nJ), dim3(64),
                       r.nU * sizeof(double), s, fsb, cK,
                       r.U_Val, accK, cb, abp, r.nR, r.nU, r.cs, r.Ulo, r.Uhi, r.start_v, r.hinv_v,
                       r.tv);
    return;
  }
  hipLaunchKernelGGL(k_adj_kick_accel, dim3((r.nR + kAccCols - 1) / kAccCols * r.nJ),
                     dim3(kAccCols * kAccGroups), 0, s, fsb, cK,
                     r.U_Val, accK, abp, r.nR, r.nU, r.cs, r.Ulo, r.Uhi, r.start_v, r.hinv_v);
  hipLaunchKernelGGL(k_adj_kick_coef, dim3(nblocks((int64_t)(r.nU + 2) * r.nR), r.nJ), dim3(kB), 0,
                     s, fsb, r.U_Val, accK, cb, r.nR, r.nU, r.cs, r.Ulo, r.Uhi, r.start_v,
                     r.hinv_v);
}

void launch_adj_kick_cols(Run &r, double *g, double *fb, const double *dMb, hipStream_t s) {
  const dim3 grid((r.nR + 63) / 64 * r.nJ);
  if (const AdjShape *a = adj_shape(r)) {
    hipLaunchKernelGGL(a->gcol, grid, dim3(64), 0, s, g, fb, r.nR, r.cs, 0, r.tv, dMb, r.Ct1,
                       r.dU, r.dR);
    return;
  }
  hipLaunchKernelGGL(k_adj_gcol_down, grid, dim3(64), 0, s, g, fb, r.nR, r.nU, r.cs, r.tv, dMb,
                     r.Ct1, r.dU, r.dR);
}

void launch_adj_scan(Run &r, const double *abp, const double *MbarIn, double *dMb, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_scan, dim3(1), dim3(kAdjScanThreads), r.nR * sizeof(double), s, abp,
                     MbarIn, r.gfac, dMb, r.nR, r.nJ);
}

void launch_adj_mass(Run &r, const double *abp, const double *MbarIn, double *dMb, double *fb,
                     bool accumulate, hipStream_t s) {
  launch_adj_scan(r, abp, MbarIn, dMb, s);
  hipLaunchKernelGGL(k_adj_moment, dim3(nblocks(r.nVal)), dim3(kB), 0, s, fb, dMb, r.Ct1, r.dU,
                     r.dR, r.nR, r.nU, r.nVal, accumulate ? 1 : 0);
}


// The drift's transposed foot table (built on the host from the folded foot
// table): for every kTile x kTile tile of the padded slices, in tile order
// (j, tile row, tile column), the slice-local indices of the points whose
// 4x4 stencil touches the tile with a non-zero weight, in increasing order,
// then kTileChain | k for the delay lines k of the slice whose back cell's
// stencil touches it (those with the back cell inside the grid).
void build_drift_tiles(Run &r, const std::vector<double> &xr, const std::vector<double> &xu,
                       std::vector<int64_t> &off, std::vector<uint32_t> &pts) {
  const int64_t nSlice = (int64_t)r.nU * r.nR;
  const int S = r.nR + 2, Mv = r.nU + 2;
  const int ntm = (Mv + kTile - 1) / kTile, ntk = (S + kTile - 1) / kTile;
  const int64_t perSlice = (int64_t)ntm * ntk, nTiles = r.nJ * perSlice;
  if (nSlice >= (int64_t)kTileChain || r.nChains >= (int64_t)kTileChain)
    throw Error(VLS_ERR_ARG, "slice or delay-line count too large for the adjoint's tables");
  off.assign(nTiles + 1, 0);
  // the delay lines of each slice, in increasing order
  std::vector<std::vector<uint32_t>> chainsOf(r.nJ);
  for (int64_t k = 0; k < r.nChains; k++)
    if (!r.hch_out[k]) chainsOf[r.hch_point[k] / nSlice].push_back((uint32_t)k);
  // the tiles of slice j that a stencil at (x, y) touches
  auto tiles = [&](int64_t j, double x, double y, uint32_t e, auto &&fn) {
    const double ix = std::trunc(x), iy = std::trunc(y);
    const int nx = (int)ix, ny = (int)iy;
    // last stencil row / column: weight BV2(0) = 0 at a fraction of 0
    const int rowHi = std::min(ny + (y - iy == 0.0 ? 2 : 3), Mv - 1);
    const int colHi = std::min(nx + (x - ix == 0.0 ? 2 : 3), S - 1);
    for (int tm = ny / kTile; tm <= rowHi / kTile; tm++)
      for (int tk = nx / kTile; tk <= colHi / kTile; tk++) fn((j * ntm + tm) * ntk + tk, e);
  };
  auto each = [&](int64_t j, auto &&fn) {
    for (int64_t p = 0; p < nSlice; p++) {
      const int64_t i = j * nSlice + p;
      if (xr[i] <= -1.0) continue;  // outside the grid or a delay line
      tiles(j, xr[i], xu[i], (uint32_t)p, fn);
    }
    for (uint32_t k : chainsOf[j]) tiles(j, r.hch_xr[k], r.hch_xu[k], kTileChain | k, fn);
  };
  // slices are independent: both passes run over the host's threads
  auto over_slices = [&](auto &&body) {
    const int nt = std::max(1, std::min<int>((int)std::thread::hardware_concurrency(), 32));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
      th.emplace_back([&, t] {
        for (int64_t j = t; j < r.nJ; j += nt) body(j);
      });
    for (auto &x : th) x.join();
  };
  over_slices([&](int64_t j) { each(j, [&](int64_t t, uint32_t) { off[t + 1]++; }); });
  for (int64_t t = 0; t < nTiles; t++) off[t + 1] += off[t];
  pts.assign(off[nTiles], 0);
  std::vector<int64_t> cur(off.begin(), off.end() - 1);
  over_slices([&](int64_t j) { each(j, [&](int64_t t, uint32_t p) { pts[cur[t]++] = p; }); });
}

}  // namespace vls
