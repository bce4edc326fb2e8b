// HIP kernels of the spherical spline time step (gfx950, FP64).
//
// Every per-point expression keeps the operation order of the reference's
// C++ (cited per kernel) and the file is compiled with -ffp-contract=off, so
// each kernel rounds exactly as the reference does; the only freedom taken is
// the order of the two reductions (moment and statistics), which is fixed
// and deterministic here.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "vls_div.hpp"
#include "vls_internal.hpp"
#include "vls_tile.hpp"

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
// the terms ((Ct1_j*dU_u)*dR_r) * avg4 over u < nU-1, summed in a fixed tree:
// serial sums over groups of 16 rows (u = 64w + 16q + t), combined
// ((G0 + G1) + G2) + G3 into a block of 64 rows, and the blocks left to
// right.  k_moment_cols takes a lane per column r of a slice (63 columns per
// block; column r + 1 arrives by a lane shift) and a wave per block of 64
// rows (loaded in two halves of 32 rows and the row after each); the blocks'
// partials meet in LDS.
__device__ __forceinline__ double lane_next(double x) {  // lane i gets lane i+1's x
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_mov_dpp(lo, 0x130 /* wave_shl:1 */, 0xf, 0xf, true);
  hi = __builtin_amdgcn_mov_dpp(hi, 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
// Stores and loads that other XCDs see without cache maintenance: system
// scope, so they go through the (per-XCD, non-coherent) L2 to memory; the
// counters below are agent-scope atomics, which the device keeps coherent.
__device__ __forceinline__ void coherent_store(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ double coherent_load(const double *p) {
  return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// this wave's stores acknowledged (s_waitcnt vmcnt(0)), before it counts itself
__device__ __forceinline__ void wait_stores() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt unconstrained
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// at most N of this wave's vector memory operations outstanding
// (s_waitcnt vmcnt(N)): the wave's memory operations complete in the order
// they were issued, so every one issued before the last N is acknowledged
template <int N>
__device__ __forceinline__ void wait_vm_upto() {
  static_assert(N >= 0 && N < 64, "vmcnt holds 6 bits");
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  // vmcnt bits 3:0 and 15:14; expcnt (6:4) and lgkmcnt (11:8) unconstrained
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// The partials P[j][r] of this block's slice j and column tile (returned:
// the tile), written by wave 0; kCoherent: with system-scope stores, which
// write through the caches to memory.
template <bool kCoherent>
__device__ __forceinline__ int moment_cols_block(const double *__restrict__ f,
                                                 const double *__restrict__ Ct1,
                                                 const double *__restrict__ dU,
                                                 const double *__restrict__ dR,
                                                 double *__restrict__ P, int nR, int nU,
                                                 double *Wl, int &jOut) {
  const int lane = threadIdx.x & 63;
  // wave-uniform: the rows u below are too, so dU[u] is one load per wave
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nX = (nR + 62) / 63;  // column tiles per slice
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int j = tile / nX;
  const int x = tile - j * nX;
  const int r = x * 63 + lane;
  const bool term = lane < 63 && r < nR - 1;
  const int nW = (nU + 63) / 64;  // blocks of 64 rows
  const double *col = f + (size_t)j * nU * nR + min(r, nR - 1);
  const double ct = Ct1[j], dr = term ? dR[r] : 0.0;
  // wave wv takes the blocks w = wv, wv + 4, ...; a block's rows arrive in
  // two halves of 32 rows (and the row after each), each half's loads all
  // at once: few enough registers for four waves per SIMD, so that the
  // blocks of a 200-slice grid are all resident together
  for (int w = wv; w < nW; w += 4) {
    double G[4];
    double aOwn = 0.0, aN = 0.0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int u0 = 64 * w + 32 * h;
      double row[33];  // row[t]: row u0 + t; row u0 itself is carried in aOwn
      if (h == 0) row[0] = u0 < nU ? col[(size_t)u0 * nR] : 0.0;
#pragma unroll
      for (int t = 1; t < 33; t++)  // rows past the last one are not loaded (no term reads them)
        row[t] = u0 + t < nU ? col[(size_t)(u0 + t) * nR] : 0.0;
      if (h == 0) {
        aOwn = row[0];
        aN = lane_next(aOwn);
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
        double acc = 0.0;
#pragma unroll
        for (int t = 16 * q; t < 16 * q + 16; t++) {
          const int u = u0 + t;
          const double b = row[t + 1];  // row u + 1
          const double bN = lane_next(b);
          if (u < nU - 1) {
            // the term ((ct dU 0.25) dR) (corner sum): get_avg's 0.25 in
            // the first product (exact scaling: the value of
            // ((ct dU) dR) (corner sum 0.25) for every sum above 2^-1020)
            const double sum4 = ((aOwn + aN) + b) + bN;
            acc += (((ct * dU[u]) * 0.25) * dr) * sum4;
          }
          aOwn = b;
          aN = bN;
        }
        G[2 * h + q] = acc;
      }
    }
    Wl[w * 64 + lane] = ((G[0] + G[1]) + G[2]) + G[3];
  }
  __syncthreads();
  if (wv == 0) {
    // the blocks left to right
    double Pj = Wl[lane];
    for (int w = 1; w < nW; w++) Pj = Pj + Wl[w * 64 + lane];
    if (lane < 63 && r < nR) {
      if constexpr (kCoherent)
        coherent_store(&P[(size_t)j * nR + r], term ? Pj : 0.0);
      else
        P[(size_t)j * nR + r] = term ? Pj : 0.0;
    }
  }
  jOut = j;
  return x;
}

__global__ void __launch_bounds__(256)
k_moment_cols(const double *__restrict__ f, const double *__restrict__ Ct1,
              const double *__restrict__ dU, const double *__restrict__ dR,
              double *__restrict__ P, int nR, int nU) {
  extern __shared__ double Wl[];  // the blocks' partials W [nW][64]
  int j;
  moment_cols_block<false>(f, Ct1, dU, dR, P, nR, nU, Wl, j);
}

// The partials of the sum of nS species planes (nVal apart): the integrand
// of int21::f(0, nS), 0 + avg_0 + avg_1 + ... per cell
// (src/integrand21.hxx:21-39), in the same tree order, a thread per (j, r).
__global__ void __launch_bounds__(kBlock)
k_moment_partial_sp(const double *__restrict__ f, const double *__restrict__ Ct1,
                    const double *__restrict__ dU, const double *__restrict__ dR,
                    double *__restrict__ P, int nR, int nU, int nJ, int nS, int64_t nVal) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (r >= nR) return;
  double Pj = 0.0;
  if (r < nR - 1) {
    const double ct = Ct1[j];
    const double dr = dR[r];
    const double *p = f + (size_t)j * nU * nR + r;
    const int nW = (nU + 63) / 64;
    for (int w = 0; w < nW; w++) {
      double G[4];
      for (int q = 0; q < 4; q++) {
        double acc = 0.0;
        for (int t = 0; t < 16; t++) {
          const int u = 64 * w + 16 * q + t;
          if (u >= nU - 1) break;
          double avg = 0.0;
          for (int k = 0; k < nS; k++) {
            const double *pk = p + (size_t)k * nVal + (size_t)u * nR;
            avg += (((pk[0] + pk[1]) + pk[nR]) + pk[nR + 1]) * 0.25;
          }
          acc += ((ct * dU[u]) * dr) * avg;
        }
        G[q] = acc;
      }
      const double W = ((G[0] + G[1]) + G[2]) + G[3];
      Pj = w == 0 ? W : Pj + W;
    }
  }
  P[(size_t)j * nR + r] = Pj;
}

// dM[r] = sum over j of the partials P[j][r] (in chunks of slices, below);
// exclusive radial prefix (gridHandler::accumulate with addSpare=1,
// src/grid_handler.hxx:350-407); acceleration M(<r) * (G/R^2 * dt)
// (src/vlasovSolver21_spline.hxx:746).  The scan runs as one block and ends
// in scan_prefix_out: one thread runs the serial prefix over r on registers,
// kS entries at a time, then the block writes M(<r) and the acceleration.
// The serial prefix of one batch of kS entries held in registers.
template <int kS>
__device__ __forceinline__ void scan_batch(double (&v)[kS], double &prev) {
#pragma unroll
  for (int k = 0; k < kS; k++) {
    const double d = v[k];
    v[k] = prev;
    prev += d;
  }
}
// sdm [nR] holds dM on entry and M(<r) on exit, sgf [nR] the factors
// G/R^2 * dt; both in LDS, written by the block before the call (the
// factors are loaded with the partials' sums, not in the write loop, where
// each pass over r would wait for its own round trip to memory).
__device__ __forceinline__ void scan_prefix_out(double *sdm, const double *sgf,
                                                double *__restrict__ Mx, double *__restrict__ acc,
                                                double *__restrict__ mTot, int nR) {
  __syncthreads();
  if (threadIdx.x == 0) {
    // batches alternate between two register sets, the next batch's LDS
    // reads issued before the current batch's adds
    constexpr int kS = 16;
    double prev = 0.0;
    double A[kS], B[kS];
    const int nFull = nR / kS;
    if (nFull > 0) {
#pragma unroll
      for (int k = 0; k < kS; k++) A[k] = sdm[k];
    }
    for (int b = 0; b < nFull; b += 2) {
      const bool hasB = b + 1 < nFull;
      if (hasB) {
#pragma unroll
        for (int k = 0; k < kS; k++) B[k] = sdm[(b + 1) * kS + k];
      }
      scan_batch(A, prev);
#pragma unroll
      for (int k = 0; k < kS; k++) sdm[b * kS + k] = A[k];
      if (hasB) {
        if (b + 2 < nFull) {
#pragma unroll
          for (int k = 0; k < kS; k++) A[k] = sdm[(b + 2) * kS + k];
        }
        scan_batch(B, prev);
#pragma unroll
        for (int k = 0; k < kS; k++) sdm[(b + 1) * kS + k] = B[k];
      }
    }
    for (int r = nFull * kS; r < nR; r++) {
      const double d = sdm[r];
      sdm[r] = prev;
      prev += d;
    }
    *mTot = prev;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nR; r += blockDim.x) {
    const double m = sdm[r];
    Mx[r] = m;
    acc[r] = m * sgf[r];
  }
}

// The sum over j runs in chunks of kJC slices: S_c = serial sum over the
// slices of chunk c, dM[r] = serial sum of the S_c.  k_moment_chunks forms
// the S_c with a block per chunk (the partials are read by many CUs at
// once) and leaves S_c in place of the chunk's first slice of P (read by the
// same thread only); k_moment_scan adds the chunks and runs the prefix.
constexpr int kScanThreads = 1024, kJC = 20, kChunkThreads = 256;
__global__ void __launch_bounds__(kChunkThreads)
k_moment_chunks(double *__restrict__ P, int nR, int nJ) {
  const int r = blockIdx.y * kChunkThreads + threadIdx.x;
  if (r >= nR) return;
  const int jb = blockIdx.x * kJC, je = min(nJ, jb + kJC);
  double v[kJC];
  // clamped rows are loaded, not added
#pragma unroll
  for (int k = 0; k < kJC; k++) v[k] = P[(size_t)min(jb + k, nJ - 1) * nR + r];
  double S = 0.0;
  if (je - jb == kJC) {
#pragma unroll
    for (int k = 0; k < kJC; k++) S += v[k];
  } else {
#pragma unroll
    for (int k = 0; k < kJC; k++)
      if (jb + k < je) S += v[k];
  }
  P[(size_t)jb * nR + r] = S;
}

__global__ void __launch_bounds__(kScanThreads)
k_moment_scan(const double *__restrict__ P, const double *__restrict__ gfac, double *__restrict__ dM,
              double *__restrict__ Mx, double *__restrict__ acc, double *__restrict__ mTot, int nR,
              int nJ) {
  extern __shared__ double sdm[];  // [nR] the scan, then [nR] the factors
  double *sgf = sdm + nR;
  const int t = threadIdx.x;
  const int nC = (nJ + kJC - 1) / kJC;
  for (int r = t; r < nR; r += kScanThreads) {
    constexpr int kB = 16;  // chunk sums in flight
    const double g = gfac[r];
    double a = 0.0;
    for (int c0 = 0; c0 < nC; c0 += kB) {
      double v[kB];
#pragma unroll
      for (int k = 0; k < kB; k++) v[k] = P[(size_t)min(c0 + k, nC - 1) * kJC * nR + r];
#pragma unroll
      for (int k = 0; k < kB; k++)
        if (c0 + k < nC) a += v[k];
    }
    dM[r] = a;
    sdm[r] = a;
    sgf[r] = g;
  }
  scan_prefix_out(sdm, sgf, Mx, acc, mTot, nR);
}

// The moment's reduction after the partials, in the launch that forms them,
// with the same sums in the same order as k_moment_chunks and k_moment_scan.
// After its partials (written by wave 0 with coherent_store), a block counts
// itself in its (chunk, column tile) counter; the block that completes a
// chunk's count forms that chunk's sums S_c for the tile's columns and
// counts itself in the last counter; the block that completes that count
// runs the scan.  Each finisher resets the counter it completed (every other
// block of the launch has counted by then), so the counters are zero again
// for the next launch; they start zeroed at allocation.  No block waits for
// another.  Partials and chunk sums pass between blocks (and XCDs) through
// memory: written with coherent_store, acknowledged (wait_stores) before
// the count, read with coherent_load after it; agent-scope fences would
// instead write back and invalidate each XCD's whole L2 in every block.
// Every thread of the block calls it; sdm holds 2 nR doubles of LDS.
struct MomentArgs {
  const double *Ct1, *dU, *dR, *gfac;
  double *P, *dM, *Mx, *acc, *mTot;
  unsigned *cnt;
};

// The block's count in its (chunk, column tile) counter, issued once its
// partials are acknowledged; the returned value o completes the chunk when
// moment_chunk_done(o, ...) holds.  lz: a zero in a VGPR that the compiler
// cannot see through, for callers that use o only much later (with an
// address it knows to be uniform the compiler makes the count wave-wide and
// waits for its result at once).
__device__ __forceinline__ unsigned moment_count(const MomentArgs &ma, int j, int x, int nR,
                                                 int lz = 0) {
  const int nX = (nR + 62) / 63, c = j / kJC;
  return __hip_atomic_fetch_add(&ma.cnt[c * nX + x + lz], 1u, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool moment_chunk_done(unsigned o, int j, int nJ) {
  const int jb = j / kJC * kJC, je = min(nJ, jb + kJC);
  return o == (unsigned)(je - jb - 1);
}

// The reduction after the block's count, in two stages (moment_finish runs
// both).  The chunk stage: if *last (LDS, for every thread of the block:
// this block completed its chunk's count), the
// chunk's sums S_c for the tile's columns and the count in the last
// counter; *last then tells whether the final stage (moment_final) falls to
// this block.
__device__ __forceinline__ void moment_chunk(const MomentArgs &ma, int j, int x, int nR, int nJ,
                                             int *last) {
  const int nX = (nR + 62) / 63, nC = (nJ + kJC - 1) / kJC;
  const int c = j / kJC, jb = c * kJC, je = min(nJ, jb + kJC);
  const int t = threadIdx.x;
  unsigned *cnt = ma.cnt;
  double *P = ma.P;
  if (!*last) return;  // block-uniform
  if (t < 64) {
    const int r = x * 63 + t;
    if (t < 63 && r < nR) {
      double v[kJC];
      // clamped rows are loaded, not added
#pragma unroll
      for (int k = 0; k < kJC; k++) v[k] = coherent_load(&P[(size_t)min(jb + k, nJ - 1) * nR + r]);
      double S = 0.0;
#pragma unroll
      for (int k = 0; k < kJC; k++)
        if (jb + k < je) S += v[k];
      coherent_store(&P[(size_t)jb * nR + r], S);
    }
    wait_stores();
    if (t == 0) {
      __hip_atomic_store(&cnt[c * nX + x], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned o =
          __hip_atomic_fetch_add(&cnt[nC * nX], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *last = o == (unsigned)(nC * nX - 1);
    }
  }
  __syncthreads();
}

// The final stage, in the block that completed the last counter: dM[r] over
// the chunks, the prefix and the acceleration.  kFR: column groups of the
// final sums in flight together (each thread holds 16 kFR loaded values:
// wide for kernels whose main loops already need more registers, 1 where
// the registers would cost occupancy).
template <int kFR>
__device__ __forceinline__ void moment_final(const MomentArgs &ma, int nR, int nJ, double *sdm) {
  const int nX = (nR + 62) / 63, nC = (nJ + kJC - 1) / kJC;
  const int t = threadIdx.x;
  unsigned *cnt = ma.cnt;
  double *P = ma.P;
  // the chunk sums of kFR of this thread's columns (r0, r0 + blockDim, ...)
  // and kB chunks are in flight together, with the columns' factors
  // G/R^2 * dt: each round trip to memory is about a microsecond with the
  // GPU otherwise idle, so the columns are not taken one after another; each
  // column still adds its chunks in order.  Clamped columns and chunks are
  // loaded, not added (loads under conditions would each cost a branch).
  double *sgf = sdm + nR;
  constexpr int kB = 16;
  const int bd = blockDim.x;
  for (int r0 = t; r0 < nR; r0 += kFR * bd) {
    double a[kFR], g[kFR];
#pragma unroll
    for (int i = 0; i < kFR; i++) {
      a[i] = 0.0;
      g[i] = ma.gfac[min(r0 + i * bd, nR - 1)];
    }
    for (int c0 = 0; c0 < nC; c0 += kB) {
      double v[kFR][kB];
#pragma unroll
      for (int i = 0; i < kFR; i++) {
        const int r = min(r0 + i * bd, nR - 1);
#pragma unroll
        for (int k = 0; k < kB; k++)
          v[i][k] = coherent_load(&P[(size_t)min(c0 + k, nC - 1) * kJC * nR + r]);
      }
#pragma unroll
      for (int i = 0; i < kFR; i++)
#pragma unroll
        for (int k = 0; k < kB; k++)
          if (c0 + k < nC) a[i] += v[i][k];
    }
#pragma unroll
    for (int i = 0; i < kFR; i++) {
      const int r = r0 + i * bd;
      if (r < nR) {
        ma.dM[r] = a[i];
        sdm[r] = a[i];
        sgf[r] = g[i];
      }
    }
  }
  if (t == 0) __hip_atomic_store(&cnt[nC * nX], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  scan_prefix_out(sdm, sgf, ma.Mx, ma.acc, ma.mTot, nR);
}

template <int kFR>
__device__ __forceinline__ void moment_finish(const MomentArgs &ma, int j, int x, int nR, int nJ,
                                              double *sdm, int *last) {
  moment_chunk(ma, j, x, nR, nJ, last);
  if (*last) moment_final<kFR>(ma, nR, nJ, sdm);  // block-uniform
}

template <int kFR>
__device__ __forceinline__ void moment_tail(const MomentArgs &ma, int j, int x, int nR, int nJ,
                                            double *sdm, int *last) {
  if (threadIdx.x < 64) {  // wave 0 wrote the partials
    wait_stores();
    if (threadIdx.x == 0) *last = moment_chunk_done(moment_count(ma, j, x, nR), j, nJ);
  }
  __syncthreads();
  moment_finish<kFR>(ma, j, x, nR, nJ, sdm, last);
}

// The three moment kernels in one launch (moment_cols_block, moment_tail).
__global__ void __launch_bounds__(256)
k_moment_fused(const double *__restrict__ f, int nR, int nU, int nJ, MomentArgs ma) {
  extern __shared__ double Wl[];  // [nW][64] partials, then [2 nR] for the scan
  __shared__ int last;
  int j;
  const int x = moment_cols_block<true>(f, ma.Ct1, ma.dU, ma.dR, ma.P, nR, nU, Wl, j);
  moment_tail<1>(ma, j, x, nR, nJ, Wl, &last);
}

// ---------------------------------------------------------------- spline
// The natural cubic B-spline solve of one padded line (create_1D with
// BT_NATURAL rows, src/localSpline.hxx:1045-1172), split in its two sweeps.
//
// Forward elimination of row k (k = 1..M-2): y_k = (rhs_k - c16*y_{k-1})/d_k,
// with y_0's contribution c16*A03; the last row then folds in y_{M-3} and
// y_{M-2}.  Back substitution: c_k = y_k - u_k*c_{k+1}, and the first row
// c_0 = (A03 - A01*c_1) - A02*c_2.

// The rows before the constant regime (1..K0-1, at most kHead of them) form a
// head whose pivots and right-hand sides are all loaded up front (the table
// through per-lane vector loads, index + lz with lz a zero the compiler
// cannot see, so that no scalar-memory wait enters the loops).  From row K0
// on the pivots are the scalars dS, rdS, uS: the steady loops load nothing
// but the data stream, kPF rows ahead.
constexpr int kHead = 16;

// The head's pivots staged in LDS, for the register line kernels whose loops
// stream their line: a global load of a table entry inside such a loop makes
// its wait (vmcnt counts every global load and store of the wave, in order)
// wait for all the line data issued before it, so that each head row would
// cost a whole memory latency.  head_load issues the loads (lane k takes row
// k, clamped to the rows a line of M entries has) before the line's own
// loads; head_store writes them to LDS once they are in, and the block's
// barrier orders them before the loops' reads.
struct HeadTab {
  double u[kHead + 1], d[kHead + 1], rd[kHead + 1];
};
struct HeadRegs {
  double u, d, rd;
};
__device__ __forceinline__ HeadRegs head_load(const ThomasAxis &ax, int lane, int lz) {
  const int last = kHead < ax.M - 2 ? kHead : ax.M - 2;
  const int k = lane < 1 ? 1 : (lane > last ? last : lane);
  return {ax.u[k + lz], ax.d[k + lz], ax.rd[k + lz]};
}
__device__ __forceinline__ void head_store(HeadTab &t, const HeadRegs &h, int lane) {
  if (lane >= 1 && lane <= kHead) {
    t.u[lane] = h.u;
    t.d[lane] = h.d;
    t.rd[lane] = h.rd;
  }
}

// Forward elimination over rows 1..M-2.  RHS(k) returns rhs_k (any k <= M-2),
// STORE(k, y) keeps y_k.  Row 1 takes y_0's contribution A03 in place of
// y_{k-1}.  Returns y_{M-3} and y_{M-2} through ym2/ym1.
// CONSUME(k, rhs_k), if given, sees each right-hand side once, in order, at
// the step that uses it.
template <int kPF, class RHS, class STORE, class CONSUME>
__device__ __forceinline__ void thomas_forward(const ThomasAxis &ax, int lz, RHS rhs, STORE store,
                                               double &ym2, double &ym1, CONSUME consume) {
  const int M = ax.M, P = ax.K0 - 1 < M - 2 ? ax.K0 - 1 : M - 2;  // head rows 1..P
  auto kc = [&](int k) { return k <= M - 2 ? k : M - 2; };
  double hr[kHead], hd[kHead], he[kHead], rb[kPF];
#pragma unroll
  for (int q = 0; q < kHead; q++) {
    hr[q] = rhs(kc(1 + q));
    hd[q] = ax.d[kc(1 + q) + lz];
    he[q] = ax.rd[kc(1 + q) + lz];
  }
#pragma unroll
  for (int q = 0; q < kPF; q++) rb[q] = rhs(kc(P + 1 + q));
  // loop constants as per-lane values (x + lz == x): their scalar loads are
  // waited for here, once, instead of inside every guarded loop body
  const double c16 = ax.c16 + (double)lz, dS = ax.dS + (double)lz, rdS = ax.rdS + (double)lz;
  ym2 = 0.0;
  ym1 = ax.A03;
#pragma unroll
  for (int q = 0; q < kHead; q++) {
    if (q < P) {
      consume(1 + q, hr[q]);
      double yk = div_known(hr[q] - c16 * ym1, hd[q], he[q]);
      store(1 + q, yk);
      ym2 = ym1;
      ym1 = yk;
    }
  }
  for (int k = P + 1; k <= M - 2; k += kPF) {
#pragma unroll
    for (int q = 0; q < kPF; q++) {  // rhs_kk sits in slot (kk-P-1) % kPF == q
      const int kk = k + q;
      if (kk <= M - 2) {
        consume(kk, rb[q]);
        double yk = div_known(rb[q] - c16 * ym1, dS, rdS);
        store(kk, yk);
        ym2 = ym1;
        ym1 = yk;
      }
      rb[q] = rhs(kc(kk + kPF));
    }
  }
}

template <int kPF, class RHS, class STORE>
__device__ __forceinline__ void thomas_forward(const ThomasAxis &ax, int lz, RHS rhs, STORE store,
                                               double &ym2, double &ym1) {
  thomas_forward<kPF>(ax, lz, rhs, store, ym2, ym1, [](int, double) {});
}

// Back substitution c_k = y_k - u_k c_{k+1} for k = M-2..1 and the first row,
// kPF rows ahead.  YLOAD(k) returns y_k, OUT(k, c) stores c_k (k = 0..M-1).
template <int kPF, class YLOAD, class OUT>
__device__ __forceinline__ void thomas_backward(const ThomasAxis &ax, int lz, double ym2, double ym1,
                                                YLOAD yload, OUT out) {
  const int M = ax.M, P = ax.K0 - 1 < M - 2 ? ax.K0 - 1 : M - 2;
  double hu[kHead];
#pragma unroll
  for (int q = 0; q < kHead; q++) hu[q] = ax.u[(1 + q <= M - 2 ? 1 + q : M - 2) + lz];
  double tt = 0.0 - ax.hi2 * ym2;
  tt = tt - ax.p * ym1;
  double cn = div_known(tt, ax.q, ax.rq);
  out(M - 1, cn);
  auto kc = [&](int k) { return k >= 1 ? k : 1; };
  double yb[kPF];
#pragma unroll
  for (int q = 0; q < kPF; q++) yb[q] = yload(kc(M - 2 - q));
  const double uS = ax.uS + (double)lz;  // see thomas_forward
  double c1 = 0.0, c2 = cn;
  // steady rows M-2 .. P+1
  for (int k = M - 2; k >= P + 1; k -= kPF) {
#pragma unroll
    for (int q = 0; q < kPF; q++) {
      const int kk = k - q;  // y_kk sits in slot (M-2-kk) % kPF == q
      if (kk >= P + 1) {
        double ck = yb[q] - uS * cn;
        out(kk, ck);
        c2 = cn;
        cn = ck;
        c1 = ck;
      }
      yb[q] = yload(kc(kk - kPF));
    }
  }
  // head rows P .. 1
#pragma unroll
  for (int q = kHead - 1; q >= 0; q--) {
    if (q < P) {
      const int kk = 1 + q;
      double ck = yload(kk) - hu[q] * cn;
      out(kk, ck);
      c2 = cn;
      cn = ck;
      c1 = ck;
    }
  }
  out(0, (ax.A03 - ax.A01 * c1) - ax.A02 * c2);
}

// Copy n consecutive doubles between global memory and an LDS tile in
// batches of kB per lane: all loads of a batch are issued before the first
// use, so their latencies overlap.  Element e of the stream maps to
// (row, k) = (e / len, e % len), tracked incrementally (no division);
// src(e, row, k) and dst(row, k, v) address the two sides.
template <int kB, class SRC, class DST>
__device__ __forceinline__ void stream_copy(int n, int len, SRC src, DST dst) {
  const int t = threadIdx.x;
  int row = 0, k = t;
  while (k >= len) {
    k -= len;
    row++;
  }
  for (int base = t; base < n; base += 64 * kB) {
    double v[kB];
    int rr[kB], kk[kB];
#pragma unroll
    for (int q = 0; q < kB; q++) {
      const int e = base + 64 * q;
      rr[q] = row;
      kk[q] = k;
      v[q] = e < n ? src(e, row, k) : 0.0;
      k += 64;
      while (k >= len) {
        k -= len;
        row++;
      }
    }
#pragma unroll
    for (int q = 0; q < kB; q++)
      if (base + 64 * q < n) dst(rr[q], kk[q], v[q]);
  }
}

// r-direction solves (create_2D first loop, src/localSpline.hxx:1248-1250).
// A block (one wave) stages TR consecutive rows of one slice in LDS with
// coalesced loads, lane t solves row t in place (position k holds rhs_k =
// f[k-1], then y_k, then c_k), and the padded rows are stored back coalesced.
// LDS layout [k][row] with an odd row stride TP >= TR: the lanes of the solve
// touch consecutive words and the staging copies stride by an odd count, so
// neither conflicts on the banks.
__global__ void __launch_bounds__(64)
k_spline_r(const double *__restrict__ f, double *__restrict__ c, int nR, int nU, int nJ,
           int64_t cs, int TR, int TP, int lzmask, ThomasAxis ax) {
  extern __shared__ double tile[];  // [M][TP]
  const int j = blockIdx.y;
  const int u0 = blockIdx.x * TR;
  const int rows = min(TR, nU - u0);
  const int t = threadIdx.x;
  const int lz = t & lzmask;  // a zero the compiler treats as per-lane
  const int M = ax.M;  // nR + 2
  const double *src = f + ((size_t)j * nU + u0) * nR;
  stream_copy<16>(
      rows * nR, nR, [&](int e, int, int) { return src[e]; },
      [&](int row, int k, double v) { tile[(k + 1) * TP + row] = v; });
  __syncthreads();
  if (t < rows) {
    double *L = tile + t;
    double ym2, ym1;
    thomas_forward<12>(
        ax, lz, [&](int k) { return L[k * TP]; }, [&](int k, double y) { L[k * TP] = y; }, ym2,
        ym1);
    thomas_backward<12>(
        ax, lz, ym2, ym1, [&](int k) { return L[k * TP]; },
        [&](int k, double v) { L[k * TP] = v; });
  }
  __syncthreads();
  double *dst = c + (size_t)j * cs + (size_t)(u0 + 1) * M;
  stream_copy<16>(
      rows * M, M, [&](int, int row, int k) { return tile[k * TP + row]; },
      [&](int row, int k, double v) { dst[row * M + k] = v; });
}

// v-direction solves of all nR+2 padded columns (create_2D second loop,
// src/localSpline.hxx:1252-1254).  Lane = column, so every global access is
// coalesced; the forward values y_k wait in LDS ([k][lane]) for the back
// substitution instead of taking a second trip through memory.  A block
// covers TC columns (TC <= 64, chosen from the LDS budget).
__global__ void __launch_bounds__(64)
k_spline_v(double *__restrict__ c, int nR, int nU, int nJ, int64_t cs, int TC, int lzmask,
           ThomasAxis ax) {
  extern __shared__ double ybuf[];  // [M][TC]
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;  // a zero the compiler treats as per-lane
  if (lane >= TC) return;  // no barriers below
  const int col = blockIdx.x * TC + lane;
  const int j = blockIdx.y;
  const int64_t st = nR + 2;
  const bool live = col < nR + 2;
  // dead lanes read a valid column and store nothing
  double *tab = c + (size_t)j * cs + (live ? col : nR + 1);
  double ym2, ym1;
  thomas_forward<24>(
      ax, lz, [&](int k) { return tab[(int64_t)k * st]; },
      [&](int k, double y) { ybuf[k * TC + lane] = y; }, ym2, ym1);
  thomas_backward<12>(
      ax, lz, ym2, ym1, [&](int k) { return ybuf[k * TC + lane]; },
      [&](int k, double v) {
        if (live) tab[(int64_t)k * st] = v;
      });
}

// Streaming line solvers for lines too long for an LDS tile of 64 lines
// (M > ~100): the forward values y_k take a trip through memory (written
// over the line, read back by the back substitution) instead of waiting on
// chip, so that no LDS per line is needed and every line of the grid is in
// flight at once; they move 4 doubles per coefficient instead of 2.  Same
// operations in the same order as thomas_forward / thomas_backward.
//
// v direction: lane = padded column, 64 per wave, every access a coalesced
// row segment, kPF rows ahead both ways; in place on c.
__global__ void __launch_bounds__(64)
k_spline_v_stream(double *__restrict__ c, int nR, int64_t cs, int lzmask, ThomasAxis ax) {
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;  // a zero the compiler treats as per-lane
  const int S = nR + 2;
  const Tile2 t = xcd_tile2((S + 63) / 64);
  const int col = t.x * 64 + lane;
  if (col >= S) return;  // no barriers below
  double *tab = c + (size_t)t.j * cs + col;
  double ym2, ym1;
  thomas_forward<16>(
      ax, lz, [&](int k) { return tab[(int64_t)k * S]; },
      [&](int k, double y) { tab[(int64_t)k * S] = y; }, ym2, ym1);
  thomas_backward<16>(
      ax, lz, ym2, ym1, [&](int k) { return tab[(int64_t)k * S]; },
      [&](int k, double v) { tab[(int64_t)k * S] = v; });
}

// r direction: lane = row, 64 consecutive rows of a slice per wave; chunks
// of kC columns through a small LDS transpose (4 rows x kC columns per pass,
// so that the global accesses are row segments).  The forward sweep reads
// f and leaves y_k in the padded coefficient rows, the back substitution
// reads them back a chunk at a time from the last column down.
__global__ void __launch_bounds__(64)
k_spline_r_stream(const double *__restrict__ f, double *__restrict__ c, int nR, int nU,
                  int64_t cs, int lzmask, ThomasAxis ax) {
  constexpr int kC = 16, kP = kC + 1;
  __shared__ double chunk[64 * kP];
  const int M = nR + 2;
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const int j = blockIdx.y;
  const int u0 = blockIdx.x * 64;
  const int rows = min(64, nU - u0);
  const int rsub = lane / kC, ksub = lane % kC;
  const double *src = f + ((size_t)j * nU + u0) * nR;
  double *dst = c + (size_t)j * cs + (size_t)(u0 + 1) * M;
  double v[kC];
  // columns [e0, e0 + kC) of a (rows x width) array into v[] of the lane's row
  auto gather = [&](const double *a, int width, int e0) {
    const int e = e0 + ksub;
    // every load of the chunk leaves before the first LDS write (the
    // scheduler would otherwise issue a few, wait for them, and go on)
    double t[kC];
#pragma unroll
    for (int q = 0; q < kC; q++) {
      const int row = min(q * 4 + rsub, rows - 1);
      t[q] = a[(size_t)row * width + min(e, width - 1)];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < kC; q++) chunk[(q * 4 + rsub) * kP + ksub] = e < width ? t[q] : 0.0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kC; i++) v[i] = chunk[lane * kP + i];
    __syncthreads();
  };
  auto scatter = [&](double *a, int width, int e0) {
#pragma unroll
    for (int i = 0; i < kC; i++) chunk[lane * kP + i] = v[i];
    __syncthreads();
    const int e = e0 + ksub;
#pragma unroll
    for (int q = 0; q < kC; q++) {
      const int row = q * 4 + rsub;
      if (row < rows && e < width) a[(size_t)row * width + e] = chunk[row * kP + ksub];
    }
    __syncthreads();
  };
  const double c16 = ax.c16 + (double)lz, dS = ax.dS + (double)lz, rdS = ax.rdS + (double)lz;
  double ym1 = ax.A03, ym2 = 0.0;
  // forward elimination: rhs_k = f[k-1], y_k to padded column k
  for (int e0 = 0; e0 < nR; e0 += kC) {
    gather(src, nR, e0);
#pragma unroll
    for (int i = 0; i < kC; i++) {
      const int k = e0 + i + 1;
      if (k <= M - 2) {
        double dv = dS, ev = rdS;
        if (k <= kHead) {
          dv = ax.d[k + lz];
          ev = ax.rd[k + lz];
        }
        const double yk = div_known(v[i] - c16 * ym1, dv, ev);
        v[i] = yk;
        ym2 = ym1;
        ym1 = yk;
      }
    }
    scatter(dst, M, e0 + 1);
  }
  // last row, back substitution and first row, chunks of padded columns
  double tt = 0.0 - ax.hi2 * ym2;
  tt = tt - ax.p * ym1;
  double cn = div_known(tt, ax.q, ax.rq);
  const double uS = ax.uS + (double)lz;
  double c1 = 0.0, c2 = cn;
  for (int e0 = ((M - 1) / kC) * kC; e0 >= 0; e0 -= kC) {
    gather(dst, M, e0);
#pragma unroll
    for (int i = kC - 1; i >= 0; i--) {
      const int k = e0 + i;
      if (k == M - 1) {
        v[i] = cn;
      } else if (k >= 1 && k <= M - 2) {
        double uv = uS;
        if (k <= kHead) uv = ax.u[k + lz];
        const double ck = v[i] - uv * cn;
        v[i] = ck;
        c2 = cn;
        cn = ck;
        c1 = ck;
      } else if (k == 0) {
        v[i] = (ax.A03 - ax.A01 * c1) - ax.A02 * c2;
      }
    }
    scatter(dst, M, e0);
  }
}

// Whole-line solve in registers, for a line length M known at compile time.
// Lane = line: the right-hand sides stream in ahead of the forward
// elimination, the elimination and the back substitution run on the register
// array y[] with no memory traffic and no branches, and the M coefficients
// are stored as they are produced.  Every line of the grid is in flight at
// once (one wave per SIMD at ~2M VGPRs), which is what a latency-bound
// recurrence needs.  Rows 1..kHead use the pivot table (whose entries from
// K0 <= kHead+1 on equal the fixed point), later rows the fixed point dS,
// rdS, uS.
//   KICK = false: the nR+2 = S padded columns (j, col); rhs_k = c[j][k][col],
//                 c_k -> c[j][k][col]      (create_2D second loop); with src
//                 non-null the right-hand sides come from src (same layout)
//   KICK = true : the nR columns of f (j, r); rhs_k = f[j][k-1][r],
//                 d_k -> d[j][k][r] (slice stride cs, row stride nR): the
//                 kick's column splines (see k_kick_line)
template <int M, bool KICK>
__global__ void __launch_bounds__(64)
k_line(const double *__restrict__ src, double *__restrict__ c, int nR, int nU, int nJ, int64_t cs,
       int lzmask, ThomasAxis ax) {
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const int nX = ((KICK ? nR : nR + 2) + 63) / 64;  // column tiles per slice
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int j = tile / nX, bx = tile - j * nX;
  const double *in;
  double *out;
  int64_t is, os;
  bool live;
  if (KICK) {
    int col = bx * 64 + lane;
    live = col < nR;
    if (!live) col = nR - 1;
    in = src + (size_t)j * nU * nR + col - nR;  // in[k * nR] = f[j][k-1][col]
    out = c + (size_t)j * cs + col;
    is = nR;
    os = nR;
  } else {
    int col = bx * 64 + lane;
    live = col < nR + 2;
    if (!live) col = nR + 1;
    in = (src ? src : c) + (size_t)j * cs + col;
    out = c + (size_t)j * cs + col;
    is = nR + 2;
    os = nR + 2;
  }
  // the line streams in kLA rows ahead of the elimination, which runs while
  // it arrives (issued all at once, the loads would wait on each other)
  constexpr int kLA = M - 2 < 48 ? M - 2 : 48;
  __shared__ HeadTab htab;
  const HeadRegs hr = head_load(ax, lane, lz);
  double y[M];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 1; k <= kLA; k++) y[k] = in[k * is];
  __builtin_amdgcn_sched_barrier(0);
  head_store(htab, hr, lane);
  // the block is one wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double c16 = ax.c16 + (double)lz, dS = ax.dS + (double)lz, rdS = ax.rdS + (double)lz;
  double prev = ax.A03;
#pragma unroll
  for (int k = 1; k <= M - 2; k++) {
    if (k + kLA <= M - 2) y[k + kLA] = in[(k + kLA) * is];
    double dv = dS, ev = rdS;
    if (k <= kHead) {
      dv = htab.d[k];
      ev = htab.rd[k];
    }
    y[k] = div_known(y[k] - c16 * prev, dv, ev);
    prev = y[k];
    __builtin_amdgcn_sched_barrier(0);
  }
  double tt = 0.0 - ax.hi2 * y[M - 3];
  tt = tt - ax.p * y[M - 2];
  double cn = div_known(tt, ax.q, ax.rq);
  if (live) out[(M - 1) * os] = cn;
  const double uS = ax.uS + (double)lz;
#pragma unroll
  for (int k = M - 2; k >= 1; k--) {
    double uv = uS;
    if (k <= kHead) uv = htab.u[k];
    y[k] = y[k] - uv * cn;
    cn = y[k];
    if (live) out[k * os] = cn;
  }
  if (live) out[0] = (ax.A03 - ax.A01 * y[1]) - ax.A02 * y[2];
}

// The kick's column splines (k_line<M, true>: d_m of each column of f into
// c, slice stride cs, row stride nR) and the moment of the same f in one pass:
// the moment needs the whole f before the kick can evaluate, but the column
// splines do not need the moment, and both read every column of f.  Tiles of
// 63 columns as in k_moment_cols: lane 63 loads (and solves) the first column
// of the next tile, which lane 62's moment terms take by a lane shift, and
// stores no partial.  The moment's terms are formed from the raw rows while
// they stream in for the elimination, in the order of k_moment_cols (serial
// sums over groups of 16 rows, ((G0 + G1) + G2) + G3 per block of 64 rows,
// blocks left to right); the reduction's chunk and final stages
// (moment_chunk, moment_final), in the blocks that complete its counters
// during their back substitution, then form dM, M(<r) and the
// acceleration.  Same bits as k_moment_cols + k_moment_chunks +
// k_moment_scan and k_line<M, true>.
template <int M>
__global__ void __launch_bounds__(64)
k_kick_solve_moment(const double *__restrict__ f, double *__restrict__ c, int nR, int nJ,
                    int64_t cs, int lzmask, ThomasAxis ax, MomentArgs ma) {
  constexpr int nU = M - 2;
  const int nX = (nR + 62) / 63;  // column tiles per slice
  constexpr int nW = (nU + 63) / 64;  // blocks of 64 rows of the moment
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int j = tile / nX, x = tile - j * nX;
  const int colRaw = x * 63 + lane;
  const int col = colRaw < nR ? colRaw : nR - 1;
  const bool owner = lane < 63 && colRaw < nR;     // stores the partial
  const bool term = lane < 63 && colRaw < nR - 1;  // has moment terms
  const double *in = f + (size_t)j * nU * nR + col - nR;  // in[k * nR] = f[j][k-1][col]
  double *out = c + (size_t)j * cs + col;
  constexpr int kLA = M - 2 < 48 ? M - 2 : 48;
  __shared__ HeadTab htab;
  __shared__ double dUs[nU];
  extern __shared__ double sdm[];  // [2 nR], the scan of the moment's last block
  __shared__ int last;
  const HeadRegs hr = head_load(ax, lane, lz);
  constexpr int nUq = (nU + 63) / 64;
  double duq[nUq];
#pragma unroll
  for (int q = 0; q < nUq; q++) duq[q] = ma.dU[min(lane + 64 * q, nU - 1)];
  const double ct = ma.Ct1[j], dr = term ? ma.dR[colRaw] : 0.0;
  double y[M];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 1; k <= kLA; k++) y[k] = in[k * nR];
  __builtin_amdgcn_sched_barrier(0);
  // the table holds (ct dU[u]) 0.25, the first product of each term (the
  // same rounding as forming it per row and lane; the moment's term is
  // ((ct dU 0.25) dR) (corner sum), as in moment_cols_block)
#pragma unroll
  for (int q = 0; q < nUq; q++)
    if (lane + 64 * q < nU) dUs[lane + 64 * q] = (ct * duq[q]) * 0.25;
  head_store(htab, hr, lane);
  // the block is one wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double c16 = ax.c16 + (double)lz, dS = ax.dS + (double)lz, rdS = ax.rdS + (double)lz;
  double prev = ax.A03;
  // moment state: the previous raw row (own and next column), the group sum,
  // the groups of the current block, the partial over finished blocks
  double aOwn = 0.0, aN = 0.0, g = 0.0, Pj = 0.0;
  double G[4] = {0.0, 0.0, 0.0, 0.0};
  double duNext = dUs[0];
#pragma unroll
  for (int k = 1; k <= M - 2; k++) {
    if (k + kLA <= M - 2) y[k + kLA] = in[(k + kLA) * nR];
    const double b = y[k];  // raw row k - 1
    const double bN = lane_next(b);
    if (k >= 2) {
      // the term of row u = k - 2 (rows u and u + 1 of columns col, col + 1)
      const int u = k - 2;
      const double du = duNext;
      if (u + 1 < nU) duNext = dUs[u + 1];
      const double sum4 = ((aOwn + aN) + b) + bN;
      if (u % 16 == 0) g = 0.0;
      g += (du * dr) * sum4;  // du = (ct dU[u]) 0.25
      // the sum is formed here, in this row's step: left to the end of the
      // kernel, where it is used, it would keep every raw row live
      asm volatile("" : "+v"(g));
      if (u % 16 == 15 || u == nU - 2) G[(u / 16) % 4] = g;
      if (u % 64 == 63 || u == nU - 2) {
        // groups of this block past the last row hold no terms
        const int qLast = (u / 16) % 4;
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (q > qLast) G[q] = 0.0;
        const double W = ((G[0] + G[1]) + G[2]) + G[3];
        Pj = u / 64 == 0 ? W : Pj + W;
        asm volatile("" : "+v"(Pj));
      }
    }
    aOwn = b;
    aN = bN;
    double dv = dS, ev = rdS;
    if (k <= kHead) {
      dv = htab.d[k];
      ev = htab.rd[k];
    }
    y[k] = div_known(b - c16 * prev, dv, ev);
    prev = y[k];
    __builtin_amdgcn_sched_barrier(0);
  }
  // blocks of 64 rows past the last term (none for the standard shapes)
#pragma unroll
  for (int w = (nU - 2) / 64 + 1; w < nW; w++) Pj = Pj + 0.0;
  double tt = 0.0 - ax.hi2 * y[M - 3];
  tt = tt - ax.p * y[M - 2];
  double cn = div_known(tt, ax.q, ax.rq);
  // The partial goes out before the back substitution, and the block counts
  // itself kCnt rows into it, once the partial's store is acknowledged: the
  // count's round trip then overlaps the remaining rows instead of following
  // the last one, and the block that completes its chunk starts the
  // reduction as soon as its own rows are out.  cn, which ends the
  // elimination chain, is pinned before the partial's store: the compiler
  // would otherwise defer the whole chain past that store's branch, keeping
  // every raw row live.
  // Every lane stores its column: lane 63 and the lanes past the last column
  // solve a column that another lane owns, bit for bit, and store the same
  // values there (unmasked, the stores let the count's result be used
  // without waiting for them).  The partial is the owner's only.
  asm volatile("" : "+v"(cn));
  if (owner) coherent_store(&ma.P[(size_t)j * nR + colRaw], term ? Pj : 0.0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  out[(M - 1) * nR] = cn;
  __builtin_amdgcn_sched_barrier(0);
  const double uS = ax.uS + (double)lz;
  constexpr int kCnt = M - 2 > 96 ? 64 : M - 2;  // rows stored before the count
  // Rows stored before the count's result is used: on long columns the
  // block that completes its chunk forms the chunk's sums there, with the
  // rest of its column still in registers, while the other blocks' back
  // substitutions (the kernel's store phase) go on; that part of the
  // reduction's chain of round trips then no longer follows the last row of
  // the last block.
  constexpr int kDec = M - 2 > 128 ? 112 : M - 2;
  constexpr bool kDeferFinal = M > 160;
  unsigned o = 0;
#pragma unroll
  for (int k = M - 2; k >= 1; k--) {
    double uv = uS;
    if (k <= kHead) uv = htab.u[k];
    y[k] = y[k] - uv * cn;
    cn = y[k];
    out[k * nR] = cn;
    if (k == M - 1 - kCnt) {
      // kCnt + 1 stores issued since the partial's: waiting until at most
      // kCnt (or 62) are outstanding waits for the partial
      __builtin_amdgcn_sched_barrier(0);
      wait_vm_upto<(kCnt < 62 ? kCnt : 62)>();
      if (lane == 0) {
        int vz;
        asm volatile("v_mov_b32 %0, 0" : "=v"(vz));
        o = moment_count(ma, j, x, nR, vz);
      }
    }
    if (k == M - 1 - kDec) {
      __builtin_amdgcn_sched_barrier(0);
      if (lane == 0) last = moment_chunk_done(o, j, nJ);
      // the block is one wave
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // Four column groups of final sums in flight where the line's
      // registers already exceed what two waves per SIMD could hold.  On
      // the longest columns the final stage follows the block's own rows:
      // run here, its loads would queue behind the other blocks' row stores
      // (at 201x201x200 kick_solve 38.8 to 38.0 us; at 129 rows the
      // remaining rows cost more than the queueing).
      if constexpr (kDeferFinal)
        moment_chunk(ma, j, x, nR, nJ, &last);
      else
        moment_finish<(M > 128 ? 4 : 1)>(ma, j, x, nR, nJ, sdm, &last);
    }
    // rows in order (moved across each other, they keep more of y live)
    __builtin_amdgcn_sched_barrier(0);
  }
  out[0] = (ax.A03 - ax.A01 * y[1]) - ax.A02 * y[2];
  if constexpr (kDeferFinal)
    if (last) moment_final<(M > 128 ? 4 : 1)>(ma, nR, nJ, sdm);  // block-uniform
}

// ---------------------------------------------------------------- fused kick
// The kick (src/vlasovSolver21_spline.hxx:750-765, evaluateNxPy
// src/localSpline.hxx:423-441) evaluates the 2D spline on r nodes only.
// There its r-direction combination (c[m][r] + 4 c[m][r+1] + c[m][r+2]) / 6
// reproduces the r-solve's right-hand sides exactly, and the r and v solves
// commute, so the combination is d_m(r), the natural v-spline coefficients
// of column r of f, and evaluateNxPy's (1/36) sum_k B_k(t) (6 d_{n+k})
// equals (1/6) sum_k B_k(t) d_{n+k}.  The kick is computed in that form: one
// v-solve per column of f, no r-solve (it agrees with the reference's 2D
// computation to rounding; the oracle's device order computes the same form).
//
// A wave takes 64 consecutive columns of a slice, in place on f.  Lane =
// column: the column streams into registers during the forward elimination,
// and the back substitution, which produces d_m from the last row down,
// keeps the last kRing rows in a ring in LDS (each lane reads only its own
// column of it).  Point (u, r) is evaluated at the step that produces row
// m = u + D, D being the smallest offset n - u over the wave's points: its
// rows are ready by then, and they are still in the ring if the offsets of
// the wave spread over at most kRing - 4 rows.  A wave whose offsets spread
// wider keeps all of d in a global scratch area and evaluates after the
// sweep.  The offsets come from the same expressions as the evaluation and
// are taken while the column arrives.  The column is read whole before any
// of its points is written.
constexpr int kKW = 64, kEB = 8;

// Long columns keep their first kYL forward values in LDS instead of VGPRs
// (they are the last ones the back substitution needs) and use a 16-row ring,
// so that the column leaves room in the 512 VGPRs of a wave for several
// evaluations in flight, within the LDS share of four waves per CU.
template <int Mv>
struct KickLineCfg {
  static constexpr int kRing = Mv > 160 ? 16 : 64;
  static constexpr int kYL = Mv > 160 ? 56 : 0;
};

template <int Mv, bool KEEP>
__global__ void __launch_bounds__(64)
k_kick_line(double *f, const double *__restrict__ U_Val, const double *__restrict__ acc,
            double *__restrict__ gsc, double Ulo, double Uhi, double start_v, double hinv_v,
            int forceScratch, int lzmask, ThomasAxis ax, double *__restrict__ dKeep, int64_t ds,
            int nR) {
  constexpr int nU = Mv - 2;
  constexpr int kRing = KickLineCfg<Mv>::kRing, kYL = KickLineCfg<Mv>::kYL;
  // ring slot s holds row m with m % kRing == s; slots kRing..kRing+2 repeat
  // slots 0..2 so that the four rows n..n+3 of a point are read from slots
  // (n % kRing) + 0..3 without wrapping
  __shared__ double ring[(kRing + 3) * 64];
  __shared__ double ylds[(kYL > 0 ? kYL : 1) * 64];
  __shared__ double Us[nU];
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int j = tile / ((nR + kKW - 1) / kKW);
  const int k = (tile - j * ((nR + kKW - 1) / kKW)) * kKW + lane;  // column of this lane
  const bool outLane = k < nR;            // writes the points f[j][.][k]
  double *fj = f + (size_t)j * nU * nR;
  const double *col = fj + (k < nR ? k : nR - 1);
  // loads in order: the U values, the acceleration, then the first kLA rows
  // of the column (below), so that the LDS copy of the U values waits only
  // for the first ones
  constexpr int nUq = (nU + 63) / 64;
  double uq[nUq];
#pragma unroll
  for (int q = 0; q < nUq; q++) uq[q] = U_Val[min(lane + 64 * q, nU - 1)];
  const double a = outLane ? acc[k] : 0.0;
  __shared__ HeadTab htab;
  const HeadRegs hr = head_load(ax, lane, lz);
  double *fo = fj + (outLane ? k : 0);
  // KEEP (the adjoint's recorded steps): the coefficients d_m of the column
  // are also stored, to dKeep[j][m][k] (slice stride ds)
  double *dk = KEEP ? dKeep + (size_t)j * ds + (outLane ? k : 0) : nullptr;
  auto keep = [&](int m, double d) {
    if (KEEP && outLane) dk[(size_t)m * nR] = d;
  };

  // forward elimination (rows 1..Mv-2, row m = point u = m - 1); the column
  // streams in kLA rows ahead of the elimination, which runs while it arrives
  constexpr int kLA = Mv - 2 < 48 ? Mv - 2 : 48;
  double y[Mv];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 1; m <= kLA; m++) y[m] = col[(m - 1) * nR];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < nUq; q++)
    if (lane + 64 * q < nU) Us[lane + 64 * q] = uq[q];
  head_store(htab, hr, lane);
  // the ring starts zeroed: a point whose target row is the last knot reads
  // the row past the padded column with weight BV2(0) = 0
#pragma unroll
  for (int s = 0; s < kRing + 3; s++) ring[s * 64 + lane] = 0.0;
  // the block is one wave: a wave-level barrier orders the LDS writes before
  // the reads of other lanes, without waiting for the column's loads
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // target-row offsets n - u of the lane's points: point u = m - 1 at step
  // m of the elimination, whose dependent chain leaves the VALU idle most of
  // the time (its U value read from LDS one step ahead)
  int dlo = 1 << 30, dhi = -(1 << 30);
  double usNext = Us[0];
  const double c16 = ax.c16 + (double)lz, dS = ax.dS + (double)lz, rdS = ax.rdS + (double)lz;
  double prev = ax.A03;
#pragma unroll
  for (int m = 1; m <= Mv - 2; m++) {
    if (m + kLA <= Mv - 2) y[m + kLA] = col[(m + kLA - 1) * nR];
    {
      const double nu = usNext + a;
      if (m < nU) usNext = Us[m];
      if (outLane && !(nu <= Ulo) && !(nu >= Uhi)) {
        const int n = (int)trunc((nu - start_v) * hinv_v);
        dlo = min(dlo, n - (m - 1));
        dhi = max(dhi, n - (m - 1));
      }
    }
    double dv = dS, ev = rdS;
    if (m <= kHead) {
      dv = htab.d[m];
      ev = htab.rd[m];
    }
    const double yk = div_known(y[m] - c16 * prev, dv, ev);
    if (m <= kYL)
      ylds[(m - 1) * 64 + lane] = yk;
    else
      y[m] = yk;
    prev = yk;
    // keep the load of row m + kLA in this step (the scheduler would
    // otherwise hoist every load to the top, where they wait on each other)
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    dlo = min(dlo, __shfl_xor(dlo, o));
    dhi = max(dhi, __shfl_xor(dhi, o));
  }
  if (dlo > dhi) dlo = dhi = 0;  // no point of the wave evaluates the spline
  // wave-uniform after the reduction: scalar registers, scalar addressing
  dlo = __builtin_amdgcn_readfirstlane(dlo);
  dhi = __builtin_amdgcn_readfirstlane(dhi);
  const int D = dlo;
  const bool useRing = !forceScratch && dhi - dlo <= kRing - 4 - (kEB - 1);
  double *gs = gsc + (size_t)tile * (Mv + 1) * 64 + lane;

  auto ring_at = [&](int row) { return ring[(row & (kRing - 1)) * 64 + lane]; };
  auto ring_put = [&](int m, double d) {  // m known at compile time
    ring[(m & (kRing - 1)) * 64 + lane] = d;
    if ((m & (kRing - 1)) < 3) ring[(kRing + (m & (kRing - 1))) * 64 + lane] = d;
  };
  auto scratch_at = [&](int row) { return gs[(size_t)row * 64]; };
  // point u of this lane's column (u wave-uniform)
  auto eval = [&](int u, auto dload) {
    const double newU = Us[u] + a;
    double val = 0;
    if (!(newU <= Ulo) && !(newU >= Uhi)) {
      double yy = (newU - start_v) * hinv_v;
      int n;
      split(yy, n, yy);
      const double d0 = dload(n), d1 = dload(n + 1), d2 = dload(n + 2), d3 = dload(n + 3);
      double result = 1.0 / 6.0;
      result *= BV2(1. - yy) * d0 + BV1(1. - yy) * d1 + BV1(yy) * d2 + BV2(yy) * d3;
      val = result;
    }
    fo[(size_t)u * nR] = val;
  };
  // The points whose steps are rows m0..m0+kEB-1, once row m0 is in: straight
  // code, so that the kEB evaluations overlap each other and the sweep.  The
  // point rows u = m0 + i - D are wave-uniform: their U values and output
  // rows are addressed from scalar registers.
  const unsigned kcol = outLane ? (unsigned)k : 0u;
  // A block is evaluated in two halves: prep_block, one block ahead, takes
  // the points' target rows and basis weights, which do not depend on the
  // spline, and finish_block reads the ring and forms the sums once the rows
  // are in, so that the weights of the next block fill the latencies of the
  // sweep and of the ring reads.
  struct Prep {
    double w0, w1, w2, w3;  // BV2(1-y), BV1(1-y), BV1(y), BV2(y)
    int at;                 // ring index of the point's first row
    bool in;
  };
  auto prep_block = [&](int m0, Prep *pp) {
    double uv[kEB];
#pragma unroll
    for (int i = 0; i < kEB; i++) uv[i] = Us[min(max(m0 + i - D, 0), nU - 1)];
    // the LDS reads of the block leave together (the scheduler would otherwise
    // issue each just before its use and wait for it)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < kEB; i++) {
      const double newU = uv[i] + a;
      const bool in = !(newU <= Ulo) && !(newU >= Uhi);
      double yy = in ? (newU - start_v) * hinv_v : 0.0;
      int n;
      split(yy, n, yy);
      pp[i].at = (n & (kRing - 1)) * 64 + lane;
      pp[i].in = in;
      pp[i].w0 = BV2(1. - yy);
      pp[i].w1 = BV1(1. - yy);
      pp[i].w2 = BV1(yy);
      pp[i].w3 = BV2(yy);
    }
  };
  auto finish_block = [&](int m0, const Prep *pp) {
    double val[kEB], g[kEB][4];
#pragma unroll
    for (int i = 0; i < kEB; i++) {
      const double *rb = ring + pp[i].at;
#pragma unroll
      for (int q = 0; q < 4; q++) g[i][q] = rb[q * 64];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < kEB; i++) {
      double result = 1.0 / 6.0;
      result *= pp[i].w0 * g[i][0] + pp[i].w1 * g[i][1] + pp[i].w2 * g[i][2] + pp[i].w3 * g[i][3];
      val[i] = pp[i].in ? result : 0.0;
    }
    if (m0 - D >= 0 && m0 + kEB - 1 - D < nU) {
      if (outLane) {
#pragma unroll
        for (int i = 0; i < kEB; i++) {
          double *row = fj + (size_t)(m0 + i - D) * nR;  // uniform
          row[kcol] = val[i];
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < kEB; i++) {
        const int u = m0 + i - D;
        if (outLane && u >= 0 && u < nU) fo[(size_t)u * nR] = val[i];
      }
    }
  };

  double tt = 0.0 - ax.hi2 * y[Mv - 3];
  tt = tt - ax.p * y[Mv - 2];
  const double cLast = div_known(tt, ax.q, ax.rq);
  const double uS = ax.uS + (double)lz;
  if (useRing) {
    // Points whose step would be above the last row have targets outside
    // the grid (n <= nU-1 < u + D): zero.
    if (outLane)
      for (int u = max(Mv - D, 0); u < nU; u++) fo[(size_t)u * nR] = 0.0;
    // back substitution, last row first; a block of kEB rows is evaluated as
    // soon as it is complete
    double cn = cLast, c2 = 0.0;
    // block b is evaluated at row m0 = Mv - (b + 1) kEB, its points are
    // u = m0 - D .. m0 - D + kEB - 1
    Prep pp[kEB];
    if (Mv - kEB >= 1) prep_block(Mv - kEB, pp);
    ring_put(Mv - 1, cn);
    keep(Mv - 1, cn);
#pragma unroll
    for (int m = Mv - 2; m >= 1; m--) {
      double uv = uS;
      if (m <= kHead) uv = htab.u[m];
      const double ym = m <= kYL ? ylds[(m - 1) * 64 + lane] : y[m];
      c2 = cn;
      cn = ym - uv * cn;
      ring_put(m, cn);
      keep(m, cn);
      if ((Mv - 1 - m) % kEB == kEB - 1) {
        finish_block(m, pp);
        if (m - kEB >= 1) prep_block(m - kEB, pp);
      }
    }
    const double d0 = (ax.A03 - ax.A01 * cn) - ax.A02 * c2;
    ring_put(0, d0);
    keep(0, d0);
    // rows 0..top not covered by a block, and the points whose step would
    // come after row 0 (D < 0); the ring then holds rows 0..kRing-1
    if (outLane) {
      const int top = Mv - 1 < kEB ? Mv - 1 : (Mv - 1) % kEB;
      for (int u = min(top - D, nU - 1); u >= 0; u--) eval(u, ring_at);
    }
  } else {
    // all of d to the scratch area, then every point
    double cn = cLast, c2 = 0.0;
    gs[(size_t)(Mv - 1) * 64] = cn;
    keep(Mv - 1, cn);
#pragma unroll
    for (int m = Mv - 2; m >= 1; m--) {
      double uv = uS;
      if (m <= kHead) uv = htab.u[m];
      const double ym = m <= kYL ? ylds[(m - 1) * 64 + lane] : y[m];
      c2 = cn;
      cn = ym - uv * cn;
      gs[(size_t)m * 64] = cn;
      keep(m, cn);
    }
    gs[0] = (ax.A03 - ax.A01 * cn) - ax.A02 * c2;
    keep(0, gs[0]);
    gs[(size_t)Mv * 64] = 0.0;  // the row past the padded column, read with weight 0
    if (outLane)
      for (int u = 0; u < nU; u++) eval(u, scratch_at);
  }
}

// ---------------------------------------------------------------- kick
// The kick for shapes without register lines, in the column form of the
// fused kick (k_kick_line): k_kick_vsolve solves the natural v-spline of
// every column of f into d ([j][m][r], slice stride ds, row stride nR; the
// forward values take a trip through d, as in k_spline_v_stream), then
// k_kick_cols evaluates each point as (1/6) sum_k B_k(t) d[n+k][r]
// (src/vlasovSolver21_spline.hxx:750-765, evaluateNxPy
// src/localSpline.hxx:423-441, in the column form).
__global__ void __launch_bounds__(64)
k_kick_vsolve(const double *__restrict__ f, double *__restrict__ d, int nR, int nU, int64_t ds,
              int lzmask, ThomasAxis ax) {
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const Tile2 t = xcd_tile2((nR + 63) / 64);
  const int col = t.x * 64 + lane;
  if (col >= nR) return;  // no barriers below
  const double *fc = f + (size_t)t.j * nU * nR + col;
  double *dc = d + (size_t)t.j * ds + col;
  double ym2, ym1;
  thomas_forward<16>(
      ax, lz, [&](int k) { return fc[(int64_t)(k - 1) * nR]; },
      [&](int k, double y) { dc[(int64_t)k * nR] = y; }, ym2, ym1);
  thomas_backward<16>(
      ax, lz, ym2, ym1, [&](int k) { return dc[(int64_t)k * nR]; },
      [&](int k, double v) { dc[(int64_t)k * nR] = v; });
}

// The kick's value of a point from the four column-spline coefficients of
// its target rows n..n+3 and its offset y in the cell (evaluateNxPy,
// src/localSpline.hxx:423-441, in the column form)
__device__ __forceinline__ double kick_value(double y, double d0, double d1, double d2, double d3) {
  double result = 1.0 / 6.0;
  result *= BV2(1. - y) * d0 + BV1(1. - y) * d1 + BV1(y) * d2 + BV2(y) * d3;
  return result;
}

// k_kick_vsolve with the moment of the same f (as k_kick_solve_moment does
// for the register lines): tiles of 63 columns, lane 63 solving the next
// tile's first column for its values only; the moment's terms are taken as
// the raw rows are used by the forward sweep, the group and block sums as
// left folds (((G0 + G1) + G2) + G3 with the groups past the last row as 0),
// then moment_tail.  Dynamic LDS: nU + 2 nR doubles.
__global__ void __launch_bounds__(64)
k_kick_vsolve_moment(const double *__restrict__ f, double *__restrict__ d, int nR, int nU, int nJ,
                     int64_t ds, int lzmask, ThomasAxis ax, MomentArgs ma) {
  extern __shared__ double mlds[];  // dU [nU], then the scan's [2 nR]
  __shared__ int last;
  double *dUs = mlds, *sdm = mlds + nU;
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const int nX = (nR + 62) / 63;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int j = tile / nX, x = tile - j * nX;
  const int colRaw = x * 63 + lane;
  const int col = colRaw < nR ? colRaw : nR - 1;
  const bool owner = lane < 63 && colRaw < nR;
  const bool term = lane < 63 && colRaw < nR - 1;
  const double *fc = f + (size_t)j * nU * nR + col;
  double *dc = d + (size_t)j * ds + col;
  const double ct = ma.Ct1[j], dr = term ? ma.dR[colRaw] : 0.0;
  // (ct dU[u]) 0.25, the first product of each term
  for (int u = lane; u < nU; u += 64) dUs[u] = (ct * ma.dU[u]) * 0.25;
  __syncthreads();
  const int nW = (nU + 63) / 64;
  double aOwn = 0.0, aN = 0.0, g = 0.0, W = 0.0, Pj = 0.0, duNext = dUs[0];
  // raw row k - 1 of the column: the term of row u = k - 2
  auto consume = [&](int k, double b) {
    const double bN = lane_next(b);
    if (k >= 2) {
      const int u = k - 2;
      const double du = duNext;
      if (u + 1 < nU) duNext = dUs[u + 1];
      const double sum4 = ((aOwn + aN) + b) + bN;
      if (u % 16 == 0) g = 0.0;
      g += (du * dr) * sum4;  // du = (ct dU[u]) 0.25
      const int q = (u / 16) % 4;
      if (u % 16 == 15 || u == nU - 2) W = q == 0 ? g : W + g;
      if (u % 64 == 63 || u == nU - 2) {
        for (int q2 = q + 1; q2 < 4; q2++) W = W + 0.0;
        Pj = u / 64 == 0 ? W : Pj + W;
      }
    }
    aOwn = b;
    aN = bN;
  };
  double ym2, ym1;
  thomas_forward<16>(
      ax, lz, [&](int k) { return fc[(int64_t)(k - 1) * nR]; },
      [&](int k, double y) {
        if (owner) dc[(int64_t)k * nR] = y;
      },
      ym2, ym1, consume);
  for (int w = (nU - 2) / 64 + 1; w < nW; w++) Pj = Pj + 0.0;
  thomas_backward<16>(
      ax, lz, ym2, ym1, [&](int k) { return dc[(int64_t)k * nR]; },
      [&](int k, double v) {
        if (owner) dc[(int64_t)k * nR] = v;
      });
  if (owner) coherent_store(&ma.P[(size_t)j * nR + colRaw], term ? Pj : 0.0);
  moment_tail<1>(ma, j, x, nR, nJ, sdm, &last);
}

// The kick's evaluation, kPts points per thread (il, il + kBlock, ... of a
// block's kPts kBlock points of a slice), whose position loads and then
// coefficient loads leave together: on large slices (513x513x256) the
// coefficients come from HBM behind the dependent position loads, and four
// points per thread keep twice the loads of two in flight at the same
// occupancy (kick_eval 374 to 357 us there; at 201x201x200, where d stays
// in the MALL, two points are 0.1 us faster).  A 1D grid of nT blocks per
// slice in the order of xcd_tile, so that the blocks of a slice, which read
// the same rows of d (each row serves four rows of points), run on one XCD
// and share its L2.  A point outside the velocity range is 0 (it reads rows
// 0..3 of its column and does not use them).

template <int kPts>
__global__ void __launch_bounds__(kBlock)
k_kick_cols(double *__restrict__ f, const double *__restrict__ d, const double *__restrict__ U_Val,
             const double *__restrict__ acc, int nR, int nU, int64_t ds, double Ulo, double Uhi,
             double start_v, double hinv_v) {
  const int nS = nU * nR;
  const int nT = (nS + kPts * kBlock - 1) / (kPts * kBlock);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int j = tile / nT;
  const int il0 = (tile - j * nT) * (kPts * kBlock) + threadIdx.x;
  if (il0 >= nS) return;
  const double *dj = d + (size_t)j * ds;
  // points past the slice repeat the first one (evaluated, not stored)
  int ils[kPts], rs[kPts];
  double newU[kPts];
#pragma unroll
  for (int h = 0; h < kPts; h++) {
    ils[h] = il0 + h * kBlock < nS ? il0 + h * kBlock : il0;
    const int u = ils[h] / nR;
    rs[h] = ils[h] - u * nR;
    newU[h] = U_Val[u] + acc[rs[h]];
  }
  double dv[kPts][4], yy[kPts];
  bool in[kPts];
#pragma unroll
  for (int h = 0; h < kPts; h++) {
    in[h] = !(newU[h] <= Ulo) && !(newU[h] >= Uhi);
    double y = in[h] ? (newU[h] - start_v) * hinv_v : 0.0;
    int n;
    split(y, n, y);
    yy[h] = y;
    // row n+3 is past the column (nU+2) only at y == 0, where its weight
    // BV2(0) is 0: the last row is read instead
    const double *dc = dj + rs[h];
    dv[h][0] = dc[(int64_t)n * nR];
    dv[h][1] = dc[(int64_t)(n + 1) * nR];
    dv[h][2] = dc[(int64_t)(n + 2) * nR];
    dv[h][3] = dc[(int64_t)min(n + 3, nU + 1) * nR];
  }
  const size_t base = (size_t)j * nS;
#pragma unroll
  for (int h = 0; h < kPts; h++)
    if (h == 0 || il0 + h * kBlock < nS)
      f[base + ils[h]] = in[h] ? kick_value(yy[h], dv[h][0], dv[h][1], dv[h][2], dv[h][3]) : 0.0;
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

// eval_c in two parts, so that several points' loads can leave together:
// the 4x4 coefficients and the offsets in the cell, then the same sum in the
// same order as eval_c.
struct CStencil {
  double v[16];  // rows Ny-1..Ny+2, columns Nx-1..Nx+2 (v[4 a + b])
  double x, y;
};
__device__ __forceinline__ void eval_c_fetch(const double *d, int64_t st, double xr, double xu,
                                             CStencil &s) {
  int Nx, Ny;
  split(xr, Nx, s.x);
  split(xu, Ny, s.y);
  Nx++;
  Ny++;
  const double *p0 = d + Nx + (int64_t)(Ny - 1) * st;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) s.v[4 * a + b] = p0[(int64_t)a * st + b - 1];
}
__device__ __forceinline__ double eval_c_sum(const CStencil &s) {
  const double x = s.x, y = s.y;
  const double *v = s.v;
  const double b2xI = BV2(1. - x);
  const double b1xI = BV1(1. - x);
  const double b2x = BV2(x);
  const double b1x = BV1(x);
  double result = 1.0 / 36.0;
  result *= BV2(1. - y) * (v[0] * b2xI + v[1] * b1xI + v[2] * b1x + v[3] * b2x) +
            BV1(1. - y) * (v[4] * b2xI + v[5] * b1xI + v[6] * b1x + v[7] * b2x) +
            BV1(y) * (v[8] * b2xI + v[9] * b1xI + v[10] * b1x + v[11] * b2x) +
            BV2(y) * (v[12] * b2xI + v[13] * b1xI + v[14] * b1x + v[15] * b2x);
  return result;
}

// drift loop, src/vlasovSolver21_spline.hxx:788-796, with the delayed central
// kernel.  The foot table carries the point's kind in xr: an evaluated foot
// has xr > -1 (trunc(xr) >= 0), kFootOut (-1) is a foot outside the grid
// (f = 0), kFootChain (-2) a point fed by a delay line whose index is in xu.
// A delay-line point takes the spline value at its back cell (or 0 when that
// cell's v is outside the grid; src/vlasovSolver21_spline.hxx:776-786) into
// the ring slot behind the front, reads the front (sphericalKernel::readAt,
// src/sphericalKernel.hxx:155-161) and, for the last species of the step,
// turns the ring by one (cellChain::roll, src/cellChain.hxx:49-56; the
// species share the ring's head).  Blocks are spread so that each
// XCD (block id mod 8) walks a contiguous range of points: the 4x4 stencils
// of neighbouring points share coefficient rows, which then stay in that
// XCD's L2.
constexpr double kFootOut = -1.0, kFootChain = -2.0;

struct ChainTables {
  const int64_t *off;
  const int32_t *len;
  int32_t *head;
  const double *bxr, *bxu;
  const uint8_t *bout;
  double *vals;  // the values of the species being drifted
  int64_t n;
  int roll;      // turn the rings (the last species of the step)
};

// Two points per thread (i and i + kBlock of a block's 2 kBlock points):
// both points' feet, then both stencils' loads, go out together, so that a
// wave keeps twice the loads in flight (a stencil's reads wait on its foot's
// read).  A point that is not evaluated fetches the stencil at (0, 0),
// inside every slice, and does not use it.  Blocks are spread so that each
// XCD (block id mod 8) walks a contiguous range of points: the 4x4 stencils
// of neighbouring points share coefficient rows, which then stay in that
// XCD's L2.
__device__ __forceinline__ void drift_chain(double *__restrict__ f, const double *cj, int stride,
                                            int64_t i, double kd, const ChainTables &ch) {
  if (!(kd >= 0) || kd >= (double)ch.n) return;  // no delay line attached
  const int64_t k = (int64_t)kd;
  const int L = ch.len[k];
  const int h = ch.head[k];
  double *v = ch.vals + ch.off[k];
  int back = h + L - 1;
  if (back >= L) back -= L;
  v[back] = ch.bout[k] ? 0.0 : eval_c(cj, stride, ch.bxr[k], ch.bxu[k]);
  f[i] = v[h];
  if (ch.roll) ch.head[k] = (h + 1 == L) ? 0 : h + 1;
}

__global__ void __launch_bounds__(kBlock)
k_drift(double *__restrict__ f, const double *__restrict__ c, const double2 *__restrict__ feet,
         int64_t nVal, int nSlice, int64_t cs, int stride, int perXcd, ChainTables ch) {
  const int b = blockIdx.x;
  const int64_t lb = (int64_t)(b & 7) * perXcd + (b >> 3);  // XCD-contiguous block
  const int64_t i0 = lb * (2 * kBlock) + threadIdx.x, i1 = i0 + kBlock;
  if (i0 >= nVal) return;
  const bool has1 = i1 < nVal;
  const double2 ft0 = feet[i0];
  const double2 ft1 = feet[has1 ? i1 : i0];
  const double *cj0 = c + (i0 / nSlice) * cs;
  const double *cj1 = c + ((has1 ? i1 : i0) / nSlice) * cs;
  const bool e0 = ft0.x > kFootOut, e1 = ft1.x > kFootOut;
  CStencil s0, s1;
  eval_c_fetch(cj0, stride, e0 ? ft0.x : 0.0, e0 ? ft0.y : 0.0, s0);
  eval_c_fetch(cj1, stride, e1 ? ft1.x : 0.0, e1 ? ft1.y : 0.0, s1);
  double r0 = eval_c_sum(s0), r1 = eval_c_sum(s1);
  // both sums here, before the branches below: left to the branch that uses
  // it, a point's stencil loads would wait for the other point's
  asm volatile("" : "+v"(r0), "+v"(r1));
  if (e0)
    f[i0] = r0;
  else if (ft0.x == kFootOut)
    f[i0] = 0;
  else
    drift_chain(f, cj0, stride, i0, ft0.y, ch);
  if (!has1) return;
  if (e1)
    f[i1] = r1;
  else if (ft1.x == kFootOut)
    f[i1] = 0;
  else
    drift_chain(f, cj1, stride, i1, ft1.y, ch);
}

// Attach delay line k to its point: xr = kFootChain, xu = k.
__global__ void __launch_bounds__(kBlock)
k_mark_chains(double2 *__restrict__ feet, const int64_t *__restrict__ point, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  feet[point[k]] = make_double2(kFootChain, (double)k);
}

// ---------------------------------------------------------------- statistics
// Per (j, r) partials of the moment (in its tree order) and, serial in u, of the integrands of
// quadrature21::integrate (src/quadrature21.hxx:57-85; integrands
// src/integrand21.hxx:82-145).
//
// With nS species (planes nVal apart) the mass, and the entropy and L2 norm of
// the statistics.txt line, take the sum of the species' cell averages
// (int21::f / entropy / L2Norm(0, nS)); the other integrands read species 0,
// as the reference's default integrands do (src/slice21Iterator.hxx:90-94).
__global__ void __launch_bounds__(kBlock)
k_stats_partial(const double *__restrict__ f, const double *__restrict__ Ct1,
                const double *__restrict__ dU, const double *__restrict__ dR,
                const double *__restrict__ R_Val, const double *__restrict__ U_Val,
                const double *__restrict__ J_Val, double *__restrict__ part, int nR, int nU,
                int nJ, int nS, int64_t nVal) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (r >= nR) return;
  double m = 0.0, ek = 0.0, s = 0.0, l1 = 0.0, l2 = 0.0, sAll = 0.0, l2All = 0.0;
  if (r < nR - 1) {
    const double ct = Ct1[j];
    const double dr = dR[r];
    const double rc = R_Val[r] + 0.5 * dr;
    const double J = J_Val[j];
    const double *p = f + (int64_t)j * nU * nR + r;
    double a0 = p[0], a1 = p[1];
    // the mass in the moment's tree order (groups of 16 rows, waves of 64),
    // the other integrands serial in u
    const int nW = (nU + 63) / 64;
    for (int wv = 0; wv < nW; wv++) {
      double G[4];
      for (int q4 = 0; q4 < 4; q4++) {
        double g = 0.0;
        for (int t = 0; t < 16; t++) {
          const int u = 64 * wv + 16 * q4 + t;
          if (u >= nU - 1) break;
          const double *q = p + (int64_t)(u + 1) * nR;
          double b0 = q[0], b1 = q[1];
          double v = (((a0 + a1) + b0) + b1) * 0.25;
          double vm = v;  // the sum over the species
          if (nS > 1) {
            vm = 0.0 + v;
            for (int k = 1; k < nS; k++) {
              const double *pk = p + (int64_t)k * nVal + (int64_t)u * nR;
              vm += (((pk[0] + pk[1]) + pk[nR]) + pk[nR + 1]) * 0.25;
            }
          }
          double du = dU[u];
          double w = (ct * du) * dr;
          double uc = U_Val[u] + 0.5 * du;
          g += w * vm;
          ek += w * (0.5 * v * (uc * uc + J * J / (rc * rc)));
          s += w * ((v <= 0) ? 0.0 : -v * log(v));
          l1 += w * fabs(v);
          l2 += w * (v * v);
          if (nS > 1) {
            sAll += w * ((vm <= 0) ? 0.0 : -vm * log(vm));
            l2All += w * (vm * vm);
          }
          a0 = b0;
          a1 = b1;
        }
        G[q4] = g;
      }
      const double Wm = ((G[0] + G[1]) + G[2]) + G[3];
      m = wv == 0 ? Wm : m + Wm;
    }
  }
  const int64_t o = (int64_t)j * nR + r, n = (int64_t)nJ * nR;
  part[o] = m;
  part[n + o] = ek;
  part[2 * n + o] = s;
  part[3 * n + o] = l1;
  part[4 * n + o] = l2;
  part[5 * n + o] = sAll;
  part[6 * n + o] = l2All;
}

// Serial sums over j (moment) and over r (integrands), then over j; the
// potential energy of vlasovSolver21::potentialEnergy
// (src/vlasovSolver21.hxx:627-642, with the gathered array shifted by one,
// src/quadrature21.hxx:36).  out = {M, E, Ek, Ep, S, L1, L2} (the species
// files' line), then the statistics.txt line: with nS species the species
// loop of vlasovSolver21::statistics (:656-688) adds that line nS times, and
// S and L2 are those of the summed species (:690-699); with one species the
// same line again.
__global__ void __launch_bounds__(kBlock)
k_stats_final(const double *__restrict__ part, const double *__restrict__ R_Val, double G,
              double *__restrict__ dM, double *__restrict__ q, double *__restrict__ out, int nR,
              int nJ, int nS) {
  const int64_t n = (int64_t)nJ * nR;
  // the radial mass in the order of the moment scan: chunks of kJC slices
  for (int r = threadIdx.x; r < nR; r += blockDim.x) {
    double s = 0.0;
    for (int j0 = 0; j0 < nJ; j0 += kJC) {
      double S = 0.0;
      for (int j = j0; j < nJ && j < j0 + kJC; j++) S += part[(int64_t)j * nR + r];
      s += S;
    }
    dM[r] = s;
  }
  for (int t = threadIdx.x; t < 6 * nJ; t += blockDim.x) {
    int k = t / nJ, j = t % nJ;
    const double *p = part + (int64_t)(k + 1) * n + (int64_t)j * nR;
    double s = 0.0;
    for (int r = 0; r < nR - 1; r++) s += p[r];
    q[t] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // shell k (R[k]..R[k+1]) meets M(<R[k]) at its midpoint: the gathered
    // array of potentialEnergy is shifted by one (addSpare=1)
    double M = 0.0 + 0.0, Ep = 0.0;
    for (int i = 1; i < nR; i++) {
      double rr = 0.5 * (R_Val[i - 1] + R_Val[i]);
      Ep -= (dM[i - 1] * G * M) / rr;
      M += dM[i - 1];
    }
    M += dM[nR - 1];
    double tot[6];
    for (int k = 0; k < 6; k++) {
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
    if (nS == 1) {
      for (int k = 0; k < 7; k++) out[7 + k] = out[k];
    } else {
      double Ep_tot = 0, M_tot = 0, Ek_tot = 0, E_tot = 0, L1_tot = 0;
      for (int k = 0; k < nS; k++) {
        Ep_tot += out[3];
        M_tot += out[0];
        Ek_tot += out[2];
        E_tot += out[1];
        L1_tot += out[5];
      }
      out[7] = M_tot;
      out[8] = E_tot;
      out[9] = Ek_tot;
      out[10] = Ep_tot;
      out[11] = tot[4];
      out[12] = L1_tot;
      out[13] = tot[5];
    }
  }
}

// Shell density and potential from the moment (the definitions of
// vls_density_potential): rho[k] = dM[k] / (vol (R[k+1]^3 - R[k]^3)) with
// vol = 4 pi / 3, and phi[i] = -G (Mx[i] / R[i] + S[i]), S the suffix sum of
// dM[k] / rbar[k] over the shells k >= i, rbar[k] = (R[k] + R[k+1]) / 2,
// added from the outermost shell inwards by one thread.
__global__ void __launch_bounds__(kBlock)
k_density_potential(const double *__restrict__ R, const double *__restrict__ dM,
                    const double *__restrict__ Mx, double G, double vol, double *__restrict__ rho,
                    double *__restrict__ phi, int nR) {
  for (int k = threadIdx.x; k < nR; k += blockDim.x)
    rho[k] = k < nR - 1
                 ? dM[k] / (vol * (R[k + 1] * R[k + 1] * R[k + 1] - R[k] * R[k] * R[k]))
                 : 0.0;
  if (threadIdx.x == 0) {
    double S = 0.0;
    for (int i = nR - 1; i >= 0; i--) {
      if (i < nR - 1) S = dM[i] / (0.5 * (R[i] + R[i + 1])) + S;
      phi[i] = -(G * (Mx[i] / R[i] + S));
    }
  }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

static void time_record(Timing *t, hipStream_t s, const char *name) {
  hipEvent_t e;
  VLS_HIP_CHECK(hipEventCreate(&e));
  VLS_HIP_CHECK(hipEventRecord(e, s));
  t->ev.push_back(e);
  t->names.push_back(name);
}

void time_mark(Timing *t, hipStream_t s, const char *name) {
  if (!t || !t->on) return;
  if (t->only.empty() || t->only == name) time_record(t, s, name);
}

void time_begin(Timing *t, hipStream_t s, const char *name) {
  if (t && t->on && !t->only.empty() && t->only == name) time_record(t, s, name);
}

// counters of k_moment_fused: one per (chunk of slices, column tile), one more
size_t moment_counters(const Run &r) {
  return (size_t)((r.nR + 62) / 63) * ((r.nJ + kJC - 1) / kJC) + 1;
}

void launch_moment(Run &r, hipStream_t s) {
  launch_moment_into(r, MomentBufs{r.f, r.P, r.dM, r.Mx, r.acc, r.mTot, r.mCnt}, s);
}

void launch_moment_into(Run &r, const MomentBufs &b, hipStream_t s) {
  dim3 g(blocks_for(r.nR), r.nJ);
  if (r.nS == 1 && b.cnt && !getenv("VLS_MOMENT_3K")) {
    const size_t lds = std::max<size_t>(64 * ((r.nU + 63) / 64), 2 * r.nR) * sizeof(double);
    const MomentArgs ma{r.Ct1, r.dU, r.dR, r.gfac, b.P, b.dM, b.Mx, b.acc, b.mTot, b.cnt};
    hipLaunchKernelGGL(k_moment_fused, dim3((r.nR + 62) / 63 * r.nJ), dim3(256), lds, s, b.f,
                       r.nR, r.nU, r.nJ, ma);
    return;
  }
  if (r.nS > 1)
    hipLaunchKernelGGL(k_moment_partial_sp, g, dim3(kBlock), 0, s, b.f, r.Ct1, r.dU, r.dR, b.P,
                       r.nR, r.nU, r.nJ, r.nS, r.nVal);
  else
    hipLaunchKernelGGL(k_moment_cols, dim3((r.nR + 62) / 63 * r.nJ), dim3(256),
                       64 * ((r.nU + 63) / 64) * sizeof(double), s, b.f, r.Ct1, r.dU,
                       r.dR, b.P, r.nR, r.nU);
  launch_moment_scan_into(r, b, s);
}

void launch_moment_scan(Run &r, hipStream_t s) {
  launch_moment_scan_into(r, MomentBufs{r.f, r.P, r.dM, r.Mx, r.acc, r.mTot, nullptr}, s);
}

void launch_moment_scan_into(Run &r, const MomentBufs &b, hipStream_t s) {
  const int nC = (r.nJ + kJC - 1) / kJC;
  hipLaunchKernelGGL(k_moment_chunks, dim3(nC, (r.nR + kChunkThreads - 1) / kChunkThreads),
                     dim3(kChunkThreads), 0, s, b.P, r.nR, r.nJ);
  hipLaunchKernelGGL(k_moment_scan, dim3(1), dim3(kScanThreads), 2 * r.nR * sizeof(double), s, b.P,
                     r.gfac, b.dM, b.Mx, b.acc, b.mTot, r.nR, r.nJ);
}

void configure_tiles(Run &r, int maxLdsPerBlock) {
  // r-sweep: TR rows of LS doubles per block; v-sweep: (nU+2) x TC doubles.
  // Prefer tiles of about 52 KB so that three blocks share a 160 KB CU.
  // three blocks per 160 KB CU
  const int budget = 163840 / 3 < maxLdsPerBlock ? 163840 / 3 : maxLdsPerBlock;
  // r-sweep: odd row stride TP (LS) of the [k][row] tile, TR = min(64, TP)
  int tp = budget / (int)((r.nR + 2) * sizeof(double));
  if (tp > 65) tp = 65;
  if (tp % 2 == 0) tp--;
  if (tp < 1) tp = 1;
  r.LS = tp;
  r.TR = tp > 64 ? 64 : tp;
  r.TC = 64;
  while (r.TC > 1 && (size_t)r.TC * (r.nU + 2) * sizeof(double) > (size_t)budget) r.TC /= 2;
  // tile overrides (used by the autotuner): VLS_TILE_R = odd row stride of the
  // r-sweep tile, VLS_TILE_V = columns per v-sweep block
  if (const char *e = getenv("VLS_TILE_R")) {
    int v = atoi(e) | 1;
    if (v >= 1 && v <= 65) {
      r.LS = v;
      r.TR = v > 64 ? 64 : v;
    }
  }
  if (const char *e = getenv("VLS_TILE_V")) {
    int v = atoi(e);
    if (v >= 1 && v <= 64) r.TC = v;
  }
  // lines too long for 64 of them in a tile: streaming solvers
  // (VLS_SPLINE_STREAM=1 / 0 forces them on / off)
  r.streamR = r.TR < 64;
  r.streamV = r.TC < 64;
  if (const char *e = getenv("VLS_SPLINE_STREAM")) r.streamR = r.streamV = atoi(e) != 0;
  const size_t ldsR = (size_t)r.LS * (r.nR + 2) * sizeof(double);
  const size_t ldsV = (size_t)r.TC * (r.nU + 2) * sizeof(double);
  if ((!r.streamR && ldsR > (size_t)maxLdsPerBlock) || (!r.streamV && ldsV > (size_t)maxLdsPerBlock))
    throw Error(VLS_ERR_ARG, "grid too large for the LDS of one workgroup");
  // the tiled solvers may use all of a workgroup's LDS (runs of an ensemble,
  // and the schedules vls_autotune tries, need different amounts)
  VLS_HIP_CHECK(hipFuncSetAttribute((const void *)k_spline_r,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, maxLdsPerBlock));
  VLS_HIP_CHECK(hipFuncSetAttribute((const void *)k_spline_v,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, maxLdsPerBlock));
}

// Row version of the register line solver with coalesced memory access.
// A wave (one block) takes up to 64 consecutive rows u0 = 64 b ... of slice
// j (rows of f, contiguous in memory).  Right-hand sides arrive in chunks of
// kC elements per row through a small LDS transpose (loads of whole row
// segments, then lane = row reads its kC values); the loads of chunks c+1
// and c+2 are in flight during the elimination of chunk c.  Coefficients
// leave the same way, a chunk at a time as the back substitution produces
// them.  Loads are clamped into the wave's rows (the extra lanes of a
// partial wave solve a copy and store nothing).
template <int M>
__global__ void __launch_bounds__(64)
k_line_rows(const double *__restrict__ f, double *__restrict__ c, int nR, int nU, int nJ,
            int64_t cs, int lzmask, ThomasAxis ax) {
  constexpr int kC = 16, kP = kC + 1;  // chunk width, LDS row pitch
  constexpr int nIn = M - 2;           // right-hand sides per line (= nR)
  constexpr int nChIn = (nIn + kC - 1) / kC, nChOut = (M + kC - 1) / kC;
  __shared__ double chunk[64 * kP];
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const int j = blockIdx.y;
  const int u0 = blockIdx.x * 64;
  const int rows = max(min(64, nU - u0), 1);
  const int rsub = lane / kC, ksub = lane % kC;  // staging role: 4 rows x kC elements per pass
  const double *src = f + ((size_t)j * nU + min(u0, nU - 1)) * nIn;
  double *dst = c + (size_t)j * cs + (size_t)(u0 + 1) * M;
  double y[M];
  __shared__ HeadTab htab;
  const HeadRegs hr = head_load(ax, lane, lz);
  // chunks of loads in flight: two (the loads of a chunk take a memory
  // latency, the elimination of a chunk far less)
  constexpr int kAh = 2;
  double tmp[kAh][kC];
  auto load_chunk = [&](int ch, double *t) {
    const int e = min(ch * kC + ksub, nIn - 1);
#pragma unroll
    for (int q = 0; q < kC; q++) {
      const int row = min(q * 4 + rsub, rows - 1);
      t[q] = src[row * nIn + e];
    }
  };
  const double c16 = ax.c16 + (double)lz, dS = ax.dS + (double)lz, rdS = ax.rdS + (double)lz;
  double prev = ax.A03;
#pragma unroll
  for (int ch = 0; ch < kAh; ch++)
    if (ch < nChIn) load_chunk(ch, tmp[ch]);
  head_store(htab, hr, lane);  // ordered before its reads by the first chunk's barrier
  // Per chunk: stage it (and issue the loads of chunk ch + kAh into the
  // registers just staged), then the elimination.
#pragma unroll
  for (int ch = 0; ch < nChIn; ch++) {
#pragma unroll
    for (int q = 0; q < kC; q++) chunk[(q * 4 + rsub) * kP + ksub] = tmp[ch % kAh][q];
    if (ch + kAh < nChIn) load_chunk(ch + kAh, tmp[ch % kAh]);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kC; i++)
      if (ch * kC + i < nIn) y[1 + ch * kC + i] = chunk[lane * kP + i];
#pragma unroll
    for (int i = 0; i < kC; i++) {
      const int k = 1 + ch * kC + i;
      if (k <= M - 2) {
        double dv = dS, ev = rdS;
        if (k <= kHead) {
          dv = htab.d[k];
          ev = htab.rd[k];
        }
        y[k] = div_known(y[k] - c16 * prev, dv, ev);
        prev = y[k];
      }
    }
    __syncthreads();
  }
  double tt = 0.0 - ax.hi2 * y[M - 3];
  tt = tt - ax.p * y[M - 2];
  y[M - 1] = div_known(tt, ax.q, ax.rq);
  const double uS = ax.uS + (double)lz;
  double cn = y[M - 1];
  // back substitution and stores, last chunk first
#pragma unroll
  for (int ch = nChOut - 1; ch >= 0; ch--) {
#pragma unroll
    for (int i = kC - 1; i >= 0; i--) {
      const int k = ch * kC + i;
      if (k >= 1 && k <= M - 2) {
        double uv = uS;
        if (k <= kHead) uv = htab.u[k];
        y[k] = y[k] - uv * cn;
        cn = y[k];
      }
    }
    if (ch == 0) y[0] = (ax.A03 - ax.A01 * y[1]) - ax.A02 * y[2];
#pragma unroll
    for (int i = 0; i < kC; i++)
      if (ch * kC + i < M) chunk[lane * kP + i] = y[ch * kC + i];
    __syncthreads();
    // all LDS reads first, so that their latencies overlap, then the stores
    const int k = ch * kC + ksub;
    double sv[kC];
#pragma unroll
    for (int q = 0; q < kC; q++) sv[q] = chunk[(q * 4 + rsub) * kP + ksub];
    __syncthreads();
    if (u0 < nU) {
      if (rows == 64 && (ch + 1) * kC <= M) {
#pragma unroll
        for (int q = 0; q < kC; q++) dst[(q * 4 + rsub) * M + k] = sv[q];
      } else {
#pragma unroll
        for (int q = 0; q < kC; q++) {
          const int row = q * 4 + rsub;
          if (row < rows && k < M) dst[row * M + k] = sv[q];
        }
      }
    }
  }
}

// The kick folded into the drift's r-solve: the right-hand sides of a row are
// the kicked values of the row's points, evaluated from the kick's column
// splines d ([j][m][r], slice stride cs, row stride nR; k_kick_solve_moment)
// as k_kick_cols evaluates them, so the kicked f never goes to memory.  The
// solved rows go to c (rows 1..nU of c's layout), which may be another
// buffer than the one d is in.  A
// wave takes 64 rows of a slice (tiles in the order of xcd_tile) and, as
// k_line_rows, a lane per row with the row in VGPRs (the first kYL forward
// values in LDS); the rows are solved a chunk of 16 columns at a time.  The
// rows m of d that a chunk's points read lie in [lo, hi], from the extreme
// points of the chunk (the target row n grows with U[u] + a[r] in both
// operands); they are staged in LDS, kWin rows at most, with the loads of
// the next chunk's window in flight during the chunk's evaluation and
// elimination.  A chunk whose window is taller (kicks that differ by more
// than ~12 cells across 16 columns) reads d directly.  Same values, in the
// same order, as k_kick_cols followed by k_line_rows.
template <int M>
struct RowsKickCfg {
  static constexpr int kYL = M > 160 ? 32 : 0;
};

template <int M>
__global__ void __launch_bounds__(64)
k_line_rows_kick(const double *__restrict__ d, double *__restrict__ c,
                 const double *__restrict__ U_Val, const double *__restrict__ acc, double Ulo,
                 double Uhi, double start_v, double hinv_v, int64_t cs, int lzmask,
                 ThomasAxis ax, int nU) {
  constexpr int nR = M - 2;
  constexpr int kC = 16, kP = kC + 1;  // chunk width, LDS row pitch
  constexpr int nChIn = (nR + kC - 1) / kC, nChOut = (M + kC - 1) / kC;
  constexpr int kWin = 80, kWQ = kWin / 4;  // window rows, window loads per lane
  constexpr int kB = 8;                     // points per batch of window reads
  const int nT = (nU + 63) / 64;            // row tiles per slice
  constexpr int kYL = RowsKickCfg<M>::kYL;
  __shared__ double chunk[64 * kP];
  __shared__ double win[kWin * kP];
  __shared__ double accs[nR];
  __shared__ double us[64];
  __shared__ double ylds[(kYL > 0 ? kYL : 1) * 64];
  __shared__ HeadTab htab;
  const int lane = threadIdx.x;
  const int lz = lane & lzmask;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int j = tile / nT;
  const int u0 = (tile - j * nT) * 64;
  const int rows = min(64, nU - u0);
  // staging role: 4 rows x kC columns per pass, a lane's points are rows
  // q*4 + rsub (q = 0..15) of column ksub of the chunk
  const int rsub = lane / kC, ksub = lane % kC;
  const double *dj = d + (size_t)j * cs;
  double *dst = c + (size_t)j * cs + (size_t)(u0 + 1) * M;
  const HeadRegs hr = head_load(ax, lane, lz);
  const double Uown = U_Val[u0 + min(lane, rows - 1)];  // extra rows: a copy of the last
  const double Ufirst = U_Val[u0], Ulast = U_Val[u0 + rows - 1];
  constexpr int nAq = (nR + 63) / 64;
  double aq[nAq];
#pragma unroll
  for (int q = 0; q < nAq; q++) aq[q] = acc[min(lane + 64 * q, nR - 1)];
  // the target row n of a point at velocity newU after the kick, with the
  // evaluation's own operations, clamped to the padded column
  auto row_of = [&](double newU) {
    double yy = (newU - start_v) * hinv_v;
    yy = yy < 0.0 ? 0.0 : (yy > (double)(nU + 1) ? (double)(nU + 1) : yy);
    return (int)trunc(yy);
  };
  // window rows [lo, hi] of chunk ch; false: taller than kWin
  auto bounds = [&](int ch, int &lo, int &hi) {
    double amin = accs[min(ch * kC, nR - 1)], amax = amin;
#pragma unroll
    for (int i = 1; i < kC; i++) {
      const double a = accs[min(ch * kC + i, nR - 1)];
      amin = a < amin ? a : amin;
      amax = a > amax ? a : amax;
    }
    lo = row_of(Ufirst + amin);
    hi = min(row_of(Ulast + amax) + 3, nU + 1);
    return hi - lo + 1 <= kWin;
  };
  auto load_win = [&](int ch, int lo, int hi, double *t) {
    const int e = min(ch * kC + ksub, nR - 1);
#pragma unroll
    for (int q = 0; q < kWQ; q++) t[q] = dj[(size_t)min(lo + q * 4 + rsub, hi) * nR + e];
  };
  auto store_win = [&](const double *t) {
#pragma unroll
    for (int q = 0; q < kWQ; q++) win[(q * 4 + rsub) * kP + ksub] = t[q];
  };
  auto wave_sync = [] {  // the block is one wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // The kicked values of this lane's points of chunk ch (k_kick_cols), into
  // the chunk's LDS transpose: per batch of kB points the targets first,
  // then all their coefficient reads, then the sums.  Points outside the
  // velocity range are 0 (their reads take row lo and are not used).
  auto kick_chunk = [&](int ch, int lo, bool inWin) {
    const int e = min(ch * kC + ksub, nR - 1);
    const double a = accs[e];
#pragma unroll
    for (int b0 = 0; b0 < 16; b0 += kB) {
      double fr[kB], dv[kB][4];
      int m0[kB], m3[kB];
      bool in[kB];
#pragma unroll
      for (int i = 0; i < kB; i++) {
        const double newU = us[(b0 + i) * 4 + rsub] + a;
        in[i] = !(newU <= Ulo) && !(newU >= Uhi);
        double yy = in[i] ? (newU - start_v) * hinv_v : 0.0;
        int n;
        split(yy, n, yy);
        fr[i] = yy;
        m0[i] = in[i] ? n : lo;
        m3[i] = in[i] ? min(n + 3, nU + 1) : lo;
      }
      if (inWin) {
#pragma unroll
        for (int i = 0; i < kB; i++) {
          const double *w = win + (m0[i] - lo) * kP + ksub;
          dv[i][0] = w[0];
          dv[i][1] = w[kP];
          dv[i][2] = w[2 * kP];
          dv[i][3] = win[(m3[i] - lo) * kP + ksub];
        }
      } else {
#pragma unroll
        for (int i = 0; i < kB; i++) {
          const double *dc = dj + e;
          dv[i][0] = dc[(int64_t)m0[i] * nR];
          dv[i][1] = dc[(int64_t)(m0[i] + 1) * nR];
          dv[i][2] = dc[(int64_t)(m0[i] + 2) * nR];
          dv[i][3] = dc[(int64_t)m3[i] * nR];
        }
      }
#pragma unroll
      for (int i = 0; i < kB; i++) {
        const double v = kick_value(fr[i], dv[i][0], dv[i][1], dv[i][2], dv[i][3]);
        chunk[((b0 + i) * 4 + rsub) * kP + ksub] = in[i] ? v : 0.0;
      }
    }
  };

#pragma unroll
  for (int q = 0; q < nAq; q++)
    if (lane + 64 * q < nR) accs[lane + 64 * q] = aq[q];
  us[lane] = Uown;
  head_store(htab, hr, lane);
  wave_sync();
  double wt[kWQ];
  int lo0, hi0, lo1 = 0, hi1 = 0;
  bool in0 = bounds(0, lo0, hi0), in1 = false;
  load_win(0, lo0, hi0, wt);
  store_win(wt);
  if (nChIn > 1) {
    in1 = bounds(1, lo1, hi1);
    load_win(1, lo1, hi1, wt);
  }
  wave_sync();
  double y[M];
  const double c16 = ax.c16 + (double)lz, dS = ax.dS + (double)lz, rdS = ax.rdS + (double)lz;
  double prev = ax.A03;
  // Software pipeline: the evaluation of chunk ch + 1 (VALU-heavy, from the
  // window in LDS into the chunk transpose) shares a step with the
  // elimination of chunk ch (a dependent chain that leaves the VALU idle),
  // whose right-hand sides are already in registers.  Each step then moves
  // chunk ch + 1 into registers and the window of ch + 2 into LDS.
  auto take_chunk = [&](int ch) {
#pragma unroll
    for (int i = 0; i < kC; i++)
      if (ch * kC + i < nR) y[1 + ch * kC + i] = chunk[lane * kP + i];
  };
  auto next_window = [&](int ch) {  // window ch into LDS, the loads of ch + 1 out
    if (ch < nChIn) {
      store_win(wt);
      lo0 = lo1;
      hi0 = hi1;
      in0 = in1;
      if (ch + 1 < nChIn) {
        in1 = bounds(ch + 1, lo1, hi1);
        load_win(ch + 1, lo1, hi1, wt);
      }
    }
  };
  kick_chunk(0, lo0, in0);
  wave_sync();
  take_chunk(0);
  next_window(1);
  wave_sync();
#pragma unroll
  for (int ch = 0; ch < nChIn; ch++) {
    if (ch + 1 < nChIn) kick_chunk(ch + 1, lo0, in0);
#pragma unroll
    for (int i = 0; i < kC; i++) {
      const int k = 1 + ch * kC + i;
      if (k <= M - 2) {
        double dv = dS, ev = rdS;
        if (k <= kHead) {
          dv = htab.d[k];
          ev = htab.rd[k];
        }
        const double yk = div_known(y[k] - c16 * prev, dv, ev);
        if (k <= kYL)
          ylds[(k - 1) * 64 + lane] = yk;
        else
          y[k] = yk;
        prev = yk;
      }
    }
    wave_sync();
    if (ch + 1 < nChIn) {
      take_chunk(ch + 1);
      next_window(ch + 2);
    }
    wave_sync();
  }
  double tt = 0.0 - ax.hi2 * (M - 3 <= kYL ? ylds[(M - 4) * 64 + lane] : y[M - 3]);
  tt = tt - ax.p * y[M - 2];
  y[M - 1] = div_known(tt, ax.q, ax.rq);
  const double uS = ax.uS + (double)lz;
  double cn = y[M - 1], c2 = 0.0;
  // back substitution and stores, last chunk first (as k_line_rows)
#pragma unroll
  for (int ch = nChOut - 1; ch >= 0; ch--) {
#pragma unroll
    for (int i = kC - 1; i >= 0; i--) {
      const int k = ch * kC + i;
      if (k >= 1 && k <= M - 2) {
        double uv = uS;
        if (k <= kHead) uv = htab.u[k];
        const double yk = k <= kYL ? ylds[(k - 1) * 64 + lane] : y[k];
        c2 = cn;
        cn = yk - uv * cn;
        y[k] = cn;
      }
    }
    if (ch == 0) y[0] = (ax.A03 - ax.A01 * cn) - ax.A02 * c2;
#pragma unroll
    for (int i = 0; i < kC; i++)
      if (ch * kC + i < M) chunk[lane * kP + i] = y[ch * kC + i];
    wave_sync();
    const int k = ch * kC + ksub;
    double sv[kC];
#pragma unroll
    for (int q = 0; q < kC; q++) sv[q] = chunk[(q * 4 + rsub) * kP + ksub];
    wave_sync();
    if (rows == 64 && (ch + 1) * kC <= M) {
#pragma unroll
      for (int q = 0; q < kC; q++) dst[(q * 4 + rsub) * M + k] = sv[q];
    } else {
#pragma unroll
      for (int q = 0; q < kC; q++) {
        const int row = q * 4 + rsub;
        if (row < rows && k < M) dst[row * M + k] = sv[q];
      }
    }
  }
}

// Register line solvers compiled for fixed line lengths: the r-direction
// kernels for the standard nR, the v-direction kernels (the drift's v-solve
// and every kick kernel) for the standard nU; a run has register lines when
// both of its axes are in the tables.  Other shapes use the LDS-tiled or
// streaming kernels.
using LineFn = void (*)(const double *, double *, int, int, int, int64_t, int, ThomasAxis);
using RowsFn = void (*)(const double *, double *, int, int, int, int64_t, int, ThomasAxis);
using KickFn = void (*)(double *, const double *, const double *, double *, double, double, double,
                        double, int, int, ThomasAxis, double *, int64_t, int);
using KickMomentFn = void (*)(const double *, double *, int, int, int64_t, int, ThomasAxis,
                              MomentArgs);
using RowsKickFn = void (*)(const double *, double *, const double *, const double *, double, double,
                            double, double, int64_t, int, ThomasAxis, int);
struct RowShape {
  int nR;
  RowsFn rows;
  RowsKickFn rowsKick;
};
struct ColShape {
  int nU;
  LineFn cols, kickCols;
  KickFn kick, kickKeep;
  KickMomentFn kickMoment;
};
#define VLS_ROW_SHAPE(R) {R, k_line_rows<(R) + 2>, k_line_rows_kick<(R) + 2>}
#define VLS_COL_SHAPE(U)                                                              \
  {U, k_line<(U) + 2, false>, k_line<(U) + 2, true>, k_kick_line<(U) + 2, false>,    \
   k_kick_line<(U) + 2, true>, k_kick_solve_moment<(U) + 2>}
static const RowShape kRowShapes[] = {
    VLS_ROW_SHAPE(17),  VLS_ROW_SHAPE(25),  VLS_ROW_SHAPE(33),  VLS_ROW_SHAPE(49),
    VLS_ROW_SHAPE(65),  VLS_ROW_SHAPE(97),  VLS_ROW_SHAPE(129), VLS_ROW_SHAPE(193),
    VLS_ROW_SHAPE(201),
};
static const ColShape kColShapes[] = {
    VLS_COL_SHAPE(13),  VLS_COL_SHAPE(17),  VLS_COL_SHAPE(33),  VLS_COL_SHAPE(49),
    VLS_COL_SHAPE(65),  VLS_COL_SHAPE(97),  VLS_COL_SHAPE(129), VLS_COL_SHAPE(193),
    VLS_COL_SHAPE(201),
};
#undef VLS_ROW_SHAPE
#undef VLS_COL_SHAPE

// the register kernels of one run's shape
struct LineShape {
  RowsFn rows;
  RowsKickFn rowsKick;
  LineFn cols, kickCols;
  KickFn kick, kickKeep;
  KickMomentFn kickMoment;
};

static const RowShape *row_shape(int nR) {
  for (const RowShape &rs : kRowShapes)
    if (rs.nR == nR) return &rs;
  return nullptr;
}
static const ColShape *col_shape(int nU) {
  for (const ColShape &cs : kColShapes)
    if (cs.nU == nU) return &cs;
  return nullptr;
}

static const LineShape *line_shape(const Run &r) {
  if (r.lineMode == 0 || (r.lineMode < 0 && getenv("VLS_NO_REGISTER_LINES"))) return nullptr;
  const RowShape *rs = row_shape(r.nR);
  const ColShape *cs = col_shape(r.nU);
  if (!rs || !cs) return nullptr;
  // one entry per (nR, nU) pair, made on first use (std::map keeps its
  // elements in place)
  static std::mutex mu;
  static std::map<std::pair<int, int>, LineShape> made;
  std::lock_guard<std::mutex> lock(mu);
  auto it = made.find({r.nR, r.nU});
  if (it == made.end())
    it = made.emplace(std::make_pair(r.nR, r.nU),
                      LineShape{rs->rows, rs->rowsKick, cs->cols, cs->kickCols, cs->kick,
                                cs->kickKeep, cs->kickMoment})
             .first;
  return &it->second;
}

bool kick_rsolve_shape(const Run &r) {
  return r.nS == 1 && row_shape(r.nR) && col_shape(r.nU);
}

// The drift's r-direction solves (rows of src -> padded rows of c) and its
// v-direction solves (in place on c), with the run's schedule.
static void launch_spline_r(Run &r, const double *src, hipStream_t s) {
  if (const LineShape *ls = line_shape(r)) {
    hipLaunchKernelGGL(ls->rows, dim3((r.nU + 63) / 64, r.nJ), dim3(64), 0, s, src, r.c, r.nR,
                       r.nU, r.nJ, r.cs, 0, r.tr);
  } else if (r.streamR) {
    hipLaunchKernelGGL(k_spline_r_stream, dim3((r.nU + 63) / 64, r.nJ), dim3(64), 0, s, src, r.c,
                       r.nR, r.nU, r.cs, 0, r.tr);
  } else {
    dim3 gr((r.nU + r.TR - 1) / r.TR, r.nJ);
    hipLaunchKernelGGL(k_spline_r, gr, dim3(64), (size_t)r.LS * (r.nR + 2) * sizeof(double), s,
                       src, r.c, r.nR, r.nU, r.nJ, r.cs, r.TR, r.LS, 0, r.tr);
  }
}

// rowsIn: the r-solved rows in c's layout (null: in c, solved in place;
// otherwise register lines only)
static void launch_spline_v(Run &r, const double *rowsIn, hipStream_t s) {
  if (const LineShape *ls = line_shape(r)) {
    hipLaunchKernelGGL(ls->cols, dim3((r.nR + 2 + 63) / 64 * r.nJ), dim3(64), 0, s, rowsIn, r.c,
                       r.nR, r.nU, r.nJ, r.cs, 0, r.tv);
  } else if (r.streamV) {
    hipLaunchKernelGGL(k_spline_v_stream, dim3((r.nR + 2 + 63) / 64 * r.nJ), dim3(64), 0, s, r.c,
                       r.nR, r.cs, 0, r.tv);
  } else {
    dim3 gv((r.nR + 2 + r.TC - 1) / r.TC, r.nJ);
    hipLaunchKernelGGL(k_spline_v, gv, dim3(64), (size_t)r.TC * (r.nU + 2) * sizeof(double), s,
                       r.c, r.nR, r.nU, r.nJ, r.cs, r.TC, 0, r.tv);
  }
}

void launch_spline(Run &r, const double *src, hipStream_t s, Timing *t) {
  time_begin(t, s, "spline_r");
  launch_spline_r(r, src, s);
  time_mark(t, s, "spline_r");
  time_begin(t, s, "spline_v");
  launch_spline_v(r, nullptr, s);
  time_mark(t, s, "spline_v");
}

// Doubles of global scratch the fused kick needs (0: no fused kick for this
// shape): one (nU+3) x 64 area per wave.
// The fused kick evaluates each column's points one after the other, so it
// pays only when the grid has enough columns to fill the GPU with waves
// (about 40k columns at 201x201x200); smaller grids run the column solves and
// the evaluation as two kernels, the evaluation a thread per point.
static bool fused_kick(const Run &r) {
  const int64_t waves = (int64_t)r.nJ * ((r.nR + kKW - 1) / kKW);
  if (r.kickMode >= 0) return r.kickMode == 1 && line_shape(r) != nullptr;
  if (getenv("VLS_FUSED_KICK")) return line_shape(r) != nullptr;
  if (!getenv("VLS_NO_KICK_MOMENT")) return false;  // the default is kick_moment's schedule
  return line_shape(r) != nullptr && waves >= 512 && !getenv("VLS_NO_FUSED_KICK");
}

size_t kick_scratch_doubles(const Run &r) {
  if (!fused_kick(r)) return 0;
  const size_t tiles = (r.nR + kKW - 1) / kKW;
  return tiles * r.nJ * (size_t)(r.nU + 3) * 64;
}

static bool kick_moment_possible(const Run &r) { return r.nS == 1 && r.mCnt; }

// The moment fused into the kick's column splines (k_kick_solve_moment on
// register lines, k_kick_vsolve_moment on the others), then the evaluation
// a thread per point: the default for one species (kickMode 2 or -1;
// VLS_NO_KICK_MOMENT turns the default off).  The adjoint's recorded steps,
// which keep the column splines, take it too (the splines go to the kept
// buffer instead of c).
static bool kick_moment(const Run &r) {
  if (!kick_moment_possible(r)) return false;
  if (r.kickMode >= 0) return r.kickMode == 2;
  return getenv("VLS_NO_KICK_MOMENT") == nullptr && !getenv("VLS_FUSED_KICK");
}

// The kick folded into the drift's r-solve (kickMode 3, or VLS_KICK_RSOLVE):
// k_kick_solve_moment into c, then k_line_rows_kick into f's buffer.  Same
// bits as kick_moment's schedule; at 201x201x200 no faster (the evaluation
// then runs at one wave per SIMD with the row solves), so not the default.
bool kick_in_rsolve(const Run &r) {
  if (!r.fRows || !line_shape(r) || !kick_moment_possible(r)) return false;
  if (r.kickMode >= 0) return r.kickMode == 3;
  return getenv("VLS_KICK_RSOLVE") != nullptr;
}

void launch_kick_rsolve(Run &r, hipStream_t s, Timing *t) {
  const LineShape *ls = line_shape(r);
  const MomentArgs ma{r.Ct1, r.dU, r.dR, r.gfac, r.P, r.dM, r.Mx, r.acc, r.mTot, r.mCnt};
  time_begin(t, s, "kick_solve");
  hipLaunchKernelGGL(ls->kickMoment, dim3((r.nR + 62) / 63 * r.nJ), dim3(64),
                     2 * (size_t)r.nR * sizeof(double), s, r.f, r.c, r.nR, r.nJ, r.cs, 0, r.tv, ma);
  time_mark(t, s, "kick_solve");
  // f is not needed again before the drift rewrites it: its buffer takes the
  // kicked rows' r-solve (the working set stays f, c and the feet)
  time_begin(t, s, "kick_spline_r");
  hipLaunchKernelGGL(ls->rowsKick, dim3((r.nU + 63) / 64 * r.nJ), dim3(64), 0, s, r.c, r.f,
                     r.U_Val, r.acc, r.Ulo, r.Uhi, r.start_v, r.hinv_v, r.cs, 0, r.tr, r.nU);
  time_mark(t, s, "kick_spline_r");
}

void launch_spline_v_timed(Run &r, hipStream_t s, Timing *t) {
  time_begin(t, s, "spline_v");
  launch_spline_v(r, r.f, s);  // the kicked rows' r-solve left them in f's buffer
  time_mark(t, s, "spline_v");
}

void launch_spline_kick(Run &r, hipStream_t s, Timing *t, bool moment, double *dKeep) {
  if (moment && kick_moment(r)) {
    // the column splines into c, or where the adjoint keeps them
    double *d = dKeep ? dKeep : r.c;
    const MomentArgs ma{r.Ct1, r.dU, r.dR, r.gfac, r.P, r.dM, r.Mx, r.acc, r.mTot, r.mCnt};
    const dim3 grid((r.nR + 62) / 63 * r.nJ);
    time_begin(t, s, "kick_solve");
    if (const LineShape *ls = line_shape(r))
      hipLaunchKernelGGL(ls->kickMoment, grid, dim3(64), 2 * (size_t)r.nR * sizeof(double), s, r.f, d,
                         r.nR, r.nJ, r.cs, 0, r.tv, ma);
    else
      hipLaunchKernelGGL(k_kick_vsolve_moment, grid, dim3(64), (size_t)(r.nU + 2 * r.nR) * sizeof(double),
                         s, r.f, d, r.nR, r.nU, r.nJ, r.cs, 0, r.tv, ma);
    time_mark(t, s, "kick_solve");
    time_begin(t, s, "kick_eval");
    launch_kick_eval(r, s, d);
    time_mark(t, s, "kick_eval");
    return;
  }
  if (moment) {
    time_begin(t, s, "moment");
    launch_moment(r, s);
    time_mark(t, s, "moment");
  }
  if (fused_kick(r) && r.gsc) {
    const LineShape *ls = line_shape(r);
    time_begin(t, s, "kick_line");
    hipLaunchKernelGGL(dKeep ? ls->kickKeep : ls->kick, dim3((r.nR + kKW - 1) / kKW * r.nJ),
                       dim3(64), 0, s, r.f, r.U_Val, r.acc, r.gsc, r.Ulo, r.Uhi, r.start_v,
                       r.hinv_v, r.kickScratch ? 1 : 0, 0, r.tv, dKeep, r.cs, r.nR);
    time_mark(t, s, "kick_line");
    return;
  }
  time_begin(t, s, "kick");
  launch_kick(r, s, dKeep ? dKeep : r.c);
  time_mark(t, s, "kick");
}

// ---------------------------------------------------------------- schedules
// vls_autotune: the drift's line solves and the kick are timed on the run's
// own data in each of their schedules and the fastest is kept.  Every
// schedule runs the same operations per line in the same order, so the
// choice changes the time of a step, never its bits.
namespace {
struct Sched {
  int lineMode, kickMode, TR, LS, TC;
  bool streamR, streamV;
  double splineUs, kickUs;
};
std::mutex g_tuneMu;
std::map<std::tuple<int, int, int, int>, Sched> g_tuned;

void apply_sched(Run &r, const Sched &c) {
  r.lineMode = c.lineMode;
  r.kickMode = c.kickMode;
  r.TR = c.TR;
  r.LS = c.LS;
  r.TC = c.TC;
  r.streamR = c.streamR;
  r.streamV = c.streamV;
}

// mean duration in microseconds of fn() on stream s over reps launches,
// after one launch to warm up
template <class F>
double time_us(hipStream_t s, F &&fn, int reps = 5) {
  hipEvent_t a, b;
  VLS_HIP_CHECK(hipEventCreate(&a));
  VLS_HIP_CHECK(hipEventCreate(&b));
  fn();
  VLS_HIP_CHECK(hipEventRecord(a, s));
  for (int k = 0; k < reps; k++) fn();
  VLS_HIP_CHECK(hipEventRecord(b, s));
  VLS_HIP_CHECK(hipEventSynchronize(b));
  float ms = 0;
  VLS_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  VLS_HIP_CHECK(hipGetLastError());
  return ms * 1e3 / reps;
}

std::string describe(const Run &r, const Sched &c, bool cached) {
  char buf[256];
  std::string sp;
  if (c.lineMode == 1) {
    sp = "register lines";
  } else {
    sp = c.streamR ? std::string("r-solve streaming") : "r-solve LDS tile of " + std::to_string(c.TR) + " rows";
    sp += c.streamV ? ", v-solve streaming" : ", v-solve LDS tile of " + std::to_string(c.TC) + " columns";
  }
  static const char *const kKick[] = {"two-kernel", "fused", "moment-fused", "r-solve-folded"};
  snprintf(buf, sizeof buf,
           "%dx%dx%d: %s, %s kick (spline %.1f us, moment, kick and r-solve %.1f us%s)",
           r.nR, r.nU, r.nJ, sp.c_str(), kKick[c.kickMode], c.splineUs, c.kickUs,
           cached ? ", cached" : "");
  return buf;
}
}  // namespace

std::string autotune(Run &r, hipStream_t s, int maxLds) {
  const auto key = std::make_tuple(r.nR, r.nU, r.nJ, maxLds);
  {
    std::lock_guard<std::mutex> lock(g_tuneMu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      apply_sched(r, it->second);
      return describe(r, it->second, true);
    }
  }
  // the drift's spline: register lines where the shape has them, else the
  // best LDS-tiled or streaming solver for each direction
  Sched best{0, 0, r.TR, r.LS, r.TC, r.streamR, r.streamV, 1e30, 0};
  Run &probe = r;  // tried in place; the best schedule is applied at the end
  probe.lineMode = 0;
  double bestR = 1e30, bestV = 1e30;
  std::vector<int> lsCands = {7, 15, 31, 63, 65};
  for (int k = -1; k < (int)lsCands.size(); k++) {
    if (k >= 0 && (size_t)lsCands[k] * (r.nR + 2) * sizeof(double) > (size_t)maxLds) continue;
    probe.streamR = k < 0;
    if (k >= 0) {
      probe.LS = lsCands[k];
      probe.TR = probe.LS > 64 ? 64 : probe.LS;
    }
    const double us = time_us(s, [&] { launch_spline_r(probe, r.f, s); });
    if (us < bestR) {
      bestR = us;
      best.streamR = probe.streamR;
      best.LS = probe.LS;
      best.TR = probe.TR;
    }
  }
  for (int tc : {-1, 8, 16, 32, 64}) {
    if (tc > 0 && (size_t)tc * (r.nU + 2) * sizeof(double) > (size_t)maxLds) continue;
    probe.streamV = tc < 0;
    if (tc > 0) probe.TC = tc;
    const double us = time_us(s, [&] { launch_spline_v(probe, nullptr, s); });
    if (us < bestV) {
      bestV = us;
      best.streamV = probe.streamV;
      best.TC = probe.TC;
    }
  }
  best.splineUs = bestR + bestV;
  probe.lineMode = 1;
  if (line_shape(probe)) {
    const double us = time_us(s, [&] { launch_spline(probe, r.f, s, nullptr); });
    if (us < best.splineUs) {
      best.lineMode = 1;
      best.splineUs = us;
    }
  }
  // the moment, the kick and the drift's r-solve (which the kick can be
  // folded into), on a copy of f (the kick works in place)
  apply_sched(r, best);
  launch_moment(r, s);
  double *f0 = r.f, *tmp = nullptr;
  // as large as f's buffer, which the kick folded into the r-solve writes
  const size_t fDoubles = r.fRows ? std::max<size_t>(r.nVal, (size_t)r.nJ * r.cs) : r.nVal;
  VLS_HIP_CHECK(hipMalloc(&tmp, fDoubles * sizeof(double)));
  r.f = tmp;
  try {
    auto time_kick = [&](int mode) {
      r.kickMode = mode;
      VLS_HIP_CHECK(hipMemcpyAsync(tmp, f0, (size_t)r.nVal * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
      return time_us(s, [&] {
        if (mode == 3) {
          launch_kick_rsolve(r, s, nullptr);
        } else {
          launch_spline_kick(r, s, nullptr, true);
          launch_spline_r(r, r.f, s);
        }
      });
    };
    best.kickMode = 0;
    best.kickUs = time_kick(0);
    if (kick_moment_possible(r)) {
      for (int mode : {2, 3}) {
        if (mode == 3 && (!r.fRows || best.lineMode != 1)) continue;
        const double us = time_kick(mode);
        if (us < best.kickUs) {
          best.kickMode = mode;
          best.kickUs = us;
        }
      }
    }
    if (best.lineMode == 1) {
      if (!r.gsc) {
        r.kickMode = 1;
        const size_t n = kick_scratch_doubles(r);
        void *p = nullptr;
        VLS_HIP_CHECK(hipMalloc(&p, n * sizeof(double)));
        r.allocs.push_back(p);
        r.allocBytes.push_back(n * sizeof(double));
        r.gsc = static_cast<double *>(p);
      }
      const double us = time_kick(1);
      if (us < best.kickUs) {
        best.kickMode = 1;
        best.kickUs = us;
      }
    }
  } catch (...) {
    r.f = f0;
    (void)hipStreamSynchronize(s);
    (void)hipFree(tmp);
    throw;
  }
  r.f = f0;
  VLS_HIP_CHECK(hipStreamSynchronize(s));
  VLS_HIP_CHECK(hipFree(tmp));
  apply_sched(r, best);
  {
    std::lock_guard<std::mutex> lock(g_tuneMu);
    g_tuned[key] = best;
  }
  return describe(r, best, false);
}

void launch_kick(Run &r, hipStream_t s, double *d) {
  if (const LineShape *ls = line_shape(r))
    hipLaunchKernelGGL(ls->kickCols, dim3((r.nR + 63) / 64 * r.nJ), dim3(64), 0, s, r.f, d, r.nR,
                       r.nU, r.nJ, r.cs, 0, r.tv);
  else
    hipLaunchKernelGGL(k_kick_vsolve, dim3((r.nR + 63) / 64 * r.nJ), dim3(64), 0, s, r.f, d, r.nR,
                       r.nU, r.cs, 0, r.tv);
  launch_kick_eval(r, s, d);
}

void launch_kick_eval(Run &r, hipStream_t s, const double *d) {
  const int nS = r.nU * r.nR;
  if (nS >= (1 << 17)) {
    const int nT = (nS + 4 * kBlock - 1) / (4 * kBlock);
    hipLaunchKernelGGL(k_kick_cols<4>, dim3(nT * r.nJ), dim3(kBlock), 0, s, r.f, d, r.U_Val,
                       r.acc, r.nR, r.nU, r.cs, r.Ulo, r.Uhi, r.start_v, r.hinv_v);
  } else {
    const int nT = (nS + 2 * kBlock - 1) / (2 * kBlock);
    hipLaunchKernelGGL(k_kick_cols<2>, dim3(nT * r.nJ), dim3(kBlock), 0, s, r.f, d, r.U_Val,
                       r.acc, r.nR, r.nU, r.cs, r.Ulo, r.Uhi, r.start_v, r.hinv_v);
  }
}

void launch_drift(Run &r, hipStream_t s) {
  ChainTables ch{r.ch_off,  r.ch_len, r.ch_head,  r.ch_xr,
                 r.ch_xu,   r.ch_out, r.ch_vals,  r.nChains, r.rollChains ? 1 : 0};
  const int64_t nb2 = (r.nVal + 2 * kBlock - 1) / (2 * kBlock);
  const int per2 = (int)((nb2 + 7) / 8);
  hipLaunchKernelGGL(k_drift, dim3(8 * per2), dim3(kBlock), 0, s, r.f, r.c,
                     (const double2 *)r.feet, r.nVal, r.nU * r.nR, r.cs, r.nR + 2, per2, ch);
}

void mark_chains(Run &r, hipStream_t s) {
  if (r.nChains == 0) return;
  hipLaunchKernelGGL(k_mark_chains, dim3(blocks_for(r.nChains)), dim3(kBlock), 0, s,
                     (double2 *)r.feet, r.ch_point, r.nChains);
}

void launch_density_potential(Run &r, double *rho, double *phi, hipStream_t s) {
  launch_density_potential_from(r, r.dM, r.Mx, rho, phi, s);
}

void launch_density_potential_from(Run &r, const double *dM, const double *Mx, double *rho,
                                   double *phi, hipStream_t s) {
  hipLaunchKernelGGL(k_density_potential, dim3(1), dim3(kBlock), 0, s, r.R_Val, dM, Mx, r.sp.G,
                     (4.0 * M_PI) / 3.0, rho, phi, r.nR);
}

void launch_stats(Run &r, hipStream_t s) {
  dim3 g(blocks_for(r.nR), r.nJ);
  hipLaunchKernelGGL(k_stats_partial, g, dim3(kBlock), 0, s, r.f, r.Ct1, r.dU, r.dR, r.R_Val,
                     r.U_Val, r.J_Val, r.st_part, r.nR, r.nU, r.nJ, r.nS, r.nVal);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kBlock), 0, s, r.st_part, r.R_Val, r.sp.G,
                     r.st_q + 6 * (int64_t)r.nJ, r.st_q, r.st_out, r.nR, r.nJ, r.nS);
}

}  // namespace vls
