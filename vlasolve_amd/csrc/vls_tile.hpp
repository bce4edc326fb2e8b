// Block order of the column kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace vls {

// Workgroups go to the 8 XCDs round-robin (block b to XCD b % 8), each XCD
// with its own L2.  xcd_tile maps block b of a 1D grid of nB blocks to a tile
// index so that consecutive tiles run on one XCD (a bijection of 0..nB-1):
// the column tiles of one slice, which share the cache lines at their edges,
// then meet in one L2.
__device__ __forceinline__ int xcd_tile(int b, int nB) {
  const int x = b & 7, q = nB >> 3, rem = nB & 7;
  return x * q + min(x, rem) + (b >> 3);
}

// (column tile, slice) of this block in a 1D grid of nX column tiles per
// slice, in the order of xcd_tile
struct Tile2 {
  int x, j;
};
__device__ __forceinline__ Tile2 xcd_tile2(int nX) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = t / nX;
  return {t - j * nX, j};
}

}  // namespace vls
