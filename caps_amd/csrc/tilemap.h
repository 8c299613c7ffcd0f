// Persistent-grid tile scheduling shared by the leaf GEMM kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace caps {

constexpr int NXCD = 8;  // gfx950: 8 XCDs, blocks dispatched round-robin

struct TileMap {
  int tiles_m, tiles_n, per_prob, total;
  int patch;  // 1: 8 x 4 patch order inside a problem
};

inline TileMap make_tilemap(int m, int tile, long long problems, bool &ok) {
  TileMap t;
  t.tiles_m = (m + tile - 1) / tile;
  t.tiles_n = t.tiles_m;
  t.per_prob = t.tiles_m * t.tiles_n;
  long long total = (long long)t.per_prob * problems;
  ok = total <= 0x7fffffffLL;
  t.total = (int)total;
  t.patch = (t.tiles_m % 8 == 0) && (t.tiles_n % 4 == 0);
  return t;
}

__device__ __forceinline__ void tile_coords(const TileMap &tm, int t, int &prob, int &ti,
                                            int &tj) {
  prob = t / tm.per_prob;
  int r = t - prob * tm.per_prob;
  if (tm.patch) {
    // patches of 8 tile-rows x 4 tile-cols, patch-rows fastest
    int patch = r >> 5, w = r & 31;
    int pr = tm.tiles_m >> 3;
    ti = (patch % pr) * 8 + (w & 7);
    tj = (patch / pr) * 4 + (w >> 3);
  } else {
    ti = r % tm.tiles_m;
    tj = r / tm.tiles_m;
  }
}

// Block b of a persistent grid of G blocks (G a multiple of NXCD) runs on XCD
// b % NXCD; give each XCD a contiguous range of the round's G tile ids.  A
// permutation of [0, G) exactly when NXCD divides G, which persistent_grid
// guarantees.
__host__ __device__ __forceinline__ int xcd_tile(int b, int G) {
  return (b % NXCD) * (G / NXCD) + b / NXCD;
}

// Index, in CAPS quadrant order (11, 21, 12, 22 at each level), of leaf
// (bi, bj) of a 2^d x 2^d quadtree of leaves.
__device__ __forceinline__ long long quad_leaf(int bi, int bj, int d) {
  long long L = 0;
  for (int l = d - 1; l >= 0; l--) L = 4 * L + (((bi >> l) & 1) + 2 * ((bj >> l) & 1));
  return L;
}

// Grid size of a persistent launch: one round of blocks_per_cu blocks per CU,
// at most enough for `total` tiles, always a positive multiple of NXCD (so
// xcd_tile is a permutation; a CU count that is not a multiple of NXCD is
// rounded down).
inline int persistent_grid(int blocks_per_cu, int cus, int total) {
  int G = (cus * blocks_per_cu / NXCD) * NXCD;
  const int need = ((total + NXCD - 1) / NXCD) * NXCD;
  if (G > need) G = need;
  if (G < NXCD) G = NXCD;
  return G;
}

}  // namespace caps
