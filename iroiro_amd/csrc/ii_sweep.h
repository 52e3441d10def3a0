// Work order of the batched network's stage kernels (ii_bnet.hip), host
// and device: the same functions run in the kernels and in the host check
// of tests/cpu/index_check.cpp.
//
// A stage runs over (16-site tile, rhs) items.  Each XCD (block mod 8 under
// round-robin dispatch) takes one eighth of the tiles, a slab of time
// slices in site order, and the waves of its blocks take the slab's items
// in turn (wave j nw + w of the XCD takes items j nw + w, then + the XCD's
// wave count, ...): at any time the XCD works on one front of consecutive
// items.  The list: the group of rg rhs, then the tile, then the rhs in the
// group; tiles of the slab t fastest, then z, then y, so that a tile's
// t-neighbours are the next items, its z-neighbours nt tiles away and its
// y-neighbours nt Lz tiles away, inside the front.
#ifndef II_SWEEP_H
#define II_SWEEP_H

#include "ii_internal.h"

namespace ii {

struct NetSweep {
  int tiles;    // 16-site tiles of the output parity
  int nxcd;     // 8: the blocks of each XCD share an eighth of the tiles; 1: one list
  int tiles_x;  // tiles per XCD (a contiguous range in site order)
  int nt;       // time slices per XCD, swept t fastest (0: site order)
  int ytiles;   // tiles per z-plane
  int ztiles;   // tiles per time slice
  int lz;       // L_z
  int nrhs;
  int rg;       // rhs per group
};
struct NetItem {
  int tile, rhs;  // rhs >= nrhs: a ragged last group's slot, skipped
};
struct NetRange {
  int t0, ntl, nit, first, step;  // the XCD's tiles, its item count, this wave's items
};

// Blocks of a stage: as many as the CUs hold at once (1024 / threads per
// block per CU: the LDS of the A operands and the 128-register budget of 4
// waves per SIMD), fewer for small problems, a multiple of 8 where there
// are at least 8.
inline void net_grid(const Geom& g, int ncu, int threads, int nrhs, int rg, NetSweep& s,
                     int* nblk) {
  s.tiles = (g.Vh + 15) / 16;
  s.nrhs = nrhs;
  s.rg = rg < 1 ? 1 : (rg > nrhs ? nrhs : rg);
  s.lz = g.L[2];
  const int per_cu = threads >= 1024 ? 1 : 1024 / threads, nw = threads / 64;
  const long need = ((long)s.tiles * nrhs + nw - 1) / nw;
  long nb = (long)ncu * per_cu;
  if (need < nb) nb = need < 1 ? 1 : need;
  if (nb >= 8) nb -= nb % 8;
  *nblk = (int)nb;
  s.nxcd = nb % 8 == 0 ? 8 : 1;
  s.tiles_x = (s.tiles + s.nxcd - 1) / s.nxcd;
  s.nt = 0;
  s.ytiles = s.ztiles = 0;
  const int plane = g.Lxh * g.L[1];  // half sites per z-plane
  if (s.nxcd == 8 && plane % 16 == 0 && g.L[3] % 8 == 0) {
    s.ytiles = plane / 16;
    s.ztiles = s.ytiles * g.L[2];
    if ((long)s.ztiles * g.L[3] == s.tiles && s.tiles_x == (g.L[3] / 8) * s.ztiles)
      s.nt = g.L[3] / 8;
  }
}

__host__ __device__ __forceinline__ NetRange net_range(const NetSweep& s, int block, int nblk,
                                                       int nw, int wave) {
  const int x = block % s.nxcd, j = block / s.nxcd;
  const int t0 = x * s.tiles_x;
  int ntl = s.tiles - t0;
  if (ntl > s.tiles_x) ntl = s.tiles_x;
  if (ntl < 0) ntl = 0;
  const int ngrp = (s.nrhs + s.rg - 1) / s.rg;
  return NetRange{t0, ntl, ngrp * ntl * s.rg, j * nw + wave, (nblk / s.nxcd) * nw};
}

__host__ __device__ __forceinline__ NetItem net_item(const NetSweep& s, const NetRange& R, int it) {
  const int per = R.ntl * s.rg;
  const int grp = it / per;
  const int rem = it - grp * per;
  const int o = rem / s.rg;
  const int rhs = grp * s.rg + (rem - o * s.rg);
  if (s.nt > 0) {
    const int tt = o % s.nt, rest = o / s.nt;
    const int zz = rest % s.lz, yi = rest / s.lz;
    return NetItem{R.t0 + tt * s.ztiles + zz * s.ytiles + yi, rhs};
  }
  return NetItem{R.t0 + o, rhs};
}

}  // namespace ii

#endif
