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
// group; the slab's tiles in (y, z) blocks of by x bz rows, blocks y-major,
// inside a block t fastest, then y, then z: a front of nt by bz tiles reads
// neighbour rows of about (nt + 2)(by + 2)(bz + 2) tiles, so the blocks are
// chosen (by the autotuner) so that a front and its halo stay in L2.
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
  int by, bz;   // (y, z) blocks of the slab's tiles (t fastest inside a block)
  int nrhs;
  int rg;       // rhs per group
};
struct NetItem {
  int tile, rhs;  // rhs >= nrhs: a ragged last group's slot, skipped
};
struct NetRange {
  int t0, ntl, nit, first, step;  // the XCD's tiles, its item count, this wave's items
};

// Blocks of a stage: as many as the CUs hold at once (per_cu blocks each),
// fewer for small problems, a multiple of 8 where there are at least 8.
inline void net_grid(const Geom& g, int ncu, int per_cu, int threads, int nrhs, int rg, int by,
                     int bz, NetSweep& s, int* nblk) {
  s.tiles = (g.Vh + 15) / 16;
  s.nrhs = nrhs;
  s.rg = rg < 1 ? 1 : (rg > nrhs ? nrhs : rg);
  s.lz = g.L[2];
  const int nw = threads / 64;
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
  // the (y, z) blocks: whole numbers of rows, else one block column per y
  s.by = 1;
  s.bz = g.L[2];
  if (s.nt > 0 && by >= 1 && bz >= 1 && s.ytiles % by == 0 && g.L[2] % bz == 0) {
    s.by = by;
    s.bz = bz;
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

// tile of position o (0 .. ntl - 1) of the XCD's tile order
__host__ __device__ __forceinline__ int net_tile(const NetSweep& s, const NetRange& R, int o) {
  if (s.nt > 0) {
    const int tt = o % s.nt;
    int rest = o / s.nt;
    const int yb = rest % s.by;
    rest /= s.by;
    const int zb = rest % s.bz;
    rest /= s.bz;
    const int nby = s.ytiles / s.by;
    const int yi = (rest % nby) * s.by + yb, zz = (rest / nby) * s.bz + zb;
    return R.t0 + tt * s.ztiles + zz * s.ytiles + yi;
  }
  return R.t0 + o;
}
__host__ __device__ __forceinline__ NetItem net_item(const NetSweep& s, const NetRange& R, int it) {
  const int per = R.ntl * s.rg;
  const int grp = it / per;
  const int rem = it - grp * per;
  const int o = rem / s.rg;
  const int rhs = grp * s.rg + (rem - o * s.rg);
  return NetItem{net_tile(s, R, o), rhs};
}

// The list of an XCD over the live rhs only (the dynamic order of the stage
// kernels, which hand the list out from a counter): live is a bit mask of
// rhs 0..63; the k-th live rhs takes the place of rhs k, in groups of rg
// (the last group holds what is left), so the list has ntl x nlive items
// and none of them is skipped.
struct NetLive {
  unsigned long long mask;
  int nlive, nit;
};
__host__ __device__ __forceinline__ NetLive net_live(const NetRange& R, unsigned long long mask) {
  const int n = __builtin_popcountll(mask);
  return NetLive{mask, n, R.ntl * n};
}
// the k-th set bit of m (k < popcount(m))
__host__ __device__ __forceinline__ int kth_bit(unsigned long long m, int k) {
  for (int i = 0; i < k; ++i) m &= m - 1;
  return __builtin_ctzll(m);
}
__host__ __device__ __forceinline__ NetItem net_item_live(const NetSweep& s, const NetRange& R,
                                                          const NetLive& L, int it) {
  const int per = R.ntl * s.rg;
  const int grp = it / per;
  const int rem = it - grp * per;
  const int left = L.nlive - grp * s.rg;
  const int rgi = left < s.rg ? left : s.rg;
  const int o = rem / rgi;
  const int k = grp * s.rg + (rem - o * rgi);
  return NetItem{net_tile(s, R, o), kth_bit(L.mask, k)};
}

// Half index of the neighbour of site s (coordinates c, either parity) along
// path dh (mu = dh / 2, forward for even dh): neighbour() of ii_internal.h
// written as s plus a step and a wrap correction that each select between 0
// and one wave-uniform value, so that no uniform value needs a vector
// register of its own.
__host__ __device__ __forceinline__ int sel0(bool c, int v) { return c ? v : 0; }
__host__ __device__ __forceinline__ int nb_of(const Geom& g, const SiteCoord& c, int s, int dh) {
  const bool odd = (c.x & 1) != 0;
  switch (dh) {
    case 0: return s + sel0(odd, 1) - sel0(odd && c.xh == g.Lxh - 1, g.Lxh);
    case 1: return s - sel0(!odd, 1) + sel0(!odd && c.xh == 0, g.Lxh);
    case 2: return s + g.sY - sel0(c.y == g.L[1] - 1, g.L[1] * g.sY);
    case 3: return s - g.sY + sel0(c.y == 0, g.L[1] * g.sY);
    case 4: return s + g.sZ - sel0(c.z == g.L[2] - 1, g.L[2] * g.sZ);
    case 5: return s - g.sZ + sel0(c.z == 0, g.L[2] * g.sZ);
    case 6: return s + g.sT - sel0(c.t == g.L[3] - 1, g.L[3] * g.sT);
    default: return s - g.sT + sel0(c.t == 0, g.L[3] * g.sT);
  }
}

}  // namespace ii

#endif
