// Host build of the library's site-index and tile arithmetic (the same
// inline functions the kernels use: decode_site, neighbour, hop_block_base,
// xcd_remap from iroiro_amd/csrc), checked exhaustively on lattices whose
// half volume is and is not a multiple of the tile sizes:
//   * neighbour() returns a site of the opposite parity inside [0, V/2),
//     one lattice step away in direction mu, and the backward neighbour of
//     the forward neighbour is the site itself; the batched network's
//     branch-free nb_of() gives the same site;
//   * every site traversal the hop autotuner may pick (contiguous with 256
//     or 128 threads per block, the XCD time-strip orders) and the 64-site
//     tiles of the multi-rhs hop cover [0, V/2) exactly once;
//   * the batched network's stage kernels take every (16-site tile, rhs)
//     item once (ii_sweep.h: per-XCD slabs, t-fastest fronts, rhs groups),
//     in the static order and in the dynamic order over the live rhs.
// Built with the host sanitizers (address, undefined behaviour) by
// tests/test_sanitizers.py; exits non-zero on the first violation.
#include "../../iroiro_amd/csrc/ii_stencil.h"
#include "../../iroiro_amd/csrc/ii_sweep.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace ii;

static int fails = 0;
#define CHECK(c, ...)                                         \
  do {                                                        \
    if (!(c)) {                                               \
      std::fprintf(stderr, "index_check: " __VA_ARGS__);     \
      std::fprintf(stderr, "\n");                             \
      if (++fails > 20) std::exit(1);                         \
    }                                                         \
  } while (0)

static Geom make_geom(const int L[4]) {
  Geom g{};
  for (int m = 0; m < 4; ++m) g.L[m] = L[m];
  g.V = L[0] * L[1] * L[2] * L[3];
  g.Vh = g.V / 2;
  g.Lxh = L[0] / 2;
  g.sY = g.Lxh;
  g.sZ = g.Lxh * L[1];
  g.sT = g.Lxh * L[1] * L[2];
  g.stride = (g.Vh + 63) / 64 * 64;
  return g;
}

static void check_cover(const Geom& g, const HopOrder& o, int bs, const char* what) {
  const int nblk = (g.Vh + bs - 1) / bs;
  std::vector<int> hit(g.Vh, 0);
  for (int b = 0; b < nblk; ++b) {
    const int base = hop_block_base(g, o, b, nblk, bs);
    CHECK(base >= 0 && base % bs == 0, "%s: block %d base %d", what, b, base);
    for (int t = 0; t < bs; ++t) {
      const int hs = base + t;
      if (hs < g.Vh) hit[hs]++;
    }
  }
  for (int hs = 0; hs < g.Vh; ++hs)
    CHECK(hit[hs] == 1, "%s (mode %d W %d bs %d) L=%dx%dx%dx%d: site %d covered %d times", what,
          o.mode, o.W, bs, g.L[0], g.L[1], g.L[2], g.L[3], hs, hit[hs]);
}

int main() {
  const int lats[][4] = {{2, 2, 2, 2},   {4, 4, 4, 4},   {6, 2, 2, 6},     {6, 4, 4, 6},
                         {4, 6, 2, 4},   {8, 8, 8, 16},  {16, 16, 16, 32}, {24, 24, 24, 48},
                         {32, 32, 32, 64}, {12, 10, 6, 8},  {8, 8, 32, 32}};
  for (const auto& L : lats) {
    const Geom g = make_geom(L);
    // neighbours
    for (int par = 0; par < 2; ++par)
      for (int hs = 0; hs < g.Vh; ++hs) {
        const SiteCoord c = decode_site(g, hs, par);
        CHECK(((c.x + c.y + c.z + c.t) & 1) == par, "parity of site %d", hs);
        const int co[4] = {c.x, c.y, c.z, c.t};
        for (int mu = 0; mu < 4; ++mu)
          for (int fw = 0; fw < 2; ++fw) {
            const int nb = neighbour(g, c, hs, mu, fw != 0);
            CHECK(nb >= 0 && nb < g.Vh, "neighbour %d of site %d out of range", nb, hs);
            if (nb < 0 || nb >= g.Vh) continue;
            const SiteCoord n = decode_site(g, nb, 1 - par);
            const int no[4] = {n.x, n.y, n.z, n.t};
            for (int k = 0; k < 4; ++k) {
              const int want = k == mu ? (co[k] + (fw ? 1 : L[k] - 1)) % L[k] : co[k];
              CHECK(no[k] == want, "L=%dx%dx%dx%d site %d mu %d fw %d: coordinate %d is %d not %d",
                    L[0], L[1], L[2], L[3], hs, mu, fw, k, no[k], want);
            }
            CHECK(neighbour(g, n, nb, mu, fw == 0) == hs, "neighbour not involutive at %d", hs);
            CHECK(nb_of(g, c, hs, 2 * mu + (fw ? 0 : 1)) == nb,
                  "L=%dx%dx%dx%d site %d parity %d mu %d fw %d: network neighbour %d, not %d",
                  L[0], L[1], L[2], L[3], hs, par, mu, fw, nb_of(g, c, hs, 2 * mu + (fw ? 0 : 1)),
                  nb);
          }
      }
    // hop traversals
    for (int bs : {256, 128}) {
      HopOrder o;
      o.bs = bs;
      check_cover(g, o, bs, "contiguous");
    }
    const int P = g.Lxh * g.L[1];
    if (P % kBlock == 0 && g.L[3] % 8 == 0)
      for (int W : {1, 2, 4, 8, 16})
        if ((g.L[3] / 8) % W == 0) check_cover(g, HopOrder{1, W, P / kBlock}, kBlock, "strip");
    check_cover(g, HopOrder(), 64, "multi-rhs tiles");
    // batched network: every (tile, rhs) item exactly once for each stage's
    // block size, rhs count and grouping, on a 256-CU and a 38-CU device
    for (int ncu : {256, 38})
      for (int per_cu : {1, 4})
        for (int nrhs : {1, 5, 12})
          for (int rg : {1, 4, 12})
           for (int shape = 0; shape < 4; ++shape) {
            const int threads = 512;
            const int bys[4] = {1, 8, 4, 2}, bzs[4] = {g.L[2], 4, 8, 2};
            NetSweep s{};
            int nblk = 0;
            net_grid(g, ncu, per_cu, threads, nrhs, rg, bys[shape], bzs[shape], s, &nblk);
            CHECK(s.tiles * 16 >= g.Vh && (s.tiles - 1) * 16 < g.Vh, "network tile count");
            CHECK(nblk >= 1, "network grid");
            std::vector<int> th((size_t)s.tiles * nrhs, 0);
            const int nw = threads / 64;
            for (int b = 0; b < nblk; ++b)
              for (int w = 0; w < nw; ++w) {
                const NetRange R = net_range(s, b, nblk, nw, w);
                for (int it = R.first; it < R.nit; it += R.step) {
                  const NetItem m = net_item(s, R, it);
                  if (m.rhs >= nrhs) continue;
                  CHECK(m.tile >= 0 && m.tile < s.tiles && m.rhs >= 0, "network item %d -> %d",
                        it, m.tile);
                  if (m.tile >= 0 && m.tile < s.tiles) th[(size_t)m.tile * nrhs + m.rhs]++;
                }
              }
            for (size_t i = 0; i < th.size(); ++i)
              CHECK(th[i] == 1, "L=%dx%dx%dx%d network item (tile %d, rhs %d) covered %d times "
                    "(ncu %d, %d blocks per CU, rg %d, nt %d, by %d, bz %d)", L[0], L[1], L[2],
                    L[3], (int)(i / nrhs), (int)(i % nrhs), th[i], ncu, per_cu, rg, s.nt, s.by,
                    s.bz);
            // the dynamic order: each XCD's list over the live rhs, handed out
            // from one counter per XCD, takes every (tile, live rhs) once
            const unsigned long long all = nrhs == 64 ? ~0ull : (1ull << nrhs) - 1;
            for (unsigned long long mask : {all, all & 0x5ull, all & ~1ull, 0ull,
                                            all & (1ull << (nrhs - 1))}) {
              std::vector<int> tl((size_t)s.tiles * nrhs, 0);
              for (int x = 0; x < s.nxcd && x < nblk; ++x) {
                const NetRange R = net_range(s, x, nblk, nw, 0);
                const NetLive Lv = net_live(R, mask);
                for (int it = 0; it < Lv.nit; ++it) {
                  const NetItem m = net_item_live(s, R, Lv, it);
                  const bool ok = m.tile >= R.t0 && m.tile < R.t0 + R.ntl && m.rhs >= 0 &&
                                  m.rhs < nrhs && ((mask >> m.rhs) & 1);
                  CHECK(ok, "dynamic network item %d -> (tile %d, rhs %d)", it, m.tile, m.rhs);
                  if (ok) tl[(size_t)m.tile * nrhs + m.rhs]++;
                }
              }
              for (size_t i = 0; i < tl.size(); ++i) {
                const int want = (mask >> (i % nrhs)) & 1 ? 1 : 0;
                CHECK(tl[i] == want, "L=%dx%dx%dx%d dynamic network item (tile %d, rhs %d) "
                      "covered %d times, want %d (mask %llx, rg %d, nblk %d)", L[0], L[1], L[2],
                      L[3], (int)(i / nrhs), (int)(i % nrhs), tl[i], want, mask, rg, nblk);
              }
            }
          }
  }
  if (fails) return 1;
  std::printf("index_check: ok\n");
  return 0;
}
