// Host build of the library's site-index and tile arithmetic (the same
// inline functions the kernels use: decode_site, neighbour, hop_block_base,
// xcd_remap from iroiro_amd/csrc), checked exhaustively on lattices whose
// half volume is and is not a multiple of the tile sizes:
//   * neighbour() returns a site of the opposite parity inside [0, V/2),
//     one lattice step away in direction mu, and the backward neighbour of
//     the forward neighbour is the site itself;
//   * every site traversal the hop autotuner may pick (contiguous with 256
//     or 128 threads per block, the XCD time-strip orders) and the 64-site
//     tiles of the multi-rhs hop cover [0, V/2) exactly once;
//   * the 16-site tiles of the batched network's stage kernels cover every
//     tile once.
// Built with the host sanitizers (address, undefined behaviour) by
// tests/test_sanitizers.py; exits non-zero on the first violation.
#include "../../iroiro_amd/csrc/ii_stencil.h"

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
                         {32, 32, 32, 64}, {12, 10, 6, 8}};
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
    // batched network: 16-site tiles, contiguous runs per block
    const int tiles = (g.Vh + 15) / 16;
    const int nb = std::min(512, tiles);
    const int tpb = (tiles + nb - 1) / nb;
    const int nblk = (tiles + tpb - 1) / tpb;
    std::vector<int> th(tiles, 0);
    for (int b = 0; b < nblk; ++b)
      for (int t = b * tpb; t < std::min(b * tpb + tpb, tiles); ++t) th[t]++;
    for (int t = 0; t < tiles; ++t) CHECK(th[t] == 1, "network tile %d covered %d times", t, th[t]);
    CHECK(tiles * 16 >= g.Vh && (tiles - 1) * 16 < g.Vh, "network tile count");
  }
  if (fails) return 1;
  std::printf("index_check: ok\n");
  return 0;
}
