// Batched mixed-precision flexible GCR on the even/odd Schur operator with
// the learned preconditioner: the propagator solve of many right-hand sides
// (the 12 columns of Qprop_EvenOdd::calc, qprop_EvenOdd.cpp:11-23) as one
// batched solve.  No counterpart in the reference, which solves each column
// with Solver_CG on Fopr_DdagD (solver_CG.cpp:29-160); the outer structure is
// Inverter_WilsonLike::invert (eoUtils.cpp:43-80): bhat = b_e - D_eo b_o, the
// Schur solve for the even part, x_o = b_o - D_oe x_e.
//
// Per restart cycle, in FP32 with every rhs in each launch (complex scalars
// per rhs, <a, b> = sum conj(a) b):
//   z_k = M r                        (the batched network, ii_bnet.hip)
//   w_k = Dhat z_k                   (two FP32 multi-rhs hops)
//   beta_j = <w_j, w_k> / |w_j|^2,   w_k -= sum_j beta_j w_j   (k - n <= j < k)
//   alpha_k = <w_k, r> / |w_k|^2,    r -= alpha_k w_k
// (classical Gram-Schmidt against the last n = northo directions of the
// cycle: one pass for the dot products, one for the update; n >= the cycle
// length is the full GCR, smaller n the truncated one, which keeps the
// residual minimal along each new direction but not over the whole cycle,
// and reads fewer vectors per direction; with n = 1 the dot pass also gives
// |w'_k|^2 and <w'_k, r> from <w_{k-1}, w_k>, |w_k|^2, <w_k, r> and
// <w_{k-1}, r>, so the orthogonalisation and the residual update share one
// pass).  The solution is never formed inside a cycle: with
// z'_k = z_k - sum_j beta_kj z'_j = sum_i T_ki z_i, the cycle's update is
// sum_i gamma_i z_i, gamma_i = sum_k alpha_k T_ki, added to the FP64
// solution in one pass at the end of the cycle.  Then the FP64 true
// residual rho = bhat - Dhat y restarts the cycle (r = rho in FP32); an rhs
// is done when |rho| <= tol |b| (|b| of the full source: with x_o
// reconstructed, this is the full-operator residual).  Inside a cycle an
// rhs stops taking directions once its recursive |r| is below 0.9 tol |b|.
//
// All scalars live on the device and every kernel reads its rhs's flags, so
// the host queues a whole cycle (optionally as one captured HIP graph) and
// reads the flags once per cycle.  With use_net = 0 the same solver runs
// with z_k = r (no preconditioner), for comparison.  A cycle that does not
// halve |rho| three times in a row for some rhs ends the flexible
// iteration; the remaining defect is solved by the batched mixed-precision
// CG on Dhat^dag Dhat (the plain-CG fallback).
#include "ii_ops.h"
#include "ii_stencil.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

namespace ii {

namespace {

constexpr int kMaxK = 16;       // directions per cycle at most
constexpr int kGThreads = 256;

struct GScal {
  double rr;              // recursive |r|^2 of the cycle
  double rho2;            // last FP64 true residual^2
  double b2;              // |b|^2 of the full source
  double tol2;            // tol^2 |b|^2
  double cyc_start;       // |rho|^2 at the start of the cycle
  double best;            // smallest |rho|^2 so far
  double w2[kMaxK];       // |w_j|^2 after orthogonalisation
  double alpha[kMaxK][2];
  double beta[kMaxK][kMaxK][2];  // beta[k][j]
  double gamma[kMaxK][2];
  int idle;               // 1: skip the direction kernels (done or cycle finished)
  int done;               // reached the tolerance
  int count;              // directions taken
  int cycles;             // restart cycles (FP64 true residuals)
  int stall;              // consecutive cycles without halving |rho|
  int nan;
  int taken;              // directions taken in the current cycle
  int pad;
};
static_assert(sizeof(GScal) % sizeof(double) == 0, "GScal is read as a double array");
constexpr int kGInts = sizeof(GScal) / sizeof(int);

__device__ __forceinline__ bool gfin(double v) { return isfinite(v); }

// sum over the nb partials of plane j of rhs blockIdx.x
__device__ __forceinline__ double gsum(const double* part, int nplanes, int j, int nb) {
  const double* p = part + ((size_t)blockIdx.x * nplanes + j) * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
  return block_sum(s);
}

// ------------------------------------------------------------ site kernels
// FP32 half fields [rhs][12][stride]; slot fields [slot][rhs][12][stride].
// <w_j, w_k> for j0 <= j < k into planes 2j, 2j+1
__global__ void __launch_bounds__(kGThreads)
k_g_dots(Geom g, const float2* __restrict__ W, size_t slot, int j0, int k,
         const GScal* __restrict__ gs, double* __restrict__ part) {
  const int rhs = blockIdx.y;
  if (gs[rhs].idle) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t he = (size_t)12 * g.stride;
  double acc[2 * kMaxK];
#pragma unroll
  for (int j = 0; j < 2 * kMaxK; ++j) acc[j] = 0.0;
  if (hs < g.Vh) {
    const float2* wk = W + k * slot + rhs * he + hs;
    for (int c = 0; c < 12; ++c) {
      const float2 b = wk[(size_t)c * g.stride];
#pragma unroll
      for (int j = 0; j < kMaxK; ++j)
        if (j >= j0 && j < k) {
          const float2 a = W[j * slot + rhs * he + (size_t)c * g.stride + hs];
          acc[2 * j] += (double)a.x * b.x + (double)a.y * b.y;
          acc[2 * j + 1] += (double)a.x * b.y - (double)a.y * b.x;
        }
    }
  }
  const int nb = gridDim.x;
#pragma unroll
  for (int j = 0; j < 2 * kMaxK; ++j)
    if (j >= 2 * j0 && j < 2 * k) {
      const double s = block_sum(acc[j]);
      if (threadIdx.x == 0) part[((size_t)rhs * 2 * kMaxK + j) * nb + blockIdx.x] = s;
    }
}

__global__ void k_g_beta(const double* __restrict__ part, int nb, int j0, int k,
                         GScal* __restrict__ gs) {
  GScal& S = gs[blockIdx.x];
  if (S.idle) return;
  for (int j = j0; j < k; ++j) {
    const double re = gsum(part, 2 * kMaxK, 2 * j, nb);
    const double im = gsum(part, 2 * kMaxK, 2 * j + 1, nb);
    if (threadIdx.x == 0) {
      S.beta[k][j][0] = re / S.w2[j];
      S.beta[k][j][1] = im / S.w2[j];
    }
  }
}

// w_k -= sum_{j0 <= j < k} beta_kj w_j; partials |w_k|^2 (plane 0), <w_k, r> (1, 2)
__global__ void __launch_bounds__(kGThreads)
k_g_orth(Geom g, float2* __restrict__ W, size_t slot, int j0, int k, const float2* __restrict__ R,
         const GScal* __restrict__ gs, double* __restrict__ part) {
  const int rhs = blockIdx.y;
  const GScal& S = gs[rhs];
  if (S.idle) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t he = (size_t)12 * g.stride;
  float br[kMaxK], bi[kMaxK];
#pragma unroll
  for (int j = 0; j < kMaxK; ++j) {
    br[j] = (j >= j0 && j < k) ? (float)S.beta[k][j][0] : 0.f;
    bi[j] = (j >= j0 && j < k) ? (float)S.beta[k][j][1] : 0.f;
  }
  double ww = 0.0, wr = 0.0, wi = 0.0;
  if (hs < g.Vh) {
    float2* wk = W + k * slot + rhs * he + hs;
    const float2* r = R + rhs * he + hs;
    for (int c = 0; c < 12; ++c) {
      float2 v = wk[(size_t)c * g.stride];
#pragma unroll
      for (int j = 0; j < kMaxK; ++j)
        if (j >= j0 && j < k) {
          const float2 a = W[j * slot + rhs * he + (size_t)c * g.stride + hs];
          v.x -= br[j] * a.x - bi[j] * a.y;
          v.y -= br[j] * a.y + bi[j] * a.x;
        }
      if (k > 0) wk[(size_t)c * g.stride] = v;
      const float2 rv = r[(size_t)c * g.stride];
      ww += (double)v.x * v.x + (double)v.y * v.y;
      wr += (double)v.x * rv.x + (double)v.y * rv.y;
      wi += (double)v.x * rv.y - (double)v.y * rv.x;
    }
  }
  const int nb = gridDim.x;
  const double s0 = block_sum(ww), s1 = block_sum(wr), s2 = block_sum(wi);
  if (threadIdx.x == 0) {
    double* p = part + (size_t)rhs * 3 * nb + blockIdx.x;
    p[0] = s0;
    p[nb] = s1;
    p[2 * (size_t)nb] = s2;
  }
}

__global__ void k_g_alpha(const double* __restrict__ part, int nb, int k, GScal* __restrict__ gs) {
  GScal& S = gs[blockIdx.x];
  if (S.idle) return;
  const double ww = gsum(part, 3, 0, nb);
  const double wr = gsum(part, 3, 1, nb);
  const double wi = gsum(part, 3, 2, nb);
  if (threadIdx.x == 0) {
    S.w2[k] = ww;
    S.alpha[k][0] = wr / ww;
    S.alpha[k][1] = wi / ww;
    if (!(ww > 0.0) || !gfin(S.alpha[k][0]) || !gfin(S.alpha[k][1])) {
      S.nan = 1;
      S.idle = 1;
      S.done = 1;
    }
  }
}

// r -= alpha_k w_k; partials |r|^2; with Rb, r also in the network's bf16
// input layout
__global__ void __launch_bounds__(kGThreads)
k_g_rupd(Geom g, const float2* __restrict__ W, size_t slot, int k, float2* __restrict__ R,
         uint32_t* __restrict__ Rb, const GScal* __restrict__ gs, double* __restrict__ part) {
  const int rhs = blockIdx.y;
  const GScal& S = gs[rhs];
  if (S.idle) return;
  const float ar = (float)S.alpha[k][0], ai = (float)S.alpha[k][1];
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t he = (size_t)12 * g.stride;
  double rr = 0.0;
  if (hs < g.Vh) {
    const float2* w = W + k * slot + rhs * he + hs;
    float2* r = R + rhs * he + hs;
    float2 nv[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      const float2 a = w[(size_t)c * g.stride];
      float2 v = r[(size_t)c * g.stride];
      v.x -= ar * a.x - ai * a.y;
      v.y -= ar * a.y + ai * a.x;
      r[(size_t)c * g.stride] = v;
      nv[c] = v;
      rr += (double)v.x * v.x + (double)v.y * v.y;
    }
    if (Rb) net_in_store(Rb, (size_t)rhs * g.stride + hs, nv);
  }
  const double s = block_sum(rr);
  if (threadIdx.x == 0) part[(size_t)rhs * gridDim.x + blockIdx.x] = s;
}

// Window 1 (each direction orthogonalised against the previous one only),
// k >= 1: one pass for every scalar the direction needs, one for the
// update.  Partials of <w_{k-1}, w_k> (planes 0, 1), |w_k|^2 (2),
// <w_k, r> (3, 4), <w_{k-1}, r> (5, 6).
__global__ void __launch_bounds__(kGThreads)
k_g_dots1(Geom g, const float2* __restrict__ W, size_t slot, int k, const float2* __restrict__ R,
          const GScal* __restrict__ gs, double* __restrict__ part) {
  const int rhs = blockIdx.y;
  if (gs[rhs].idle) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t he = (size_t)12 * g.stride;
  double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (hs < g.Vh) {
    const float2* wp = W + (k - 1) * slot + rhs * he + hs;
    const float2* wk = W + k * slot + rhs * he + hs;
    const float2* r = R + rhs * he + hs;
    for (int c = 0; c < 12; ++c) {
      const float2 a = wp[(size_t)c * g.stride], b = wk[(size_t)c * g.stride],
                   v = r[(size_t)c * g.stride];
      acc[0] += (double)a.x * b.x + (double)a.y * b.y;
      acc[1] += (double)a.x * b.y - (double)a.y * b.x;
      acc[2] += (double)b.x * b.x + (double)b.y * b.y;
      acc[3] += (double)b.x * v.x + (double)b.y * v.y;
      acc[4] += (double)b.x * v.y - (double)b.y * v.x;
      acc[5] += (double)a.x * v.x + (double)a.y * v.y;
      acc[6] += (double)a.x * v.y - (double)a.y * v.x;
    }
  }
  const int nb = gridDim.x;
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const double sj = block_sum(acc[j]);
    if (threadIdx.x == 0) part[((size_t)rhs * 7 + j) * nb + blockIdx.x] = sj;
  }
}

// beta = <w_{k-1}, w_k> / |w_{k-1}|^2; for w' = w_k - beta w_{k-1}:
// |w'|^2 = |w_k|^2 - |<w_{k-1}, w_k>|^2 / |w_{k-1}|^2 and
// <w', r> = <w_k, r> - conj(beta) <w_{k-1}, r> (the sums are FP64 over exact
// products of the stored FP32 values); alpha = <w', r> / |w'|^2
__global__ void k_g_ab1(const double* __restrict__ part, int nb, int k, GScal* __restrict__ gs) {
  GScal& S = gs[blockIdx.x];
  if (S.idle) return;
  double v[7];
  for (int j = 0; j < 7; ++j) v[j] = gsum(part, 7, j, nb);
  if (threadIdx.x == 0) {
    const double wp = S.w2[k - 1];
    const double br = v[0] / wp, bi = v[1] / wp;
    const double ww = v[2] - (v[0] * v[0] + v[1] * v[1]) / wp;
    const double wr = v[3] - (br * v[5] + bi * v[6]);
    const double wi = v[4] - (br * v[6] - bi * v[5]);
    S.beta[k][k - 1][0] = br;
    S.beta[k][k - 1][1] = bi;
    S.w2[k] = ww;
    S.alpha[k][0] = wr / ww;
    S.alpha[k][1] = wi / ww;
    if (!(ww > 0.0) || !gfin(br) || !gfin(bi) || !gfin(S.alpha[k][0]) || !gfin(S.alpha[k][1])) {
      S.nan = 1;
      S.idle = 1;
      S.done = 1;
    }
  }
}

// w_k -= beta w_{k-1}, r -= alpha w_k (the new w_k); partials |r|^2; with Rb,
// r also in the network's bf16 input layout
__global__ void __launch_bounds__(kGThreads)
k_g_upd1(Geom g, float2* __restrict__ W, size_t slot, int k, float2* __restrict__ R,
         uint32_t* __restrict__ Rb, const GScal* __restrict__ gs, double* __restrict__ part) {
  const int rhs = blockIdx.y;
  const GScal& S = gs[rhs];
  if (S.idle) return;
  const float br = (float)S.beta[k][k - 1][0], bi = (float)S.beta[k][k - 1][1];
  const float ar = (float)S.alpha[k][0], ai = (float)S.alpha[k][1];
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t he = (size_t)12 * g.stride;
  double rr = 0.0;
  if (hs < g.Vh) {
    const float2* wp = W + (k - 1) * slot + rhs * he + hs;
    float2* wk = W + k * slot + rhs * he + hs;
    float2* r = R + rhs * he + hs;
    float2 nv[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      const float2 a = wp[(size_t)c * g.stride];
      float2 w = wk[(size_t)c * g.stride];
      w.x -= br * a.x - bi * a.y;
      w.y -= br * a.y + bi * a.x;
      wk[(size_t)c * g.stride] = w;
      float2 v = r[(size_t)c * g.stride];
      v.x -= ar * w.x - ai * w.y;
      v.y -= ar * w.y + ai * w.x;
      r[(size_t)c * g.stride] = v;
      nv[c] = v;
      rr += (double)v.x * v.x + (double)v.y * v.y;
    }
    if (Rb) net_in_store(Rb, (size_t)rhs * g.stride + hs, nv);
  }
  const double s = block_sum(rr);
  if (threadIdx.x == 0) part[(size_t)rhs * gridDim.x + blockIdx.x] = s;
}

// the recursive |r|^2 and the end of this rhs's cycle
__global__ void k_g_rr(const double* __restrict__ part, int nb, GScal* __restrict__ gs) {
  GScal& S = gs[blockIdx.x];
  if (S.idle) return;
  const double rr = gsum(part, 1, 0, nb);
  if (threadIdx.x == 0) {
    S.rr = rr;
    S.count += 1;
    S.taken += 1;
    if (!gfin(rr)) {
      S.nan = 1;
      S.idle = 1;
      S.done = 1;
    } else if (rr < 0.81 * S.tol2) {
      S.idle = 1;
    }
  }
}

// start of a cycle: clear the cycle's coefficients
__global__ void k_g_cycle(GScal* __restrict__ gs) {
  GScal& S = gs[blockIdx.x];
  for (int i = threadIdx.x; i < kMaxK * 2; i += blockDim.x) {
    (&S.alpha[0][0])[i] = 0.0;
    (&S.gamma[0][0])[i] = 0.0;
  }
  for (int i = threadIdx.x; i < kMaxK * kMaxK * 2; i += blockDim.x) (&S.beta[0][0][0])[i] = 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    S.idle = S.done;
    S.cyc_start = S.rho2;
    S.taken = 0;
  }
}

// gamma_i = sum_k alpha_k T_ki with z'_k = sum_i T_ki z_i,
// T_ki = delta_ki - sum_{j: i <= j < k} beta_kj T_ji (one thread per rhs)
__global__ void k_g_coef(int m, GScal* __restrict__ gs) {
  GScal& S = gs[blockIdx.x];
  if (threadIdx.x != 0 || S.done) return;
  double T[kMaxK][kMaxK][2];
  for (int k = 0; k < m; ++k)
    for (int i = 0; i < m; ++i) {
      double tr = (i == k) ? 1.0 : 0.0, ti = 0.0;
      if (i < k)
        for (int j = i; j < k; ++j) {
          const double br = S.beta[k][j][0], bi = S.beta[k][j][1];
          tr -= br * T[j][i][0] - bi * T[j][i][1];
          ti -= br * T[j][i][1] + bi * T[j][i][0];
        }
      T[k][i][0] = tr;
      T[k][i][1] = ti;
    }
  for (int i = 0; i < m; ++i) {
    double gr = 0.0, gi = 0.0;
    for (int k = i; k < m; ++k) {
      const double ar = S.alpha[k][0], ai = S.alpha[k][1];
      gr += ar * T[k][i][0] - ai * T[k][i][1];
      gi += ar * T[k][i][1] + ai * T[k][i][0];
    }
    S.gamma[i][0] = gr;
    S.gamma[i][1] = gi;
  }
}

// y += sum_i gamma_i z_i (FP64 solution, FP32 directions) over the
// directions this rhs took in the cycle (the others' z slots were not
// written for it)
__global__ void __launch_bounds__(kGThreads)
k_g_fold(Geom g, const float2* __restrict__ Z, size_t slot, int mm, double2* __restrict__ y,
         const GScal* __restrict__ gs) {
  const int rhs = blockIdx.y;
  const GScal& S = gs[rhs];
  if (S.done) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const int m = min(mm, S.taken);
  const size_t he = (size_t)12 * g.stride;
  double gr[kMaxK], gi[kMaxK];
#pragma unroll
  for (int i = 0; i < kMaxK; ++i) {
    gr[i] = i < m ? S.gamma[i][0] : 0.0;
    gi[i] = i < m ? S.gamma[i][1] : 0.0;
  }
  for (int c = 0; c < 12; ++c) {
    double2 v = y[rhs * he + (size_t)c * g.stride + hs];
#pragma unroll
    for (int i = 0; i < kMaxK; ++i)
      if (i < m) {
        const float2 a = Z[i * slot + rhs * he + (size_t)c * g.stride + hs];
        v.x += gr[i] * a.x - gi[i] * a.y;
        v.y += gr[i] * a.y + gi[i] * a.x;
      }
    y[rhs * he + (size_t)c * g.stride + hs] = v;
  }
}

// after the FP64 true residual: done / stall tests, restart of the recursion
__global__ void k_g_rho(const double* __restrict__ part, int nb, GScal* __restrict__ gs) {
  GScal& S = gs[blockIdx.x];
  if (S.done) return;
  const double* p = part + (size_t)blockIdx.x * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    S.rho2 = s;
    S.rr = s;
    S.cycles += 1;
    if (!gfin(s)) {
      S.nan = 1;
      S.done = 1;
    } else if (s <= S.tol2) {
      S.done = 1;
    } else if (s < 0.25 * S.best) {
      S.best = s;
      S.stall = 0;
    } else {
      S.stall += 1;
    }
    S.idle = S.done;
  }
}

__global__ void k_g_init(const double* __restrict__ bb, const double* __restrict__ rr, double tol,
                         int nr, GScal* __restrict__ gs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nr) return;
  GScal z{};
  z.b2 = bb[k];
  z.tol2 = tol * tol * z.b2;
  z.rho2 = z.rr = z.best = z.cyc_start = rr[k];
  z.done = !(z.rho2 > z.tol2);
  z.idle = z.done;
  gs[k] = z;
}

// r (FP32) = rho (FP64) for the rhs not done
__global__ void __launch_bounds__(kGThreads)
k_g_to32(Geom g, const double2* __restrict__ a, float2* __restrict__ b, uint32_t* __restrict__ bb,
         const GScal* __restrict__ gs) {
  const int rhs = blockIdx.y;
  if (gs[rhs].done) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const size_t he = (size_t)12 * g.stride;
  float2 nv[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) {
    const double2 v = a[rhs * he + (size_t)c * g.stride + hs];
    nv[c] = make_float2((float)v.x, (float)v.y);
    b[rhs * he + (size_t)c * g.stride + hs] = nv[c];
  }
  if (bb) net_in_store(bb, (size_t)rhs * g.stride + hs, nv);
}

// z = r (no preconditioner)
__global__ void __launch_bounds__(kGThreads)
k_g_copy32(Geom g, const float2* __restrict__ r, float2* __restrict__ z,
           const GScal* __restrict__ gs) {
  const int rhs = blockIdx.y;
  if (gs[rhs].idle) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const size_t he = (size_t)12 * g.stride;
  for (int c = 0; c < 12; ++c)
    z[rhs * he + (size_t)c * g.stride + hs] = r[rhs * he + (size_t)c * g.stride + hs];
}

struct Buf {
  void* p = nullptr;
  size_t n = 0;
};
struct FGState {
  Buf Z, W, R, Rb, T, y, be, bo, bh, rho, t64, part, gs, aux;
  GScal* host = nullptr;
  int host_n = 0;
};

ii_status grow(ii_context* ctx, Buf& b, size_t bytes) {
  if (b.n >= bytes) return II_OK;
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.n = 0;
  II_HIP(ctx, hipMalloc(&b.p, bytes));
  b.n = bytes;
  return II_OK;
}

}  // namespace

void fgcr_release(void* p) {
  FGState* s = static_cast<FGState*>(p);
  if (!s) return;
  for (Buf* b : {&s->Z, &s->W, &s->R, &s->Rb, &s->T, &s->y, &s->be, &s->bo, &s->bh, &s->rho, &s->t64,
                 &s->part, &s->gs, &s->aux})
    if (b->p) (void)hipFree(b->p);
  if (s->host) (void)hipHostFree(s->host);
  delete s;
}

// Batched flexible GCR for the rhs of src (all at once); use_net selects the
// learned preconditioner.  outs: one entry per rhs.
ii_status fgcr_batched(ii_context* ctx, void** state, double kappa, const ii_prop_params* prm,
                       const ii_field* src, ii_field* sol, ii_solver_output* outs) {
  const Geom& g = ctx->g;
  const int nr = src->nrhs;
  const bool use_net = prm->use_learned != 0;
  const int m = std::max(1, std::min(prm->nkrylov > 0 ? prm->nkrylov : 10, kMaxK));
  // orthogonalisation window: the last nor directions of the cycle
  const int nor = prm->northo > 0 ? prm->northo : 1;
  const double tol = prm->tol > 0 ? prm->tol : 1e-10;
  const int maxiter = prm->maxiter > 0 ? prm->maxiter : 100000;
  if (!ctx->gauge_lp[II_PREC_SINGLE]) II_CHECK(build_lowprec_gauge(ctx, II_PREC_SINGLE));
  if (!*state) *state = new FGState();
  FGState& S = *static_cast<FGState*>(*state);
  const size_t he = half_elems(g);
  const size_t slot = (size_t)nr * he;  // float2 per slot field
  II_CHECK(grow(ctx, S.Z, m * slot * sizeof(float2)));
  II_CHECK(grow(ctx, S.W, m * slot * sizeof(float2)));
  II_CHECK(grow(ctx, S.R, slot * sizeof(float2)));
  if (use_net) II_CHECK(grow(ctx, S.Rb, (size_t)nr * g.stride * 12 * sizeof(uint32_t)));
  II_CHECK(grow(ctx, S.T, slot * sizeof(float2)));
  for (Buf* b : {&S.y, &S.be, &S.bo, &S.bh, &S.rho, &S.t64})
    II_CHECK(grow(ctx, *b, nr * he * sizeof(double2)));
  const int nbs = blocks_for(g.Vh, kGThreads);
  const int nbh = hop_dot_parts(g, nr, II_PREC_DOUBLE);
  II_CHECK(grow(ctx, S.part, sizeof(double) * (size_t)nr *
                                 std::max<size_t>((size_t)2 * kMaxK * nbs, (size_t)nbh + 64)));
  II_CHECK(grow(ctx, S.gs, sizeof(GScal) * nr));
  II_CHECK(grow(ctx, S.aux, sizeof(double) * (3 * nr + 8)));
  if (S.host_n < nr) {
    if (S.host) (void)hipHostFree(S.host);
    S.host = nullptr;
    S.host_n = 0;
    II_HIP(ctx, hipHostMalloc((void**)&S.host, sizeof(GScal) * nr));
    S.host_n = nr;
  }
  float2* Z = (float2*)S.Z.p;
  float2* W = (float2*)S.W.p;
  float2* R = (float2*)S.R.p;
  uint32_t* Rb = use_net ? (uint32_t*)S.Rb.p : nullptr;  // r as the network's bf16 input
  float2* T = (float2*)S.T.p;
  double2 *y = (double2*)S.y.p, *be = (double2*)S.be.p, *bo = (double2*)S.bo.p,
          *bh = (double2*)S.bh.p, *rho = (double2*)S.rho.p, *t64 = (double2*)S.t64.p;
  double* part = (double*)S.part.p;
  GScal* gs = (GScal*)S.gs.p;
  double* aux = (double*)S.aux.p;  // |b_e|^2, |b_o|^2, |bhat|^2, the scalar 1.0
  double* one = aux + 3 * nr;
  const dim3 sgrid(nbs, nr);

  if (use_net) II_CHECK(bnet_prepare(ctx, nr));
  // the FP32 and FP64 multi-rhs hops' launch shapes, before the cycle graph
  if (ctx->mrhs_nrhs[II_PREC_SINGLE] != nr) II_CHECK(tune_mrhs(ctx, II_PREC_SINGLE, nr));
  if (ctx->mrhs_nrhs[II_PREC_DOUBLE] != nr) II_CHECK(tune_mrhs(ctx, II_PREC_DOUBLE, nr));
  II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  // bhat = b_e + kappa H_eo b_o (eoUtils.cpp:54-57), |b|^2 of the full sources
  for (int k = 0; k < nr; ++k) {
    const double2* base = src->data + (size_t)k * 2 * he;
    II_CHECK(copy64(ctx, be + k * he, base));
    II_CHECK(copy64(ctx, bo + k * he, base + he));
  }
  II_CHECK(dot64(ctx, be, be, aux, nr));
  II_CHECK(dot64(ctx, bo, bo, aux + nr, nr));
  HopSpec h;
  h.out_par = 0;
  h.a = kappa;
  h.xpay = true;
  h.nrhs = nr;
  II_CHECK(hop64(ctx, h, bo, bh, be, nullptr));
  II_CHECK(dot64(ctx, bh, bh, aux + 2 * nr, nr));
  {
    std::vector<double> hb(3 * nr + 1);
    II_HIP(ctx, hipMemcpyAsync(hb.data(), aux, 3 * nr * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nr; ++k) hb[k] += hb[nr + k];
    hb[3 * nr] = 1.0;
    II_HIP(ctx, hipMemcpyAsync(aux, hb.data(), nr * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
    II_HIP(ctx, hipMemcpyAsync(one, &hb[3 * nr], sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
  }
  hipLaunchKernelGGL(k_g_init, dim3((nr + 63) / 64), dim3(64), 0, ctx->stream, aux, aux + 2 * nr,
                     tol, nr, gs);
  II_CHECK(zero64(ctx, y, nr));
  hipLaunchKernelGGL(k_g_to32, sgrid, dim3(kGThreads), 0, ctx->stream, g, bh, R, Rb, gs);
  II_HIP(ctx, hipGetLastError());

  const int* idle = &gs[0].idle;
  const int* donef = &gs[0].done;
  // one restart cycle: m directions, the fold into y and the FP64 true residual
  auto cycle = [&]() -> ii_status {
    hipLaunchKernelGGL(k_g_cycle, dim3(nr), dim3(64), 0, ctx->stream, gs);
    for (int k = 0; k < m; ++k) {
      float2* zk = Z + k * slot;
      float2* wk = W + k * slot;
      if (use_net) {
        II_CHECK(bnet_apply(ctx, nr, Rb, zk, idle, kGInts));
      } else {
        hipLaunchKernelGGL(k_g_copy32, sgrid, dim3(kGThreads), 0, ctx->stream, g, R, zk, gs);
      }
      // w = Dhat z = z - kappa^2 H_eo H_oe z (FP32, r12 links)
      HopSpec a;
      a.nrhs = nr;
      a.skip = idle;
      a.skip_stride = kGInts;
      a.out_par = 1;
      II_CHECK(hop32(ctx, a, zk, T, nullptr, nullptr));
      a.out_par = 0;
      a.a = -kappa * kappa;
      a.xpay = true;
      II_CHECK(hop32(ctx, a, T, wk, zk, nullptr));
      const int j0 = std::max(0, k - nor);
      if (nor == 1 && k > 0) {
        // window 1: all scalars from one pass, then one update pass
        hipLaunchKernelGGL(k_g_dots1, sgrid, dim3(kGThreads), 0, ctx->stream, g, W, slot, k, R,
                           gs, part);
        hipLaunchKernelGGL(k_g_ab1, dim3(nr), dim3(256), 0, ctx->stream, part, nbs, k, gs);
        hipLaunchKernelGGL(k_g_upd1, sgrid, dim3(kGThreads), 0, ctx->stream, g, W, slot, k, R, Rb,
                           gs, part);
      } else {
        if (k > 0) {
          hipLaunchKernelGGL(k_g_dots, sgrid, dim3(kGThreads), 0, ctx->stream, g, W, slot, j0, k,
                             gs, part);
          hipLaunchKernelGGL(k_g_beta, dim3(nr), dim3(256), 0, ctx->stream, part, nbs, j0, k, gs);
        }
        hipLaunchKernelGGL(k_g_orth, sgrid, dim3(kGThreads), 0, ctx->stream, g, W, slot, j0, k, R,
                           gs, part);
        hipLaunchKernelGGL(k_g_alpha, dim3(nr), dim3(256), 0, ctx->stream, part, nbs, k, gs);
        hipLaunchKernelGGL(k_g_rupd, sgrid, dim3(kGThreads), 0, ctx->stream, g, W, slot, k, R, Rb,
                           gs, part);
      }
      hipLaunchKernelGGL(k_g_rr, dim3(nr), dim3(256), 0, ctx->stream, part, nbs, gs);
      II_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_g_coef, dim3(nr), dim3(64), 0, ctx->stream, m, gs);
    hipLaunchKernelGGL(k_g_fold, sgrid, dim3(kGThreads), 0, ctx->stream, g, Z, slot, m, y, gs);
    // rho = bhat - Dhat y (FP64) with |rho|^2 in the second hop's epilogue
    HopSpec b;
    b.nrhs = nr;
    b.skip = donef;
    b.skip_stride = kGInts;
    b.out_par = 1;
    II_CHECK(hop64(ctx, b, y, t64, nullptr, nullptr));
    b.out_par = 0;
    b.a = -kappa * kappa;
    b.xpay = true;
    b.epi = kEpiRUpd;
    b.coef = one;
    b.coef_stride = 0;
    b.partials = part;
    II_CHECK(hop64(ctx, b, t64, rho, y, bh));
    hipLaunchKernelGGL(k_g_rho, dim3(nr), dim3(1024), 0, ctx->stream, part, nbh, gs);
    hipLaunchKernelGGL(k_g_to32, sgrid, dim3(kGThreads), 0, ctx->stream, g, rho, R, Rb, gs);
    II_HIP(ctx, hipGetLastError());
    return II_OK;
  };
  hipGraphExec_t gexec = nullptr;
  struct GraphGuard {
    hipGraphExec_t* e;
    ~GraphGuard() {
      if (*e) (void)hipGraphExecDestroy(*e);
    }
  } gguard{&gexec};
  int dirs = 0;
  bool all_done = false, stalled = false;
  std::vector<int> stalled_rhs;
  while (!all_done && dirs < maxiter) {
    if (prm->use_graph) {
      if (!gexec) {
        hipGraph_t graph = nullptr;
        II_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ii_status cst = cycle();
        hipError_t ee = hipStreamEndCapture(ctx->stream, &graph);
        II_CHECK(cst);
        if (ee != hipSuccess) return set_error(ctx, II_ERR_DEVICE, "graph capture failed");
        hipError_t ie = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return set_error(ctx, II_ERR_DEVICE, "graph instantiation failed");
      }
      II_HIP(ctx, hipGraphLaunch(gexec, ctx->stream));
    } else {
      II_CHECK(cycle());
    }
    dirs += m;
    II_HIP(ctx, hipMemcpyAsync(S.host, gs, sizeof(GScal) * nr, hipMemcpyDeviceToHost, ctx->stream));
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    all_done = true;
    stalled_rhs.clear();
    for (int k = 0; k < nr; ++k) {
      const GScal& H = S.host[k];
      if (H.nan)
        return set_error(ctx, II_ERR_NAN, "flexible GCR: non-finite scalar (rhs " +
                                              std::to_string(k) + ", direction " +
                                              std::to_string(H.count) + ")");
      if (!H.done) {
        all_done = false;
        if (H.stall >= 3) stalled_rhs.push_back(k);
      }
    }
    if (!stalled_rhs.empty()) {
      stalled = true;
      break;
    }
  }
  int fb_its = 0;
  if (!all_done) {
    // plain-CG fallback on the remaining defect of the unfinished rhs:
    // Dhat e = rho with the batched mixed CG (a full field whose even part
    // is rho and odd part zero has bhat = rho), y += e
    ii_field tmp_src, tmp_sol;
    tmp_src.ctx = tmp_sol.ctx = ctx;
    tmp_src.subset = tmp_sol.subset = II_FULL;
    tmp_src.nrhs = tmp_sol.nrhs = nr;
    void *ps = nullptr, *px = nullptr;
    II_CHECK(workspace(ctx, 84, (size_t)nr * 2 * he * sizeof(double2), &ps));
    II_CHECK(workspace(ctx, 85, (size_t)nr * 2 * he * sizeof(double2), &px));
    tmp_src.data = (double2*)ps;
    tmp_sol.data = (double2*)px;
    II_CHECK(zero64(ctx, tmp_src.data, 2 * nr));
    for (int k = 0; k < nr; ++k)
      if (!S.host[k].done) II_CHECK(copy64(ctx, tmp_src.data + (size_t)k * 2 * he, rho + k * he));
    ii_prop_params fp = *prm;
    fp.use_learned = 0;
    fp.method = II_METHOD_CG;
    fp.precision = II_PREC_SINGLE;
    fp.batched = 1;
    // the defect's own tolerance: the target relative to |rho|
    double worst = 0.0;
    for (int k = 0; k < nr; ++k)
      if (!S.host[k].done) worst = std::max(worst, std::sqrt(S.host[k].rho2 / S.host[k].b2));
    fp.tol = 0.5 * tol / std::max(worst, 1e-300);
    std::vector<ii_solver_output> fo(nr);
    II_CHECK(propagator(ctx, kappa, &fp, &tmp_src, &tmp_sol, fo.data()));
    for (int k = 0; k < nr; ++k)
      if (!S.host[k].done) {
        II_CHECK(axpby64(ctx, 1.0, tmp_sol.data + (size_t)k * 2 * he, 1.0, y + k * he));
        fb_its = std::max(fb_its, fo[k].iterations);
      }
    (void)stalled;
  }
  // x_o = b_o + kappa H_oe y (eoUtils.cpp:67)
  h = HopSpec();
  h.out_par = 1;
  h.a = kappa;
  h.xpay = true;
  h.nrhs = nr;
  II_CHECK(hop64(ctx, h, y, t64, bo, nullptr));
  for (int k = 0; k < nr; ++k) {
    double2* base = sol->data + (size_t)k * 2 * he;
    II_CHECK(copy64(ctx, base, y + k * he));
    II_CHECK(copy64(ctx, base + he, t64 + k * he));
  }
  II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  II_HIP(ctx, hipMemcpyAsync(S.host, gs, sizeof(GScal) * nr, hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  for (int k = 0; k < nr; ++k) {
    ii_solver_output& o = outs[k];
    const GScal& H = S.host[k];
    o.iterations = H.count;
    o.inner_iterations = H.done ? 0 : fb_its;
    o.restarts = H.cycles;
    o.recursion_restarts = 0;
    o.diff = std::sqrt(H.rho2 / H.b2);
    o.timing_ms = ms;
    o.true_residual = -1.0;
  }
  return II_OK;
}

}  // namespace ii
