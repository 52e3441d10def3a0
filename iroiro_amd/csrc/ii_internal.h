// Internal declarations of the gfx950 Wilson even/odd solve library.
//
// Device layouts (all "site-major SoA": the half-lattice site index is the
// fastest index, so a wavefront of 64 consecutive sites reads 64 consecutive
// complex numbers of one component per load):
//   half spinor  [12 = 3*spin + colour][stride] complex, stride >= V/2
//   multi-rhs    [rhs][12][stride]
//   full field   [rhs][parity][12][stride]
//   FP64 gauge   [parity][mu][9 = 3*row + col][stride] complex, where
//                parity is the parity of the site the link starts from.
// The half-site index of a site (x,y,z,t) in its parity is
//   hs = x/2 + (Lx/2)*(y + Ly*(z + Lz*t)),
// i.e. ascending lexicographic order inside each parity, the order of the
// reference's esec/osec lists (lib/Geometry/siteIndex_EvenOdd.cpp:21-62).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include "../../include/iroiro_hip.h"

namespace ii {

constexpr int kBlock = 256;        // threads per block of the site kernels
constexpr int kRedBlocksMax = 2048;  // partial sums per reduction

struct Geom {
  int L[4];
  int Lxh;
  int Vh;       // sites per parity
  int V;
  int sY, sZ, sT;  // half-index strides of y, z, t
  int stride;      // per-component stride of a half field (multiple of 64)
};

// Site traversal of the single-rhs hop, chosen per lattice by the
// autotuner (ii_dslash.hip).  mode 0: the XCD that runs block b (b % 8)
// takes a contiguous range of half sites.  mode 1: XCD x takes the time
// slices [x*Lt/8, (x+1)*Lt/8) and walks them in strips of W slices, z-planes
// outer and t inner, so that the t- and z-neighbour planes of a block were
// read a few planes earlier into the same L2.  bpp: blocks per (z,t) plane.
// bs: block size of the contiguous order.
struct HopOrder {
  int mode = 0;
  int W = 4;
  int bpp = 1;
  int bs = 256;  // threads per block of the non-DOT FP64 hops (128 or 256)
};

// Solver scalars kept on the device so that iterations need no host sync.
struct CGScalars {
  double rr;       // |r|^2 of the current iterate
  double rr_new;
  double pap;      // <p, A p>
  double alpha;
  double beta;
  double snorm;    // 1/|b|^2
  double prec;     // stopping threshold on rr*snorm
  double diff;
  int done;        // set once the stopping test passed
  int iter;        // loop index at which it passed (reference counting)
  int count;       // iterations executed so far
  int nan;         // a non-finite scalar appeared
};

struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace ii

struct ii_context {
  int device = 0;
  hipStream_t stream = nullptr;
  ii::Geom g{};
  std::string err;
  // gauge fields
  double2* gauge64 = nullptr;     // FP64 r18, [2][4][9][stride]
  bool gauge_valid = false;
  // derived copies of the gauge field, rebuilt lazily after it changes:
  // [II_PREC_SINGLE..II_PREC_BF16] the reconstruct-12 links of the hops,
  // [kGaugePairs] the bfloat16 transport matrices of the batched network
  static constexpr int kGaugeCopies = 5;
  static constexpr int kGaugePairs = 4;
  void* gauge_lp[kGaugeCopies] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int tbc_sign = 1;               // sign of the t-links on the last time slice (r12)
  bool r12_ok = false;
  // workspaces, by slot index (see ii::workspace)
  static constexpr int kSlots = 96;
  ii::Buffer work[kSlots];
  double* partials = nullptr;     // reduction partial sums
  size_t partials_len = 0;
  ii::CGScalars* scal = nullptr;  // device solver scalars (array)
  ii::CGScalars* scal_host = nullptr;  // pinned mirror
  int nscal = 0;
  void* host_pin = nullptr;       // pinned host block of the multishift scalars
  size_t host_pin_len = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // autotuned launch shapes
  ii::HopOrder hop_order;
  ii::HopOrder mrhs_order;  // tile order of the multi-rhs hop (64-site tiles)
  // multi-rhs hop launch shape per storage precision (index II_PREC_*): rhs
  // per block and tile order, chosen by tune_mrhs for mrhs_nrhs rhs (0: not
  // yet; the default is 12 rhs per block in the contiguous order)
  int mrhs_rb[4] = {0, 0, 0, 0};
  int mrhs_nrhs[4] = {0, 0, 0, 0};
  ii::HopOrder mrhs_ord[4];
  std::string mrhs_report[4];
  bool hop_tuned = false;
  std::string tune_report;
  // learned preconditioner state (owned by ii_learned.hip) and the batched
  // network and flexible solver workspaces (owned by ii_bnet.hip)
  void* learned = nullptr;
  void* bnet = nullptr;
  void* fgcr = nullptr;
  // live device fields of this context: destroying the context releases
  // their memory and detaches them, so a later ii_field_destroy is safe
  std::set<ii_field*> fields;
};

struct ii_field {
  ii_context* ctx = nullptr;
  int subset = II_EVEN;
  int nrhs = 1;
  double2* data = nullptr;  // FP64 device layout
};

namespace ii {

// ---------------------------------------------------------------- errors
ii_status set_error(ii_context* ctx, ii_status st, const std::string& msg);
#define II_HIP(ctx, call)                                                  \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return ::ii::set_error((ctx), II_ERR_DEVICE,                         \
                             std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define II_CHECK(expr)                 \
  do {                                 \
    ii_status s_ = (expr);             \
    if (s_ != II_OK) return s_;        \
  } while (0)

// Grows the workspace slot `slot` to at least `bytes` (never shrinks, so
// steady-state calls allocate nothing).
ii_status workspace(ii_context* ctx, int slot, size_t bytes, void** ptr);
ii_status ensure_partials(ii_context* ctx, size_t n);
ii_status ensure_scalars(ii_context* ctx, int n);
// Pinned host block of at least `bytes` owned by the context (grows, never
// shrinks), so solves allocate no pinned memory per call.
ii_status ensure_host_pin(ii_context* ctx, size_t bytes, void** ptr);
ii_status require_gauge(ii_context* ctx);
// Entry of every API call on a context: selects its device and clears a
// stale HIP error left by an earlier, already reported failure.
ii_status enter(ii_context* ctx);

inline int blocks_for(long n, int block = kBlock) {
  return (int)((n + block - 1) / block);
}
inline size_t half_elems(const Geom& g) { return (size_t)12 * g.stride; }

// ---------------------------------------------------------------- complex
template <typename R> struct Cx;
template <> struct Cx<double> { using T = double2; };
template <> struct Cx<float> { using T = float2; };

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) r, decltype(C::x) i) {
  C c; c.x = r; c.y = i; return c;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
// a + i b, a - i b
template <typename C> __device__ __forceinline__ C cadd_i(C a, C b) { return cmk<C>(a.x - b.y, a.y + b.x); }
template <typename C> __device__ __forceinline__ C csub_i(C a, C b) { return cmk<C>(a.x + b.y, a.y - b.x); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }
// acc += u * w
template <typename C> __device__ __forceinline__ void cmac(C& acc, C u, C w) {
  acc.x = fma(u.x, w.x, acc.x); acc.x = fma(-u.y, w.y, acc.x);
  acc.y = fma(u.x, w.y, acc.y); acc.y = fma(u.y, w.x, acc.y);
}
// acc += conj(u) * w
template <typename C> __device__ __forceinline__ void cmacc(C& acc, C u, C w) {
  acc.x = fma(u.x, w.x, acc.x); acc.x = fma(u.y, w.y, acc.x);
  acc.y = fma(u.x, w.y, acc.y); acc.y = fma(-u.y, w.x, acc.y);
}

// ---------------------------------------------------------------- geometry
struct SiteCoord {
  int x, y, z, t, xh, row;
};

__host__ __device__ __forceinline__ SiteCoord decode_site(const Geom& g, int hs, int par) {
  SiteCoord c;
  c.row = hs / g.Lxh;
  c.xh = hs - c.row * g.Lxh;
  int zt = c.row / g.L[1];
  c.y = c.row - zt * g.L[1];
  c.t = zt / g.L[2];
  c.z = zt - c.t * g.L[2];
  c.x = 2 * c.xh + ((c.y + c.z + c.t + par) & 1);
  return c;
}

// Half index (in the opposite parity) of the neighbour of site c in
// direction mu, forward (fw) or backward; periodic wrap as on one process
// (lib/Communicator/communicator_single.cpp:45-66).
__host__ __device__ __forceinline__ int neighbour(const Geom& g, const SiteCoord& c, int hs, int mu, bool fw) {
  switch (mu) {
    case 0: {
      int xn = fw ? (c.x + 1 == g.L[0] ? 0 : c.x + 1) : (c.x == 0 ? g.L[0] - 1 : c.x - 1);
      return c.row * g.Lxh + (xn >> 1);
    }
    case 1:
      return fw ? (c.y == g.L[1] - 1 ? hs - (g.L[1] - 1) * g.sY : hs + g.sY)
                : (c.y == 0 ? hs + (g.L[1] - 1) * g.sY : hs - g.sY);
    case 2:
      return fw ? (c.z == g.L[2] - 1 ? hs - (g.L[2] - 1) * g.sZ : hs + g.sZ)
                : (c.z == 0 ? hs + (g.L[2] - 1) * g.sZ : hs - g.sZ);
    default:
      return fw ? (c.t == g.L[3] - 1 ? hs - (g.L[3] - 1) * g.sT : hs + g.sT)
                : (c.t == 0 ? hs + (g.L[3] - 1) * g.sT : hs - g.sT);
  }
}

// ---------------------------------------------------------------- reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sum; result valid in thread 0.  blockDim.x multiple of 64,
// at most 1024.
template <typename T>
__device__ __forceinline__ T block_sum(T v) {
  __shared__ T red_smem[16];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red_smem[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
    int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) s += red_smem[i];
  }
  __syncthreads();
  return s;
}

}  // namespace ii
