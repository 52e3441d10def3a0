// Fopr_DdagD::mult for callers that cannot allocate device memory.
//
// ii_fopr_ddagd_mult stages its host arrays and temporaries through the
// context's workspaces, which grow with hipMalloc on first use.  Code under
// HIP graph capture, tight serving loops and harnesses that forbid
// allocation cannot take that path.  ii_fopr_ddagd_mult_noalloc computes the
// same A = Dhat^dag Dhat (lib/Fopr/fopr.h:146-153) with the same kernels, but
// every device buffer beyond the caller's context (the host-layout staging
// field, the packed input, output and the two hop temporaries) is carved out
// of one file-scope device array `scratch` of kScratchBytes.
//
// Each GPU holds its own copy of `scratch` at its own address, alive until
// the process ends.  The address is taken once per device with
// hipGetSymbolAddress when the first context on that device is created; the
// array's name is only a host-side stand-in and is never passed to a kernel
// or a copy.  Every call on one GPU shares that GPU's array and leaves the
// last call's contents in it: calls that overlap on one GPU (two contexts of
// one device used from two threads) race on it.
#include "ii_ops.h"

#include <cstdio>
#include <mutex>
#include <string>

namespace ii {

namespace {

// 32 MiB: the staging field plus four FP64 half fields of 192 B per site,
// i.e. half volumes up to 34,944 sites (e.g. 16^4, 8^3 x 64)
constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// bytes of scratch the call needs on lattice g: the host-layout staging field
// (24 doubles per site) and four packed FP64 half fields
size_t noalloc_bytes(const Geom& g) {
  return align256((size_t)24 * g.Vh * sizeof(double)) +
         4 * align256(half_elems(g) * sizeof(double2));
}

}  // namespace

// Takes the device address of `scratch` on the context's device (current
// when this runs), once per device.
ii_status noalloc_setup(ii_context* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_error(ctx, II_ERR_BAD_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[ctx->device]) return II_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_error(ctx, II_ERR_DEVICE,
                     std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[ctx->device] = p;
  return II_OK;
}

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_fopr_ddagd_mult_noalloc(ii_context* ctx, double kappa, double* out,
                                     const double* in) {
  if (!ctx || !out || !in) return II_ERR_BAD_ARG;
  const Geom& g = ctx->g;
  const size_t need = noalloc_bytes(g);
  if (need > kScratchBytes) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "ii_fopr_ddagd_mult_noalloc: this lattice needs %zu bytes of scratch, the "
                  "maximum is %zu (half volume above ~34,900 sites); use ii_fopr_ddagd_mult",
                  need, kScratchBytes);
    return set_error(ctx, II_ERR_BAD_ARG, buf);
  }
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices)
      base = (unsigned char*)g_scratch_ptr[ctx->device];
  }
  if (!base) return set_error(ctx, II_ERR_STATE, "scratch address of this device not set up");
  const size_t hb = align256((size_t)24 * g.Vh * sizeof(double));
  const size_t fb = align256(half_elems(g) * sizeof(double2));
  double* stage = (double*)base;
  double2* a = (double2*)(base + hb);
  double2* b = (double2*)(base + hb + fb);
  double2* q = (double2*)(base + hb + 2 * fb);
  double2* t = (double2*)(base + hb + 3 * fb);
  const size_t nbytes = (size_t)24 * g.Vh * sizeof(double);
  II_HIP(ctx, hipMemcpyAsync(stage, in, nbytes, hipMemcpyHostToDevice, ctx->stream));
  II_CHECK(pack_half_spinor(ctx, stage, a));
  // q = Dhat a = a - kappa^2 H_eo H_oe a; b = Dhat^dag q (dirac_wilson_EvenOdd.hpp:79-92).
  // The hop traversal is the context's current one; the autotuner (which
  // allocates) is not run from here.
  HopSpec h;
  h.no_tune = true;
  h.out_par = 1;
  II_CHECK(hop64(ctx, h, a, t, nullptr, nullptr));
  h.out_par = 0;
  h.a = -kappa * kappa;
  h.xpay = true;
  II_CHECK(hop64(ctx, h, t, q, a, nullptr));
  h = HopSpec();
  h.no_tune = true;
  h.out_par = 1;
  h.g5in = true;
  II_CHECK(hop64(ctx, h, q, t, nullptr, nullptr));
  h.out_par = 0;
  h.a = -kappa * kappa;
  h.xpay = true;
  h.g5in = false;
  h.g5out = true;
  II_CHECK(hop64(ctx, h, t, b, q, nullptr));
  II_CHECK(unpack_half_spinor(ctx, b, stage));
  II_HIP(ctx, hipMemcpyAsync(out, stage, nbytes, hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

}  // extern "C"
