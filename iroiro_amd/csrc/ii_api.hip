// C ABI entry points: context, gauge field, fields and operator
// applications on host arrays.  Solvers live in ii_solvers.hip.
#include "ii_ops.h"

#include <cstring>
#include <memory>

namespace ii {

ii_status set_error(ii_context* ctx, ii_status st, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return st;
}

ii_status workspace(ii_context* ctx, int slot, size_t bytes, void** ptr) {
  if (slot < 0 || slot >= ii_context::kSlots)
    return set_error(ctx, II_ERR_BAD_ARG, "internal: workspace slot out of range");
  Buffer& b = ctx->work[slot];
  if (b.bytes < bytes) {
    if (b.ptr) {
      II_HIP(ctx, hipStreamSynchronize(ctx->stream));
      II_HIP(ctx, hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    II_HIP(ctx, hipMalloc(&b.ptr, bytes));
    // zero once so that padding sites of half fields read as 0
    II_HIP(ctx, hipMemsetAsync(b.ptr, 0, bytes, ctx->stream));
    b.bytes = bytes;
  }
  *ptr = b.ptr;
  return II_OK;
}

ii_status ensure_partials(ii_context* ctx, size_t n) {
  if (ctx->partials_len < n) {
    if (ctx->partials) {
      II_HIP(ctx, hipStreamSynchronize(ctx->stream));
      II_HIP(ctx, hipFree(ctx->partials));
    }
    size_t len = n < 65536 ? 65536 : n;
    II_HIP(ctx, hipMalloc(&ctx->partials, len * sizeof(double)));
    ctx->partials_len = len;
  }
  return II_OK;
}

ii_status ensure_host_pin(ii_context* ctx, size_t bytes, void** ptr) {
  if (ctx->host_pin_len < bytes) {
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->host_pin) II_HIP(ctx, hipHostFree(ctx->host_pin));
    ctx->host_pin = nullptr;
    ctx->host_pin_len = 0;
    II_HIP(ctx, hipHostMalloc(&ctx->host_pin, bytes));
    ctx->host_pin_len = bytes;
  }
  *ptr = ctx->host_pin;
  return II_OK;
}

ii_status ensure_scalars(ii_context* ctx, int n) {
  if (ctx->nscal < n) {
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->scal) II_HIP(ctx, hipFree(ctx->scal));
    if (ctx->scal_host) II_HIP(ctx, hipHostFree(ctx->scal_host));
    int len = n < 64 ? 64 : n;
    II_HIP(ctx, hipMalloc(&ctx->scal, len * sizeof(CGScalars)));
    II_HIP(ctx, hipHostMalloc(&ctx->scal_host, len * sizeof(CGScalars)));
    ctx->nscal = len;
  }
  return II_OK;
}

ii_status enter(ii_context* ctx) {
  if (!ctx) return II_ERR_BAD_ARG;
  II_HIP(ctx, hipSetDevice(ctx->device));
  (void)hipGetLastError();
  return II_OK;
}

ii_status require_gauge(ii_context* ctx) {
  if (!ctx) return II_ERR_BAD_ARG;
  if (!ctx->gauge_valid)
    return set_error(ctx, II_ERR_STATE,
                     "no valid gauge field on the device (ii_gauge_upload after "
                     "update_internal_state)");
  return II_OK;
}

// Host layout staging: copies n doubles from a host array into the staging
// workspace (slot 0).
ii_status stage_in(ii_context* ctx, const double* host, size_t n, double** dev) {
  void* p = nullptr;
  II_CHECK(workspace(ctx, 0, n * sizeof(double), &p));
  II_HIP(ctx, hipMemcpyAsync(p, host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  *dev = (double*)p;
  return II_OK;
}
ii_status stage_out(ii_context* ctx, double* host, const double* dev, size_t n) {
  II_HIP(ctx, hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}
// A second staging area for outputs (slot 1).
ii_status stage_buf(ii_context* ctx, size_t n, double** dev) {
  void* p = nullptr;
  II_CHECK(workspace(ctx, 1, n * sizeof(double), &p));
  *dev = (double*)p;
  return II_OK;
}

// FP64 half-field workspaces: slots 2.. (nrhs = 1).
ii_status half64(ii_context* ctx, int k, double2** p) {
  void* v = nullptr;
  II_CHECK(workspace(ctx, 2 + k, half_elems(ctx->g) * sizeof(double2), &v));
  *p = (double2*)v;
  return II_OK;
}

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_context_create(ii_context** out, const int dims[4], int device) {
  if (!out || !dims) return II_ERR_BAD_ARG;
  *out = nullptr;
  for (int m = 0; m < 4; ++m)
    if (dims[m] < 2) return II_ERR_BAD_ARG;
  if (dims[0] % 2) return II_ERR_BAD_ARG;
  long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  if (V / 2 > (1L << 30)) return II_ERR_BAD_ARG;
  std::unique_ptr<ii_context> ctx(new ii_context());
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) return II_ERR_DEVICE;
  Geom& g = ctx->g;
  for (int m = 0; m < 4; ++m) g.L[m] = dims[m];
  g.V = (int)V;
  g.Vh = (int)(V / 2);
  g.Lxh = dims[0] / 2;
  g.sY = g.Lxh;
  g.sZ = g.Lxh * dims[1];
  g.sT = g.Lxh * dims[1] * dims[2];
  g.stride = (g.Vh + 63) / 64 * 64;
  // kernels address each array with one buffer descriptor and 32-bit
  // offsets: an FP64 half spinor of one rhs (12 components) is the largest
  // such span of the solvers (gauge fields take one descriptor per parity
  // and direction, 9 components), so V/2 may reach about 22M sites (64^3 x
  // 128 has 16.8M).  The network paths check their own, smaller limits.
  if ((unsigned long long)12 * g.stride * sizeof(double2) >= (1ull << 32)) return II_ERR_BAD_ARG;
  // default-priority stream (no priority streams on this platform)
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return II_ERR_DEVICE;
  if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess)
    return II_ERR_DEVICE;
  size_t gbytes = (size_t)2 * 4 * 9 * g.stride * sizeof(double2);
  if (hipMalloc(&ctx->gauge64, gbytes) != hipSuccess) return II_ERR_DEVICE;
  if (hipMemset(ctx->gauge64, 0, gbytes) != hipSuccess) return II_ERR_DEVICE;
  if (noalloc_setup(ctx.get()) != II_OK) return II_ERR_DEVICE;
  *out = ctx.release();
  return II_OK;
}

ii_status ii_context_destroy(ii_context* ctx) {
  if (!ctx) return II_ERR_BAD_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  learned_release(ctx);
  bnet_release(ctx);
  fgcr_release(ctx->fgcr);
  ctx->fgcr = nullptr;
  for (ii_field* f : ctx->fields) {
    (void)hipFree(f->data);
    f->data = nullptr;
    f->ctx = nullptr;  // detached: ii_field_destroy only frees the handle
  }
  ctx->fields.clear();
  for (auto& b : ctx->work)
    if (b.ptr) (void)hipFree(b.ptr);
  for (int i = 0; i < ii_context::kGaugeCopies; ++i)
    if (ctx->gauge_lp[i]) (void)hipFree(ctx->gauge_lp[i]);
  if (ctx->gauge64) (void)hipFree(ctx->gauge64);
  if (ctx->partials) (void)hipFree(ctx->partials);
  if (ctx->scal) (void)hipFree(ctx->scal);
  if (ctx->scal_host) (void)hipHostFree(ctx->scal_host);
  if (ctx->host_pin) (void)hipHostFree(ctx->host_pin);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return II_OK;
}

const char* ii_last_error(const ii_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

ii_status ii_device_info(ii_context* ctx, char* name, size_t name_len, double* hbm_gib) {
  if (!ctx) return II_ERR_BAD_ARG;
  hipDeviceProp_t prop;
  II_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
  if (name && name_len) {
    std::string s = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    std::strncpy(name, s.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (hbm_gib) *hbm_gib = prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
  return II_OK;
}

ii_status ii_gauge_upload(ii_context* ctx, const double* u) {
  if (!ctx || !u) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  double* d = nullptr;
  II_CHECK(stage_in(ctx, u, (size_t)18 * 4 * ctx->g.V, &d));
  II_CHECK(pack_gauge(ctx, d));
  for (int i = 0; i < ii_context::kGaugeCopies; ++i)
    if (ctx->gauge_lp[i]) {
      II_HIP(ctx, hipStreamSynchronize(ctx->stream));
      (void)hipFree(ctx->gauge_lp[i]);
      ctx->gauge_lp[i] = nullptr;
    }
  ctx->r12_ok = false;
  ctx->tbc_sign = 1;
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->gauge_valid = true;
  return II_OK;
}

ii_status ii_gauge_invalidate(ii_context* ctx) {
  if (!ctx) return II_ERR_BAD_ARG;
  ctx->gauge_valid = false;
  return II_OK;
}

ii_status ii_gauge_plaquette(ii_context* ctx, double* plaq) {
  if (!plaq) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  return plaquette64(ctx, plaq);
}

ii_status ii_gauge_download(ii_context* ctx, double* u) {
  if (!u) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  double* d = nullptr;
  size_t n = (size_t)18 * 4 * ctx->g.V;
  II_CHECK(stage_buf(ctx, n, &d));
  II_CHECK(unpack_gauge(ctx, d));
  return stage_out(ctx, u, d, n);
}

// ------------------------------------------------------------ operators
static ii_status single_hop(ii_context* ctx, double kappa, double* out, const double* in,
                            int out_par, bool dag) {
  if (!out || !in) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  double *din = nullptr, *dout = nullptr;
  double2 *a = nullptr, *b = nullptr;
  II_CHECK(stage_in(ctx, in, (size_t)24 * g.Vh, &din));
  II_CHECK(stage_buf(ctx, (size_t)24 * g.Vh, &dout));
  II_CHECK(half64(ctx, 0, &a));
  II_CHECK(half64(ctx, 1, &b));
  II_CHECK(pack_half_spinor(ctx, din, a));
  HopSpec h;
  h.out_par = out_par;
  h.a = -kappa;
  h.g5in = dag;
  h.g5out = dag;
  II_CHECK(hop64(ctx, h, a, b, nullptr, nullptr));
  II_CHECK(unpack_half_spinor(ctx, b, dout));
  return stage_out(ctx, out, dout, (size_t)24 * g.Vh);
}

ii_status ii_dirac_mult_eo(ii_context* ctx, double kappa, double* out, const double* in) {
  return single_hop(ctx, kappa, out, in, 0, false);
}
ii_status ii_dirac_mult_oe(ii_context* ctx, double kappa, double* out, const double* in) {
  return single_hop(ctx, kappa, out, in, 1, false);
}
ii_status ii_dirac_mult_eo_dag(ii_context* ctx, double kappa, double* out, const double* in) {
  return single_hop(ctx, kappa, out, in, 1, true);
}
ii_status ii_dirac_mult_oe_dag(ii_context* ctx, double kappa, double* out, const double* in) {
  return single_hop(ctx, kappa, out, in, 0, true);
}

static ii_status even_op(ii_context* ctx, double kappa, double* out, const double* in, int which) {
  if (!out || !in) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  double *din = nullptr, *dout = nullptr;
  double2 *a, *b, *t, *q;
  II_CHECK(stage_in(ctx, in, (size_t)24 * g.Vh, &din));
  II_CHECK(stage_buf(ctx, (size_t)24 * g.Vh, &dout));
  II_CHECK(half64(ctx, 0, &a));
  II_CHECK(half64(ctx, 1, &b));
  II_CHECK(half64(ctx, 2, &t));
  II_CHECK(half64(ctx, 3, &q));
  II_CHECK(pack_half_spinor(ctx, din, a));
  if (which == 0) II_CHECK(dhat64(ctx, kappa, a, b, t));
  else if (which == 1) II_CHECK(dhat_dag64(ctx, kappa, a, b, t));
  else II_CHECK(ddagd64(ctx, kappa, a, b, q, t));
  II_CHECK(unpack_half_spinor(ctx, b, dout));
  return stage_out(ctx, out, dout, (size_t)24 * g.Vh);
}

ii_status ii_dirac_mult(ii_context* ctx, double kappa, double* out, const double* in) {
  return even_op(ctx, kappa, out, in, 0);
}
ii_status ii_dirac_mult_dag(ii_context* ctx, double kappa, double* out, const double* in) {
  return even_op(ctx, kappa, out, in, 1);
}
ii_status ii_fopr_ddagd_mult(ii_context* ctx, double kappa, double* out, const double* in) {
  return even_op(ctx, kappa, out, in, 2);
}

ii_status ii_dirac_gamma5(ii_context* ctx, double* out, const double* in, long nsites) {
  if (!ctx || !out || !in || nsites < 0) return II_ERR_BAD_ARG;
  // a pure spin permutation (wilsonLikeUtils.cpp:42-49); no device work
  for (long s = 0; s < nsites; ++s) {
    double tmp[24];
    std::memcpy(tmp, in + 24 * s + 12, 12 * sizeof(double));
    std::memcpy(tmp + 12, in + 24 * s, 12 * sizeof(double));
    std::memcpy(out + 24 * s, tmp, 24 * sizeof(double));
  }
  return II_OK;
}

ii_status ii_dirac_full_mult(ii_context* ctx, double kappa, double* out, const double* in,
                             int dag) {
  if (!out || !in) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  double *din = nullptr, *dout = nullptr;
  double2 *e, *o, *we, *wo;
  II_CHECK(stage_in(ctx, in, (size_t)24 * g.V, &din));
  II_CHECK(stage_buf(ctx, (size_t)24 * g.V, &dout));
  II_CHECK(half64(ctx, 0, &e));
  II_CHECK(half64(ctx, 1, &o));
  II_CHECK(half64(ctx, 2, &we));
  II_CHECK(half64(ctx, 3, &wo));
  II_CHECK(pack_full_spinor(ctx, din, e, o));
  // D f = f - kappa H f per parity (mult_full, dirac_wilson.cpp:115-118);
  // D^dag = g5 D g5 folds g5 into the hop and leaves the identity term.
  HopSpec h;
  h.a = -kappa;
  h.xpay = true;
  h.g5in = h.g5out = dag != 0;
  h.out_par = 0;
  II_CHECK(hop64(ctx, h, o, we, e, nullptr));
  h.out_par = 1;
  II_CHECK(hop64(ctx, h, e, wo, o, nullptr));
  II_CHECK(unpack_full_spinor(ctx, we, wo, dout));
  return stage_out(ctx, out, dout, (size_t)24 * g.V);
}

// ------------------------------------------------------------ fields
ii_status ii_field_create(ii_context* ctx, int subset, int nrhs, ii_field** f) {
  if (!ctx || !f || nrhs < 1 || subset < II_EVEN || subset > II_FULL) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  std::unique_ptr<ii_field> fl(new ii_field());
  fl->ctx = ctx;
  fl->subset = subset;
  fl->nrhs = nrhs;
  size_t n = half_elems(ctx->g) * nrhs * (subset == II_FULL ? 2 : 1);
  II_HIP(ctx, hipMalloc(&fl->data, n * sizeof(double2)));
  II_HIP(ctx, hipMemsetAsync(fl->data, 0, n * sizeof(double2), ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->fields.insert(fl.get());
  *f = fl.release();
  return II_OK;
}

ii_status ii_field_destroy(ii_field* f) {
  if (!f) return II_ERR_BAD_ARG;
  if (f->ctx) {
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    (void)hipFree(f->data);
    f->ctx->fields.erase(f);
  }
  delete f;
  return II_OK;
}

ii_status ii_field_upload(ii_field* f, int rhs, const double* host) {
  if (!f || !f->ctx || !host || rhs < 0 || rhs >= f->nrhs) return II_ERR_BAD_ARG;
  ii_context* ctx = f->ctx;
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  size_t he = half_elems(g);
  double* d = nullptr;
  if (f->subset == II_FULL) {
    II_CHECK(stage_in(ctx, host, (size_t)24 * g.V, &d));
    double2* base = f->data + (size_t)rhs * 2 * he;
    II_CHECK(pack_full_spinor(ctx, d, base, base + he));
  } else {
    II_CHECK(stage_in(ctx, host, (size_t)24 * g.Vh, &d));
    II_CHECK(pack_half_spinor(ctx, d, f->data + (size_t)rhs * he));
  }
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

ii_status ii_field_download(const ii_field* f, int rhs, double* host) {
  if (!f || !f->ctx || !host || rhs < 0 || rhs >= f->nrhs) return II_ERR_BAD_ARG;
  ii_context* ctx = f->ctx;
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  size_t he = half_elems(g);
  double* d = nullptr;
  if (f->subset == II_FULL) {
    II_CHECK(stage_buf(ctx, (size_t)24 * g.V, &d));
    const double2* base = f->data + (size_t)rhs * 2 * he;
    II_CHECK(unpack_full_spinor(ctx, base, base + he, d));
    return stage_out(ctx, host, d, (size_t)24 * g.V);
  }
  II_CHECK(stage_buf(ctx, (size_t)24 * g.Vh, &d));
  II_CHECK(unpack_half_spinor(ctx, f->data + (size_t)rhs * he, d));
  return stage_out(ctx, host, d, (size_t)24 * g.Vh);
}

ii_status ii_field_point_sources(ii_field* f, const int site[4]) {
  if (!f || !f->ctx || !site || f->subset != II_FULL || f->nrhs != 12) return II_ERR_BAD_ARG;
  ii_context* ctx = f->ctx;
  const Geom& g = ctx->g;
  for (int m = 0; m < 4; ++m)
    if (site[m] < 0 || site[m] >= g.L[m]) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  size_t he = half_elems(g);
  II_HIP(ctx, hipMemsetAsync(f->data, 0, he * 2 * 12 * sizeof(double2), ctx->stream));
  int par = (site[0] + site[1] + site[2] + site[3]) & 1;
  int lex = site[0] + g.L[0] * (site[1] + g.L[1] * (site[2] + g.L[2] * site[3]));
  int hs = lex >> 1;
  double2 one = make_double2(1.0, 0.0);
  for (int s = 0; s < 4; ++s)
    for (int c = 0; c < 3; ++c) {
      int rhs = 3 * s + c;  // Qprop_EvenOdd column order (qprop_EvenOdd.cpp:13-21)
      double2* dst = f->data + (size_t)rhs * 2 * he + (size_t)par * he +
                     (size_t)(3 * s + c) * g.stride + hs;
      II_HIP(ctx, hipMemcpyAsync(dst, &one, sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
    }
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

}  // extern "C"
