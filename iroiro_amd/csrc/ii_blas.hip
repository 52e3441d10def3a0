// BLAS-1, reductions, layout packing and gauge utilities for gfx950.
//
// Reductions follow the reference's Field::operator* / norm semantics
// (Re<a,b> = sum over all real components, lib/Fields/field.inl:58-62,
// include/field.h:74-78): per-block partial sums with a wave shuffle and an
// LDS stage, then one block sums the partials in a fixed order, so results
// are deterministic run to run.
#include "ii_ops.h"
#include "ii_stencil.h"

namespace ii {

// ------------------------------------------------------------ reductions
__global__ void __launch_bounds__(kBlock)
k_dot64(const double2* __restrict__ a, const double2* __restrict__ b, size_t n,
        double* __restrict__ partials) {
  const size_t off = (size_t)blockIdx.y * n;
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double2 x = a[off + i], y = b[off + i];
    s = fma(x.x, y.x, s);
    s = fma(x.y, y.y, s);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock)
k_finalize(const double* __restrict__ partials, int nblk, double* __restrict__ dst) {
  const double* p = partials + (size_t)blockIdx.x * nblk;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += p[i];
  s = block_sum(s);
  if (threadIdx.x == 0) dst[blockIdx.x] = s;
}

int red_blocks(size_t n) {
  long b = blocks_for((long)n);
  return (int)(b < kRedBlocksMax ? b : kRedBlocksMax);
}

ii_status finalize_sums(ii_context* ctx, const double* partials, int nblk, double* dst,
                        int nrhs) {
  hipLaunchKernelGGL(k_finalize, dim3(nrhs), dim3(kBlock), 0, ctx->stream, partials, nblk, dst);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

ii_status dot64(ii_context* ctx, const double2* a, const double2* b, double* dst, int nrhs) {
  size_t n = half_elems(ctx->g);
  int nb = red_blocks(n);
  II_CHECK(ensure_partials(ctx, (size_t)nb * nrhs));
  hipLaunchKernelGGL(k_dot64, dim3(nb, nrhs), dim3(kBlock), 0, ctx->stream, a, b, n,
                     ctx->partials);
  II_HIP(ctx, hipGetLastError());
  return finalize_sums(ctx, ctx->partials, nb, dst, nrhs);
}

// ------------------------------------------------------------ elementwise
__global__ void __launch_bounds__(kBlock)
k_axpby64(double a, const double2* __restrict__ x, double b, double2* __restrict__ y, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double2 xv = x[i], yv = y[i];
    y[i] = make_double2(fma(a, xv.x, b * yv.x), fma(a, xv.y, b * yv.y));
  }
}

ii_status axpby64(ii_context* ctx, double a, const double2* x, double b, double2* y, int nrhs) {
  size_t n = half_elems(ctx->g) * nrhs;
  hipLaunchKernelGGL(k_axpby64, dim3(blocks_for((long)n)), dim3(kBlock), 0, ctx->stream, a, x, b,
                     y, n);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

ii_status copy64(ii_context* ctx, double2* dst, const double2* src, int nrhs) {
  II_HIP(ctx, hipMemcpyAsync(dst, src, half_elems(ctx->g) * nrhs * sizeof(double2),
                             hipMemcpyDeviceToDevice, ctx->stream));
  return II_OK;
}

__global__ void __launch_bounds__(kBlock) k_zero64(double2* __restrict__ y, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = make_double2(0.0, 0.0);
}

ii_status zero64(ii_context* ctx, double2* dst, int nrhs) {
  size_t n = half_elems(ctx->g) * nrhs;
  hipLaunchKernelGGL(k_zero64, dim3(blocks_for((long)n)), dim3(kBlock), 0, ctx->stream, dst, n);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// gamma5 as the spin permutation 0<->2, 1<->3 (wilsonLikeUtils.cpp:42-49)
__global__ void __launch_bounds__(kBlock)
k_gamma5_64(const double2* __restrict__ in, double2* __restrict__ out, int stride, int vh) {
  int hs = blockIdx.x * blockDim.x + threadIdx.x;
  size_t off = (size_t)blockIdx.y * 12 * stride;
  if (hs >= vh) return;
  double2 v[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = in[off + (size_t)k * stride + hs];
#pragma unroll
  for (int k = 0; k < 12; ++k) out[off + (size_t)((k + 6) % 12) * stride + hs] = v[k];
}

ii_status gamma5_64(ii_context* ctx, const double2* in, double2* out, int nrhs) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_gamma5_64, dim3(blocks_for(g.Vh), nrhs), dim3(kBlock), 0, ctx->stream, in,
                     out, g.stride, g.Vh);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// ------------------------------------------------------------ conversions
__global__ void __launch_bounds__(kBlock)
k_64_to_32(const double2* __restrict__ s, float2* __restrict__ d, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_float2((float)s[i].x, (float)s[i].y);
}
__global__ void __launch_bounds__(kBlock)
k_32_to_64(const float2* __restrict__ s, double2* __restrict__ d, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_double2((double)s[i].x, (double)s[i].y);
}
ii_status convert_64_to_32(ii_context* ctx, const double2* src, float2* dst, int nrhs) {
  size_t n = half_elems(ctx->g) * nrhs;
  hipLaunchKernelGGL(k_64_to_32, dim3(blocks_for((long)n)), dim3(kBlock), 0, ctx->stream, src,
                     dst, n);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
ii_status convert_32_to_64(ii_context* ctx, const float2* src, double2* dst, int nrhs) {
  size_t n = half_elems(ctx->g) * nrhs;
  hipLaunchKernelGGL(k_32_to_64, dim3(blocks_for((long)n)), dim3(kBlock), 0, ctx->stream, src,
                     dst, n);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// ------------------------------------------------------------ packing
// Host Format_F half field [hs][24] (already copied to the device) ->
// device [12][stride].
__global__ void __launch_bounds__(kBlock)
k_pack_half(const double* __restrict__ src, double2* __restrict__ dst, int vh, int stride) {
  int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= vh) return;
#pragma unroll
  for (int k = 0; k < 12; ++k)
    dst[(size_t)k * stride + hs] = make_double2(src[24 * (size_t)hs + 2 * k],
                                                src[24 * (size_t)hs + 2 * k + 1]);
}
__global__ void __launch_bounds__(kBlock)
k_unpack_half(const double2* __restrict__ src, double* __restrict__ dst, int vh, int stride) {
  int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= vh) return;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    double2 v = src[(size_t)k * stride + hs];
    dst[24 * (size_t)hs + 2 * k] = v.x;
    dst[24 * (size_t)hs + 2 * k + 1] = v.y;
  }
}
// Lexicographic full field [site][24]; site s of parity p has half index
// s/2 (Lx even), the esec/osec position of siteIndex_EvenOdd.cpp:40-45.
__global__ void __launch_bounds__(kBlock)
k_pack_full(const double* __restrict__ src, double2* __restrict__ ev, double2* __restrict__ od,
            Geom g) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x = s % g.L[0], r = s / g.L[0];
  int y = r % g.L[1]; r /= g.L[1];
  int z = r % g.L[2], t = r / g.L[2];
  int par = (x + y + z + t) & 1;
  double2* d = par ? od : ev;
  int hs = s >> 1;
#pragma unroll
  for (int k = 0; k < 12; ++k)
    d[(size_t)k * g.stride + hs] = make_double2(src[24 * (size_t)s + 2 * k],
                                                src[24 * (size_t)s + 2 * k + 1]);
}
__global__ void __launch_bounds__(kBlock)
k_unpack_full(const double2* __restrict__ ev, const double2* __restrict__ od,
              double* __restrict__ dst, Geom g) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x = s % g.L[0], r = s / g.L[0];
  int y = r % g.L[1]; r /= g.L[1];
  int z = r % g.L[2], t = r / g.L[2];
  int par = (x + y + z + t) & 1;
  const double2* d = par ? od : ev;
  int hs = s >> 1;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    double2 v = d[(size_t)k * g.stride + hs];
    dst[24 * (size_t)s + 2 * k] = v.x;
    dst[24 * (size_t)s + 2 * k + 1] = v.y;
  }
}

ii_status pack_half_spinor(ii_context* ctx, const double* src, double2* dst) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_pack_half, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, src, dst,
                     g.Vh, g.stride);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
ii_status unpack_half_spinor(ii_context* ctx, const double2* src, double* dst) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_unpack_half, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, src,
                     dst, g.Vh, g.stride);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
ii_status pack_full_spinor(ii_context* ctx, const double* src, double2* ev, double2* od) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_pack_full, dim3(blocks_for(g.V)), dim3(kBlock), 0, ctx->stream, src, ev,
                     od, g);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
ii_status unpack_full_spinor(ii_context* ctx, const double2* ev, const double2* od, double* dst) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_unpack_full, dim3(blocks_for(g.V)), dim3(kBlock), 0, ctx->stream, ev, od,
                     dst, g);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// Host Format_G [mu][site][3][3] -> device [parity][mu][9][stride].
__global__ void __launch_bounds__(kBlock)
k_pack_gauge(const double* __restrict__ src, double2* __restrict__ dst, Geom g, int to_host) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int mu = blockIdx.y;
  if (s >= g.V) return;
  int x = s % g.L[0], r = s / g.L[0];
  int y = r % g.L[1]; r /= g.L[1];
  int z = r % g.L[2], t = r / g.L[2];
  int par = (x + y + z + t) & 1;
  int hs = s >> 1;
  size_t hoff = 18 * ((size_t)s + (size_t)g.V * mu);
  size_t doff = (size_t)((par * 4 + mu) * 9) * g.stride + hs;
  if (to_host) {
    double* h = const_cast<double*>(src);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      double2 v = dst[doff + (size_t)k * g.stride];
      h[hoff + 2 * k] = v.x;
      h[hoff + 2 * k + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k)
      dst[doff + (size_t)k * g.stride] = make_double2(src[hoff + 2 * k], src[hoff + 2 * k + 1]);
  }
}

ii_status pack_gauge_field(ii_context* ctx, const double* src, double2* dst) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_pack_gauge, dim3(blocks_for(g.V), 4), dim3(kBlock), 0, ctx->stream, src,
                     dst, g, 0);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
ii_status unpack_gauge_field(ii_context* ctx, const double2* src, double* dst) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_pack_gauge, dim3(blocks_for(g.V), 4), dim3(kBlock), 0, ctx->stream, dst,
                     const_cast<double2*>(src), g, 1);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
ii_status pack_gauge(ii_context* ctx, const double* src) {
  return pack_gauge_field(ctx, src, ctx->gauge64);
}
ii_status unpack_gauge(ii_context* ctx, double* dst) {
  return unpack_gauge_field(ctx, ctx->gauge64, dst);
}

// Reconstruct-12 copies.  stat[0]: links whose third row matches neither
// sign of conj(row0 x row1); stat[1]: bulk links with sign -1; stat[2]:
// last-slice t-links with sign +1; stat[3]: last-slice t-links with sign -1.
template <typename T>
__device__ __forceinline__ T to_store(double2 v);
template <> __device__ __forceinline__ float2 to_store<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}
template <> __device__ __forceinline__ __half2 to_store<__half2>(double2 v) {
  return __floats2half2_rn((float)v.x, (float)v.y);
}
template <> __device__ __forceinline__ __hip_bfloat162 to_store<__hip_bfloat162>(double2 v) {
  __hip_bfloat162 r;
  r.x = __float2bfloat16((float)v.x);
  r.y = __float2bfloat16((float)v.y);
  return r;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_build_r12(const double2* __restrict__ src, T* __restrict__ dst, Geom g,
            unsigned int* __restrict__ stat) {
  int hs = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;  // parity*4 + mu
  if (hs >= g.Vh) return;
  int par = pm >> 2, mu = pm & 3;
  const double2* b = src + (size_t)(pm * 9) * g.stride + hs;
  double2 U[3][3];
#pragma unroll
  for (int k = 0; k < 9; ++k) U[k / 3][k % 3] = b[(size_t)k * g.stride];
  T* d = dst + (size_t)(pm * 6) * g.stride + hs;
#pragma unroll
  for (int k = 0; k < 6; ++k) d[(size_t)k * g.stride] = to_store<T>(U[k / 3][k % 3]);
  double2 R[3][3];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k / 3][k % 3] = U[k / 3][k % 3];
  r12_third_row(R, false);
  double ep = 0.0, em = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double dx = U[2][j].x - R[2][j].x, dy = U[2][j].y - R[2][j].y;
    ep += dx * dx + dy * dy;
    dx = U[2][j].x + R[2][j].x; dy = U[2][j].y + R[2][j].y;
    em += dx * dx + dy * dy;
  }
  SiteCoord c = decode_site(g, hs, par);
  bool bdry = (mu == 3) && (c.t == g.L[3] - 1);
  const double tol = 1e-10;
  if (ep > tol && em > tol) atomicAdd(&stat[0], 1u);
  else if (!bdry && em <= tol && ep > tol) atomicAdd(&stat[1], 1u);
  else if (bdry && ep <= tol) atomicAdd(&stat[2], 1u);
  else if (bdry && em <= tol) atomicAdd(&stat[3], 1u);
}

ii_status build_lowprec_gauge(ii_context* ctx, int prec) {
  const Geom& g = ctx->g;
  size_t elems = (size_t)2 * 4 * 6 * g.stride;
  size_t esz = prec == II_PREC_SINGLE ? sizeof(float2) : sizeof(__half2);
  if (!ctx->gauge_lp[prec]) II_HIP(ctx, hipMalloc(&ctx->gauge_lp[prec], elems * esz));
  void* statp = nullptr;
  II_CHECK(workspace(ctx, 31, 4 * sizeof(unsigned int), &statp));
  unsigned int* stat = (unsigned int*)statp;
  II_HIP(ctx, hipMemsetAsync(stat, 0, 4 * sizeof(unsigned int), ctx->stream));
  dim3 grid(blocks_for(g.Vh), 8);
  if (prec == II_PREC_SINGLE)
    hipLaunchKernelGGL(k_build_r12<float2>, grid, dim3(kBlock), 0, ctx->stream, ctx->gauge64,
                       (float2*)ctx->gauge_lp[prec], g, stat);
  else if (prec == II_PREC_HALF)
    hipLaunchKernelGGL(k_build_r12<__half2>, grid, dim3(kBlock), 0, ctx->stream, ctx->gauge64,
                       (__half2*)ctx->gauge_lp[prec], g, stat);
  else
    hipLaunchKernelGGL(k_build_r12<__hip_bfloat162>, grid, dim3(kBlock), 0, ctx->stream,
                       ctx->gauge64, (__hip_bfloat162*)ctx->gauge_lp[prec], g, stat);
  II_HIP(ctx, hipGetLastError());
  unsigned int h[4];
  II_HIP(ctx, hipMemcpyAsync(h, stat, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h[0] != 0 || h[1] != 0 || (h[2] != 0 && h[3] != 0)) {
    (void)hipFree(ctx->gauge_lp[prec]);
    ctx->gauge_lp[prec] = nullptr;
    return set_error(ctx, II_ERR_STATE,
                     "gauge field is not SU(3) up to a uniform time-boundary sign; "
                     "reconstruct-12 storage is not applicable (" + std::to_string(h[0]) +
                         " non-SU(3) links)");
  }
  ctx->tbc_sign = h[3] != 0 ? -1 : 1;
  ctx->r12_ok = true;
  return II_OK;
}

// ------------------------------------------------------------ plaquette
__device__ __forceinline__ void ld_link(const double2* __restrict__ u, const Geom& g, int par,
                                        int mu, int hs, double2 U[3][3]) {
  const double2* b = u + (size_t)((par * 4 + mu) * 9) * g.stride + hs;
#pragma unroll
  for (int k = 0; k < 9; ++k) U[k / 3][k % 3] = b[(size_t)k * g.stride];
}
__device__ __forceinline__ void mm3(double2 C[3][3], const double2 A[3][3], const double2 B[3][3],
                                    bool adjB) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (adjB) cmacc(s, B[j][k], A[i][k]);
        else cmac(s, A[i][k], B[k][j]);
      }
      C[i][j] = s;
    }
}

__global__ void __launch_bounds__(kBlock)
k_plaquette(const double2* __restrict__ u, Geom g, double* __restrict__ partials) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (s < g.V) {
    int hs = s >> 1;
    int x = s % g.L[0], r = s / g.L[0];
    int y = r % g.L[1]; r /= g.L[1];
    int z = r % g.L[2], t = r / g.L[2];
    int par = (x + y + z + t) & 1;
    SiteCoord c = decode_site(g, hs, par);
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu) {
        double2 A[3][3], B[3][3], T1[3][3], T2[3][3];
        ld_link(u, g, par, mu, hs, A);
        ld_link(u, g, 1 - par, nu, neighbour(g, c, hs, mu, true), B);
        mm3(T1, A, B, false);
        ld_link(u, g, 1 - par, mu, neighbour(g, c, hs, nu, true), B);
        mm3(T2, T1, B, true);
        ld_link(u, g, par, nu, hs, B);
        mm3(T1, T2, B, true);
        acc += (T1[0][0].x + T1[1][1].x + T1[2][2].x) / 3.0;
      }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

ii_status plaquette64(ii_context* ctx, double* plaq) {
  const Geom& g = ctx->g;
  int nb = blocks_for(g.V);
  II_CHECK(ensure_partials(ctx, (size_t)nb + 1));
  hipLaunchKernelGGL(k_plaquette, dim3(nb), dim3(kBlock), 0, ctx->stream, ctx->gauge64, g,
                     ctx->partials);
  II_HIP(ctx, hipGetLastError());
  II_CHECK(finalize_sums(ctx, ctx->partials, nb, ctx->partials + nb, 1));
  double h = 0.0;
  II_HIP(ctx, hipMemcpyAsync(&h, ctx->partials + nb, sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *plaq = h / (6.0 * g.V);
  return II_OK;
}

// ------------------------------------------------------------ FP64 operators
ii_status dhat64(ii_context* ctx, double kappa, const double2* f, double2* out, double2* t,
                 int nrhs) {
  HopSpec h1;
  h1.out_par = 1;
  h1.a = 1.0;
  h1.nrhs = nrhs;
  II_CHECK(hop64(ctx, h1, f, t, nullptr, nullptr));
  HopSpec h2;
  h2.out_par = 0;
  h2.a = -kappa * kappa;
  h2.xpay = true;
  h2.nrhs = nrhs;
  return hop64(ctx, h2, t, out, f, nullptr);
}

ii_status dhat_dag64(ii_context* ctx, double kappa, const double2* f, double2* out, double2* t,
                     const double2* dv, double* partials, int nrhs) {
  HopSpec h1;
  h1.out_par = 1;
  h1.a = 1.0;
  h1.g5in = true;
  h1.nrhs = nrhs;
  II_CHECK(hop64(ctx, h1, f, t, nullptr, nullptr));
  HopSpec h2;
  h2.out_par = 0;
  h2.a = -kappa * kappa;
  h2.xpay = true;
  h2.g5out = true;
  h2.nrhs = nrhs;
  h2.epi = dv != nullptr ? kEpiDot : kEpiStore;
  h2.partials = partials;
  return hop64(ctx, h2, t, out, f, dv);
}

ii_status ddagd64(ii_context* ctx, double kappa, const double2* f, double2* out, double2* q,
                  double2* t, const double2* dv, double* partials, int nrhs) {
  II_CHECK(dhat64(ctx, kappa, f, q, t, nrhs));
  return dhat_dag64(ctx, kappa, q, out, t, dv, partials, nrhs);
}

}  // namespace ii
