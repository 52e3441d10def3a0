// Gauge configuration files: the formats of GaugeGlobal::initialize
// (lib/Main/gaugeGlobal.cpp:133-180) "Binary", "NERSC" and "ILDG", read into
// and written from the device gauge field.
//
// The host reads or writes the file bytes and the text/record headers; the
// device does the per-link work on the raw bytes: byte order, 32/64-bit
// conversion, the third row of 3x2-compressed links, the reorder between
// the file order and the site-major SoA layout, the NERSC checksum, the
// plaquette and link trace.
//
// File data order (all three formats): sites in lexicographic order
// (x fastest), per site the 4 directions, per direction the 3x3 matrix row
// by row as (re, im) pairs -- the reference's format(gsite, in, ex) =
// in + 18 (ex + 4 gsite) (binary_reader.cpp:20-22, NERSC_reader.cpp:24-26,
// ILDG_reader.cpp:18-20).  Byte order, as the reference has it on a
// little-endian host:
//   Binary  read: big-endian doubles (BinaryReader::read byte-swaps,
//                 binary_reader.cpp:24-35); write: host-order doubles
//                 (BinaryWriter::write, binary_writer.cpp:19-22, does not
//                 swap), so the reference cannot read back its own Binary
//                 files on such a host; kept as is;
//   NERSC   big-endian IEEE64BIG or IEEE32(BIG), "4D_SU3_GAUGE_3x3" or
//           "4D_SU3_GAUGE" (first two rows; third = conj(row0 x row1),
//           NERSC_reader.cpp:218-240); written as 3x3 IEEE64BIG with the
//           header lines of NERSC_writer.cpp:55-76;
//   ILDG    LIME records; the "ildg-binary-data" record holds big-endian
//           doubles (ILDG_reader.cpp:22-54); written as a "lime-test-text"
//           record and the data record (ILDG_writer.cpp:22-74).
#include "ii_ops.h"
#include "ii_stencil.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace ii {

__device__ __forceinline__ uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }
__device__ __forceinline__ uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// one link (site s, direction mu) from the raw file words to the device field
__global__ void __launch_bounds__(kBlock)
k_io_decode(const unsigned char* __restrict__ raw, double2* __restrict__ u, Geom g, int bytes,
            int nreal, int big_endian) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (s >= g.V) return;
  const size_t base = ((size_t)s * 4 + mu) * nreal;
  double m[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) {
    if (k >= nreal) break;
    if (bytes == 8) {
      uint64_t w;
      __builtin_memcpy(&w, raw + (base + k) * 8, 8);
      if (big_endian) w = bswap64(w);
      double d;
      __builtin_memcpy(&d, &w, 8);
      m[k] = d;
    } else {
      uint32_t w;
      __builtin_memcpy(&w, raw + (base + k) * 4, 4);
      if (big_endian) w = bswap32(w);
      float f;
      __builtin_memcpy(&f, &w, 4);
      m[k] = (double)f;
    }
  }
  if (nreal == 12) {
    // third row = conj(row0 x row1), the reference's expression order
    m[12] = m[2] * m[10] - m[4] * m[8] - m[3] * m[11] + m[5] * m[9];
    m[13] = m[4] * m[9] - m[2] * m[11] + m[5] * m[8] - m[3] * m[10];
    m[14] = m[4] * m[6] - m[0] * m[10] - m[5] * m[7] + m[1] * m[11];
    m[15] = m[0] * m[11] - m[4] * m[7] + m[1] * m[10] - m[5] * m[6];
    m[16] = m[0] * m[8] - m[2] * m[6] - m[1] * m[9] + m[3] * m[7];
    m[17] = m[2] * m[7] - m[0] * m[9] + m[3] * m[6] - m[1] * m[8];
  }
  const int x = s % g.L[0], r = s / g.L[0];
  const int y = r % g.L[1], r2 = r / g.L[1];
  const int z = r2 % g.L[2], t = r2 / g.L[2];
  const int par = (x + y + z + t) & 1;
  const size_t doff = (size_t)((par * 4 + mu) * 9) * g.stride + (s >> 1);
#pragma unroll
  for (int k = 0; k < 9; ++k) u[doff + (size_t)k * g.stride] = make_double2(m[2 * k], m[2 * k + 1]);
}

// device field -> 64-bit file words (big-endian or host order)
__global__ void __launch_bounds__(kBlock)
k_io_encode(const double2* __restrict__ u, unsigned char* __restrict__ raw, Geom g,
            int big_endian) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (s >= g.V) return;
  const int x = s % g.L[0], r = s / g.L[0];
  const int y = r % g.L[1], r2 = r / g.L[1];
  const int z = r2 % g.L[2], t = r2 / g.L[2];
  const int par = (x + y + z + t) & 1;
  const size_t doff = (size_t)((par * 4 + mu) * 9) * g.stride + (s >> 1);
  const size_t base = ((size_t)s * 4 + mu) * 18;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double2 v = u[doff + (size_t)k * g.stride];
    uint64_t a, b;
    __builtin_memcpy(&a, &v.x, 8);
    __builtin_memcpy(&b, &v.y, 8);
    if (big_endian) {
      a = bswap64(a);
      b = bswap64(b);
    }
    __builtin_memcpy(raw + (base + 2 * k) * 8, &a, 8);
    __builtin_memcpy(raw + (base + 2 * k + 1) * 8, &b, 8);
  }
}

// sum of the 32-bit words of nbytes (a multiple of 4) as stored, modulo 2^32
__global__ void __launch_bounds__(kBlock)
k_io_checksum(const uint32_t* __restrict__ w, size_t nwords, uint32_t* __restrict__ partial) {
  uint32_t acc = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
       i += (size_t)gridDim.x * blockDim.x)
    acc += w[i];
  __shared__ uint32_t sm[kBlock];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

// Re tr U summed per block (Staples::link_trace, staples.cpp:24-33)
__global__ void __launch_bounds__(kBlock)
k_link_trace(const double2* __restrict__ u, Geom g, double* __restrict__ partials) {
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (hs < g.Vh) {
    for (int pm = 0; pm < 8; ++pm) {
      const double2* b = u + (size_t)pm * 9 * g.stride + hs;
      acc += b[0].x + b[4 * (size_t)g.stride].x + b[8 * (size_t)g.stride].x;
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

namespace {

constexpr int kCheckBlocks = 1024;
constexpr int kWsRaw = 6, kWsChk = 7;

ii_status checksum_dev(ii_context* ctx, const void* dev, size_t nbytes, uint32_t* out) {
  void* p = nullptr;
  II_CHECK(workspace(ctx, kWsChk, kCheckBlocks * sizeof(uint32_t), &p));
  hipLaunchKernelGGL(k_io_checksum, dim3(kCheckBlocks), dim3(kBlock), 0, ctx->stream,
                     (const uint32_t*)dev, nbytes / 4, (uint32_t*)p);
  II_HIP(ctx, hipGetLastError());
  std::vector<uint32_t> h(kCheckBlocks);
  II_HIP(ctx, hipMemcpyAsync(h.data(), p, kCheckBlocks * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  uint32_t s = 0;
  for (uint32_t v : h) s += v;
  *out = s;
  return II_OK;
}

ii_status link_trace_dev(ii_context* ctx, double* out) {
  const Geom& g = ctx->g;
  const int nb = blocks_for(g.Vh);
  II_CHECK(ensure_partials(ctx, (size_t)nb + 1));
  hipLaunchKernelGGL(k_link_trace, dim3(nb), dim3(kBlock), 0, ctx->stream, ctx->gauge64, g,
                     ctx->partials);
  II_HIP(ctx, hipGetLastError());
  II_CHECK(finalize_sums(ctx, ctx->partials, nb, ctx->partials + nb, 1));
  double h = 0.0;
  II_HIP(ctx, hipMemcpyAsync(&h, ctx->partials + nb, sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = h / (3.0 * 4.0 * g.V);
  return II_OK;
}

struct FileCloser {
  FILE* f;
  ~FileCloser() { if (f) std::fclose(f); }
};

// pinned host staging, freed on scope exit
struct Pinned {
  void* p = nullptr;
  ~Pinned() { if (p) (void)hipHostFree(p); }
};

// the NERSC header as NERSCReader::header parses it (NERSC_reader.cpp:48-72):
// one "KEY = VALUE" per line between BEGIN_HEADER and END_HEADER
bool nersc_header(FILE* f, std::map<std::string, std::string>& kv, std::string& err) {
  char line[1024];
  if (!std::fgets(line, sizeof line, f) || std::strcmp(line, "BEGIN_HEADER\n") != 0) {
    err = "NERSC type header not found";
    return false;
  }
  for (;;) {
    if (!std::fgets(line, sizeof line, f)) {
      err = "NERSC header without END_HEADER";
      return false;
    }
    if (std::strcmp(line, "END_HEADER\n") == 0) return true;
    std::string l(line);
    if (!l.empty() && l.back() == '\n') l.pop_back();
    const size_t k = l.find_first_of(" =");
    if (k == std::string::npos) continue;
    const std::string key = l.substr(0, k);
    size_t v = l.find('=', k);
    v = (v == std::string::npos) ? l.size() : v + 1;
    while (v < l.size() && l[v] == ' ') ++v;
    kv[key] = l.substr(v);
  }
}

// ------------------------------------------------------------------ LIME
// Record header (144 bytes, big-endian): magic 0x456789ab (4), version 1
// (2), flags (2: MB = 0x8000, ME = 0x4000), data length (8), type (128,
// NUL padded); the data follows, padded with zeros to a multiple of 8.
constexpr uint32_t kLimeMagic = 0x456789abu;

void put_be(unsigned char* p, uint64_t v, int n) {
  for (int i = n - 1; i >= 0; --i) {
    p[i] = (unsigned char)(v & 0xFF);
    v >>= 8;
  }
}
uint64_t get_be(const unsigned char* p, int n) {
  uint64_t v = 0;
  for (int i = 0; i < n; ++i) v = (v << 8) | p[i];
  return v;
}

bool lime_write_record(FILE* f, const char* type, bool mb, bool me, const void* data,
                       uint64_t len) {
  unsigned char h[144];
  std::memset(h, 0, sizeof h);
  put_be(h, kLimeMagic, 4);
  put_be(h + 4, 1, 2);
  put_be(h + 6, (mb ? 0x8000u : 0u) | (me ? 0x4000u : 0u), 2);
  put_be(h + 8, len, 8);
  std::strncpy((char*)h + 16, type, 127);
  if (std::fwrite(h, 1, 144, f) != 144) return false;
  if (len && std::fwrite(data, 1, len, f) != len) return false;
  static const unsigned char zeros[8] = {0};
  const uint64_t pad = (8 - len % 8) % 8;
  return pad == 0 || std::fwrite(zeros, 1, pad, f) == pad;
}

}  // namespace
}  // namespace ii

using namespace ii;

namespace {

ii_status gauge_read(ii_context* ctx, const char* path, int format, int check,
                     ii_io_info* info) {
  if (!ctx || !path || format < II_IO_BINARY || format > II_IO_ILDG) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot open ") + path);
  FileCloser fc{f};
  ii_io_info inf;
  std::memset(&inf, 0, sizeof inf);
  inf.bits = 64;
  bool big = true;
  std::map<std::string, std::string> kv;
  if (format == II_IO_NERSC) {
    std::string err;
    if (!nersc_header(f, kv, err)) return set_error(ctx, II_ERR_BAD_ARG, err);
    const std::string dt = kv["DATATYPE"], fp = kv["FLOATING_POINT"];
    if (dt == "4D_SU3_GAUGE") inf.is3x2 = 1;
    else if (dt == "4D_SU3_GAUGE_3x3") inf.is3x2 = 0;
    else return set_error(ctx, II_ERR_BAD_ARG, "NERSC DATATYPE " + dt + " not supported");
    if (fp == "IEEE32" || fp == "IEEE32BIG") inf.bits = 32;
    else if (fp == "IEEE64BIG") inf.bits = 64;
    else return set_error(ctx, II_ERR_BAD_ARG, "NERSC FLOATING_POINT " + fp + " not supported");
    for (int m = 0; m < 4; ++m) {
      const std::string key = "DIMENSION_" + std::to_string(m + 1);
      if (std::atoi(kv[key].c_str()) != g.L[m])
        return set_error(ctx, II_ERR_BAD_ARG,
                         "NERSC file dimension " + key + " does not match the lattice");
    }
    inf.header_plaquette = std::atof(kv["PLAQUETTE"].c_str());
    inf.header_link_trace = std::atof(kv["LINK_TRACE"].c_str());
    inf.header_checksum = (unsigned)std::strtoul(kv["CHECKSUM"].c_str(), nullptr, 16);
  } else if (format == II_IO_ILDG) {
    // walk the records to "ildg-binary-data" (ILDG_reader.cpp:44-54); an
    // "ildg-format" record announcing <precision>32</precision> selects
    // 32-bit data (the reference assumes 64).  Record lengths come from the
    // file, so each is checked against the bytes the file still holds; only
    // the small ildg-format body is read, every other record is skipped.
    if (fseeko(f, 0, SEEK_END) != 0)
      return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot seek in ") + path);
    const uint64_t fsize = (uint64_t)ftello(f);
    if (fseeko(f, 0, SEEK_SET) != 0)
      return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot seek in ") + path);
    constexpr uint64_t kMaxFormatRecord = 1u << 20;
    for (;;) {
      unsigned char h[144];
      if (std::fread(h, 1, 144, f) != 144)
        return set_error(ctx, II_ERR_BAD_ARG, "ILDG: no ildg-binary-data record");
      if ((uint32_t)get_be(h, 4) != kLimeMagic)
        return set_error(ctx, II_ERR_BAD_ARG, "ILDG: bad LIME magic number");
      const uint64_t len = get_be(h + 8, 8);
      const uint64_t here = (uint64_t)ftello(f);
      if (here > fsize || len > fsize - here)
        return set_error(ctx, II_ERR_BAD_ARG, "ILDG: record length beyond the end of the file");
      char type[129];
      std::memcpy(type, h + 16, 128);
      type[128] = 0;
      if (std::strncmp(type, "ildg-binary-data", 16) == 0) {
        const uint64_t want = (uint64_t)g.V * 4 * 18 * (inf.bits / 8);
        if (len != want)
          return set_error(ctx, II_ERR_BAD_ARG,
                           "ILDG: ildg-binary-data holds " + std::to_string(len) +
                               " bytes, the lattice needs " + std::to_string(want));
        break;
      }
      if (std::strncmp(type, "ildg-format", 11) == 0 && len <= kMaxFormatRecord) {
        std::string xml(len, '\0');
        if (len && std::fread(&xml[0], 1, len, f) != len)
          return set_error(ctx, II_ERR_BAD_ARG, "ILDG: truncated record");
        const size_t p = xml.find("<precision>");
        if (p != std::string::npos && std::atoi(xml.c_str() + p + 11) == 32) inf.bits = 32;
      } else if (fseeko(f, (off_t)len, SEEK_CUR) != 0) {
        return set_error(ctx, II_ERR_BAD_ARG, "ILDG: truncated record");
      }
      const uint64_t pad = (8 - len % 8) % 8;
      if (pad && fseeko(f, (off_t)pad, SEEK_CUR) != 0)
        return set_error(ctx, II_ERR_BAD_ARG, "ILDG: truncated record");
    }
  }
  const int nreal = inf.is3x2 ? 12 : 18;
  const int bytes = inf.bits / 8;
  const size_t nbytes = (size_t)g.V * 4 * nreal * bytes;
  Pinned hp;
  II_HIP(ctx, hipHostMalloc(&hp.p, nbytes));
  if (std::fread(hp.p, 1, nbytes, f) != nbytes)
    return set_error(ctx, II_ERR_BAD_ARG, std::string("short read from ") + path);
  void* raw = nullptr;
  II_CHECK(workspace(ctx, kWsRaw, nbytes, &raw));
  II_HIP(ctx, hipMemcpyAsync(raw, hp.p, nbytes, hipMemcpyHostToDevice, ctx->stream));
  if (format == II_IO_NERSC) II_CHECK(checksum_dev(ctx, raw, nbytes, &inf.checksum));
  hipLaunchKernelGGL(k_io_decode, dim3(blocks_for(g.V), 4), dim3(kBlock), 0, ctx->stream,
                     (const unsigned char*)raw, ctx->gauge64, g, bytes, nreal, big ? 1 : 0);
  II_HIP(ctx, hipGetLastError());
  // the field is new: drop the low-precision copies as ii_gauge_upload does
  for (int i = 0; i < ii_context::kGaugeCopies; ++i)
    if (ctx->gauge_lp[i]) {
      II_HIP(ctx, hipStreamSynchronize(ctx->stream));
      (void)hipFree(ctx->gauge_lp[i]);
      ctx->gauge_lp[i] = nullptr;
    }
  ctx->r12_ok = false;
  ctx->tbc_sign = 1;
  ctx->gauge_valid = true;
  II_CHECK(plaquette64(ctx, &inf.plaquette));
  II_CHECK(link_trace_dev(ctx, &inf.link_trace));
  if (info) *info = inf;
  if (check && format == II_IO_NERSC) {
    // NERSCReader::check (NERSC_reader.cpp:172-183)
    const double tol = inf.bits == 32 ? 1e-8 : 1e-15;
    if (!(std::fabs(inf.plaquette - inf.header_plaquette) < tol)) {
      char buf[200];
      std::snprintf(buf, sizeof buf, "NERSC plaquette check failed: %.17g read, %.17g in header",
                    inf.plaquette, inf.header_plaquette);
      return set_error(ctx, II_ERR_STATE, buf);
    }
  }
  return II_OK;
}

ii_status gauge_write(ii_context* ctx, const char* path, int format, ii_io_info* info) {
  if (!path || format < II_IO_BINARY || format > II_IO_ILDG) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  const size_t nbytes = (size_t)g.V * 4 * 18 * 8;
  void* raw = nullptr;
  II_CHECK(workspace(ctx, kWsRaw, nbytes, &raw));
  const bool big = format != II_IO_BINARY;
  hipLaunchKernelGGL(k_io_encode, dim3(blocks_for(g.V), 4), dim3(kBlock), 0, ctx->stream,
                     (const double2*)ctx->gauge64, (unsigned char*)raw, g, big ? 1 : 0);
  II_HIP(ctx, hipGetLastError());
  ii_io_info inf;
  std::memset(&inf, 0, sizeof inf);
  inf.bits = 64;
  II_CHECK(plaquette64(ctx, &inf.plaquette));
  II_CHECK(link_trace_dev(ctx, &inf.link_trace));
  if (format == II_IO_NERSC) II_CHECK(checksum_dev(ctx, raw, nbytes, &inf.checksum));
  Pinned hp;
  II_HIP(ctx, hipHostMalloc(&hp.p, nbytes));
  II_HIP(ctx, hipMemcpyAsync(hp.p, raw, nbytes, hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot create ") + path);
  FileCloser fc{f};
  bool ok = true;
  if (format == II_IO_NERSC) {
    // NERSCWriter::header (NERSC_writer.cpp:55-76)
    ok = std::fprintf(f,
                      "BEGIN_HEADER\nHDR_VERSION = 1.0\nDATATYPE = 4D_SU3_GAUGE_3x3\n"
                      "STORAGE_FORMAT = 1.0\nDIMENSION_1 = %d\nDIMENSION_2 = %d\n"
                      "DIMENSION_3 = %d\nDIMENSION_4 = %d\nLINK_TRACE = %1.15g\n"
                      "PLAQUETTE = %1.15g\nCHECKSUM = %x\nENSEMBLE_ID = %s\n"
                      "SEQUENCE_NUMBER = %d\nFLOATING_POINT = IEEE64BIG\nEND_HEADER\n",
                      g.L[0], g.L[1], g.L[2], g.L[3], inf.link_trace, inf.plaquette,
                      inf.checksum, "JLQCD_null", 0) > 0;
    inf.header_plaquette = inf.plaquette;
    inf.header_link_trace = inf.link_trace;
    inf.header_checksum = inf.checksum;
  }
  if (format == II_IO_ILDG) {
    static const char msg[] = "LIME test message";
    ok = lime_write_record(f, "lime-test-text", true, false, msg, std::strlen(msg)) &&
         lime_write_record(f, "ildg-binary-data", false, true, hp.p, nbytes);
  } else {
    ok = ok && std::fwrite(hp.p, 1, nbytes, f) == nbytes;
  }
  if (!ok) return set_error(ctx, II_ERR_BAD_ARG, std::string("write to ") + path + " failed");
  if (info) *info = inf;
  return II_OK;
}

}  // namespace

extern "C" {

// File contents decide allocation sizes and string parsing, so a C++
// exception (bad_alloc, length_error) must not cross the C boundary.
ii_status ii_gauge_read(ii_context* ctx, const char* path, int format, int check,
                        ii_io_info* info) {
  try {
    return gauge_read(ctx, path, format, check, info);
  } catch (const std::exception& e) {
    return set_error(ctx, II_ERR_BAD_ARG, std::string("gauge read: ") + e.what());
  }
}

ii_status ii_gauge_write(ii_context* ctx, const char* path, int format, ii_io_info* info) {
  try {
    return gauge_write(ctx, path, format, info);
  } catch (const std::exception& e) {
    return set_error(ctx, II_ERR_BAD_ARG, std::string("gauge write: ") + e.what());
  }
}

}  // extern "C"
