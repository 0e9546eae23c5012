// conv.hip -- Convolution::Engine (fused three-stage form) and the fused
// LoadToFold pipeline (unpack + convolution + detection + fold in three
// kernels).
//
// Reference: Convolution::transformation (Signal/General/Convolution.C:646-813),
// CUDA::ConvolutionEngine::perform (ConvolutionCUDA.cu:567-770), the LoadToFold
// chain (Signal/Pulsar/LoadToFold1.C:109-258).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>

#include "dspx_internal.hpp"
#include "fb4.hpp"
#include "fold_impl.hpp"
#include "onepass.hpp"
#include "plain.hpp"
#include "stageb_mfma.hpp"
#include "stages.hpp"

namespace dspx {

// N = N1 * N2 with N2 <= 4096 (one LDS row) and N1 <= 1024 (column tile).
static bool split_n(uint32_t N, uint32_t& N1, uint32_t& N2, uint32_t& m)
{
  if (N == 0 || (N & (N - 1))) return false;
  m = 0;
  while ((1u << m) < N) m++;
  if (m < 12 || m > 22) return false;
  if (m >= 16) {
    N1 = 1u << (m - 12);
    N2 = 4096;
  } else {
    N1 = 16;
    N2 = 1u << (m - 4);
  }
  return true;
}

// Everything the three stages share for one geometry.
struct ConvCore {
  dspx_ctx* ctx = nullptr;
  uint32_t nchan = 0, npol = 0, N = 0, N1 = 0, N2 = 0, m = 0;
  uint32_t nfilt_pos = 0, nfilt_neg = 0, step = 0;
  uint32_t chirp_mode = 0;
  bool onepass = false;  // the one-kernel form (onepass.hpp) instead of stages A/B/C
  dspx_dedisp_desc dd{};
  DevBuf tw2;      // lo[2^h] then hi[N >> h]
  int h = 0;
  DevBuf chan_par; // coeff[nchan], cf[nchan], fdelay[nchan]
  DevBuf table;    // [nchan][N1][N2]
  DevBuf work;     // [nchan][npart][npol][N], unless work_ext is set
  DevBuf* work_ext = nullptr;  // shared work buffer (plans of a batch)
  uint32_t work_npart = 0;
  double binwidth = 0, half_cw = 0;
  float snorm = 0;
  bool have_table = false;
  // complex values per thread in the row transform (16 or 8; the faster
  // of the two measured for each row length); a tuning parameter
  // (environment DSPX_ROW_E overrides the default)
  int row_e = 16;
  // the matrix-core form of 4096-point rows (stageb_mfma.hpp): the largest
  // magnitude of an unpacked input component (0: unknown, float input, the
  // form is not used) and whether it is selected
  float in_max = 0.f;
  bool row_mfma = false;
  // real-sampled input (k_stageB_real): W_2N^{k1 + N1*k2} at [k1][k2]
  bool nyquist = false;
  DevBuf splitw;

  void init(dspx_ctx* c, uint32_t nchan_, uint32_t npol_, uint32_t N_, uint32_t pos, uint32_t neg,
            uint32_t mode, const dspx_dedisp_desc& d, bool one = false)
  {
    ctx = c;
    onepass = one;
    nchan = nchan_;
    npol = npol_;
    N = N_;
    nfilt_pos = pos;
    nfilt_neg = neg;
    chirp_mode = mode;
    dd = d;
    DSPX_REQUIRE(npol == 1 || npol == 2, DSPX_E_PARAM, "npol=%u", npol);
    DSPX_REQUIRE(nchan >= 1, DSPX_E_PARAM, "nchan=0");
    if (onepass) {
      DSPX_REQUIRE(N && !(N & (N - 1)) && N >= (1u << kOnePassMinLog2) && N <= (1u << kOnePassMaxLog2),
                   DSPX_E_UNSUPPORTED, "nfft=%u: the one-pass form supports powers of two 2^%d..2^%d", N,
                   kOnePassMinLog2, kOnePassMaxLog2);
      m = 0;
      while ((1u << m) < N) m++;
      N1 = 1;
      N2 = N;
    } else {
      DSPX_REQUIRE(split_n(N, N1, N2, m), DSPX_E_UNSUPPORTED,
                   "nfft=%u: the three-stage form supports powers of two 2^12..2^22", N);
    }
    DSPX_REQUIRE(uint64_t(pos) + neg < N, DSPX_E_STATE, "error nfft=%u < nfilt=%u", N, pos + neg);
    step = N - pos - neg;
    DSPX_REQUIRE(mode <= DSPX_CHIRP_TRUTH, DSPX_E_PARAM, "chirp_mode=%u", mode);
    DSPX_REQUIRE(!onepass || mode != DSPX_CHIRP_TABLE, DSPX_E_UNSUPPORTED, "the one-pass form evaluates the chirp on chip");
    // two-level twiddle table W_N^m = lo[m & (2^h-1)] * hi[m >> h]
    h = (int(m) + 1) / 2;
    if (onepass) h = int(m);  // unused: the one-pass form takes the context's tables
    uint32_t nlo = 1u << h, nhi = N >> h;
    std::vector<float2> t(nlo + nhi);
    for (uint32_t i = 0; i < nlo; i++) {
      double a = -2.0 * M_PI * double(i) / double(N);
      t[i] = make_float2(float(std::cos(a)), float(std::sin(a)));
    }
    for (uint32_t i = 0; i < nhi; i++) {
      double a = -2.0 * M_PI * double(uint64_t(i) << h) / double(N);
      t[nlo + i] = make_float2(float(std::cos(a)), float(std::sin(a)));
    }
    tw2.reserve(t.size() * sizeof(float2));
    DSPX_HIP(hipMemcpy(tw2.p, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
    snorm = float(1.0 / std::sqrt(double(N)));  // Convolution.C:401-414
    row_e = N2 == 512 || N2 == 1024 ? 8 : 16;
    if (const char* e = std::getenv("DSPX_ROW_E")) row_e = std::atoi(e) == 8 ? 8 : 16;
    if (N2 < 256 * 2 && row_e == 8) row_e = 16;
    if (const char* e = std::getenv("DSPX_STAGEB")) row_mfma = std::string(e) == "mfma";
    if (mode != DSPX_CHIRP_TABLE) {
      DSPX_REQUIRE(d.nchan == nchan, DSPX_E_PARAM, "dedisp.nchan=%u != nchan=%u", d.nchan, nchan);
      DSPX_REQUIRE(d.bandwidth != 0.0 && d.centre_frequency > 0.0, DSPX_E_PARAM, "invalid band");
      double chanwidth = d.bandwidth / double(nchan);
      binwidth = chanwidth / double(N);
      half_cw = 0.5 * chanwidth;
      std::vector<double> pc(3 * nchan, 0.0);
      for (uint32_t ic = 0; ic < nchan; ic++) {
        double cf = channel_centre(d.centre_frequency, d.bandwidth, nchan, d.dc_centred, ic);
        pc[ic] = chirp_coeff(d.bandwidth, d.dispersion_measure, cf);
        pc[nchan + ic] = cf;
        if (d.interchan)
          pc[2 * nchan + ic] =
              interchan_fractional_delay(d.centre_frequency, d.bandwidth, nchan, d.dispersion_measure, cf);
      }
      chan_par.reserve(pc.size() * sizeof(double));
      DSPX_HIP(hipMemcpy(chan_par.p, pc.data(), pc.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }

  void set_table(const float* host)
  {
    // transpose bin k = k1 + N1*k2 to [k1][k2] so stage B reads rows
    std::vector<float2> t(size_t(nchan) * N);
    const float2* hp = reinterpret_cast<const float2*>(host);
    for (uint32_t ic = 0; ic < nchan; ic++)
      for (uint32_t k1 = 0; k1 < N1; k1++)
        for (uint32_t k2 = 0; k2 < N2; k2++)
          t[(size_t(ic) * N1 + k1) * N2 + k2] = hp[size_t(ic) * N + k1 + size_t(N1) * k2];
    table.reserve(t.size() * sizeof(float2));
    DSPX_HIP(hipMemcpy(table.p, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
    have_table = true;
  }

  DevBuf& wbuf() { return work_ext ? *work_ext : work; }
  const DevBuf& wbuf() const { return work_ext ? *work_ext : work; }

  void reserve_work(uint32_t npart)
  {
    if (onepass) return;  // no work buffer
    wbuf().reserve(size_t(nchan) * npart * npol * N * sizeof(float2));
    work_npart = npart;
  }

  ConvShape shape(uint32_t npart) const
  {
    ConvShape s;
    s.work = wbuf().as<float2>();
    s.N1 = N1;
    s.N2 = N2;
    s.log2N = m;
    s.nchan = nchan;
    s.npol = npol;
    s.npart = npart;
    s.step = step;
    s.nfilt_pos = nfilt_pos;
    s.tw.lo = tw2.as<float2>();
    s.tw.hi = tw2.as<float2>() + (1u << h);
    s.tw.h = h;
    return s;
  }

  ChirpArgs chirp() const
  {
    ChirpArgs c;
    c.mode = chirp_mode;
    c.snorm = snorm;
    c.binwidth = binwidth;
    c.half_cw = half_cw;
    c.coeff = chan_par.as<double>();
    c.cf = chan_par.as<double>() ? chan_par.as<double>() + nchan : nullptr;
    c.fdelay = chan_par.as<double>() ? chan_par.as<double>() + 2 * nchan : nullptr;
    c.table = table.as<float2>();
    c.nyquist = nyquist ? 1u : 0u;
    return c;
  }

  // Real-sampled input: the split twiddles of k_stageB_real.
  void set_nyquist()
  {
    DSPX_REQUIRE(!onepass, DSPX_E_UNSUPPORTED, "real-sampled input in the one-pass form");
    nyquist = true;
    std::vector<float2> t(static_cast<size_t>(N));
    for (uint32_t k1 = 0; k1 < N1; k1++)
      for (uint32_t k2 = 0; k2 < N2; k2++) {
        const double a = -M_PI * double(k1 + uint64_t(N1) * k2) / double(N);
        t[size_t(k1) * N2 + k2] = make_float2(float(std::cos(a)), float(std::sin(a)));
      }
    splitw.reserve(t.size() * sizeof(float2));
    DSPX_HIP(hipMemcpy(splitw.p, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
  }

  uint32_t ncb() const { return N2 / (kColTile / N1); }

  template <class LD> void stage_a(const LD& ld, uint32_t npart)
  {
    ConvShape s = shape(npart);
    uint64_t nb = uint64_t(ncb()) * npart * nchan;
    DSPX_REQUIRE(nb < (1ull << 31), DSPX_E_UNSUPPORTED, "grid too large");
    const float2* twc = twiddles(ctx, N1);
    switch (N1) {
#define CASE_A(n) \
  case n: k_stageA<n, LD><<<unsigned(nb), kColThreads, 0, ctx->stream>>>(ld, s, twc); break;
      CASE_A(16) CASE_A(32) CASE_A(64) CASE_A(128) CASE_A(256) CASE_A(512) CASE_A(1024)
#undef CASE_A
      default: fail(DSPX_E_UNSUPPORTED, "N1=%u", N1);
    }
    DSPX_HIP(hipGetLastError());
  }

  void stage_b(uint32_t npart)
  {
    DSPX_REQUIRE(chirp_mode != DSPX_CHIRP_TABLE || have_table, DSPX_E_STATE,
                 "no response: call dspx_conv_set_response first");
    ConvShape s = shape(npart);
    ChirpArgs c = chirp();
    // parts per workgroup: as many as keep >= 2048 workgroups in flight,
    // and at least enough to give each of the rw rows in flight a row
    const uint32_t rw = uint32_t(stage_b_rows(int(N2), row_e));
    uint64_t rows = uint64_t(N1) * nchan;
    uint32_t npg = rows >= 2048 ? 1u : uint32_t(std::min<uint64_t>(npart, (2048 + rows - 1) / rows));
    uint32_t ppg = std::max((npart + npg - 1) / npg, std::min(npart, (rw + npol - 1) / npol));
    npg = (npart + ppg - 1) / ppg;
    uint64_t nb = rows * npg;
    DSPX_REQUIRE(nb < (1ull << 31), DSPX_E_UNSUPPORTED, "grid too large");
    const float2* twr = twiddles(ctx, N2);
    if (nyquist) {
      // row pairs (k1, N1 - k1), rows 0 and N1/2 alone
      const uint64_t nbr = uint64_t(N1 / 2 + 1) * nchan * npg;
      DSPX_REQUIRE(nbr < (1ull << 31), DSPX_E_UNSUPPORTED, "grid too large");
      switch (N2) {
#define CASE_BR(n) \
  case n: k_stageB_real<n><<<unsigned(nbr), n / 16, 0, ctx->stream>>>(s, c, twr, splitw.as<float2>(), ppg); break;
        CASE_BR(256) CASE_BR(512) CASE_BR(1024) CASE_BR(2048) CASE_BR(4096)
#undef CASE_BR
        default: fail(DSPX_E_UNSUPPORTED, "N2=%u", N2);
      }
      DSPX_HIP(hipGetLastError());
      return;
    }
    if (row_mfma && N2 == 4096 && in_max > 0.f) {
      // scale on load so that no complex magnitude exceeds 46341 (its FP16
      // components then stay below 65504 through every pass): stage A's
      // N1-point column sums reach N1 * sqrt(2) * in_max at most
      const double peak = double(N1) * std::sqrt(2.0) * double(in_max);
      const int p = int(std::ceil(std::log2(peak / 46341.0)));
      const int q = int(m) / 2;
      MfmaScale sc{std::ldexp(1.0f, -p), std::ldexp(1.0f, q), std::ldexp(1.0f, p + 24 - q)};
      k_stageB_mfma<<<unsigned(nb), 256, 0, ctx->stream>>>(s, c, twr, ppg, sc);
      DSPX_HIP(hipGetLastError());
      return;
    }
#define CASE_B(n, e)                                                                      \
  case n * 100 + e: {                                                                     \
    constexpr int RW = stage_b_rows(n, e);                                                \
    k_stageB<n, e, RW><<<unsigned(nb), RW * (n / e), 0, ctx->stream>>>(s, c, twr, ppg);   \
  } break;
    switch (N2 * 100 + row_e) {
      CASE_B(256, 16) CASE_B(512, 16) CASE_B(1024, 16) CASE_B(2048, 16) CASE_B(4096, 16)
      CASE_B(512, 8) CASE_B(1024, 8) CASE_B(2048, 8) CASE_B(4096, 8)
#undef CASE_B
      default: fail(DSPX_E_UNSUPPORTED, "N2=%u", N2);
    }
    DSPX_HIP(hipGetLastError());
  }

  // ZH: the fold applies the zap mask and counts per-channel hits (fold_lane)
  template <int MODE, bool ZH = false> void stage_c(uint32_t npart, const FilterbankArgs& fb, const FoldArgs& fa)
  {
    static_assert(!ZH || MODE == kStageCFold, "the zap mask is a fold option");
    ConvShape s = shape(npart);
    uint32_t G = MODE == kStageCFold ? fa.G : ncb();
    uint64_t nb = uint64_t(G) * npart * nchan;
    DSPX_REQUIRE(nb < (1ull << 31), DSPX_E_UNSUPPORTED, "grid too large");
    size_t shm = kColTile * sizeof(float2);
    if (MODE == kStageCFold) shm += size_t(fa.nslot) * fa.hrow * sizeof(float);
    DSPX_REQUIRE(shm <= 160 * 1024, DSPX_E_UNSUPPORTED, "LDS request %zu too large", shm);
    const float2* twc = twiddles(ctx, N1);
    // The next tile's prefetch (64 registers) only where it costs no
    // occupancy: where the LDS already holds the CU to two workgroups (the
    // prefetching kernels take 178-210 VGPRs, two waves per SIMD; without it
    // 114-146, three or four)
    const bool pf = MODE == kStageCFold && !ZH && N1 <= 64 && 3 * shm > 160 * 1024;
    switch (N1) {
#define LAUNCH_C(n, nv, p, w)                                                                           \
  do {                                                                                                   \
    /* the whole LDS allowed once per kernel (not a stream operation, so \
       also outside any graph capture that follows) */                                                   \
    static bool lds_attr = false;                                                                        \
    if (!lds_attr) {                                                                                     \
      allow_full_lds((const void*) k_stageC<n, MODE, nv, p, ZH, w>);                                     \
      lds_attr = true;                                                                                   \
    }                                                                                                    \
    k_stageC<n, MODE, nv, p, ZH, w><<<unsigned(nb), kColThreads, shm, ctx->stream>>>(s, fb, fa, twc);    \
  } while (0)
// the narrow-bin ownership only matters at 32 and 64 points (FoldOwn)
#define LAUNCH_CW(n, nv, p)                                          \
  do {                                                               \
    if constexpr (MODE == kStageCFold && (n == 32 || n == 64)) {     \
      if (fa.narrow)                                                 \
        LAUNCH_C(n, nv, p, true);                                    \
      else                                                           \
        LAUNCH_C(n, nv, p, false);                                   \
    } else {                                                         \
      LAUNCH_C(n, nv, p, true);                                      \
    }                                                                \
  } while (0)
#define LAUNCH_C2(n, nv)                                     \
  do {                                                       \
    if constexpr (MODE == kStageCFold && n <= 64 && !ZH) {   \
      if (pf)                                                \
        LAUNCH_CW(n, nv, true);                              \
      else                                                   \
        LAUNCH_CW(n, nv, false);                             \
    } else {                                                 \
      LAUNCH_CW(n, nv, false);                               \
    }                                                        \
  } while (0)
#define CASE_C(n)                                    \
  case n: {                                          \
    if (MODE == kStageCFilterbank || fa.nval == 1)   \
      LAUNCH_C2(n, 1);                               \
    else if (fa.nval == 2)                           \
      LAUNCH_C2(n, 2);                               \
    else                                             \
      LAUNCH_C2(n, 4);                               \
  } break;
      CASE_C(16) CASE_C(32) CASE_C(64) CASE_C(128) CASE_C(256) CASE_C(512) CASE_C(1024)
#undef CASE_C
#undef LAUNCH_C2
#undef LAUNCH_CW
#undef LAUNCH_C
      default: fail(DSPX_E_UNSUPPORTED, "N1=%u", N1);
    }
    DSPX_HIP(hipGetLastError());
  }

  // The one-kernel form over a block's windows (onepass.hpp); fold mode
  // needs fa's histogram geometry and key plan, filterbank mode fb.
  template <int MODE, class LD, bool ZH = false>
  void run_onepass(const LD& ld, const OnePassArgs& a, const FilterbankArgs& fb, const FoldArgs& fa)
  {
    static_assert(!ZH || MODE == kStageCFold, "the zap mask is a fold option");
    DSPX_REQUIRE(onepass, DSPX_E_INTERNAL, "not a one-pass plan");
    const uint64_t nb = uint64_t(a.nchan) * a.ngrp;
    DSPX_REQUIRE(nb >= 1 && nb < (1ull << 31), DSPX_E_UNSUPPORTED, "grid of %llu workgroups",
                 (unsigned long long) nb);
    size_t shm = onepass_lds_bytes(N);
    if (MODE == kStageCFold) shm += size_t(fa.nslot) * fa.hrow * sizeof(float);
    DSPX_REQUIRE(shm <= 160 * 1024, DSPX_E_UNSUPPORTED, "one-pass LDS request %zu too large", shm);
    // a window's word offsets j * tstride (j < N) are 32-bit products
    DSPX_REQUIRE(uint64_t(N) * ld.tstride < (1ull << 32), DSPX_E_UNSUPPORTED, "sample stride %llu",
                 (unsigned long long) ld.tstride);
    ChirpArgs c = chirp();
    const float2* tw = twiddles(ctx, int(N));
#define LAUNCH_OP(n, nv)                                                                                        \
  do {                                                                                                          \
    static bool lds_attr = false;                                                                               \
    if (!lds_attr) {                                                                                            \
      allow_full_lds((const void*) k_onepass<n, MODE, nv, LD, ZH>);                                             \
      lds_attr = true;                                                                                          \
    }                                                                                                           \
    k_onepass<n, MODE, nv, LD, ZH><<<unsigned(nb), OnePass<n>::NT, shm, ctx->stream>>>(ld, a, c, fb, fa, tw);   \
  } while (0)
#define CASE_OP(n)                                     \
  case n:                                              \
    if (MODE == kStageCFilterbank || fa.nval == 1)     \
      LAUNCH_OP(n, 1);                                 \
    else if (fa.nval == 2)                             \
      LAUNCH_OP(n, 2);                                 \
    else                                               \
      LAUNCH_OP(n, 4);                                 \
    break;
    switch (N) {
      CASE_OP(128) CASE_OP(256) CASE_OP(512) CASE_OP(1024) CASE_OP(2048) CASE_OP(4096) CASE_OP(8192)
      default: fail(DSPX_E_UNSUPPORTED, "one-pass nfft=%u", N);
    }
#undef CASE_OP
#undef LAUNCH_OP
    DSPX_HIP(hipGetLastError());
  }
};

// Raw TFP [ndat][nchan] words (one word = both pols of one sample) to
// channel-major [nchan][ndat] words, so that stage A reads each channel's
// time samples contiguously.  64 x 64 word tiles through LDS.  Full tiles
// with 16-byte aligned rows move 16-byte vectors: a 16-lane group reads 256
// contiguous bytes of one time row and writes 256 contiguous bytes of one
// channel row.  Edge tiles and unaligned layouts move single words.
template <class W>
__global__ __launch_bounds__(256) void k_tfp_to_ftp(const W* __restrict__ in, uint64_t ndat, uint32_t nchan,
                                                    W* __restrict__ out, uint64_t out_cstride, uint64_t out_off,
                                                    int vec)
{
  __shared__ W tile[64][65];
  const uint64_t t0 = uint64_t(blockIdx.x) * 64;
  const uint32_t c0 = blockIdx.y * 64;
  constexpr int VW = 16 / sizeof(W);  // words per vector
  constexpr int VPR = 64 / VW;        // vectors per tile row
  constexpr int RPI = 256 / VPR;      // tile rows per pass
  if (vec && t0 + 64 <= ndat && c0 + 64 <= nchan) {
    const uint32_t x = threadIdx.x % VPR, y = threadIdx.x / VPR;
#pragma unroll
    for (int k = 0; k < 64 / RPI; k++) {
      const uint32_t i = y + RPI * k;
      const uint4 v = *reinterpret_cast<const uint4*>(in + (t0 + i) * nchan + c0 + x * VW);
      const W* w = reinterpret_cast<const W*>(&v);
#pragma unroll
      for (int j = 0; j < VW; j++) tile[i][x * VW + j] = w[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 64 / RPI; k++) {
      const uint32_t i = y + RPI * k;
      uint4 v;
      W* w = reinterpret_cast<W*>(&v);
#pragma unroll
      for (int j = 0; j < VW; j++) w[j] = tile[x * VW + j][i];
      *reinterpret_cast<uint4*>(out + uint64_t(c0 + i) * out_cstride + out_off + t0 + x * VW) = v;
    }
    return;
  }
  const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
  for (uint32_t k = 0; k < 16; k++) {
    const uint32_t i = ty + 4 * k;
    uint64_t t = t0 + i;
    uint32_t c = c0 + tx;
    if (t < ndat && c < nchan) tile[i][tx] = in[t * nchan + c];
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < 16; k++) {
    const uint32_t i = ty + 4 * k;
    uint32_t c = c0 + i;
    uint64_t t = t0 + tx;
    if (t < ndat && c < nchan) out[uint64_t(c) * out_cstride + out_off + t] = tile[tx][i];
  }
}

// TFP words -> channel-major words: out[c * out_cstride + out_off + t] (in
// words of wbytes, one sample of one channel, all pols).
void tfp_to_ftp(dspx_ctx* ctx, const void* in, uint64_t ndat, uint32_t nchan, uint32_t wbytes, void* out,
                uint64_t out_cstride, uint64_t out_off)
{
  DSPX_REQUIRE((ndat + 63) / 64 < (1ull << 31), DSPX_E_UNSUPPORTED, "ndat=%llu", (unsigned long long) ndat);
  dim3 grid(unsigned((ndat + 63) / 64), (nchan + 63) / 64);
  // vectors need every row start (input times, output channels) 16-byte aligned
  const uint64_t vw = 16 / wbytes;
  const int vec = (reinterpret_cast<uintptr_t>(in) % 16 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0 &&
                   nchan % vw == 0 && out_cstride % vw == 0 && out_off % vw == 0)
                      ? 1
                      : 0;
  switch (wbytes) {
#define CASE_W(b, T)                                                                                                \
  case b:                                                                                                           \
    k_tfp_to_ftp<T><<<grid, 256, 0, ctx->stream>>>((const T*) in, ndat, nchan, (T*) out, out_cstride, out_off, vec); \
    break;
    CASE_W(2, uint16_t) CASE_W(4, uint32_t) CASE_W(8, uint64_t)
#undef CASE_W
    default: fail(DSPX_E_INTERNAL, "word size %u", wbytes);
  }
  DSPX_HIP(hipGetLastError());
}

// Per-channel copy of n words: dst[c * dcs + doff + t] = src[c * scs + soff + t].
template <class W>
__global__ void k_copy_chan(const W* __restrict__ src, uint64_t scs, uint64_t soff, W* __restrict__ dst, uint64_t dcs,
                            uint64_t doff, uint64_t n)
{
  uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t c = blockIdx.y;
  if (t < n) dst[uint64_t(c) * dcs + doff + t] = src[uint64_t(c) * scs + soff + t];
}

void copy_chan(dspx_ctx* ctx, const void* src, uint64_t scs, uint64_t soff, void* dst, uint64_t dcs, uint64_t doff,
               uint64_t n, uint32_t nchan, uint32_t wbytes)
{
  if (!n) return;
  DSPX_REQUIRE((n + 255) / 256 < (1ull << 31) && nchan <= 65535, DSPX_E_UNSUPPORTED, "copy too large");
  dim3 grid(unsigned((n + 255) / 256), nchan);
  switch (wbytes) {
#define CASE_W(b, T)                                                                                               \
  case b:                                                                                                          \
    k_copy_chan<T><<<grid, 256, 0, ctx->stream>>>((const T*) src, scs, soff, (T*) dst, dcs, doff, n); \
    break;
    CASE_W(2, uint16_t) CASE_W(4, uint32_t) CASE_W(8, uint64_t)
#undef CASE_W
    default: fail(DSPX_E_INTERNAL, "word size %u", wbytes);
  }
  DSPX_HIP(hipGetLastError());
}

// zero-DM copy (Convolution.C:792-799)
__global__ void k_zero_dm(const float* __restrict__ in, uint64_t ics, uint64_t ips, float* __restrict__ out,
                          uint64_t ocs, uint64_t ops, uint32_t npol, uint32_t step, uint32_t nfilt_pos,
                          uint64_t nout)
{
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t cp = blockIdx.y;
  if (i >= nout) return;
  uint32_t chan = cp / npol, pol = cp % npol;
  uint64_t p = i / step, j = i % step;
  const float* x = in + chan * ics + pol * ips + 2 * (p * step + nfilt_pos + j);
  float* o = out + chan * ocs + pol * ops + 2 * i;
  *reinterpret_cast<float2*>(o) = *reinterpret_cast<const float2*>(x);
}

// Sum the fused stage C partial histograms into the profiles, in a fixed
// order (the result does not depend on scheduling):
//   k_fold_reduce_g: part[chan][p][slot][e] = sum over groups g
//   k_fold_reduce_p: prof[isub][chan][e] += sum over parts p whose slots hold isub
//                    (the one-pass form's workgroups of windows take the place
//                    of parts here)
__global__ void k_fold_reduce_g(const float* __restrict__ partial, uint32_t G, uint32_t nslot_hl, uint64_t n,
                                float* __restrict__ out)
{
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // (chan, p, slot, e)
  if (i >= n) return;
  uint64_t cp = i / nslot_hl, r = i % nslot_hl;
  const float* b = partial + cp * G * nslot_hl + r;
  float acc = 0.f;
#pragma unroll 8
  for (uint32_t g = 0; g < G; g++) acc += b[uint64_t(g) * nslot_hl];
  out[i] = acc;
}

//   (ZH: the rows also hold nbin hit counts after the nvb profile values;
//   each unit's count, an integer below 2^24 and so exact in a float, is
//   added as an integer into hits[isub][chan][bin])
template <bool ZH>
__global__ void k_fold_reduce_p(const float* __restrict__ part, const uint32_t* __restrict__ meta, uint32_t meta_sf,
                                uint32_t meta_off, uint32_t meta_parts, uint32_t npart, uint32_t nslot, uint32_t hl,
                                uint32_t nvb, uint32_t nchan, float* __restrict__ prof, uint32_t* __restrict__ hits)
{
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t chan = blockIdx.y;
  uint32_t isub = blockIdx.z;
  if (e >= hl) return;
  const uint32_t* sub_first = meta + meta_sf;
  const uint32_t k0 = meta[meta_off + isub], k1 = meta[meta_off + isub + 1];
  if (k0 == k1) return;
  if (ZH && e >= nvb) {
    if (!hits) return;
    uint32_t acc = 0;
    for (uint32_t k = k0; k < k1; k++) {
      const uint32_t p = meta[meta_parts + k];
      acc += uint32_t(part[((uint64_t(chan) * npart + p) * nslot + (isub - sub_first[p])) * hl + e]);
    }
    const uint32_t nbin = hl - nvb;
    hits[(uint64_t(isub) * nchan + chan) * nbin + (e - nvb)] += acc;
    return;
  }
  float acc = 0.f;
#pragma unroll 8
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t p = meta[meta_parts + k];
    acc += part[((uint64_t(chan) * npart + p) * nslot + (isub - sub_first[p])) * hl + e];
  }
  prof[(uint64_t(isub) * nchan + chan) * nvb + e] += acc;
}

// The SpectralKurtosis mask on an FPT complex series of nout samples per
// row (the plain form's convolution output): sample idat of channel c is
// zeroed where mask[idat / M][c] is set or idat >= nblk*M.
__global__ void k_zap_fpt(float* __restrict__ x, uint64_t chan_stride, uint64_t pol_stride, uint32_t npol,
                          uint64_t nout, ZapArgs z)
{
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nout) return;
  const uint32_t chan = blockIdx.y;
  const uint64_t k = i / z.M;
  if (k < z.nblk && !z.mask[k * z.nchan + chan]) return;
  for (uint32_t pol = 0; pol < npol; pol++)
    *reinterpret_cast<float2*>(x + chan * chan_stride + pol * pol_stride + 2 * i) = make_float2(0.f, 0.f);
}

// The fold plan in the order stage C meets it, evaluated from the plan's
// segments: for each part p and each output t of its transform, the LDS
// histogram key ((isub - sub_first[p]) << 16) | ibin of output sample
// p*step + t - nfilt_pos, or ~0 where the sample is discarded (the overlap)
// or not folded.  Indexed like the work array, so a thread of stage C reads
// its consecutive columns' keys with one vector load.  The same pass counts
// the plan's hits (Fold.C:827-836): each thread takes kKeyRun/4 groups of 4
// consecutive outputs, group q of thread t at offset 4*(q*256 + t) of the
// workgroup's chunk, so that each vector store of a wave writes 1 KiB in one
// piece.  Runs of equal keys are added into an LDS histogram of the part's
// slots, once per wave where the wave's groups share one key, and the
// workgroup adds its non-zero bins into hits[isub][ibin].  Within a segment
// the bin never decreases, so a group whose outputs lie in one segment and
// whose first and last keys agree evaluates two samples, not 4.
constexpr int kKeyRun = 16;
constexpr int kKeyChunk = 256 * kKeyRun;  // outputs per workgroup, inside one part
__global__ __launch_bounds__(256) void k_fused_keys(const FoldSegment* __restrict__ seg, uint32_t nseg, double nbin_d,
                                                    uint32_t nbin, const uint32_t* __restrict__ sub_first,
                                                    uint32_t nslot, uint32_t nfft, uint32_t step, uint32_t nfilt_pos,
                                                    uint64_t n, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ hits)
{
  extern __shared__ uint32_t h[];  // nslot * nbin
  // the segments from the one holding the workgroup's first kept output:
  // kSegCache of them in LDS (a 4096-output chunk spans a few)
  constexpr uint32_t kSegCache = 64;
  __shared__ FoldSegment sseg[kSegCache];
  __shared__ uint32_t s_si0;
  __shared__ uint64_t s_jend;  // first sample past the cached segments
  const uint64_t c0 = uint64_t(blockIdx.x) * kKeyChunk;
  const uint32_t p = uint32_t(c0 / nfft);
  const uint32_t s0 = sub_first[p];
  const uint32_t hn = nslot * nbin;
  for (uint32_t k = threadIdx.x; k < hn; k += blockDim.x) h[k] = 0;
  uint32_t si0 = 0;
  if (nseg <= kSegCache) {
    // every segment fits: one load round, no search
    if (threadIdx.x < nseg) sseg[threadIdx.x] = seg[threadIdx.x];
    if (threadIdx.x == 0) s_jend = ~uint64_t(0);
  } else {
    if (threadIdx.x == 0) {
      const uint32_t tc = uint32_t(c0 - uint64_t(p) * nfft);
      const uint32_t tf = max(tc, nfilt_pos);
      const uint64_t idat0 = uint64_t(p) * step + (tf < nfilt_pos + step ? tf - nfilt_pos : 0);
      s_si0 = seg_find(seg, nseg, idat0);
    }
    __syncthreads();
    si0 = s_si0;
    const uint32_t nc = min(kSegCache, nseg - si0);
    if (threadIdx.x < nc) sseg[threadIdx.x] = seg[si0 + threadIdx.x];
    if (threadIdx.x == 0) s_jend = si0 + nc < nseg ? seg[si0 + nc].j0 : ~uint64_t(0);
  }
  __syncthreads();
  const uint32_t ns = min(kSegCache, nseg - si0);
  const uint64_t jbeg = sseg[0].j0, jend = s_jend;
  // the thread's current segment, held in registers: its outputs are
  // consecutive samples and nearly always lie in one segment
  FoldSegment sg{};
  uint64_t jlo = 1, jnext = 0;  // sg covers [jlo, jnext); empty at first
  auto locate = [&](uint64_t idat) {
    if (idat >= jlo && idat < jnext) return;
    if (idat >= jbeg && idat < jend) {
      uint32_t lo = 0, hi = ns;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sseg[mid].j0 <= idat)
          lo = mid;
        else
          hi = mid;
      }
      sg = sseg[lo];
      jnext = lo + 1 < ns ? sseg[lo + 1].j0 : jend;
    } else {
      const uint32_t g = seg_find(seg, nseg, idat);
      sg = seg[g];
      jnext = g + 1 < nseg ? seg[g + 1].j0 : ~uint64_t(0);
    }
    jlo = sg.j0;
  };
  auto slot_key = [&](uint32_t pl) {
    return pl == 0xffffffffu ? 0xffffffffu : (((pl >> 16) - s0) << 16) | (pl & 0xffffu);
  };
  constexpr uint32_t kNone = 0xffffffffu;
  auto add_hits = [&](uint32_t key, uint32_t cnt) {
    if (key != kNone && (key >> 16) < nslot) atomicAdd(&h[(key >> 16) * nbin + (key & 0xffffu)], cnt);
  };
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < kKeyRun / 4; q++) {
    const uint64_t i0 = c0 + 4 * (uint64_t(q) * 256 + threadIdx.x);
    // the group's part (transforms shorter than a chunk put several parts in
    // one chunk, all with the same sub_first: they share a one-pass group)
    const uint32_t pq = nfft >= uint32_t(kKeyChunk) ? p : uint32_t(i0 / nfft);
    uint32_t kv[4];
    // Common case: the group's 4 outputs kept and in one segment, where the
    // bin is non-decreasing, so equal keys at both ends mean one key.
    uint32_t ukey = kNone;  // the group's one key, else kNone
    bool uniform = false;
    const uint32_t tt0 = uint32_t(i0 - uint64_t(pq) * nfft);
    if (i0 + 4 <= n && tt0 >= nfilt_pos && tt0 - nfilt_pos + 3 < step) {
      const uint64_t ia = uint64_t(pq) * step + (tt0 - nfilt_pos), ib = ia + 3;
      locate(ia);
      if (ib < jnext) {
        const uint32_t ka = seg_key(sg, ia, nbin_d);
        if (ka == seg_key(sg, ib, nbin_d)) {
          ukey = slot_key(ka);
          uniform = true;
        }
      }
    }
    if (uniform) {
#pragma unroll
      for (int k = 0; k < 4; k++) kv[k] = ukey;
    } else {
      uint32_t cur = kNone, cnt = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint64_t i = i0 + k;
        const uint32_t t = uint32_t(i - uint64_t(pq) * nfft);
        uint32_t key = kNone;
        if (i < n && t >= nfilt_pos && t - nfilt_pos < step) {
          const uint64_t idat = uint64_t(pq) * step + (t - nfilt_pos);
          locate(idat);
          key = slot_key(seg_key(sg, idat, nbin_d));
        }
        kv[k] = key;
        if (key != cur) {
          add_hits(cur, cnt);
          cur = key;
          cnt = 0;
        }
        cnt++;
      }
      add_hits(cur, cnt);
    }
    // the uniform groups' hits: the lanes whose key is lane 0's add it once
    const uint32_t lead = __builtin_amdgcn_readfirstlane(ukey);
    const bool same = uniform && ukey == lead;
    const uint64_t m = __ballot(same);
    if (same) {
      if (lane == uint32_t(__ffsll((long long) m) - 1)) add_hits(lead, 4u * uint32_t(__popcll(m)));
    } else if (uniform) {
      add_hits(ukey, 4u);
    }
    if (i0 + 4 <= n) {
      *reinterpret_cast<uint4*>(keys + i0) = make_uint4(kv[0], kv[1], kv[2], kv[3]);
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (i0 + k < n) keys[i0 + k] = kv[k];
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < hn; k += blockDim.x)
    if (h[k]) atomicAdd(&hits[uint64_t(s0 + k / nbin) * nbin + (k % nbin)], h[k]);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v)
{
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_add_u32(const uint32_t* __restrict__ src, uint64_t n, uint32_t* __restrict__ dst)
{
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += src[i];
}

}  // namespace dspx

using namespace dspx;

// ------------------------------------------------------- Convolution engine

struct dspx_conv {
  CtxRef hold;
  dspx_ctx* ctx = nullptr;
  dspx_conv_desc d{};
  ConvCore core;
  PlainConv plain;
};

extern "C" {

int dspx_conv_create(dspx_ctx* ctx, const dspx_conv_desc* d, dspx_conv** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && d && out, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipSetDevice(ctx->device));
    auto* c = new dspx_conv;
    c->hold.reset(ctx);
    c->ctx = ctx;
    c->d = *d;
    try {
      DSPX_REQUIRE(d->nyquist <= 1, DSPX_E_PARAM, "nyquist=%u", d->nyquist);
      if (d->form == DSPX_FORM_PLAIN) {
        DSPX_REQUIRE(!d->nyquist, DSPX_E_UNSUPPORTED, "real-sampled input in the plain form");
        c->plain.init(ctx, d->nchan, d->npol, d->nfft, d->nfilt_pos, d->nfilt_neg, d->chirp_mode, d->dedisp);
      } else if (d->form == DSPX_FORM_FUSED || d->form == DSPX_FORM_TWOLEVEL) {
        c->core.init(ctx, d->nchan, d->npol, d->nfft, d->nfilt_pos, d->nfilt_neg, d->chirp_mode, d->dedisp);
        if (d->nyquist) c->core.set_nyquist();
      } else {
        fail(DSPX_E_PARAM, "form=%u", d->form);
      }
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

void dspx_conv_destroy(dspx_conv* c)
{
  if (!c) return;
  (void) hipSetDevice(c->ctx->device);
  (void) hipStreamSynchronize(c->ctx->stream);
  delete c;
}

int dspx_conv_set_response(dspx_conv* c, const float* host)
{
  dspx_ctx* ctx = c ? c->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(c && host, DSPX_E_PARAM, "null argument");
    if (c->d.form == DSPX_FORM_PLAIN)
      c->plain.set_table(host);
    else
      c->core.set_table(host);
  });
}

int dspx_conv_perform(dspx_conv* c, const float* d_in, uint64_t ics, uint64_t ips, uint64_t ndat, float* d_out,
                      uint64_t ocs, uint64_t ops, float* d_zdm, uint32_t npart)
{
  dspx_ctx* ctx = c ? c->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(c && d_in && d_out, DSPX_E_PARAM, "null argument");
    const uint32_t N = c->d.nfft, nov = c->d.nfilt_pos + c->d.nfilt_neg, step = N - nov;
    // real-sampled input: parts of 2N real samples every 2*step
    uint64_t max_part = c->d.nyquist ? (ndat < 2ull * N ? 0 : (ndat - 2ull * nov) / (2ull * step))
                                     : (ndat < N ? 0 : (ndat - nov) / step);
    DSPX_REQUIRE(!(c->d.nyquist && d_zdm), DSPX_E_UNSUPPORTED, "zero-DM copy of real-sampled input");
    DSPX_REQUIRE(npart <= max_part, DSPX_E_PARAM, "npart=%u > %llu parts available in ndat=%llu", npart,
                 (unsigned long long) max_part, (unsigned long long) ndat);
    if (npart == 0) return;
    if (c->d.form == DSPX_FORM_PLAIN) {
      c->plain.perform(d_in, ics, ips, d_out, ocs, ops, npart);
    } else {
      ConvCore& k = c->core;
      k.reserve_work(npart);
      LoadF32 ld{d_in, ics, ips, k.npol};
      k.stage_a(ld, npart);
      k.stage_b(npart);
      FilterbankArgs fb{d_out, ocs, ops};
      k.stage_c<kStageCFilterbank>(npart, fb, FoldArgs{});
    }
    if (d_zdm) {
      uint64_t nout = uint64_t(npart) * step;
      dim3 grid(unsigned((nout + 255) / 256), c->d.nchan * c->d.npol);
      k_zero_dm<<<grid, 256, 0, ctx->stream>>>(d_in, ics, ips, d_zdm, ocs, ops, c->d.npol, step, c->d.nfilt_pos,
                                               nout);
      DSPX_HIP(hipGetLastError());
    }
  });
}

}  // extern "C"

// ------------------------------------------------------------ fused LoadToFold

// Large per-block scratch of the fused form: owned by one plan, or shared by
// the plans of a batch, which run one after another on one stream.
struct LtfScratch {
  DevBuf work;      // ConvCore work buffer
  DevBuf ftp;       // channel-major copy of the raw block (nchan > 1)
  DevBuf partial;   // [nchan][npart][G][nslot][nval*nbin]
  DevBuf partial2;  // [nchan][npart][nslot][nval*nbin]
};

struct dspx_ltf {
  CtxRef hold;
  dspx_ctx* ctx = nullptr;
  dspx_ltf_desc d{};
  uint64_t npart = 0, ndat_out = 0;
  uint32_t step = 0;
  uint32_t nval = 1, npol_out = 1, ndim_out = 1;
  ConvCore core;
  PlainConv plain;
  FoldPlan fold;
  bool have_fold = false;
  // fused fold: partial-histogram geometry, the per-part key plan and the
  // per-block metadata [sub_first: npart][off: nsub+1][parts: off[nsub]]
  // (the parts whose slots hold sub-integration isub are parts[off[isub] ..
  // off[isub+1]), in increasing order)
  uint32_t G = 1, nslot = 1;
  // the same for a fold with the zap mask and per-channel hits (a longer
  // histogram row); zh_ok: it fits in LDS and every unit's hit counts stay
  // below 2^24
  uint32_t G_zh = 1;
  bool zh_ok = false;
  // one-pass form (N <= 2^13): windows per workgroup and workgroups per
  // channel of the current block; the reduction's sub_first[] per workgroup
  bool onepass = false;
  uint32_t ppg = 0, ngrp = 0;
  uint32_t meta_sf = 0;
  DevBuf keys;       // [npart][nfft] histogram key of each transform output, ~0 for discarded or unfolded
  DevBuf foldmeta;
  uint32_t meta_off = 0, meta_parts = 0;  // word offsets of off[] and parts[] in foldmeta
  DevBuf plan_hits;  // [nsub][nbin] hits of the current fold plan
  HostStage meta_stage;
  LtfScratch own;
  LtfScratch* ext = nullptr;
  LtfScratch& sc() { return ext ? *ext : own; }
  // stream mode (InputBuffering): the unprocessed tail of the previous call
  // (channel-major raw words) is prepended to the next block
  bool streaming = false;
  uint64_t ncarry = 0, consumed = 0;
  uint64_t cap = 0;        // samples per channel row of the assembled input
  uint64_t carry_cap = 0;  // samples per channel row of the carry buffer
  DevBuf carry;
  // plain form intermediates
  DevBuf unpacked, filtered, detected;
  // timing
  bool timing = false;
  bool nstage_open = false;
  int nstage = 0;
  const char* stage_name[8] = {};
};

namespace {

void ltf_record(dspx_ltf* L, int /*unused*/)
{
  // event k closes stage k-1; event 0 opens the first stage
  if (L->timing) DSPX_HIP(hipEventRecord(L->ctx->ev[L->nstage + (L->nstage_open ? 1 : 0)], L->ctx->stream));
  L->nstage_open = true;
}

// Calls f(loader) with a loader over the raw block.  Multichannel TFP input
// is first transposed to channel-major order (one extra 2*B_in bytes per
// sample) so that stage A reads each channel contiguously.
template <class F> void with_loader(dspx_ltf* L, const void* raw, F&& f)
{
  const auto& d = L->d;
  uint64_t wbytes = uint64_t(d.npol) * 2 * d.nbit / 8;  // one sample of one channel, all pols
  uint64_t tstride = uint64_t(d.nchan) * d.npol * 2, cstride = uint64_t(d.npol) * 2;  // in elements
  const void* src = raw;
  if (d.nchan > 1) {
    L->sc().ftp.reserve(d.ndat * d.nchan * wbytes);
    tfp_to_ftp(L->ctx, raw, d.ndat, d.nchan, uint32_t(wbytes), L->sc().ftp.p, d.ndat, 0);
    src = L->sc().ftp.p;
    tstride = uint64_t(d.npol) * 2;
    cstride = d.ndat * d.npol * 2;
    ltf_record(L, 1);
    if (L->nstage < 7) L->stage_name[L->nstage++] = "transpose_tfp_to_ftp";
  }
  if (d.nbit == 8) {
    LoadI8 ld{static_cast<const uint8_t*>(src), tstride, cstride, d.npol, d.offset_binary, d.scale};
    f(ld);
  } else {
    LoadI16 ld{static_cast<const int16_t*>(src), tstride, cstride, d.npol, d.offset_binary};
    f(ld);
  }
}

// Windows per one-pass workgroup (a multiple of the windows in flight and of
// the windows in a key-plan chunk) and
// workgroups per channel for a block of np windows: about four rounds of the
// workgroups the device holds at once, as few as that allows (each
// workgroup evaluates its channel's chirp once and writes one histogram).
void onepass_groups(const dspx_ltf* L, uint64_t np, uint32_t& ppg, uint32_t& ngrp)
{
  // whole batches of windows in flight, and whole chunks of the key plan
  // (whose workgroups assume one sub_first per chunk)
  const uint32_t wf = std::max(onepass_wf(L->d.nfft), std::max<uint32_t>(1, uint32_t(kKeyChunk) / L->d.nfft));
  const size_t lds = onepass_lds_bytes(L->d.nfft) + size_t(L->nval) * L->d.nbin * sizeof(float);
  const uint64_t per_cu = std::max<uint64_t>(1, (160 * 1024) / lds);
  const uint64_t target = uint64_t(std::max(L->ctx->ncu, 1)) * per_cu * 4;
  const uint64_t units = std::max<uint64_t>(1, (np + wf - 1) / wf);
  const uint64_t g = std::min<uint64_t>(units, std::max<uint64_t>(1, (target + L->d.nchan - 1) / L->d.nchan));
  const uint64_t per = (units + g - 1) / g * wf;
  DSPX_REQUIRE(per < (1ull << 31), DSPX_E_UNSUPPORTED, "too many windows per channel");
  ppg = uint32_t(per);
  ngrp = uint32_t(std::max<uint64_t>(1, (np + per - 1) / per));
}

// Whether the one-pass form holds an N-point transform's exchange rows and a
// two-slot fold histogram of nval*nbin values in the LDS.
bool onepass_fits(uint32_t nfft, uint32_t nval, uint32_t nbin)
{
  if (!nfft || (nfft & (nfft - 1)) || nfft < (1u << kOnePassMinLog2) || nfft > (1u << kOnePassMaxLog2)) return false;
  return onepass_lds_bytes(nfft) + 2 * size_t(nval) * nbin * sizeof(float) <= 160 * 1024;
}

}  // namespace

namespace {

// Builds a plan; with ext, the large scratch is the caller's (shared by the
// plans of a batch) and grows to this plan's needs.
dspx_ltf* ltf_create(dspx_ctx* ctx, const dspx_ltf_desc* d, LtfScratch* ext)
{
  dspx_ltf* result = nullptr;
  {
    DSPX_REQUIRE(d->nbit == 8 || d->nbit == 16, DSPX_E_PARAM, "nbit=%u", d->nbit);
    DSPX_REQUIRE(d->npol == 1 || d->npol == 2, DSPX_E_PARAM, "npol=%u", d->npol);
    DSPX_REQUIRE(d->state <= DSPX_STOKES, DSPX_E_PARAM, "state=%u", d->state);
    DSPX_REQUIRE(d->state <= DSPX_PPQQ || d->npol == 2, DSPX_E_STATE, "Coherence/Stokes need npol=2");
    DSPX_REQUIRE(d->state <= DSPX_PPQQ || d->ndim == 1 || d->ndim == 2 || d->ndim == 4, DSPX_E_STATE,
                 "invalid ndim=%u", d->ndim);
    DSPX_REQUIRE(d->chirp_mode == DSPX_CHIRP_PARITY || d->chirp_mode == DSPX_CHIRP_TRUTH, DSPX_E_PARAM,
                 "chirp_mode=%u (on-chip modes only)", d->chirp_mode);
    DSPX_REQUIRE(d->nbin >= 1 && d->nbin <= 65535, DSPX_E_PARAM, "nbin=%u", d->nbin);
    DSPX_HIP(hipSetDevice(ctx->device));
    auto* L = new dspx_ltf;
    L->hold.reset(ctx);
    L->ctx = ctx;
    L->d = *d;
    L->ext = ext;
    if (ext) L->core.work_ext = &ext->work;
    try {
      uint32_t nov = d->nfilt_pos + d->nfilt_neg;
      DSPX_REQUIRE(nov < d->nfft, DSPX_E_STATE, "error nfft=%u < nfilt=%u", d->nfft, nov);
      L->step = d->nfft - nov;
      L->npart = d->ndat < d->nfft ? 0 : (d->ndat - nov) / L->step;
      L->ndat_out = L->npart * L->step;
      switch (d->state) {
        case DSPX_INTENSITY: L->nval = 1; L->npol_out = 1; L->ndim_out = 1; break;
        case DSPX_PPQQ: L->nval = d->npol; L->npol_out = d->npol; L->ndim_out = 1; break;
        default: L->nval = 4; L->ndim_out = d->ndim; L->npol_out = 4 / d->ndim; break;
      }
      if (d->form == DSPX_FORM_PLAIN) {
        L->plain.init(ctx, d->nchan, d->npol, d->nfft, d->nfilt_pos, d->nfilt_neg, d->chirp_mode, d->dedisp);
      } else {
        DSPX_REQUIRE(d->form == DSPX_FORM_FUSED || d->form == DSPX_FORM_TWOLEVEL || d->form == DSPX_FORM_ONEPASS,
                     DSPX_E_PARAM, "form=%u", d->form);
        // the fused form: one kernel up to 4096 points where the transform
        // and the fold histogram fit the LDS, else the three stages
        const bool fits = onepass_fits(d->nfft, L->nval, d->nbin);
        DSPX_REQUIRE(d->form != DSPX_FORM_ONEPASS || fits, DSPX_E_UNSUPPORTED,
                     "nfft=%u with %u x %u fold values: outside the one-pass form", d->nfft, L->nval, d->nbin);
        L->onepass = d->form == DSPX_FORM_ONEPASS ||
                     (d->form == DSPX_FORM_FUSED && fits && d->nfft <= (1u << kOnePassFusedMaxLog2));
        L->core.init(ctx, d->nchan, d->npol, d->nfft, d->nfilt_pos, d->nfilt_neg, d->chirp_mode, d->dedisp,
                     L->onepass);
        // largest unpacked component (LoadI8: |(v + 0.5) * scale| <= 127.5 |scale|; LoadI16: 32767.5)
        L->core.in_max = d->nbit == 8 ? 127.5f * std::fabs(d->scale) : 32767.5f;
        DSPX_REQUIRE(L->npart < (1u << 30), DSPX_E_UNSUPPORTED, "npart too large");
        DSPX_REQUIRE(L->onepass || L->npart < (1u << 20), DSPX_E_UNSUPPORTED, "npart too large");
        L->core.reserve_work(uint32_t(L->npart));
      }
    } catch (...) {
      delete L;
      throw;
    }
    result = L;
  }
  return result;
}

// Reserves the scratch an execute of L needs (so that an execute captured
// into a graph allocates nothing).
void ltf_reserve(dspx_ltf* L)
{
  const auto& d = L->d;
  if (d.form == DSPX_FORM_PLAIN || !L->npart) return;
  uint64_t wbytes = uint64_t(d.npol) * 2 * d.nbit / 8;
  if (d.nchan > 1) L->sc().ftp.reserve(d.ndat * d.nchan * wbytes);
  // the per-part reduction of stage C's column groups (fold_finish): the
  // three-stage form with more than one group per part only (one-pass
  // workgroups write per-unit partials that need no such pass)
  if (L->have_fold && !L->onepass) {
    const uint64_t per = uint64_t(d.nchan) * L->npart * L->nslot * d.nbin * sizeof(float);
    uint64_t need = L->G > 1 ? per * L->nval : 0;
    if (L->zh_ok && L->G_zh > 1) need = std::max<uint64_t>(need, per * (L->nval + 1));
    if (need) L->sc().partial2.reserve(need);
  }
}


// Per-block set-up of the fused fold, after L->fold holds this block's
// segments: the sub-integration slots of each part (of each workgroup's
// windows in the one-pass form), the partial-histogram geometry, the
// metadata upload, the key plan and the plan's hits.  Nothing here waits for
// the stream.
void fused_fold_setup(dspx_ltf* L)
{
  dspx_ctx* ctx = L->ctx;
  const uint64_t np = L->npart;
  // Units whose histograms the reduction sums: parts (three stages) or the
  // one-pass workgroups of ppg windows.  Each unit's slots start at the first
  // sub-integration its kept outputs meet.
  uint64_t per_unit = 1;
  if (L->onepass) {
    onepass_groups(L, np, L->ppg, L->ngrp);
    per_unit = L->ppg;
    DSPX_REQUIRE(L->d.nfft >= uint32_t(kKeyChunk) || (uint64_t(L->ppg) * L->d.nfft) % kKeyChunk == 0, DSPX_E_INTERNAL,
                 "one-pass groups of %u windows split key-plan chunks", L->ppg);
  }
  const uint64_t nunit = (np + per_unit - 1) / per_unit;
  std::vector<uint32_t> sfu(nunit);
  uint32_t nslot = 1;
  for (uint64_t u = 0; u < nunit; u++) {
    uint32_t first, last;
    const uint64_t p0 = u * per_unit, p1 = std::min(np, p0 + per_unit);
    L->fold.sub_range(p0 * L->step, p1 * L->step, first, last);
    if (first == 0xffffffffu) {
      first = 0;
      last = 0;
    }
    sfu[u] = first;
    nslot = std::max(nslot, last - first + 1);
  }
  L->nslot = nslot;
  const size_t hist_bytes = size_t(nslot) * L->nval * L->d.nbin * sizeof(float);
  const size_t hist_bytes_zh = size_t(nslot) * (L->nval + 1) * L->d.nbin * sizeof(float);
  const size_t tile_bytes = L->onepass ? onepass_lds_bytes(L->d.nfft) : kColTile * sizeof(float2);
  DSPX_REQUIRE(tile_bytes + hist_bytes <= 160 * 1024, DSPX_E_UNSUPPORTED,
               "fold histogram of %zu bytes does not fit in LDS", hist_bytes);
  L->zh_ok = tile_bytes + hist_bytes_zh <= 160 * 1024 && uint64_t(per_unit) * L->step < (1u << 24);
  const uint32_t nsub = L->fold.nsub;
  // units whose slots hold isub: sfu[u] <= isub < sfu[u] + nslot
  std::vector<uint32_t> off(nsub + 1, 0), units;
  for (uint32_t isub = 0; isub < nsub; isub++) {
    for (uint64_t u = 0; u < nunit; u++)
      if (isub >= sfu[u] && isub < sfu[u] + nslot) units.push_back(uint32_t(u));
    off[isub + 1] = uint32_t(units.size());
  }
  // metadata words: [sub_first of each part: np][off: nsub+1][units][sub_first of each unit: nunit]
  L->meta_off = uint32_t(np);
  L->meta_parts = uint32_t(np + nsub + 1);
  L->meta_sf = uint32_t(np + nsub + 1 + units.size());
  const size_t words = size_t(L->meta_sf) + nunit;
  L->foldmeta.reserve(words * sizeof(uint32_t));
  int slot;
  auto* h = static_cast<uint32_t*>(L->meta_stage.acquire(words * sizeof(uint32_t), slot));
  for (uint64_t p = 0; p < np; p++) h[p] = sfu[p / per_unit];
  std::copy(off.begin(), off.end(), h + L->meta_off);
  std::copy(units.begin(), units.end(), h + L->meta_parts);
  std::copy(sfu.begin(), sfu.end(), h + L->meta_sf);
  L->meta_stage.upload(slot, L->foldmeta.p, words * sizeof(uint32_t), ctx->stream);

  const uint64_t units_all = uint64_t(L->d.nchan) * nunit;
  if (L->onepass) {
    L->G = L->G_zh = 1;
    L->sc().partial.reserve(units_all * (L->zh_ok ? hist_bytes_zh : hist_bytes));
  } else {
    // Column groups per (channel, part) in stage C.  Every stage C workgroup
    // zeroes, writes and has reduced one histogram: with large histograms the
    // grid is one round of the workgroups the device holds at once
    // (LDS-bound), with small ones four rounds, for balance.  The count is
    // then lowered to give every group the same number of column tiles (128
    // tiles over 118 groups would leave most groups one tile and a few two,
    // each group with its own histogram).
    const uint32_t ncb = L->core.ncb();
    auto groups = [&](size_t hb) {
      const uint64_t per_cu = std::max<uint64_t>(1, (160 * 1024) / (kColTile * sizeof(float2) + hb));
      const uint64_t slots = uint64_t(std::max(ctx->ncu, 1)) * per_cu;
      const uint64_t target = hb > 16384 ? slots : 4 * slots;
      const uint64_t g0 = std::min<uint64_t>(ncb, std::max<uint64_t>(1, (target + units_all - 1) / units_all));
      const uint64_t tiles_per_group = (ncb + g0 - 1) / g0;
      return uint32_t((ncb + tiles_per_group - 1) / tiles_per_group);
    };
    L->G = groups(hist_bytes);
    L->G_zh = groups(hist_bytes_zh);
    L->sc().partial.reserve(units_all * std::max(L->G * hist_bytes, L->zh_ok ? L->G_zh * hist_bytes_zh : 0));
  }

  // key plan and hits
  const uint64_t n = np * L->d.nfft;
  const uint64_t nh = uint64_t(nsub) * L->d.nbin;
  DSPX_REQUIRE(L->d.nfft % 4 == 0 && (L->d.nfft % kKeyChunk == 0 || L->onepass), DSPX_E_INTERNAL, "nfft=%u",
               L->d.nfft);
  const uint64_t nkb = (n + kKeyChunk - 1) / kKeyChunk;
  DSPX_REQUIRE(nkb < (1ull << 31), DSPX_E_UNSUPPORTED, "key plan too long");
  L->keys.reserve(n * sizeof(uint32_t));
  L->plan_hits.reserve(std::max<uint64_t>(nh, 1) * sizeof(uint32_t));
  k_fill_u32<<<unsigned((nh + 255) / 256), 256, 0, ctx->stream>>>(L->plan_hits.as<uint32_t>(), nh, 0u);
  DSPX_HIP(hipGetLastError());
  const size_t shm = size_t(nslot) * L->d.nbin * sizeof(uint32_t);
  static bool lds_attr = false;
  if (!lds_attr) {
    allow_full_lds((const void*) k_fused_keys);
    lds_attr = true;
  }
  k_fused_keys<<<unsigned(nkb), 256, shm, ctx->stream>>>(
      L->fold.seg.as<FoldSegment>(), L->fold.nseg, double(L->d.nbin), L->d.nbin, L->foldmeta.as<uint32_t>(), nslot,
      L->d.nfft, L->step, L->d.nfilt_pos, n, L->keys.as<uint32_t>(), L->plan_hits.as<uint32_t>());
  DSPX_HIP(hipGetLastError());
}

// zap: the SpectralKurtosis mask of a fold with per-channel hits, or null
FoldArgs fold_args(dspx_ltf* L, const ZapArgs* zap = nullptr)
{
  FoldArgs fa;
  fa.keys = L->keys.as<uint32_t>();
  fa.partial = L->sc().partial.as<float>();
  fa.nbin = L->d.nbin;
  fa.nval = L->nval;
  fa.ndim = L->ndim_out;
  fa.nslot = L->nslot;
  fa.G = zap ? L->G_zh : L->G;
  fa.state = L->d.state;
  fa.npol_out = L->npol_out;
  for (uint32_t i = 0; i < 4; i++) fa.voff[i] = (i / L->ndim_out) * L->d.nbin * L->ndim_out + (i % L->ndim_out);
  fa.hrow = (L->nval + (zap ? 1 : 0)) * L->d.nbin;
  fa.zap = zap ? *zap : ZapArgs{};
  // samples per bin 1 / (turns per sample * bins per turn) below 64
  fa.narrow = L->fold.pps * L->d.nbin > 1.0 / 64 ? 1u : 0u;
  return fa;
}

// The profile reduction over the units' partial histograms (parts, or
// one-pass workgroups) and the hits: the plan's, or with zh the per-channel
// counts the fold made.
void fold_finish(dspx_ltf* L, uint32_t np, float* d_prof, uint32_t* d_hits, int& ev, bool zh = false)
{
  dspx_ctx* ctx = L->ctx;
  const auto& d = L->d;
  const uint32_t nunit = L->onepass ? L->ngrp : np;
  const uint32_t nvb = L->nval * d.nbin;
  const uint32_t hl = zh ? nvb + d.nbin : nvb;
  const uint32_t G = zh ? L->G_zh : L->G;
  uint64_t n2 = uint64_t(d.nchan) * nunit * L->nslot * hl;
  // with one group the partials already are [chan][unit][slot][e]
  const float* per_unit = L->sc().partial.as<float>();
  if (G > 1) {
    L->sc().partial2.reserve(n2 * sizeof(float));
    k_fold_reduce_g<<<unsigned((n2 + 255) / 256), 256, 0, ctx->stream>>>(
        L->sc().partial.as<float>(), G, L->nslot * hl, n2, L->sc().partial2.as<float>());
    DSPX_HIP(hipGetLastError());
    per_unit = L->sc().partial2.as<float>();
  }
  dim3 grid((hl + 255) / 256, d.nchan, L->fold.nsub);
  if (zh)
    k_fold_reduce_p<true><<<grid, 256, 0, ctx->stream>>>(per_unit, L->foldmeta.as<uint32_t>(), L->meta_sf,
                                                         L->meta_off, L->meta_parts, nunit, L->nslot, hl, nvb,
                                                         d.nchan, d_prof, d_hits);
  else
    k_fold_reduce_p<false><<<grid, 256, 0, ctx->stream>>>(per_unit, L->foldmeta.as<uint32_t>(), L->meta_sf,
                                                          L->meta_off, L->meta_parts, nunit, L->nslot, hl, nvb,
                                                          d.nchan, d_prof, nullptr);
  DSPX_HIP(hipGetLastError());
  if (d_hits && !zh) {
    const uint64_t nh = uint64_t(L->fold.nsub) * d.nbin;
    k_add_u32<<<unsigned((nh + 255) / 256), 256, 0, ctx->stream>>>(L->plan_hits.as<uint32_t>(), nh, d_hits);
    DSPX_HIP(hipGetLastError());
  }
  ltf_record(L, ev++);
  L->stage_name[L->nstage++] = "fold_reduce_hits";
}

// Stages B and C, the profile reduction and the hits of a fold execute
// (stage A has run on np parts); zap: the SpectralKurtosis mask, or null.
void fused_fold_after_a(dspx_ltf* L, uint32_t np, float* d_prof, uint32_t* d_hits, int& ev,
                        const ZapArgs* zap = nullptr)
{
  ConvCore& k = L->core;
  k.stage_b(np);
  ltf_record(L, ev++);
  L->stage_name[L->nstage++] = "stageB_rowfft_chirp";
  if (zap)
    k.stage_c<kStageCFold, true>(np, FilterbankArgs{}, fold_args(L, zap));
  else
    k.stage_c<kStageCFold>(np, FilterbankArgs{}, fold_args(L));
  ltf_record(L, ev++);
  L->stage_name[L->nstage++] = "stageC_colifft_detect_fold";
  fold_finish(L, np, d_prof, d_hits, ev, zap != nullptr);
}

OnePassArgs onepass_args(const dspx_ltf* L, uint32_t np, uint32_t ppg, uint32_t ngrp)
{
  return OnePassArgs{L->d.nchan, L->d.npol, np, L->step, L->d.nfilt_pos, ppg, ngrp};
}

// The whole fold execute of np windows over the loader ld: the three stages
// or the one-pass kernel, then the reduction (stage timing events as the
// stages complete).
template <class LD>
void fold_exec(dspx_ltf* L, const LD& ld, uint32_t np, float* d_prof, uint32_t* d_hits, int& ev,
               const ZapArgs* zap = nullptr)
{
  if (L->onepass) {
    const OnePassArgs oa = onepass_args(L, np, L->ppg, L->ngrp);
    if (zap)
      L->core.run_onepass<kStageCFold, LD, true>(ld, oa, FilterbankArgs{}, fold_args(L, zap));
    else
      L->core.run_onepass<kStageCFold>(ld, oa, FilterbankArgs{}, fold_args(L));
    ltf_record(L, ev++);
    L->stage_name[L->nstage++] = "onepass_unpack_fft_chirp_ifft_detect_fold";
    fold_finish(L, np, d_prof, d_hits, ev, zap != nullptr);
  } else {
    L->core.stage_a(ld, np);
    ltf_record(L, ev++);
    L->stage_name[L->nstage++] = "stageA_unpack_colfft";
    fused_fold_after_a(L, np, d_prof, d_hits, ev, zap);
  }
}

// The convolution of np windows into the FPT filterbank output.
template <class LD> void filterbank_exec(dspx_ltf* L, const LD& ld, uint32_t np, float* d_out)
{
  const uint64_t nout = uint64_t(np) * L->step;
  FilterbankArgs fb{d_out, uint64_t(L->d.npol) * nout * 2, nout * 2};
  if (L->onepass) {
    uint32_t ppg, ngrp;
    onepass_groups(L, np, ppg, ngrp);
    L->core.run_onepass<kStageCFilterbank>(ld, onepass_args(L, np, ppg, ngrp), fb, FoldArgs{});
  } else {
    L->core.stage_a(ld, np);
    L->core.stage_b(np);
    L->core.stage_c<kStageCFilterbank>(np, fb, FoldArgs{});
  }
}
}  // namespace

extern "C" {

int dspx_ltf_create(dspx_ctx* ctx, const dspx_ltf_desc* d, dspx_ltf** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && d && out, DSPX_E_PARAM, "null argument");
    *out = ltf_create(ctx, d, nullptr);
  });
}

void dspx_ltf_destroy(dspx_ltf* L)
{
  if (!L) return;
  (void) hipSetDevice(L->ctx->device);
  (void) hipStreamSynchronize(L->ctx->stream);
  delete L;
}

int dspx_ltf_geometry(const dspx_ltf* L, uint64_t* npart, uint64_t* ndat_out)
{
  if (!L || !npart || !ndat_out) return DSPX_E_PARAM;
  *npart = L->npart;
  *ndat_out = L->ndat_out;
  return DSPX_OK;
}

int dspx_ltf_set_fold(dspx_ltf* L, double pps, const dspx_fold_chunk* chunks, uint32_t nchunk)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L && chunks, DSPX_E_PARAM, "null argument");
    const bool fused = L->d.form != DSPX_FORM_PLAIN;
    L->have_fold = false;  // until the whole plan is in place
    L->fold.build(ctx, L->ndat_out, pps, L->d.nbin, chunks, nchunk, !fused);
    if (fused && L->npart) fused_fold_setup(L);
    L->have_fold = true;
  });
}

}  // extern "C"

namespace {

// One block through the fold plan; zap: the SpectralKurtosis mask with
// per-channel hits of zeroed data, or null.
void ltf_execute(dspx_ltf* L, const void* d_raw, float* d_prof, uint32_t* d_hits, const ZapArgs* zap)
{
  dspx_ctx* ctx = L->ctx;
  DSPX_REQUIRE(L->have_fold, DSPX_E_STATE, "no fold plan: call dspx_ltf_set_fold");
  const auto& d = L->d;
  if (!L->npart) return;
  int ev = 0;
  L->nstage = 0;
  L->nstage_open = false;
  ltf_record(L, ev++);
  if (d.form != DSPX_FORM_PLAIN) {
    uint32_t np = uint32_t(L->npart);
    with_loader(L, d_raw, [&](auto& ld) { fold_exec(L, ld, np, d_prof, d_hits, ev, zap); });
  } else {
    // plain: unpack -> hipFFT convolution -> (mask) -> detect -> fold, each a pass
    uint64_t nin = d.ndat, nout = L->ndat_out;
    L->unpacked.reserve(size_t(d.nchan) * d.npol * nin * sizeof(float2));
    L->filtered.reserve(size_t(d.nchan) * d.npol * nout * sizeof(float2));
    L->detected.reserve(size_t(d.nchan) * L->nval * nout * sizeof(float));
    unpack(ctx, d_raw, d.nbit, d.offset_binary, d.nchan, d.npol, nin, d.scale, L->unpacked.as<float>(),
           uint64_t(d.npol) * nin * 2, nin * 2);
    ltf_record(L, ev++);
    L->stage_name[L->nstage++] = "unpack";
    L->plain.perform(L->unpacked.as<float>(), uint64_t(d.npol) * nin * 2, nin * 2, L->filtered.as<float>(),
                     uint64_t(d.npol) * nout * 2, nout * 2, uint32_t(L->npart));
    ltf_record(L, ev++);
    L->stage_name[L->nstage++] = "convolution_hipfft";
    if (zap) {
      dim3 grid(unsigned((nout + 255) / 256), d.nchan);
      k_zap_fpt<<<grid, 256, 0, ctx->stream>>>(L->filtered.as<float>(), uint64_t(d.npol) * nout * 2, nout * 2,
                                               d.npol, nout, *zap);
      DSPX_HIP(hipGetLastError());
    }
    uint64_t ops = nout * L->ndim_out;  // floats per output pol
    uint64_t ocs = uint64_t(L->npol_out) * ops;
    detect(ctx, d.state, L->ndim_out, L->filtered.as<float>(), uint64_t(d.npol) * nout * 2, nout * 2, d.nchan,
           d.npol, nout, L->detected.as<float>(), ocs, ops);
    ltf_record(L, ev++);
    L->stage_name[L->nstage++] = "detect";
    if (zap) {
      L->fold.fold_fpt(L->detected.as<float>(), ocs, ops, d.nchan, L->npol_out, L->ndim_out, d_prof, d_hits,
                       d.nchan);
    } else {
      L->fold.fold_fpt(L->detected.as<float>(), ocs, ops, d.nchan, L->npol_out, L->ndim_out, d_prof);
      if (d_hits) L->fold.hits(d_hits);
    }
    ltf_record(L, ev++);
    L->stage_name[L->nstage++] = "fold_hits";
  }
}

}  // namespace

extern "C" {

int dspx_ltf_execute(dspx_ltf* L, const void* d_raw, float* d_prof, uint32_t* d_hits)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L && d_raw && d_prof, DSPX_E_PARAM, "null argument");
    ltf_execute(L, d_raw, d_prof, d_hits, nullptr);
  });
}

int dspx_ltf_execute_masked(dspx_ltf* L, const void* d_raw, const uint8_t* d_mask, uint32_t M, uint64_t nblk,
                            float* d_prof, uint32_t* d_hits)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L && d_raw && d_mask && d_prof, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(M >= 2, DSPX_E_PARAM, "M=%u < 2", M);
    DSPX_REQUIRE(nblk * M <= L->ndat_out, DSPX_E_PARAM, "%llu blocks of %u exceed the %llu output samples",
                 (unsigned long long) nblk, M, (unsigned long long) L->ndat_out);
    DSPX_REQUIRE(L->ndat_out + L->d.nfft < (1ull << 31), DSPX_E_UNSUPPORTED,
                 "block of %llu output samples too long for the zap mask", (unsigned long long) L->ndat_out);
    if (L->d.form != DSPX_FORM_PLAIN && L->have_fold && L->npart)
      DSPX_REQUIRE(L->zh_ok, DSPX_E_UNSUPPORTED,
                   "fold histogram with per-channel hit counts does not fit in LDS (nbin=%u, %u slots)", L->d.nbin,
                   L->nslot);
    ZapArgs z{d_mask, M, uint32_t(nblk), L->d.nchan, ~0ull / M + 1};
    ltf_execute(L, d_raw, d_prof, d_hits, &z);
  });
}

int dspx_ltf_filterbank(dspx_ltf* L, const void* d_raw, float* d_out)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L && d_raw && d_out, DSPX_E_PARAM, "null argument");
    const auto& d = L->d;
    if (!L->npart) return;
    uint64_t nout = L->ndat_out;
    L->nstage = 0;
    L->nstage_open = false;
    if (d.form != DSPX_FORM_PLAIN) {
      uint32_t np = uint32_t(L->npart);
      with_loader(L, d_raw, [&](auto& ld) { filterbank_exec(L, ld, np, d_out); });
    } else {
      uint64_t nin = d.ndat;
      L->unpacked.reserve(size_t(d.nchan) * d.npol * nin * sizeof(float2));
      unpack(ctx, d_raw, d.nbit, d.offset_binary, d.nchan, d.npol, nin, d.scale, L->unpacked.as<float>(),
             uint64_t(d.npol) * nin * 2, nin * 2);
      L->plain.perform(L->unpacked.as<float>(), uint64_t(d.npol) * nin * 2, nin * 2, d_out,
                       uint64_t(d.npol) * nout * 2, nout * 2, uint32_t(L->npart));
    }
  });
}

// ------------------------------------------------------------ stream mode
// InputBuffering (InputBuffering.C:66-180, TimeSeries::prepend,
// TimeSeries.C:514-533): each call's input is the carried tail of the
// previous call followed by d.ndat new samples; the parts that fit are
// processed and the rest (the overlap and any remainder, fewer than
// overlap + step samples) is carried on.  The input is assembled
// channel-major in the transpose buffer (the carry copied to the head of each
// channel's row, the new block transposed after it), so stage A reads one
// contiguous row per channel as in a single block.

int dspx_ltf_stream_reset(dspx_ltf* L)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L, DSPX_E_PARAM, "null argument");
    const auto& d = L->d;
    DSPX_REQUIRE(d.form != DSPX_FORM_PLAIN, DSPX_E_UNSUPPORTED, "stream mode runs the fused form");
    DSPX_REQUIRE(d.ndat >= d.nfft, DSPX_E_PARAM, "stream blocks of %llu samples shorter than nfft=%u",
                 (unsigned long long) d.ndat, d.nfft);
    const uint64_t nov = uint64_t(d.nfilt_pos) + d.nfilt_neg;
    L->carry_cap = nov + L->step;
    L->cap = d.ndat + L->carry_cap;
    const uint64_t npmax = (L->cap - nov) / L->step;
    DSPX_REQUIRE(npmax < (1u << 20), DSPX_E_UNSUPPORTED, "npart too large");
    L->core.reserve_work(uint32_t(npmax));
    const uint64_t wbytes = uint64_t(d.npol) * 2 * d.nbit / 8;
    L->sc().ftp.reserve(L->cap * d.nchan * wbytes);
    L->carry.reserve(L->carry_cap * d.nchan * wbytes);
    L->streaming = true;
    L->ncarry = 0;
    L->consumed = 0;
    L->npart = 0;
    L->ndat_out = 0;
    L->have_fold = false;
  });
}

int dspx_ltf_stream_next(dspx_ltf* L, uint64_t* npart, uint64_t* ndat_out, uint64_t* t_first)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L && npart && ndat_out && t_first, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(L->streaming, DSPX_E_STATE, "not streaming (dspx_ltf_stream_reset)");
    const auto& d = L->d;
    const uint64_t nov = uint64_t(d.nfilt_pos) + d.nfilt_neg;
    const uint64_t ntot = L->ncarry + d.ndat;
    L->npart = ntot >= d.nfft ? (ntot - nov) / L->step : 0;
    L->ndat_out = L->npart * L->step;
    L->have_fold = false;  // the fold plan is per call
    *npart = L->npart;
    *ndat_out = L->ndat_out;
    *t_first = L->consumed + d.nfilt_pos;  // Convolution.C:611
  });
}

}  // extern "C"

namespace {

// Assembles carry + block, runs f(loader) for stage A, then (after the caller's
// remaining stages, run inside f's scope by the caller) updates the carry.
template <class F> void stream_call(dspx_ltf* L, const void* raw, F&& f)
{
  dspx_ctx* ctx = L->ctx;
  const auto& d = L->d;
  DSPX_REQUIRE(L->streaming, DSPX_E_STATE, "not streaming (dspx_ltf_stream_reset)");
  const uint64_t nov = uint64_t(d.nfilt_pos) + d.nfilt_neg;
  const uint64_t ntot = L->ncarry + d.ndat;
  const uint64_t np = ntot >= d.nfft ? (ntot - nov) / L->step : 0;
  DSPX_REQUIRE(np == L->npart, DSPX_E_STATE, "call dspx_ltf_stream_next before each stream call");
  const uint32_t wbytes = uint32_t(uint64_t(d.npol) * 2 * d.nbit / 8);
  void* ftp = L->sc().ftp.p;
  copy_chan(ctx, L->carry.p, L->carry_cap, 0, ftp, L->cap, 0, L->ncarry, d.nchan, wbytes);
  if (d.nchan > 1) {
    tfp_to_ftp(ctx, raw, d.ndat, d.nchan, wbytes, ftp, L->cap, L->ncarry);
  } else {
    DSPX_HIP(hipMemcpyAsync(static_cast<char*>(ftp) + L->ncarry * wbytes, raw, d.ndat * wbytes,
                            hipMemcpyDeviceToDevice, ctx->stream));
  }
  const uint64_t tstride = uint64_t(d.npol) * 2, cstride = L->cap * d.npol * 2;  // in elements
  if (d.nbit == 8) {
    LoadI8 ld{static_cast<const uint8_t*>(ftp), tstride, cstride, d.npol, d.offset_binary, d.scale};
    f(ld, uint32_t(np));
  } else {
    LoadI16 ld{static_cast<const int16_t*>(ftp), tstride, cstride, d.npol, d.offset_binary};
    f(ld, uint32_t(np));
  }
  const uint64_t done = np * L->step;
  const uint64_t rem = ntot - done;
  DSPX_REQUIRE(rem <= L->carry_cap, DSPX_E_INTERNAL, "carry of %llu samples", (unsigned long long) rem);
  copy_chan(ctx, ftp, L->cap, done, L->carry.p, L->carry_cap, 0, rem, d.nchan, wbytes);
  L->ncarry = rem;
  L->consumed += done;
}

}  // namespace

extern "C" {

int dspx_ltf_stream_execute(dspx_ltf* L, const void* d_raw, float* d_prof, uint32_t* d_hits)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L && d_raw && d_prof, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(L->npart == 0 || L->have_fold, DSPX_E_STATE, "no fold plan for this call: dspx_ltf_set_fold");
    L->nstage = 0;
    L->nstage_open = false;
    int ev = 0;
    stream_call(L, d_raw, [&](auto& ld, uint32_t np) {
      if (!np) return;
      ltf_record(L, ev++);
      fold_exec(L, ld, np, d_prof, d_hits, ev);
    });
  });
}

int dspx_ltf_stream_filterbank(dspx_ltf* L, const void* d_raw, float* d_out)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L && d_raw && d_out, DSPX_E_PARAM, "null argument");
    L->nstage = 0;
    L->nstage_open = false;
    stream_call(L, d_raw, [&](auto& ld, uint32_t np) {
      if (!np) return;
      filterbank_exec(L, ld, np, d_out);
    });
  });
}

int dspx_ltf_set_timing(dspx_ltf* L, int enable)
{
  if (!L) return DSPX_E_PARAM;
  L->timing = enable != 0;
  return DSPX_OK;
}

int dspx_ltf_stage_times(dspx_ltf* L, float* ms, const char** names, int cap, int* n)
{
  dspx_ctx* ctx = L ? L->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(L && n, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(L->timing, DSPX_E_STATE, "timing not enabled");
    DSPX_HIP(hipEventSynchronize(ctx->ev[L->nstage]));
    int k = std::min(cap, L->nstage);
    for (int i = 0; i < k; i++) {
      float t = 0;
      DSPX_HIP(hipEventElapsedTime(&t, ctx->ev[i], ctx->ev[i + 1]));
      if (ms) ms[i] = t;
      if (names) names[i] = L->stage_name[i];
    }
    *n = L->nstage;
  });
}

}  // extern "C"

// ------------------------------------------------------------------- batch
// Many beams of uneven shape (channels, DM, FFT length, 8/16-bit input,
// fold bins) through the fused path: one LoadToFold plan per beam, one set of
// large scratch buffers shared by all of them (sized for the largest beam;
// the plans run one after another on the context's stream), the sequence
// captured once into a HIP graph and replayed.  The FFT length of a beam can
// be chosen by timing candidate lengths (autotune_nfft) instead of the
// reference's cost model (optimize_fft.c:64-135).

struct dspx_batch {
  CtxRef hold;
  dspx_ctx* ctx = nullptr;
  LtfScratch sc;
  std::vector<dspx_ltf*> beams;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf tune_raw, tune_prof, tune_hits;
  // Branches: beams of branch k > 0 run on a context (and stream) of their
  // own with scratch of their own; an execute forks from the batch stream
  // and joins it again (one graph with parallel branches when captured).
  std::vector<dspx_ctx*> branch_ctx;  // [0] = ctx (not owned)
  std::vector<std::unique_ptr<LtfScratch>> branch_sc;
  std::vector<hipEvent_t> branch_ev;  // [0] fork, [k] join of branch k

  ~dspx_batch()
  {
    if (exec) (void) hipGraphExecDestroy(exec);
    if (graph) (void) hipGraphDestroy(graph);
    for (auto* L : beams) delete L;
    for (auto e : branch_ev)
      if (e) (void) hipEventDestroy(e);
    for (size_t k = 1; k < branch_ctx.size(); k++) dspx_ctx_destroy(branch_ctx[k]);
  }
};

namespace {

// The descriptor with its overlap filled in for FFT length nfft (0: the
// reference's rule, Response.C:553-588, raised to at least the one-pass
// form's shortest transform).
dspx_ltf_desc with_geometry(const dspx_ltf_desc& d0, uint32_t nfft)
{
  dspx_dedisp_desc dd = d0.dedisp;
  dd.nfft = nfft;
  dspx_geometry g{};
  int rc = geometry(&dd, &g);
  DSPX_REQUIRE(rc == DSPX_OK, rc, "geometry of a beam (nfft=%u) failed", nfft);
  if (nfft == 0 && g.nfft < (1u << kOnePassMinLog2)) {
    // below the shortest transform built: the next supported length (as a
    // user-set length would be, Response::set_frequency_resolution,
    // Response.C:74-80, 242-250)
    dd.nfft = 1u << kOnePassMinLog2;
    rc = geometry(&dd, &g);
    DSPX_REQUIRE(rc == DSPX_OK, rc, "geometry of a beam (nfft=%u) failed", dd.nfft);
  }
  dspx_ltf_desc d = d0;
  d.nfft = g.nfft;
  d.nfilt_pos = g.impulse_pos;
  d.nfilt_neg = g.impulse_neg;
  return d;
}

// Seconds per output sample of the fused LoadToFold at FFT length nfft: the
// fastest of three timed executions of the beam's whole block, zero-filled
// (the kernels' run time does not depend on the sample values), so that
// grids, tails and cache footprints are those of the replay.  (Timing a
// 2^20-sample slice instead let lengths within a few per cent of each other
// trade places from run to run.)
double time_per_sample(dspx_batch* B, const dspx_ltf_desc& dn, double* block_s = nullptr)
{
  dspx_ctx* ctx = B->ctx;
  dspx_ltf_desc d = dn;
  if (d.ndat < d.nfft) return 1e30;
  std::unique_ptr<dspx_ltf> L(ltf_create(ctx, &d, &B->sc));
  if (!L->npart) return 1e30;
  dspx_fold_chunk ch{0, L->ndat_out, 0.0, 0, 0};
  L->fold.build(ctx, L->ndat_out, 1.0 / 97.0e3, d.nbin, &ch, 1, false);
  L->have_fold = true;
  fused_fold_setup(L.get());
  ltf_reserve(L.get());
  const uint64_t raw_bytes = d.ndat * d.nchan * d.npol * 2 * (d.nbit / 8);
  B->tune_raw.reserve(raw_bytes);
  DSPX_HIP(hipMemsetAsync(B->tune_raw.p, 0, raw_bytes, ctx->stream));
  B->tune_prof.reserve(size_t(d.nchan) * L->nval * d.nbin * sizeof(float));
  B->tune_hits.reserve(size_t(d.nbin) * sizeof(uint32_t));
  auto run = [&] {
    int rc = dspx_ltf_execute(L.get(), B->tune_raw.p, B->tune_prof.as<float>(), B->tune_hits.as<uint32_t>());
    DSPX_REQUIRE(rc == DSPX_OK, rc, "autotune execute: %s", ctx->last_error.c_str());
  };
  run();
  for (int i = 0; i < 3; i++) {
    DSPX_HIP(hipEventRecord(ctx->ev[i], ctx->stream));
    run();
  }
  DSPX_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
  DSPX_HIP(hipEventSynchronize(ctx->ev[3]));
  float best = 1e30f;
  for (int i = 0; i < 3; i++) {
    float ms = 0;
    DSPX_HIP(hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]));
    best = std::min(best, ms);
  }
  if (block_s) *block_s = double(best) * 1e-3;
  return double(best) * 1e-3 / double(L->ndat_out);
}

// Candidate lengths: the reference's choice and the next three powers of two
// (longer transforms waste less on the overlap), and, where the reference's
// choice runs in one pass, the shorter one-pass lengths whose step is at
// least the overlap (a 256-point window runs about 1.7 times as many samples
// per second through the one-pass kernel as a 512-point one: two passes per
// transform against three), within the fused forms' range; each runs in the
// form the plan picks for it (one pass where it fits).
// block_s: the block's time at the chosen length (0 if none was timed).
uint32_t autotune_nfft(dspx_batch* B, const dspx_ltf_desc& d0, double& block_s)
{
  dspx_ltf_desc ref = with_geometry(d0, 0);
  uint32_t best = ref.nfft;
  double best_t = 1e30;
  block_s = 0;
  const uint64_t nov = uint64_t(ref.nfilt_pos) + ref.nfilt_neg;
  uint32_t lo = ref.nfft;
  while (lo <= (1u << kOnePassMaxLog2) && lo / 2 >= (1u << kOnePassMinLog2) && lo / 2 >= 2 * nov) lo /= 2;
  for (uint32_t n = lo; n <= (1u << 22) && n <= 8 * ref.nfft; n *= 2) {
    dspx_ltf_desc d = with_geometry(d0, n);
    if (d.nfilt_pos + d.nfilt_neg >= n || d.ndat < n) continue;
    double bs = 0;
    double t = time_per_sample(B, d, &bs);
    if (t < best_t) {
      best_t = t;
      best = n;
      block_s = bs;
    }
  }
  return best;
}

}  // namespace

extern "C" {

int dspx_batch_create(dspx_ctx* ctx, const dspx_ltf_desc* beams, uint32_t nbeam, uint32_t flags, dspx_batch** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && beams && out && nbeam, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE((flags & ~uint32_t(DSPX_BATCH_AUTOTUNE | DSPX_BATCH_BRANCH_MASK)) == 0, DSPX_E_PARAM, "flags=%u",
                 flags);
    DSPX_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<dspx_batch> B(new dspx_batch);
    B->hold.reset(ctx);
    B->ctx = ctx;
    uint32_t nbranch = (flags & DSPX_BATCH_BRANCH_MASK) >> DSPX_BATCH_BRANCH_SHIFT;
    if (!nbranch) nbranch = 1;
    nbranch = std::min(nbranch, nbeam);
    std::vector<dspx_ltf_desc> descs(nbeam);
    std::vector<double> timed(nbeam, 0.0);  // the autotuner's time of each beam's block
    for (uint32_t i = 0; i < nbeam; i++) {
      const dspx_ltf_desc& d0 = beams[i];
      DSPX_REQUIRE(d0.form != DSPX_FORM_PLAIN, DSPX_E_PARAM, "beam %u: the batch runs the fused forms", i);
      dspx_ltf_desc d = d0;
      if (d0.nfft == 0) {
        uint32_t n = (flags & DSPX_BATCH_AUTOTUNE) ? autotune_nfft(B.get(), d0, timed[i]) : 0;
        d = with_geometry(d0, n);
      }
      descs[i] = d;
    }
    // branches: the main context, then contexts with streams of their own
    B->branch_ctx.push_back(ctx);
    B->branch_sc.emplace_back(nullptr);
    for (uint32_t k = 1; k < nbranch; k++) {
      dspx_ctx* c = nullptr;
      int rc = dspx_ctx_create(ctx->device, nullptr, &c);
      DSPX_REQUIRE(rc == DSPX_OK, rc, "branch context %u", k);
      B->branch_ctx.push_back(c);
      B->branch_sc.emplace_back(new LtfScratch);
    }
    for (uint32_t k = 0; k <= nbranch && nbranch > 1; k++) {
      hipEvent_t e = nullptr;
      DSPX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      B->branch_ev.push_back(e);
    }
    // beams to branches, longest first to the least loaded: a beam's cost is
    // its block's time as the autotuner measured it where every beam was
    // timed, else its samples times log2 of its length
    std::vector<uint32_t> order(nbeam), branch(nbeam, 0);
    std::vector<double> cost(nbeam), load(nbranch, 0.0);
    const bool all_timed = std::all_of(timed.begin(), timed.end(), [](double t) { return t > 0; });
    for (uint32_t i = 0; i < nbeam; i++) {
      order[i] = i;
      cost[i] = all_timed ? timed[i]
                          : double(descs[i].ndat) * descs[i].nchan * descs[i].npol * std::log2(double(descs[i].nfft));
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    for (uint32_t i : order) {
      const uint32_t k = uint32_t(std::min_element(load.begin(), load.end()) - load.begin());
      branch[i] = k;
      load[k] += cost[i];
    }
    for (uint32_t i = 0; i < nbeam; i++) {
      const uint32_t k = branch[i];
      LtfScratch* sc = k ? B->branch_sc[k].get() : &B->sc;
      B->beams.push_back(ltf_create(B->branch_ctx[k], &descs[i], sc));
    }
    B->tune_raw.release();
    B->tune_prof.release();
    B->tune_hits.release();
    *out = B.release();
  });
}

void dspx_batch_destroy(dspx_batch* B)
{
  if (!B) return;
  (void) hipSetDevice(B->ctx->device);
  for (auto* c : B->branch_ctx) (void) hipStreamSynchronize(c->stream);
  delete B;
}

int dspx_batch_beam(const dspx_batch* B, uint32_t ibeam, dspx_ltf_desc* desc, uint64_t* npart, uint64_t* ndat_out)
{
  if (!B || ibeam >= B->beams.size()) return DSPX_E_PARAM;
  const dspx_ltf* L = B->beams[ibeam];
  if (desc) {
    *desc = L->d;
    desc->form = L->onepass ? DSPX_FORM_ONEPASS : DSPX_FORM_TWOLEVEL;  // the form the plan runs
  }
  if (npart) *npart = L->npart;
  if (ndat_out) *ndat_out = L->ndat_out;
  return DSPX_OK;
}

int dspx_batch_set_fold(dspx_batch* B, uint32_t ibeam, double pps, const dspx_fold_chunk* chunks, uint32_t nchunk)
{
  dspx_ctx* ctx = B ? B->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(B && ibeam < B->beams.size(), DSPX_E_PARAM, "beam %u", ibeam);
    DSPX_REQUIRE(!B->exec, DSPX_E_STATE, "fold plans are fixed once the batch is captured");
    int rc = dspx_ltf_set_fold(B->beams[ibeam], pps, chunks, nchunk);
    DSPX_REQUIRE(rc == DSPX_OK, rc, "%s", B->beams[ibeam]->ctx->last_error.c_str());
    ltf_reserve(B->beams[ibeam]);
  });
}

}  // extern "C"

namespace {

// Every beam's execute: on the batch stream, or forked into the branches'
// streams after the work already on the batch stream and joined back.
int batch_run(dspx_batch* B, const void* const* d_raw, float* const* d_prof, uint32_t* const* d_hits,
              std::string& err)
{
  const size_t nb = B->branch_ctx.size();
  if (nb > 1) {
    DSPX_HIP(hipEventRecord(B->branch_ev[0], B->ctx->stream));
    for (size_t k = 1; k < nb; k++) DSPX_HIP(hipStreamWaitEvent(B->branch_ctx[k]->stream, B->branch_ev[0], 0));
  }
  int rc = DSPX_OK;
  for (size_t i = 0; i < B->beams.size() && rc == DSPX_OK; i++) {
    rc = dspx_ltf_execute(B->beams[i], d_raw[i], d_prof[i], d_hits ? d_hits[i] : nullptr);
    if (rc != DSPX_OK) err = "beam " + std::to_string(i) + ": " + B->beams[i]->ctx->last_error;
  }
  if (nb > 1) {
    // joined even after a failure, so that a capture can end
    for (size_t k = 1; k < nb; k++) {
      DSPX_HIP(hipEventRecord(B->branch_ev[k], B->branch_ctx[k]->stream));
      DSPX_HIP(hipStreamWaitEvent(B->ctx->stream, B->branch_ev[k], 0));
    }
  }
  return rc;
}

}  // namespace

extern "C" {

int dspx_batch_execute(dspx_batch* B, const void* const* d_raw, float* const* d_prof, uint32_t* const* d_hits)
{
  dspx_ctx* ctx = B ? B->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(B && d_raw && d_prof, DSPX_E_PARAM, "null argument");
    std::string err;
    int rc = batch_run(B, d_raw, d_prof, d_hits, err);
    DSPX_REQUIRE(rc == DSPX_OK, rc, "%s", err.c_str());
  });
}

int dspx_batch_capture(dspx_batch* B, const void* const* d_raw, float* const* d_prof, uint32_t* const* d_hits)
{
  dspx_ctx* ctx = B ? B->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(B && d_raw && d_prof, DSPX_E_PARAM, "null argument");
    for (size_t i = 0; i < B->beams.size(); i++) {
      DSPX_REQUIRE(B->beams[i]->have_fold, DSPX_E_STATE, "beam %zu has no fold plan", i);
      DSPX_REQUIRE(!B->beams[i]->timing, DSPX_E_STATE, "timing is not captured");
      ltf_reserve(B->beams[i]);
    }
    if (B->exec) {
      DSPX_HIP(hipGraphExecDestroy(B->exec));
      B->exec = nullptr;
    }
    if (B->graph) {
      DSPX_HIP(hipGraphDestroy(B->graph));
      B->graph = nullptr;
    }
    for (auto* c : B->branch_ctx) DSPX_HIP(hipStreamSynchronize(c->stream));
    DSPX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    std::string err;
    int rc = batch_run(B, d_raw, d_prof, d_hits, err);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    DSPX_REQUIRE(rc == DSPX_OK, rc, "capture: %s", err.c_str());
    DSPX_HIP(e);
    B->graph = g;
    DSPX_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  });
}

int dspx_batch_replay(dspx_batch* B)
{
  dspx_ctx* ctx = B ? B->ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(B, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(B->exec, DSPX_E_STATE, "nothing captured (dspx_batch_capture)");
    DSPX_HIP(hipGraphLaunch(B->exec, ctx->stream));
  });
}

}  // extern "C"

namespace dspx {

// The convolving filterbank's four-step form: stage A over the FPT float
// input (windows in_step = nsub * nkeep apart), the rows' forward transform,
// response product and sub-band backward transforms (k_stageB_fb), and stage
// C's backward column transform into the kept samples of each sub-band.
struct Fb4 {
  ConvCore core;
  uint32_t nsub = 0, nkeep = 0, F2 = 0;
  uint64_t in_step = 0;
};

Fb4* fb4_create(dspx_ctx* ctx, const dspx_fb_desc& d)
{
  const uint64_t n = uint64_t(d.nsub) * d.fres;
  uint32_t N1 = 0, N2 = 0, m = 0;
  if (n > (1u << 22) || !split_n(uint32_t(n), N1, N2, m) || m < 13 || d.fres % N1) return nullptr;
  if (!fb4_rows_ok(N2, d.fres / N1)) return nullptr;
  std::unique_ptr<Fb4> f(new Fb4);
  f->core.init(ctx, d.nchan, d.npol, uint32_t(n), 0, 0, DSPX_CHIRP_TABLE, dspx_dedisp_desc{});
  f->nsub = d.nsub;
  f->nkeep = d.fres - d.nfilt_pos - d.nfilt_neg;
  f->F2 = d.fres / N1;
  f->in_step = uint64_t(d.nsub) * f->nkeep;
  DSPX_REQUIRE(f->in_step < (1ull << 32), DSPX_E_UNSUPPORTED, "window step");
  f->core.step = uint32_t(f->in_step);  // stage A's window starts
  f->core.nfilt_pos = d.nfilt_pos;      // stage C's first kept sample of a sub-band
  return f.release();
}

void fb4_destroy(Fb4* f) { delete f; }

void fb4_set_response(Fb4* f, const float* host) { f->core.set_table(host); }

void fb4_perform(Fb4* f, const float* d_in, uint64_t ics, uint64_t ips, uint64_t npart, float* d_out, uint64_t ocs,
                 uint64_t ops, float* d_zdm)
{
  ConvCore& k = f->core;
  DSPX_REQUIRE(k.have_table, DSPX_E_STATE, "no response: call dspx_fb_set_response first");
  // parts per pass: a work buffer of at most 2 GiB
  const uint64_t per_part = uint64_t(k.nchan) * k.npol * k.N * sizeof(float2);
  const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(npart, (2ull << 30) / per_part));
  for (uint64_t p0 = 0; p0 < npart; p0 += chunk) {
    const uint32_t np = uint32_t(std::min(chunk, npart - p0));
    k.reserve_work(np);
    LoadF32 ld{d_in + 2 * p0 * f->in_step, ics, ips, k.npol};
    k.stage_a(ld, np);
    // rows per workgroup as stage B of the convolution: parts grouped so
    // that >= 2048 workgroups are in flight where the rows alone are fewer
    const uint64_t rows = uint64_t(k.N1) * k.nchan;
    uint32_t npg = rows >= 2048 ? 1u : uint32_t(std::min<uint64_t>(np, (2048 + rows - 1) / rows));
    const uint32_t ppg = (np + npg - 1) / npg;
    npg = (np + ppg - 1) / ppg;
    const uint64_t nb = rows * npg;
    DSPX_REQUIRE(nb < (1ull << 31), DSPX_E_UNSUPPORTED, "grid too large");
    // the dedispersed output, then (stage B overwrote stage A's rows) the
    // zero-DM output from the windows transformed again, with a null table
    // (a unit response) in stage B
    for (int pass = 0; pass < (d_zdm ? 2 : 1); pass++) {
      if (pass) k.stage_a(ld, np);
      ChirpArgs c = k.chirp();
      if (pass) c.table = nullptr;
      fb4_stage_b(k.N2, f->F2, unsigned(nb), k.ctx->stream, k.shape(np), c, twiddles(k.ctx, int(k.N2)),
                  twiddles(k.ctx, int(f->F2)), ppg, f->nsub);
      FilterbankArgs fb{(pass ? d_zdm : d_out) + 2 * p0 * f->nkeep, ocs, ops};
      fb.f2 = f->F2;
      fb.lf2 = uint32_t(__builtin_ctz(f->F2));
      fb.nsub = f->nsub;
      fb.nkeep = f->nkeep;
      k.stage_c<kStageCFilterbank>(np, fb, FoldArgs{});
    }
  }
}

}  // namespace dspx
