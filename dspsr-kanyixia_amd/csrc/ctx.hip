// ctx.hip -- context, twiddle tables, unpack and detection engines.
#include <cmath>
#include <cstring>

#include "dspx_internal.hpp"
#include "stages.hpp"

namespace dspx {

void fail(int code, const char* fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error{code, std::string(buf)};
}

// ------------------------------------------------------------------ unpack
// GenericEightBitUnpackerCUDA.cu:23-45 / SixteenBitUnpacker.C:56-78.
// One thread per (time sample, channel): both pols of a TFP word, written as
// FPT float2 (consecutive threads -> consecutive time samples of one channel).
template <class LD>
__global__ void k_unpack(LD ld, uint64_t ndat, float* __restrict__ out, uint64_t chan_stride,
                         uint64_t pol_stride, uint32_t npol)
{
  uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t chan = blockIdx.y;
  if (t >= ndat) return;
  float2 a, b;
  ld(chan, t, a, b);
  float* o = out + uint64_t(chan) * chan_stride + 2 * t;
  *reinterpret_cast<float2*>(o) = a;
  if (npol == 2) *reinterpret_cast<float2*>(o + pol_stride) = b;
}

// Multichannel input: a 32-channel x 64-sample tile is read with consecutive
// threads on consecutive channels (the TFP word order) and written with
// consecutive threads on consecutive samples (the FPT order), through LDS.
template <class LD>
__global__ __launch_bounds__(256) void k_unpack_tile(LD ld, uint32_t nchan, uint64_t ndat, float* __restrict__ out,
                                                     uint64_t chan_stride, uint64_t pol_stride, uint32_t npol)
{
  __shared__ typename LD::Word tile[32][65];
  const uint32_t c0 = blockIdx.y * 32;
  const uint64_t t0 = uint64_t(blockIdx.x) * 64;
  const int tid = threadIdx.x;
  const uint32_t cl = c0 + (tid & 31);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int tt = (tid >> 5) + 8 * i;
    if (cl < nchan && t0 + tt < ndat) tile[tid & 31][tt] = ld.word(cl, t0 + tt);
  }
  __syncthreads();
  const int tt = tid & 63;
  const uint64_t t = t0 + tt;
  if (t >= ndat) return;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int cc = (tid >> 6) + 4 * i;
    if (c0 + cc >= nchan) break;
    typename LD::Word w = tile[cc][tt];
    float* o = out + uint64_t(c0 + cc) * chan_stride + 2 * t;
    *reinterpret_cast<float2*>(o) = ld.pol(w, 0);
    if (npol == 2) *reinterpret_cast<float2*>(o + pol_stride) = ld.pol(w, 1);
  }
}

template <class LD>
static void launch_unpack(dspx_ctx* ctx, const LD& ld, uint32_t nchan, uint32_t npol, uint64_t ndat, float* out,
                          uint64_t cs, uint64_t ps)
{
  if (nchan >= 8) {
    dim3 grid(unsigned((ndat + 63) / 64), (nchan + 31) / 32);
    k_unpack_tile<<<grid, 256, 0, ctx->stream>>>(ld, nchan, ndat, out, cs, ps, npol);
  } else {
    dim3 grid(unsigned((ndat + 255) / 256), nchan);
    k_unpack<<<grid, 256, 0, ctx->stream>>>(ld, ndat, out, cs, ps, npol);
  }
}

void unpack(dspx_ctx* ctx, const void* raw, uint32_t nbit, uint32_t offset_binary, uint32_t nchan,
            uint32_t npol, uint64_t ndat, float scale, float* out, uint64_t cs, uint64_t ps)
{
  DSPX_REQUIRE(npol == 1 || npol == 2, DSPX_E_PARAM, "npol=%u", npol);
  DSPX_REQUIRE(nchan >= 1 && nchan <= 65535, DSPX_E_PARAM, "nchan=%u", nchan);
  if (!ndat) return;
  DSPX_REQUIRE((ndat + 63) / 64 < (1ull << 31), DSPX_E_UNSUPPORTED, "ndat=%llu too large", (unsigned long long) ndat);
  if (nbit == 8) {
    LoadI8 ld{static_cast<const uint8_t*>(raw), uint64_t(nchan) * npol * 2, uint64_t(npol) * 2, npol, offset_binary,
              scale};
    launch_unpack(ctx, ld, nchan, npol, ndat, out, cs, ps);
  } else if (nbit == 16) {
    LoadI16 ld{static_cast<const int16_t*>(raw), uint64_t(nchan) * npol * 2, uint64_t(npol) * 2, npol, offset_binary};
    launch_unpack(ctx, ld, nchan, npol, ndat, out, cs, ps);
  } else {
    fail(DSPX_E_PARAM, "nbit=%u (8 or 16 supported)", nbit);
  }
  DSPX_HIP(hipGetLastError());
}

// --------------------------------------------------------------- detection
__global__ void k_detect(uint32_t state, uint32_t ndim, const float* __restrict__ in, uint64_t ics,
                         uint64_t ips, uint32_t npol, uint64_t ndat, float* __restrict__ out,
                         uint64_t ocs, uint64_t ops)
{
  uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t chan = blockIdx.y;
  if (t >= ndat) return;
  const float* x = in + uint64_t(chan) * ics + 2 * t;
  float2 a = *reinterpret_cast<const float2*>(x);
  float2 b = npol == 2 ? *reinterpret_cast<const float2*>(x + ips) : make_float2(0.f, 0.f);
  float d[4];
  int nv = detect_values(state, npol, a, b, d);
  float* o = out + uint64_t(chan) * ocs;
  if (state == DSPX_INTENSITY) {
    o[t] = d[0];
  } else if (state == DSPX_PPQQ) {
    o[t] = d[0];
    o[ops + t] = d[1];
  } else {
    for (int i = 0; i < nv; i++) o[(i / ndim) * ops + t * ndim + (i % ndim)] = d[i];
  }
}

void detect(dspx_ctx* ctx, uint32_t state, uint32_t ndim, const float* in, uint64_t ics, uint64_t ips,
            uint32_t nchan, uint32_t npol, uint64_t ndat, float* out, uint64_t ocs, uint64_t ops)
{
  DSPX_REQUIRE(state <= DSPX_STOKES, DSPX_E_PARAM, "state=%u", state);
  DSPX_REQUIRE(npol == 1 || npol == 2, DSPX_E_PARAM, "npol=%u", npol);
  DSPX_REQUIRE(state <= DSPX_PPQQ || npol == 2, DSPX_E_STATE,
               "invalid npol=%u for Coherence/Stokes formation", npol);  // Detection.C:626-631
  DSPX_REQUIRE(state <= DSPX_PPQQ || ndim == 1 || ndim == 2 || ndim == 4, DSPX_E_STATE,
               "invalid ndim=%u", ndim);  // Detection.C:640-651
  DSPX_REQUIRE(nchan >= 1 && nchan <= 65535, DSPX_E_PARAM, "nchan=%u", nchan);
  if (!ndat) return;
  dim3 grid(unsigned((ndat + 255) / 256), nchan);
  k_detect<<<grid, 256, 0, ctx->stream>>>(state, ndim, in, ics, ips, npol, ndat, out, ocs, ops);
  DSPX_HIP(hipGetLastError());
}

void ctx_free(dspx_ctx* c)
{
  (void) hipSetDevice(c->device);
  if (c->stream) (void) hipStreamSynchronize(c->stream);
  for (auto& e : c->ev)
    if (e) (void) hipEventDestroy(e);
  if (c->d_tw) (void) hipFree(c->d_tw);
  if (c->own_stream && c->stream) (void) hipStreamDestroy(c->stream);
  delete c;
}

}  // namespace dspx

using namespace dspx;

extern "C" {

const char* dspx_version(void) { return "dspx 0.1 gfx950"; }

int dspx_ctx_create(int device, void* stream, dspx_ctx** out)
{
  return guard(nullptr, [&] {
    DSPX_REQUIRE(out, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipSetDevice(device));
    auto* c = new dspx_ctx;
    c->device = device;
    try {
      DSPX_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device));
      if (stream) {
        c->stream = static_cast<hipStream_t>(stream);
      } else {
        DSPX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
      }
      // exp(-2 pi i m / L) for L = 2, 4, ..., 16384 at offset L, then the
      // power records of each span (layout in fft_core.hpp), all in FP64
      // rounded once to float
      const int total = kTwTotal;
      std::vector<float2> h(total, make_float2(0.f, 0.f));
      auto w = [](int64_t m, int64_t L) {
        double a = -2.0 * M_PI * double(m % L) / double(L);
        return make_float2(float(std::cos(a)), float(std::sin(a)));
      };
      for (int L = 2; L <= kTwMaxL; L *= 2)
        for (int m = 0; m < L; m++) h[L + m] = w(m, L);
      for (int S = 2; S <= kTwMaxL; S *= 2)
        for (int k = 0; k < S / 2; k++)
          for (int p = 0; p < 4; p++) h[kTwPowBase + 4 * (S / 2 + k) + p] = w(int64_t(k) << p, S);
      DSPX_HIP(hipMalloc(&c->d_tw, total * sizeof(float2)));
      DSPX_HIP(hipMemcpy(c->d_tw, h.data(), total * sizeof(float2), hipMemcpyHostToDevice));
      for (auto& e : c->ev) DSPX_HIP(hipEventCreate(&e));
      fold_static_setup(device);
    } catch (...) {
      ctx_free(c);
      throw;
    }
    *out = c;
  });
}

void dspx_ctx_destroy(dspx_ctx* c)
{
  if (!c) return;
  // the work queued so far completes before the call returns; the context
  // itself lives on while engines made on it hold references
  (void) hipSetDevice(c->device);
  if (c->stream) (void) hipStreamSynchronize(c->stream);
  CtxRef::drop(c);
}

const char* dspx_last_error(const dspx_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

int dspx_ctx_synchronize(dspx_ctx* c)
{
  return guard(c, [&] {
    DSPX_REQUIRE(c, DSPX_E_PARAM, "null context");
    DSPX_HIP(hipStreamSynchronize(c->stream));
  });
}

void* dspx_ctx_stream(dspx_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int dspx_dedisp_geometry(const dspx_dedisp_desc* d, dspx_geometry* g) { return geometry(d, g); }

int dspx_sample_delays(const dspx_dedisp_desc* d, int64_t* delay, double* frac, int64_t* applied,
                       int64_t* zero_delay, uint64_t* total_delay)
{
  return sample_delays(d, delay, frac, applied, zero_delay, total_delay);
}

int dspx_unpack(dspx_ctx* ctx, const void* d_raw, uint32_t nbit, uint32_t offset_binary, uint32_t nchan,
                uint32_t npol, uint64_t ndat, float scale, float* d_out, uint64_t cs, uint64_t ps)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && d_raw && d_out, DSPX_E_PARAM, "null argument");
    unpack(ctx, d_raw, nbit, offset_binary, nchan, npol, ndat, scale, d_out, cs, ps);
  });
}

int dspx_detect(dspx_ctx* ctx, uint32_t state, uint32_t ndim, const float* d_in, uint64_t ics, uint64_t ips,
                uint32_t nchan, uint32_t npol, uint64_t ndat, float* d_out, uint64_t ocs, uint64_t ops)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && d_in && d_out, DSPX_E_PARAM, "null argument");
    detect(ctx, state, ndim, d_in, ics, ips, nchan, npol, ndat, d_out, ocs, ops);
  });
}

}  // extern "C"
