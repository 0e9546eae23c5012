// delay.hip -- SampleDelay::Engine: per-(channel, pol) sample offsets.
//
// Reference: SampleDelay::transformation (Signal/General/SampleDelay.C:
// 246-300) and CUDA::SampleDelayEngine (Signal/General/SampleDelayCUDA.cu:
// 24-160): output sample i of (chan, pol) is input sample i + applied delay,
// for the output_ndat samples that every channel has.  A memory-bound copy:
// each thread moves one float4 (two complex samples) where the rows allow.
#include "dspx_internal.hpp"

namespace dspx {

// grid (ceil(nfloat / (256 * 4)), nchan * npol); nfloat = ndat_out * ndim
__global__ __launch_bounds__(256) void k_retard(const float* __restrict__ in, uint64_t ics, uint64_t ips,
                                                float* __restrict__ out, uint64_t ocs, uint64_t ops, uint32_t npol,
                                                uint32_t ndim, const int64_t* __restrict__ applied, uint64_t nfloat)
{
  const uint32_t cp = blockIdx.y;
  const uint32_t chan = cp / npol, pol = cp % npol;
  const float* x = in + chan * ics + pol * ips + uint64_t(applied[cp]) * ndim;
  float* y = out + chan * ocs + pol * ops;
  const uint64_t i = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
  // 16-byte moves where both rows allow them, else 8-byte ones (complex
  // rows start 8-byte aligned whatever the delay and the row length); the
  // alignment is uniform over the workgroup
  const uintptr_t al = reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y);
  if (i + 4 <= nfloat) {
    if ((al & 15) == 0) {
      *reinterpret_cast<float4*>(y + i) = *reinterpret_cast<const float4*>(x + i);
      return;
    }
    if ((al & 7) == 0) {
      const float2 a = *reinterpret_cast<const float2*>(x + i), b = *reinterpret_cast<const float2*>(x + i + 2);
      *reinterpret_cast<float2*>(y + i) = a;
      *reinterpret_cast<float2*>(y + i + 2) = b;
      return;
    }
  }
  for (uint64_t k = i; k < i + 4 && k < nfloat; k++) y[k] = x[k];
}

}  // namespace dspx

using namespace dspx;

struct dspx_delay {
  CtxRef hold;
  dspx_ctx* ctx = nullptr;
  uint32_t nchan = 0, npol = 0;
  DevBuf applied;  // int64 [nchan * npol]
};

extern "C" {

int dspx_delay_create(dspx_ctx* ctx, uint32_t nchan, uint32_t npol, const int64_t* applied, dspx_delay** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && applied && out, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(nchan >= 1 && npol >= 1 && uint64_t(nchan) * npol <= 65535, DSPX_E_PARAM, "nchan=%u npol=%u",
                 nchan, npol);
    for (uint64_t i = 0; i < uint64_t(nchan) * npol; i++)
      DSPX_REQUIRE(applied[i] >= 0, DSPX_E_STATE, "delay of channel %llu is negative",
                   (unsigned long long) (i / npol));  // SampleDelayCUDA.cu:100-104
    DSPX_HIP(hipSetDevice(ctx->device));
    auto* D = new dspx_delay;
    D->hold.reset(ctx);
    D->ctx = ctx;
    D->nchan = nchan;
    D->npol = npol;
    try {
      D->applied.reserve(size_t(nchan) * npol * sizeof(int64_t));
      DSPX_HIP(hipMemcpy(D->applied.p, applied, size_t(nchan) * npol * sizeof(int64_t), hipMemcpyHostToDevice));
    } catch (...) {
      delete D;
      throw;
    }
    *out = D;
  });
}

void dspx_delay_destroy(dspx_delay* D)
{
  if (!D) return;
  (void) hipSetDevice(D->ctx->device);
  (void) hipStreamSynchronize(D->ctx->stream);
  delete D;
}

int dspx_delay_retard(dspx_delay* D, const float* d_in, uint64_t ics, uint64_t ips, uint32_t ndim, uint64_t ndat_out,
                      float* d_out, uint64_t ocs, uint64_t ops)
{
  return guard(D ? D->ctx : nullptr, [&] {
    DSPX_REQUIRE(D && d_in && d_out, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(ndim >= 1 && ndim <= 4, DSPX_E_PARAM, "ndim=%u", ndim);
    const uint64_t nfloat = ndat_out * ndim;
    if (!nfloat) return;
    const uint64_t nb = (nfloat + 1023) / 1024;
    DSPX_REQUIRE(nb < (1ull << 31), DSPX_E_UNSUPPORTED, "block too long");
    dim3 grid(unsigned(nb), D->nchan * D->npol);
    k_retard<<<grid, 256, 0, D->ctx->stream>>>(d_in, ics, ips, d_out, ocs, ops, D->npol, ndim,
                                                D->applied.as<int64_t>(), nfloat);
    DSPX_HIP(hipGetLastError());
  });
}

}  // extern "C"
