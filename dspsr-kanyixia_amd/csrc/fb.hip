// fb.hip -- convolving filterbank engine: one kernel per window for
// 1024-4096-point windows, the four-step form (three kernels, conv.hip) for
// 2^13-2^22-point windows whose sub-bands hold at least 8*N1 bins, the
// FFT-library form otherwise.
//
// Reference: Filterbank::prepare / resize_output (Filterbank.C:75-200,
// 405-460) and FilterbankEngineCPU::perform (FilterbankEngineCPU.C:184-370):
// per input channel, pol and part, the nsub*fres-point forward transform of
// the window at ipart*in_step, the response product, the fres-point backward
// transform of each sub-band's contiguous slice, and the kept samples
// [nfilt_pos, nfilt_pos + nkeep) written to output channel ichan*nsub + isub.
// The common window lengths run as one fused kernel per window; the others
// are batched into library transforms, with the response product and the
// overlap discard as kernels of their own.
#include <cstdlib>

#include <hipfft/hipfft.h>

#include <algorithm>
#include <memory>
#include <string>

#include "dspx_internal.hpp"
#include "fft_core.hpp"

namespace dspx {

#define DSPX_FB_FFT(call)                                                                   \
  do {                                                                                      \
    hipfftResult r_ = (call);                                                               \
    if (r_ != HIPFFT_SUCCESS) ::dspx::fail(DSPX_E_HIP, "%s failed: hipfft %d", #call, int(r_)); \
  } while (0)

// Batch unit u = (chan * npol + pol) * npart + part; scratch row b = u - u0.
// grid (ceil(nfft/512), nb): two complex samples per thread, moved as 16
// bytes where the window and the scratch row both allow it (uniform over
// the workgroup), else as two 8-byte moves.
__global__ void k_fb_windows(const float* __restrict__ in, uint64_t ics, uint64_t ips, uint32_t npol,
                             uint64_t nfft, uint64_t in_step, uint32_t npart, uint64_t u0,
                             float2* __restrict__ scratch)
{
  const uint64_t k = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 2;
  if (k >= nfft) return;
  const uint64_t b = blockIdx.y, u = u0 + b;
  const uint64_t p = u % npart, cp = u / npart;
  const uint64_t chan = cp / npol, pol = cp % npol;
  const float* x = in + chan * ics + pol * ips + 2 * p * in_step;
  float2* y = scratch + b * nfft;
  const uintptr_t al = reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y);
  if (k + 2 <= nfft && (al & 15) == 0) {
    *reinterpret_cast<float4*>(y + k) = *reinterpret_cast<const float4*>(x + 2 * k);
  } else {
    y[k] = *reinterpret_cast<const float2*>(x + 2 * k);
    if (k + 1 < nfft) y[k + 1] = *reinterpret_cast<const float2*>(x + 2 * k + 2);
  }
}

// Response::operate (Response.C:653-726): d * f per bin, the response of the
// unit's input channel.
__global__ void k_fb_multiply(float2* __restrict__ scratch, const float2* __restrict__ resp, uint64_t nfft,
                              uint32_t npol, uint32_t npart, uint64_t u0)
{
  uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nfft) return;
  uint64_t b = blockIdx.y;
  uint64_t chan = ((u0 + b) / npart) / npol;
  float2 d = scratch[b * nfft + k];
  float2 f = resp[chan * nfft + k];
  scratch[b * nfft + k] = make_float2(f.x * d.x - f.y * d.y, f.y * d.x + f.x * d.y);
}

// grid (ceil(nsub * nkeep / 256), nb): x runs over (sub-band, kept sample)
// of one unit, so that short sub-bands still fill the workgroups
__global__ void k_fb_discard(const float2* __restrict__ scratch, uint64_t nfft, uint32_t fres, uint32_t nsub,
                             uint32_t npol, uint32_t npart, uint64_t u0, uint32_t nkeep, uint32_t nfilt_pos,
                             float* __restrict__ out, uint64_t ocs, uint64_t ops)
{
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nsub * nkeep) return;
  const uint32_t isub = idx / nkeep, m = idx % nkeep;
  const uint64_t b = blockIdx.y;
  uint64_t u = u0 + b;
  uint64_t p = u % npart, cp = u / npart;
  uint64_t chan = cp / npol, pol = cp % npol;
  float* o = out + (chan * nsub + isub) * ocs + pol * ops + 2 * (p * nkeep + m);
  *reinterpret_cast<float2*>(o) = scratch[b * nfft + uint64_t(isub) * fres + nfilt_pos + m];
}

// The whole filterbank of one unit in one workgroup of NT = NFFT/16 threads,
// 16 complex values per thread: the window loaded in the forward plan's
// first-pass order, the NFFT-point forward transform in registers and LDS,
// the response product on the way into the batched-column layout (bin k =
// s*FRES + m at column s, row m), NSUB = NFFT/FRES inverse FRES-point
// transforms of those columns, and the kept rows [nfilt_pos, nfilt_pos +
// nkeep) of every column written with consecutive lanes on consecutive
// output samples.  Same transforms and product as the library form above,
// without its four HBM round trips of the window.  ZDM: the sub-bands of
// the spectrum without the response are transformed back too, into the
// zero-DM output zdm (same layout as out), from the forward transform kept
// in registers (FilterbankEngineCPU.C:266-300).
template <int NFFT, int FRES, bool ZDM = false>
__global__ __launch_bounds__(NFFT / 16) void k_fb_fused(const float* __restrict__ in, uint64_t ics, uint64_t ips,
                                                        uint32_t npol, uint64_t in_step, uint32_t npart,
                                                        const float2* __restrict__ resp,
                                                        const float2* __restrict__ tw_f,
                                                        const float2* __restrict__ tw_b, uint32_t nkeep,
                                                        uint32_t nfilt_pos, float* __restrict__ out, uint64_t ocs,
                                                        uint64_t ops, float* __restrict__ zdm)
{
  constexpr int NT = NFFT / 16, E = 16, NSUB = NFFT / FRES, RS = NSUB + 1;
  using ROW = RowLayout<NFFT, NT>;
  using COL = ColLayout<FRES, NSUB, RS, NT>;
  constexpr int LROW = NFFT + NFFT / 16, LCOL = FRES * RS;
  __shared__ float2 lds[LROW > LCOL ? LROW : LCOL];
  const int tid = threadIdx.x;
  const uint32_t u = blockIdx.x;  // fewer than 2^31 units
  const uint32_t p = u % npart, cp = u / npart;
  const uint32_t chan = cp / npol, pol = cp % npol;
  const float2* x = reinterpret_cast<const float2*>(in + chan * ics + pol * ips + 2 * uint64_t(p) * in_step);

  float2 v[E];
  constexpr int R0 = FirstRadix<NFFT, E, -1>::v;
#pragma unroll
  for (int q = 0; q < E / R0; q++)
#pragma unroll
    for (int r = 0; r < R0; r++) v[q * R0 + r] = x[tid + q * NT + r * (NFFT / R0)];
  fft_regs<ROW, NFFT, 1, NT, -1>(lds, v, tw_f, tid);

  // the forward transform's last pass leaves bin g + r*NFFT/RL in v[q*RL + r]
  constexpr int RL = LastRadix<NFFT, E, -1>::v;
  const float2* f = resp + uint64_t(chan) * NFFT;
  float2 v0[ZDM ? E : 1];
  if constexpr (ZDM)
#pragma unroll
    for (int e = 0; e < E; e++) v0[e] = v[e];
  // the spectrum's sub-bands (times the response, or not) transformed back
  // and their kept samples stored at dst
  auto backward = [&](float2* vv, bool product, float* dst) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E / RL; q++)
#pragma unroll
      for (int r = 0; r < RL; r++) {
        const int k = tid + q * NT + r * (NFFT / RL);
        const float2 d = vv[q * RL + r];
        if (product) {
          const float2 w = f[k];
          lds[COL::idx(k % FRES, k / FRES)] = make_float2(w.x * d.x - w.y * d.y, w.y * d.x + w.x * d.y);
        } else {
          lds[COL::idx(k % FRES, k / FRES)] = d;
        }
      }
    __syncthreads();
    lds_load_pass<COL, FRES, NSUB, NT, FirstRadix<FRES, E, +1>::v>(lds, vv, tid);
    fft_regs<COL, FRES, NSUB, NT, +1>(lds, vv, tw_b, tid);
    constexpr int NP = RadixPlan<FRES, E>::n;
    __syncthreads();
    lds_store_pass<COL, FRES, NSUB, NT, InvRadix<FRES, E, NP - 1>::v, InvNs<FRES, E, NP - 1>::v>(lds, vv, tid);
    __syncthreads();
    float* o = dst + uint64_t(chan) * NSUB * ocs + pol * ops + 2 * uint64_t(p) * nkeep;
#pragma unroll
    for (int s = 0; s < NSUB; s++)
      for (uint32_t m = tid; m < nkeep; m += NT)
        *reinterpret_cast<float2*>(o + s * ocs + 2 * m) = lds[COL::idx(int(nfilt_pos + m), s)];
  };
  backward(v, true, out);
  if constexpr (ZDM) backward(v0, false, zdm);
}

}  // namespace dspx

using namespace dspx;

struct dspx_fb {
  CtxRef hold;
  dspx_ctx* ctx = nullptr;
  dspx_fb_desc d{};
  uint64_t nfft = 0, in_step = 0;
  uint32_t nkeep = 0;
  DevBuf resp, scratch;
  bool have_resp = false;
  hipfftHandle fwd = 0, bwd = 0;
  bool have_plan = false;
  uint64_t plan_batch = 0;
  uint32_t last_batches = 0;  // library-form batches of the last perform (0: one kernel)
  Fb4* fb4 = nullptr;         // the four-step form, where the window's shape has one

  ~dspx_fb()
  {
    fb4_destroy(fb4);
    if (have_plan) {
      hipfftDestroy(fwd);
      hipfftDestroy(bwd);
    }
  }
};

extern "C" {

int dspx_fb_create(dspx_ctx* ctx, const dspx_fb_desc* desc, dspx_fb** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && desc && out, DSPX_E_PARAM, "null argument");
    const dspx_fb_desc& d = *desc;
    DSPX_REQUIRE(d.nchan >= 1 && d.npol >= 1 && d.npol <= 2 && d.nsub >= 1 && d.fres >= 1, DSPX_E_PARAM,
                 "invalid shape");
    DSPX_REQUIRE(uint64_t(d.nfilt_pos) + d.nfilt_neg < d.fres, DSPX_E_STATE, "nfilt=%u >= freq_res=%u",
                 d.nfilt_pos + d.nfilt_neg, d.fres);  // Filterbank.C:425-428 analogue
    DSPX_REQUIRE(uint64_t(d.nchan) * d.nsub <= 65535 * 64, DSPX_E_UNSUPPORTED, "too many output channels");
    DSPX_HIP(hipSetDevice(ctx->device));
    auto* F = new dspx_fb;
    F->hold.reset(ctx);
    F->ctx = ctx;
    F->d = d;
    F->nfft = uint64_t(d.nsub) * d.fres;
    F->nkeep = d.fres - d.nfilt_pos - d.nfilt_neg;
    F->in_step = uint64_t(d.nsub) * F->nkeep;  // nsamp_fft - nsamp_overlap (Filterbank.C:154-192)
    std::unique_ptr<dspx_fb> hold(F);
    F->fb4 = fb4_create(ctx, d);
    *out = hold.release();
  });
}

void dspx_fb_destroy(dspx_fb* F)
{
  if (!F) return;
  (void) hipSetDevice(F->ctx->device);
  (void) hipStreamSynchronize(F->ctx->stream);
  delete F;
}

int dspx_fb_set_response(dspx_fb* F, const float* host)
{
  return guard(F ? F->ctx : nullptr, [&] {
    DSPX_REQUIRE(F && host, DSPX_E_PARAM, "null argument");
    const size_t bytes = size_t(F->d.nchan) * F->nfft * sizeof(float2);
    F->resp.reserve(bytes);
    DSPX_HIP(hipMemcpy(F->resp.p, host, bytes, hipMemcpyHostToDevice));
    if (F->fb4) fb4_set_response(F->fb4, host);
    F->have_resp = true;
  });
}

int dspx_fb_geometry(const dspx_fb* F, uint64_t ndat, uint64_t* npart, uint64_t* ndat_out)
{
  if (!F || !npart || !ndat_out) return DSPX_E_PARAM;
  const uint64_t novl = uint64_t(F->d.nfilt_pos + F->d.nfilt_neg) * F->d.nsub;
  *npart = ndat > novl ? (ndat - novl) / F->in_step : 0;  // Filterbank.C:423-424
  if (*npart && (*npart - 1) * F->in_step + F->nfft > ndat) (*npart)--;
  *ndat_out = *npart * F->nkeep;
  return DSPX_OK;
}

}  // extern "C"

namespace {

// d_zdm: the zero-DM output (the sub-bands without the response), same
// layout as d_out, or null
int fb_perform(dspx_fb* F, const float* d_in, uint64_t ics, uint64_t ips, uint64_t ndat, float* d_out, uint64_t ocs,
               uint64_t ops, float* d_zdm, bool zdm_call)
{
  return guard(F ? F->ctx : nullptr, [&] {
    DSPX_REQUIRE(F && d_in && d_out && (d_zdm || !zdm_call), DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(F->have_resp, DSPX_E_STATE, "no response: dspx_fb_set_response");
    uint64_t npart, nout;
    dspx_fb_geometry(F, ndat, &npart, &nout);
    if (!npart) return;
    DSPX_REQUIRE(npart < (1u << 31), DSPX_E_UNSUPPORTED, "npart");
    const dspx_fb_desc& d = F->d;
    const uint64_t nunit = uint64_t(d.nchan) * d.npol * npart;
    DSPX_REQUIRE(nunit < (1ull << 31), DSPX_E_UNSUPPORTED, "too many windows: %llu", (unsigned long long) nunit);
    hipStream_t st = F->ctx->stream;
    F->last_batches = 0;

    // One fused kernel for the window lengths it is built for (1024-4096
    // points in sub-bands of 32-512), the four-step form for the longer
    // windows it is built for, the library form otherwise or where the
    // environment asks for it (DSPX_FB_FORM=library).
    const char* form = std::getenv("DSPX_FB_FORM");
    const bool library = form && std::string(form) == "library";
#define FB_FUSED(n, m)                                                                                        \
  if (!library && F->nfft == n && d.fres == m) {                                                              \
    if (d_zdm)                                                                                                \
      k_fb_fused<n, m, true><<<unsigned(nunit), n / 16, 0, st>>>(                                            \
          d_in, ics, ips, d.npol, F->in_step, uint32_t(npart), F->resp.as<float2>(), twiddles(F->ctx, n),     \
          twiddles(F->ctx, m), F->nkeep, d.nfilt_pos, d_out, ocs, ops, d_zdm);                                \
    else                                                                                                      \
      k_fb_fused<n, m><<<unsigned(nunit), n / 16, 0, st>>>(d_in, ics, ips, d.npol, F->in_step,                \
                                                           uint32_t(npart), F->resp.as<float2>(),             \
                                                           twiddles(F->ctx, n), twiddles(F->ctx, m),          \
                                                           F->nkeep, d.nfilt_pos, d_out, ocs, ops, nullptr);  \
    DSPX_HIP(hipGetLastError());                                                                              \
    return;                                                                                                   \
  }
    FB_FUSED(1024, 32) FB_FUSED(1024, 64) FB_FUSED(1024, 128) FB_FUSED(1024, 256) FB_FUSED(1024, 512)
    FB_FUSED(2048, 32) FB_FUSED(2048, 64) FB_FUSED(2048, 128) FB_FUSED(2048, 256) FB_FUSED(2048, 512)
    FB_FUSED(4096, 32) FB_FUSED(4096, 64) FB_FUSED(4096, 128) FB_FUSED(4096, 256) FB_FUSED(4096, 512)
#undef FB_FUSED
    if (!library && F->fb4) {
      fb4_perform(F->fb4, d_in, ics, ips, npart, d_out, ocs, ops, d_zdm);
      return;
    }

    const uint64_t per = F->nfft * sizeof(float2);
    // units per batch: scratch that stays in the 256 MB last-level cache
    // between the passes (128 MB measured fastest of 16..256 MB on 128
    // channels x 2^20 samples in 2048-point windows; environment
    // DSPX_FB_SCRATCH_MB overrides), at most 65535 (grid.y), spread evenly
    // over the batches so that the last one is not mostly padding
    uint64_t cap = 128ull << 20;
    if (const char* e = std::getenv("DSPX_FB_SCRATCH_MB")) cap = std::max<uint64_t>(1, std::strtoull(e, nullptr, 10)) << 20;
    const uint64_t maxb = std::max<uint64_t>(1, std::min<uint64_t>({nunit, cap / per, 65535}));
    const uint64_t nbat = (nunit + maxb - 1) / maxb;
    const uint64_t batch = (nunit + nbat - 1) / nbat;
    const uint64_t nkb = (uint64_t(d.nsub) * F->nkeep + 255) / 256;
    DSPX_REQUIRE(uint64_t(d.nsub) * F->nkeep < (1ull << 32) && nkb < (1ull << 31), DSPX_E_UNSUPPORTED,
                 "nsub=%u too large", d.nsub);
    if (!F->have_plan || F->plan_batch != batch) {
      if (F->have_plan) {
        hipfftDestroy(F->fwd);
        hipfftDestroy(F->bwd);
        F->have_plan = false;
      }
      int n = int(F->nfft), m = int(d.fres);
      DSPX_FB_FFT(hipfftPlanMany(&F->fwd, 1, &n, nullptr, 1, n, nullptr, 1, n, HIPFFT_C2C, int(batch)));
      DSPX_FB_FFT(hipfftPlanMany(&F->bwd, 1, &m, nullptr, 1, m, nullptr, 1, m, HIPFFT_C2C, int(batch * d.nsub)));
      DSPX_FB_FFT(hipfftSetStream(F->fwd, F->ctx->stream));
      DSPX_FB_FFT(hipfftSetStream(F->bwd, F->ctx->stream));
      F->have_plan = true;
      F->plan_batch = batch;
    }
    F->scratch.reserve(batch * per);
    F->last_batches = uint32_t(nbat);
    float2* s = F->scratch.as<float2>();
    auto* sc = reinterpret_cast<hipfftComplex*>(s);
    // each batch's windows transformed forward; then the product with the
    // response, the backward transforms and the discard into d_out; with a
    // zero-DM output the windows are transformed again and their sub-bands
    // transformed back without the product, into d_zdm
    for (uint64_t u0 = 0; u0 < nunit; u0 += batch) {
      const uint64_t nb = std::min(batch, nunit - u0);
      for (int pass = 0; pass < (d_zdm ? 2 : 1); pass++) {
        if (nb != batch)  // the unused rows of the last batch stay finite
          DSPX_HIP(hipMemsetAsync(s + nb * F->nfft, 0, (batch - nb) * per, st));
        dim3 g0(unsigned((F->nfft + 511) / 512), unsigned(nb));
        k_fb_windows<<<g0, 256, 0, st>>>(d_in, ics, ips, d.npol, F->nfft, F->in_step, uint32_t(npart), u0, s);
        DSPX_HIP(hipGetLastError());
        DSPX_FB_FFT(hipfftExecC2C(F->fwd, sc, sc, HIPFFT_FORWARD));
        if (pass == 0) {
          dim3 g1(unsigned((F->nfft + 255) / 256), unsigned(nb));
          k_fb_multiply<<<g1, 256, 0, st>>>(s, F->resp.as<float2>(), F->nfft, d.npol, uint32_t(npart), u0);
          DSPX_HIP(hipGetLastError());
        }
        DSPX_FB_FFT(hipfftExecC2C(F->bwd, sc, sc, HIPFFT_BACKWARD));
        dim3 g2(static_cast<unsigned>(nkb), static_cast<unsigned>(nb));
        k_fb_discard<<<g2, 256, 0, st>>>(s, F->nfft, d.fres, d.nsub, d.npol, uint32_t(npart), u0, F->nkeep,
                                         d.nfilt_pos, pass == 0 ? d_out : d_zdm, ocs, ops);
        DSPX_HIP(hipGetLastError());
      }
    }
  });
}

}  // namespace

extern "C" {

int dspx_fb_perform(dspx_fb* F, const float* d_in, uint64_t ics, uint64_t ips, uint64_t ndat, float* d_out,
                    uint64_t ocs, uint64_t ops)
{
  return fb_perform(F, d_in, ics, ips, ndat, d_out, ocs, ops, nullptr, false);
}

int dspx_fb_perform_zdm(dspx_fb* F, const float* d_in, uint64_t ics, uint64_t ips, uint64_t ndat, float* d_out,
                        uint64_t ocs, uint64_t ops, float* d_zdm)
{
  return fb_perform(F, d_in, ics, ips, ndat, d_out, ocs, ops, d_zdm, true);
}

int dspx_fb_batches(const dspx_fb* F, uint32_t* nbatch)
{
  if (!F || !nbatch) return DSPX_E_PARAM;
  *nbatch = F->last_batches;
  return DSPX_OK;
}

}  // extern "C"
