// plain.hip -- hipFFT-based plain FP32 convolution (see plain.hpp).
#include <algorithm>
#include <cmath>

#include "plain.hpp"
#include "stages.hpp"

namespace dspx {

#define DSPX_FFT(call)                                                                 \
  do {                                                                                 \
    hipfftResult r_ = (call);                                                          \
    if (r_ != HIPFFT_SUCCESS) ::dspx::fail(DSPX_E_HIP, "%s failed: hipfft %d", #call, int(r_)); \
  } while (0)

// response in FFT-bin order from the on-chip chirp (same arithmetic as stage B)
__global__ void k_build_resp(ChirpArgs c, uint32_t N, float2* __restrict__ resp)
{
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t chan = blockIdx.y;
  if (k >= N) return;
  resp[uint64_t(chan) * N + k] = chirp_onchip(c, chan, k, N);
}

// scratch[(cp * npart + p) * N + k] = in[cp][p*step + k]
__global__ void k_windows(const float* __restrict__ in, uint64_t ics, uint64_t ips, uint32_t npol, uint32_t N,
                          uint32_t step, uint32_t npart, uint32_t cp0, float2* __restrict__ scratch)
{
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bp = blockIdx.y;  // (cp - cp0) * npart + p
  if (k >= N) return;
  uint32_t p = bp % npart, cp = cp0 + bp / npart;
  uint32_t chan = cp / npol, pol = cp % npol;
  const float* x = in + chan * ics + pol * ips + 2 * (uint64_t(p) * step + k);
  scratch[uint64_t(bp) * N + k] = *reinterpret_cast<const float2*>(x);
}

__global__ void k_multiply(float2* __restrict__ scratch, const float2* __restrict__ resp, uint32_t N,
                           uint32_t npol, uint32_t npart, uint32_t cp0)
{
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bp = blockIdx.y;
  if (k >= N) return;
  uint32_t chan = (cp0 + bp / npart) / npol;
  float2 d = scratch[uint64_t(bp) * N + k];
  float2 f = resp[uint64_t(chan) * N + k];
  // Response::operate (Response.C:709-721)
  scratch[uint64_t(bp) * N + k] = make_float2(f.x * d.x - f.y * d.y, f.y * d.x + f.x * d.y);
}

__global__ void k_discard(const float2* __restrict__ scratch, uint32_t N, uint32_t npol, uint32_t npart,
                          uint32_t cp0, uint32_t step, uint32_t nfilt_pos, float* __restrict__ out, uint64_t ocs,
                          uint64_t ops)
{
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bp = blockIdx.y;
  if (j >= step) return;
  uint32_t p = bp % npart, cp = cp0 + bp / npart;
  uint32_t chan = cp / npol, pol = cp % npol;
  float* o = out + chan * ocs + pol * ops + 2 * (uint64_t(p) * step + j);
  *reinterpret_cast<float2*>(o) = scratch[uint64_t(bp) * N + nfilt_pos + j];
}

void PlainConv::init(dspx_ctx* c, uint32_t nchan_, uint32_t npol_, uint32_t N_, uint32_t pos, uint32_t neg,
                     uint32_t mode, const dspx_dedisp_desc& d)
{
  ctx = c;
  nchan = nchan_;
  npol = npol_;
  N = N_;
  nfilt_pos = pos;
  nfilt_neg = neg;
  chirp_mode = mode;
  DSPX_REQUIRE(npol == 1 || npol == 2, DSPX_E_PARAM, "npol=%u", npol);
  DSPX_REQUIRE(N >= 2 && uint64_t(pos) + neg < N, DSPX_E_STATE, "error nfft=%u < nfilt=%u", N, pos + neg);
  DSPX_REQUIRE(nchan >= 1 && nchan <= 65535, DSPX_E_PARAM, "nchan=%u", nchan);
  step = N - pos - neg;
  resp.reserve(size_t(nchan) * N * sizeof(float2));
  if (mode != DSPX_CHIRP_TABLE) {
    DSPX_REQUIRE(d.nchan == nchan, DSPX_E_PARAM, "dedisp.nchan=%u != nchan=%u", d.nchan, nchan);
    double chanwidth = d.bandwidth / double(nchan);
    std::vector<double> pc(3 * nchan, 0.0);
    for (uint32_t ic = 0; ic < nchan; ic++) {
      double cf = channel_centre(d.centre_frequency, d.bandwidth, nchan, d.dc_centred, ic);
      pc[ic] = chirp_coeff(d.bandwidth, d.dispersion_measure, cf);
      pc[nchan + ic] = cf;
      if (d.interchan)
        pc[2 * nchan + ic] =
            interchan_fractional_delay(d.centre_frequency, d.bandwidth, nchan, d.dispersion_measure, cf);
    }
    DevBuf par;
    par.reserve(pc.size() * sizeof(double));
    DSPX_HIP(hipMemcpy(par.p, pc.data(), pc.size() * sizeof(double), hipMemcpyHostToDevice));
    ChirpArgs ca;
    ca.mode = mode;
    ca.snorm = float(1.0 / std::sqrt(double(N)));
    ca.binwidth = chanwidth / double(N);
    ca.half_cw = 0.5 * chanwidth;
    ca.coeff = par.as<double>();
    ca.cf = par.as<double>() + nchan;
    ca.fdelay = par.as<double>() + 2 * nchan;
    ca.table = nullptr;
    dim3 grid((N + 255) / 256, nchan);
    k_build_resp<<<grid, 256, 0, ctx->stream>>>(ca, N, resp.as<float2>());
    DSPX_HIP(hipGetLastError());
    DSPX_HIP(hipStreamSynchronize(ctx->stream));
    have_resp = true;
  }
}

void PlainConv::set_table(const float* host)
{
  DSPX_HIP(hipMemcpy(resp.p, host, size_t(nchan) * N * sizeof(float2), hipMemcpyHostToDevice));
  have_resp = true;
}

void PlainConv::perform(const float* in, uint64_t ics, uint64_t ips, float* out, uint64_t ocs, uint64_t ops,
                        uint32_t npart)
{
  DSPX_REQUIRE(have_resp, DSPX_E_STATE, "no response: call dspx_conv_set_response first");
  // batch as many (chan, pol) units as fit a 1 GiB scratch
  uint32_t ncp = nchan * npol;
  uint64_t per_cp = uint64_t(npart) * N * sizeof(float2);
  uint32_t cp_batch = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(ncp, (1ull << 30) / per_cp)));
  uint32_t batch = cp_batch * npart;
  if (!have_plan || plan_batch != batch) {
    if (have_plan) hipfftDestroy(plan);
    have_plan = false;
    int n = int(N);
    DSPX_FFT(hipfftPlanMany(&plan, 1, &n, nullptr, 1, int(N), nullptr, 1, int(N), HIPFFT_C2C, int(batch)));
    DSPX_FFT(hipfftSetStream(plan, ctx->stream));
    have_plan = true;
    plan_batch = batch;
  }
  scratch.reserve(uint64_t(batch) * N * sizeof(float2));
  float2* s = scratch.as<float2>();
  for (uint32_t cp0 = 0; cp0 < ncp; cp0 += cp_batch) {
    uint32_t ncur = std::min(cp_batch, ncp - cp0);
    uint32_t nb = ncur * npart;
    if (nb != batch) {
      // tail: zero the unused windows so the full-batch transform stays finite
      DSPX_HIP(hipMemsetAsync(s + uint64_t(nb) * N, 0, uint64_t(batch - nb) * N * sizeof(float2), ctx->stream));
    }
    dim3 g1((N + 255) / 256, nb);
    k_windows<<<g1, 256, 0, ctx->stream>>>(in, ics, ips, npol, N, step, npart, cp0, s);
    DSPX_HIP(hipGetLastError());
    DSPX_FFT(hipfftExecC2C(plan, reinterpret_cast<hipfftComplex*>(s), reinterpret_cast<hipfftComplex*>(s),
                           HIPFFT_FORWARD));
    k_multiply<<<g1, 256, 0, ctx->stream>>>(s, resp.as<float2>(), N, npol, npart, cp0);
    DSPX_HIP(hipGetLastError());
    DSPX_FFT(hipfftExecC2C(plan, reinterpret_cast<hipfftComplex*>(s), reinterpret_cast<hipfftComplex*>(s),
                           HIPFFT_BACKWARD));
    dim3 g2((step + 255) / 256, nb);
    k_discard<<<g2, 256, 0, ctx->stream>>>(s, N, npol, npart, cp0, step, nfilt_pos, out, ocs, ops);
    DSPX_HIP(hipGetLastError());
  }
}

PlainConv::~PlainConv()
{
  if (have_plan) hipfftDestroy(plan);
}

}  // namespace dspx
