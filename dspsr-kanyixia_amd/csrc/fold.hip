// fold.hip -- Fold::Engine: device bin plan, hits and phase-bin accumulation.
//
// Reference: Fold::fold (Signal/Pulsar/Fold.C:662-949), CUDA::FoldEngine
// (Signal/Pulsar/FoldCUDA.cu:25-704).  The reference GPU engine builds its bin
// plan serially on the host and copies it every block; here the host only
// walks binade runs (binplan.cpp) and the device evaluates every sample.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "dspx_internal.hpp"
#include "fft_core.hpp"
#include "fold_impl.hpp"

namespace dspx {

__global__ void k_binplan(const FoldSegment* __restrict__ seg, uint32_t nseg, uint64_t ndat,
                          double nbin_d, uint32_t* __restrict__ out)
{
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ndat) return;
  out[i] = seg_key(seg[seg_find(seg, nseg, i)], i, nbin_d);
}

// hits[isub][bin] += count (Fold.C:827-836 for unzapped data).  A workgroup
// covers 4096 consecutive samples.
constexpr int kHitsChunk = 4096;
__global__ __launch_bounds__(256) void k_hits(const uint32_t* __restrict__ plan, uint64_t ndat,
                                              uint32_t nbin, uint32_t* __restrict__ hits)
{
  extern __shared__ uint32_t h[];  // 2 * nbin
  uint64_t i0 = uint64_t(blockIdx.x) * kHitsChunk;
  uint64_t i1 = min(ndat, i0 + kHitsChunk);
  for (uint32_t k = threadIdx.x; k < 2 * nbin; k += blockDim.x) h[k] = 0;
  __shared__ uint32_t sub0;
  if (threadIdx.x == 0) {
    uint32_t s = 0xffffffffu;
    for (uint64_t i = i0; i < i1; i++)
      if (plan[i] != 0xffffffffu) {
        s = plan[i] >> 16;
        break;
      }
    sub0 = s;
  }
  __syncthreads();
  for (uint64_t base = i0; base < i1; base += blockDim.x) {
    uint64_t i = base + threadIdx.x;
    uint32_t key = i < i1 ? plan[i] : 0xffffffffu;
    // count as a float sum: exact for counts below 2^24
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
    if (__all(key == k0)) {
      if (k0 == 0xffffffffu) continue;
      if ((threadIdx.x & 63u) == 63) {
        uint32_t sub = k0 >> 16, bin = k0 & 0xffffu;
        if (sub - sub0 < 2)
          atomicAdd(&h[(sub - sub0) * nbin + bin], 64u);
        else
          atomicAdd(&hits[uint64_t(sub) * nbin + bin], 64u);
      }
    } else if (key != 0xffffffffu) {
      uint32_t sub = key >> 16, bin = key & 0xffffu;
      if (sub - sub0 < 2)
        atomicAdd(&h[(sub - sub0) * nbin + bin], 1u);
      else
        atomicAdd(&hits[uint64_t(sub) * nbin + bin], 1u);
    }
  }
  __syncthreads();
  if (sub0 == 0xffffffffu) return;
  for (uint32_t k = threadIdx.x; k < 2 * nbin; k += blockDim.x)
    if (h[k]) atomicAdd(&hits[uint64_t(sub0 + k / nbin) * nbin + (k % nbin)], h[k]);
}

// Detected FPT input -> prof[isub][chan][pol][bin*ndim + idim].  A workgroup
// folds a contiguous run of samples of one (chan, pol) into an LDS copy of
// the (at most two) sub-integrations the run meets, summing each wavefront's
// samples first when they share a bin, then adds the non-zero bins into the
// profile.
constexpr int kFoldChunk = 8192;
__global__ __launch_bounds__(256) void k_fold_fpt(const float* __restrict__ in, uint64_t chan_stride,
                                                  uint64_t pol_stride, uint32_t npol, uint32_t ndim,
                                                  const uint32_t* __restrict__ plan, uint64_t ndat,
                                                  uint32_t nbin, uint32_t nchan, uint32_t nchunk,
                                                  float* __restrict__ prof)
{
  extern __shared__ float hf[];  // 2 * nbin * ndim
  uint32_t bid = blockIdx.x;
  uint32_t chunk = bid % nchunk;
  bid /= nchunk;
  uint32_t pol = bid % npol;
  uint32_t chan = bid / npol;
  uint64_t i0 = uint64_t(chunk) * kFoldChunk;
  uint64_t i1 = min(ndat, i0 + kFoldChunk);
  const uint32_t hl = nbin * ndim;
  for (uint32_t k = threadIdx.x; k < 2 * hl; k += blockDim.x) hf[k] = 0.f;
  __shared__ uint32_t sub0;
  if (threadIdx.x == 0) {
    uint32_t s = 0xffffffffu;
    for (uint64_t i = i0; i < i1; i++)
      if (plan[i] != 0xffffffffu) {
        s = plan[i] >> 16;
        break;
      }
    sub0 = s;
  }
  __syncthreads();
  const float* x = in + uint64_t(chan) * chan_stride + uint64_t(pol) * pol_stride;
  const uint64_t prof_sub = uint64_t(nchan) * npol * hl;
  float* pc = prof + (uint64_t(chan) * npol + pol) * hl;
  for (uint64_t base = i0; base < i1; base += blockDim.x) {
    uint64_t i = base + threadIdx.x;
    uint32_t key = i < i1 ? plan[i] : 0xffffffffu;
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
    const bool uniform = __all(key == k0);
    for (uint32_t d = 0; d < ndim; d++) {
      float v = (key != 0xffffffffu) ? x[i * ndim + d] : 0.f;
      uint32_t k = key;
      if (uniform) {
        if (k0 == 0xffffffffu) break;
        v = wave_sum_lane63(v);
        if ((threadIdx.x & 63u) != 63) continue;
        k = k0;
      } else if (key == 0xffffffffu) {
        continue;
      }
      uint32_t sub = k >> 16, bin = k & 0xffffu;
      if (sub - sub0 < 2)
        atomicAdd(&hf[(sub - sub0) * hl + bin * ndim + d], v);
      else
        atomicAdd(&pc[uint64_t(sub) * prof_sub + bin * ndim + d], v);
    }
  }
  __syncthreads();
  if (sub0 == 0xffffffffu) return;
  for (uint32_t k = threadIdx.x; k < 2 * hl; k += blockDim.x)
    if (hf[k] != 0.f) atomicAdd(&pc[uint64_t(sub0 + k / hl) * prof_sub + (k % hl)], hf[k]);
}

void FoldPlan::build(dspx_ctx* c, uint64_t ndat_, double pps_, uint32_t nbin_,
                     const dspx_fold_chunk* chunks, uint32_t nchunk, bool sample_plan)
{
  ctx = c;
  ndat = ndat_;
  pps = pps_;
  nbin = nbin_;
  DSPX_REQUIRE(nbin >= 1 && nbin <= 65535, DSPX_E_PARAM, "nbin=%u outside 1..65535", nbin);
  DSPX_REQUIRE(pps > 0.0 && pps < 0.5, DSPX_E_PARAM,
               "phase_per_sample=%g outside (0, 0.5)", pps);
  DSPX_REQUIRE(nchunk >= 1, DSPX_E_PARAM, "no fold chunks");
  std::vector<dspx_fold_chunk> cs(chunks, chunks + nchunk);
  std::sort(cs.begin(), cs.end(),
            [](const dspx_fold_chunk& a, const dspx_fold_chunk& b) { return a.idat_start < b.idat_start; });
  std::vector<FoldSegment> segs;
  uint64_t cursor = 0;
  nsub = 0;
  for (auto& ch : cs) {
    DSPX_REQUIRE(ch.idat_start >= cursor, DSPX_E_PARAM, "fold chunks overlap at sample %llu",
                 (unsigned long long) ch.idat_start);
    DSPX_REQUIRE(ch.idat_start + ch.ndat <= ndat, DSPX_E_PARAM,
                 "fold chunk [%llu, +%llu) beyond ndat=%llu", (unsigned long long) ch.idat_start,
                 (unsigned long long) ch.ndat, (unsigned long long) ndat);
    DSPX_REQUIRE(ch.isub < 0xffff, DSPX_E_PARAM, "isub=%u too large", ch.isub);
    if (ch.idat_start > cursor) segs.push_back(FoldSegment{cursor, 0, 0, 0, 0xffffffffu});
    if (ch.ndat) fold_segments(ch.phi, pps, ch.idat_start, ch.ndat, ch.isub, segs);
    cursor = ch.idat_start + ch.ndat;
    nsub = std::max(nsub, ch.isub + 1);
  }
  if (cursor < ndat || segs.empty()) segs.push_back(FoldSegment{cursor, 0, 0, 0, 0xffffffffu});
  nseg = uint32_t(segs.size());
  seg.reserve(segs.size() * sizeof(FoldSegment));
  int slot;
  void* h = stage.acquire(segs.size() * sizeof(FoldSegment), slot);
  std::memcpy(h, segs.data(), segs.size() * sizeof(FoldSegment));
  stage.upload(slot, seg.p, segs.size() * sizeof(FoldSegment), ctx->stream);
  have_sample_plan = sample_plan;
  if (sample_plan && ndat) {
    plan.reserve(ndat * sizeof(uint32_t));
    unsigned nb = unsigned((ndat + 255) / 256);
    k_binplan<<<nb, 256, 0, ctx->stream>>>(seg.as<FoldSegment>(), nseg, ndat, double(nbin),
                                           plan.as<uint32_t>());
    DSPX_HIP(hipGetLastError());
  }
  // host copy of the chunk table for the fused path's sub-integration slots
  chunk_table = cs;
}

void FoldPlan::hits(uint32_t* d_hits)
{
  if (!ndat) return;
  DSPX_REQUIRE(have_sample_plan, DSPX_E_INTERNAL, "hits without a sample plan");
  unsigned nb = unsigned((ndat + kHitsChunk - 1) / kHitsChunk);
  k_hits<<<nb, 256, 2 * nbin * sizeof(uint32_t), ctx->stream>>>(plan.as<uint32_t>(), ndat, nbin,
                                                                   d_hits);
  DSPX_HIP(hipGetLastError());
}

void FoldPlan::fold_fpt(const float* d_in, uint64_t chan_stride, uint64_t pol_stride, uint32_t nchan,
                        uint32_t npol, uint32_t ndim, float* d_prof)
{
  if (!ndat) return;
  DSPX_REQUIRE(have_sample_plan, DSPX_E_INTERNAL, "fold without a sample plan");
  DSPX_REQUIRE(ndim >= 1 && ndim <= 4, DSPX_E_PARAM, "ndim=%u", ndim);
  size_t shm = size_t(2) * nbin * ndim * sizeof(float);
  DSPX_REQUIRE(shm <= 64 * 1024, DSPX_E_UNSUPPORTED, "nbin*ndim=%u too large for LDS", nbin * ndim);
  uint32_t nchunk = uint32_t((ndat + kFoldChunk - 1) / kFoldChunk);
  uint64_t nb = uint64_t(nchunk) * npol * nchan;
  DSPX_REQUIRE(nb < (1ull << 31), DSPX_E_UNSUPPORTED, "grid too large");
  k_fold_fpt<<<unsigned(nb), 256, shm, ctx->stream>>>(d_in, chan_stride, pol_stride, npol, ndim,
                                                      plan.as<uint32_t>(), ndat, nbin, nchan, nchunk,
                                                      d_prof);
  DSPX_HIP(hipGetLastError());
}

}  // namespace dspx

using namespace dspx;

struct dspx_fold {
  FoldPlan plan;
};

extern "C" {

int dspx_fold_create(dspx_ctx* ctx, uint64_t ndat, double pps, uint32_t nbin,
                     const dspx_fold_chunk* chunks, uint32_t nchunk, dspx_fold** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && out && chunks, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipSetDevice(ctx->device));
    auto* f = new dspx_fold;
    try {
      f->plan.build(ctx, ndat, pps, nbin, chunks, nchunk);
    } catch (...) {
      delete f;
      throw;
    }
    *out = f;
  });
}

void dspx_fold_destroy(dspx_fold* f) { delete f; }

int dspx_fold_binplan(dspx_fold* f, uint32_t* d_out)
{
  dspx_ctx* ctx = f ? f->plan.ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(f && d_out, DSPX_E_PARAM, "null argument");
    if (f->plan.ndat)
      DSPX_HIP(hipMemcpyAsync(d_out, f->plan.plan.p, f->plan.ndat * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, ctx->stream));
  });
}

int dspx_fold_hits(dspx_fold* f, uint32_t* d_hits)
{
  dspx_ctx* ctx = f ? f->plan.ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(f && d_hits, DSPX_E_PARAM, "null argument");
    f->plan.hits(d_hits);
  });
}

int dspx_fold_execute(dspx_fold* f, const float* d_in, uint64_t in_chan_stride, uint64_t in_pol_stride,
                      uint32_t nchan, uint32_t npol, uint32_t ndim, float* d_prof)
{
  dspx_ctx* ctx = f ? f->plan.ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(f && d_in && d_prof, DSPX_E_PARAM, "null argument");
    f->plan.fold_fpt(d_in, in_chan_stride, in_pol_stride, nchan, npol, ndim, d_prof);
  });
}

}  // extern "C"
