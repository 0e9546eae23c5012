// fold.hip -- Fold::Engine: device bin plan, hits and phase-bin accumulation.
//
// Reference: Fold::fold (Signal/Pulsar/Fold.C:662-973), CUDA::FoldEngine
// (Signal/Pulsar/FoldCUDA.cu:25-704).  The reference GPU engine builds its bin
// plan serially on the host and copies it every block; here the host only
// walks binade runs (segments.hpp) and the device evaluates every sample.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

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

// First sub-integration folded in samples [i0, i1), or ~0.
template <class KEYS> __device__ uint32_t first_sub(const KEYS& keys, uint64_t i0, uint64_t i1)
{
  for (uint64_t i = i0; i < i1; i++) {
    const uint32_t k = keys(i);
    if (k != 0xffffffffu) return k >> 16;
  }
  return 0xffffffffu;
}

// hits[isub][bin] += count (Fold.C:827-836 without zeroed data).  A
// workgroup covers 4096 consecutive samples.
constexpr int kHitsChunk = 4096;
template <class KEYS>
__global__ __launch_bounds__(256) void k_hits(KEYS keys, uint64_t ndat, uint32_t nbin, uint32_t* __restrict__ hits)
{
  extern __shared__ uint32_t h[];  // 2 * nbin
  uint64_t i0 = uint64_t(blockIdx.x) * kHitsChunk;
  uint64_t i1 = min(ndat, i0 + kHitsChunk);
  for (uint32_t k = threadIdx.x; k < 2 * nbin; k += blockDim.x) h[k] = 0;
  __shared__ uint32_t sub0;
  if (threadIdx.x == 0) sub0 = first_sub(keys, i0, i1);
  __syncthreads();
  for (uint64_t base = i0; base < i1; base += blockDim.x) {
    uint64_t i = base + threadIdx.x;
    uint32_t key = i < i1 ? keys(i) : 0xffffffffu;
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
// profile.  With zhits (zeroed data, Fold.C:893-930) the pol-0 workgroups
// also count, per bin, the samples whose value 0 is non-zero into
// zhits[isub][hchan][bin] (hchan = chan when hits_nchan = nchan, else 0) and
// add their number into *nfolded.
constexpr int kFoldChunk = 8192;
template <class KEYS>
__global__ __launch_bounds__(256) void k_fold_fpt(const float* __restrict__ in, uint64_t chan_stride,
                                                  uint64_t pol_stride, uint32_t npol, uint32_t ndim, KEYS keys,
                                                  uint64_t ndat, uint32_t nbin, uint32_t nchan, uint32_t nchunk,
                                                  float* __restrict__ prof, uint32_t* __restrict__ zhits,
                                                  uint32_t hits_nchan, unsigned long long* __restrict__ nfolded)
{
  extern __shared__ float hf[];  // 2 * nbin * ndim floats, then (zhits) 2 * nbin counts
  uint32_t bid = blockIdx.x;
  uint32_t chunk = bid % nchunk;
  bid /= nchunk;
  uint32_t pol = bid % npol;
  uint32_t chan = bid / npol;
  uint64_t i0 = uint64_t(chunk) * kFoldChunk;
  uint64_t i1 = min(ndat, i0 + kFoldChunk);
  const uint32_t hl = nbin * ndim;
  const bool count = zhits && pol == 0;
  uint32_t* hz = reinterpret_cast<uint32_t*>(hf + 2 * hl);
  for (uint32_t k = threadIdx.x; k < 2 * hl; k += blockDim.x) hf[k] = 0.f;
  if (count)
    for (uint32_t k = threadIdx.x; k < 2 * nbin; k += blockDim.x) hz[k] = 0;
  __shared__ uint32_t sub0;
  __shared__ uint32_t nz;
  if (threadIdx.x == 0) {
    sub0 = first_sub(keys, i0, i1);
    nz = 0;
  }
  __syncthreads();
  const float* x = in + uint64_t(chan) * chan_stride + uint64_t(pol) * pol_stride;
  const uint64_t prof_sub = uint64_t(nchan) * npol * hl;
  float* pc = prof + (uint64_t(chan) * npol + pol) * hl;
  const uint64_t hz_sub = uint64_t(hits_nchan) * nbin;
  uint32_t* hzg = count ? zhits + uint64_t(hits_nchan == nchan ? chan : 0) * nbin : nullptr;
  uint32_t my_nz = 0;
  for (uint64_t base = i0; base < i1; base += blockDim.x) {
    uint64_t i = base + threadIdx.x;
    uint32_t key = i < i1 ? keys(i) : 0xffffffffu;
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
    const bool uniform = __all(key == k0);
    if (uniform && k0 == 0xffffffffu) continue;
    for (uint32_t d = 0; d < ndim; d++) {
      float v = (key != 0xffffffffu) ? x[i * ndim + d] : 0.f;
      if (count && d == 0) {
        // samples zapped to zero are folded but not counted (Fold.C:911-920)
        const bool hit = key != 0xffffffffu && v != 0.f;
        my_nz += hit ? 1u : 0u;
        if (uniform) {
          const uint32_t nh = uint32_t(__popcll(__ballot(hit)));
          if ((threadIdx.x & 63u) == 63 && nh) {
            uint32_t sub = k0 >> 16, bin = k0 & 0xffffu;
            if (sub - sub0 < 2)
              atomicAdd(&hz[(sub - sub0) * nbin + bin], nh);
            else
              atomicAdd(&hzg[uint64_t(sub) * hz_sub + bin], nh);
          }
        } else if (hit) {
          uint32_t sub = key >> 16, bin = key & 0xffffu;
          if (sub - sub0 < 2)
            atomicAdd(&hz[(sub - sub0) * nbin + bin], 1u);
          else
            atomicAdd(&hzg[uint64_t(sub) * hz_sub + bin], 1u);
        }
      }
      uint32_t k = key;
      if (uniform) {
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
  if (count && my_nz) atomicAdd(&nz, my_nz);
  __syncthreads();
  if (sub0 == 0xffffffffu) return;
  for (uint32_t k = threadIdx.x; k < 2 * hl; k += blockDim.x)
    if (hf[k] != 0.f) atomicAdd(&pc[uint64_t(sub0 + k / hl) * prof_sub + (k % hl)], hf[k]);
  if (count) {
    for (uint32_t k = threadIdx.x; k < 2 * nbin; k += blockDim.x)
      if (hz[k]) atomicAdd(&hzg[uint64_t(sub0 + k / nbin) * hz_sub + (k % nbin)], hz[k]);
    if (threadIdx.x == 0 && nz && nfolded) atomicAdd(nfolded, (unsigned long long) nz);
  }
}

void FoldPlan::build(dspx_ctx* c, uint64_t ndat_, double pps_, uint32_t nbin_,
                     const dspx_fold_chunk* chunks, uint32_t nchunk, bool sample_plan, const FoldWeights* weights)
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
  WeightView wv;
  if (weights && weights->ndatperweight) {
    wv.w = weights->w.data();
    wv.nweights = weights->w.size();
    wv.ndatperweight = weights->ndatperweight;
    wv.weight_idat = weights->weight_idat;
  }
  std::vector<FoldSegment> segs;
  plan_segments(ndat, pps, cs.data(), uint32_t(cs.size()), wv, [&](const FoldSegment& s) { segs.push_back(s); },
                nsub, nfold);
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

// Launches of the two fold kernels over any key source (shared by the
// planned engine and the form without a per-sample plan).
template <class KEYS> void launch_hits(hipStream_t st, const KEYS& keys, uint64_t ndat, uint32_t nbin, uint32_t* d_hits)
{
  if (!ndat) return;
  unsigned nb = unsigned((ndat + kHitsChunk - 1) / kHitsChunk);
  k_hits<KEYS><<<nb, 256, 2 * nbin * sizeof(uint32_t), st>>>(keys, ndat, nbin, d_hits);
  DSPX_HIP(hipGetLastError());
}

template <class KEYS>
void launch_fold(hipStream_t st, const KEYS& keys, uint64_t ndat, uint32_t nbin, const float* d_in, uint64_t chan_stride,
                 uint64_t pol_stride, uint32_t nchan, uint32_t npol, uint32_t ndim, float* d_prof, uint32_t* zhits,
                 uint32_t hits_nchan, unsigned long long* nfolded)
{
  if (!ndat) return;
  DSPX_REQUIRE(ndim >= 1 && ndim <= 4, DSPX_E_PARAM, "ndim=%u", ndim);
  DSPX_REQUIRE(!zhits || hits_nchan == 1 || hits_nchan == nchan, DSPX_E_PARAM,
               "hits_nchan=%u must be 1 or nchan=%u", hits_nchan, nchan);
  size_t shm = size_t(2) * nbin * ndim * sizeof(float) + (zhits ? size_t(2) * nbin * sizeof(uint32_t) : 0);
  DSPX_REQUIRE(shm <= 64 * 1024, DSPX_E_UNSUPPORTED, "nbin*ndim=%u too large for LDS", nbin * ndim);
  uint32_t nchunk = uint32_t((ndat + kFoldChunk - 1) / kFoldChunk);
  uint64_t nb = uint64_t(nchunk) * npol * nchan;
  DSPX_REQUIRE(nb < (1ull << 31), DSPX_E_UNSUPPORTED, "grid too large");
  k_fold_fpt<KEYS><<<unsigned(nb), 256, shm, st>>>(d_in, chan_stride, pol_stride, npol, ndim, keys, ndat, nbin,
                                                   nchan, nchunk, d_prof, zhits, hits_nchan, nfolded);
  DSPX_HIP(hipGetLastError());
}

void FoldPlan::hits(uint32_t* d_hits)
{
  DSPX_REQUIRE(have_sample_plan, DSPX_E_INTERNAL, "hits without a sample plan");
  launch_hits(ctx->stream, PlanKeys{plan.as<uint32_t>()}, ndat, nbin, d_hits);
}

void FoldPlan::fold_fpt(const float* d_in, uint64_t chan_stride, uint64_t pol_stride, uint32_t nchan,
                        uint32_t npol, uint32_t ndim, float* d_prof, uint32_t* zhits, uint32_t hits_nchan,
                        unsigned long long* nfolded)
{
  DSPX_REQUIRE(have_sample_plan, DSPX_E_INTERNAL, "fold without a sample plan");
  launch_fold(ctx->stream, PlanKeys{plan.as<uint32_t>()}, ndat, nbin, d_in, chan_stride, pol_stride, nchan, npol,
              ndim, d_prof, zhits, hits_nchan, nfolded);
}

// ------------------------------------------------- fold without allocation
// dspx_fold_static keeps its segment table in this array: one copy per GPU,
// held until the process ends, shared by every call on that GPU.  Host code
// reaches it only through the address hipGetSymbolAddress gives for the
// device (static_ptr), never through the array's name.
constexpr uint32_t kStaticSegMax = 32768;  // 1.25 MiB
static __device__ __align__(256) FoldSegment scratch[kStaticSegMax];

constexpr int kMaxDevices = 64;
static FoldSegment* static_ptr[kMaxDevices] = {};
static std::mutex static_mu;
// host staging of the segments (a plain array: the call allocates nothing)
static FoldSegment static_host[kStaticSegMax];

// Called by dspx_ctx_create with the context's device current.
void fold_static_setup(int device)
{
  DSPX_REQUIRE(device >= 0 && device < kMaxDevices, DSPX_E_UNSUPPORTED, "device %d", device);
  std::lock_guard<std::mutex> lock(static_mu);
  if (static_ptr[device]) return;
  void* p = nullptr;
  DSPX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  static_ptr[device] = static_cast<FoldSegment*>(p);
}

void fold_static(dspx_ctx* ctx, uint64_t ndat, double pps, uint32_t nbin, const dspx_fold_chunk* chunks,
                 uint32_t nchunk, const float* d_in, uint64_t ics, uint64_t ips, uint32_t nchan, uint32_t npol,
                 uint32_t ndim, float* d_prof, uint32_t* d_hits)
{
  DSPX_REQUIRE(nbin >= 1 && nbin <= 65535, DSPX_E_PARAM, "nbin=%u outside 1..65535", nbin);
  DSPX_REQUIRE(pps > 0.0 && pps < 0.5, DSPX_E_PARAM, "phase_per_sample=%g outside (0, 0.5)", pps);
  DSPX_REQUIRE(nchunk >= 1, DSPX_E_PARAM, "no fold chunks");
  FoldSegment* dev = static_ptr[ctx->device];
  DSPX_REQUIRE(dev, DSPX_E_STATE, "no segment table on device %d", ctx->device);
  std::lock_guard<std::mutex> lock(static_mu);
  uint32_t nseg = 0, nsub = 0;
  uint64_t nfold = 0;
  bool over = false;
  plan_segments(ndat, pps, chunks, nchunk, WeightView{},
                [&](const FoldSegment& s) {
                  if (nseg < kStaticSegMax)
                    static_host[nseg] = s;
                  else
                    over = true;
                  nseg++;
                },
                nsub, nfold);
  DSPX_REQUIRE(!over, DSPX_E_UNSUPPORTED, "the plan needs %u segments; at most %u without allocation", nseg,
               kStaticSegMax);
  // stream-ordered after the previous call's kernels that read the table;
  // the host array is reused once the copy has been made
  DSPX_HIP(hipMemcpyAsync(dev, static_host, nseg * sizeof(FoldSegment), hipMemcpyHostToDevice, ctx->stream));
  DSPX_HIP(hipStreamSynchronize(ctx->stream));
  const SegKeys keys{dev, nseg, double(nbin)};
  if (d_hits) launch_hits(ctx->stream, keys, ndat, nbin, d_hits);
  if (d_in) launch_fold(ctx->stream, keys, ndat, nbin, d_in, ics, ips, nchan, npol, ndim, d_prof, nullptr, 1, nullptr);
}

}  // namespace dspx

using namespace dspx;

struct dspx_fold {
  CtxRef hold;
  FoldPlan plan;
  FoldWeights weights;
  bool zeroed = false;
};

namespace {

void fold_plan(dspx_fold* f, dspx_ctx* ctx, uint64_t ndat, double pps, uint32_t nbin, const dspx_fold_chunk* chunks,
               uint32_t nchunk, const dspx_fold_opts* opts)
{
  f->zeroed = opts && opts->zeroed_samples;
  f->weights.w.clear();
  f->weights.ndatperweight = 0;
  f->weights.weight_idat = 0;
  // with zeroed data the weights mark nothing (Fold.C:728, 802)
  if (opts && opts->ndatperweight && !f->zeroed) {
    DSPX_REQUIRE(opts->weights && opts->nweights, DSPX_E_PARAM, "ndatperweight=%u without weights",
                 opts->ndatperweight);
    f->weights.w.assign(opts->weights, opts->weights + opts->nweights);
    f->weights.ndatperweight = opts->ndatperweight;
    f->weights.weight_idat = opts->weight_idat;
  }
  f->plan.build(ctx, ndat, pps, nbin, chunks, nchunk, true, &f->weights);
}

}  // namespace

extern "C" {

int dspx_fold_create(dspx_ctx* ctx, uint64_t ndat, double pps, uint32_t nbin,
                     const dspx_fold_chunk* chunks, uint32_t nchunk, dspx_fold** out)
{
  return dspx_fold_create_opts(ctx, ndat, pps, nbin, chunks, nchunk, nullptr, out);
}

int dspx_fold_create_opts(dspx_ctx* ctx, uint64_t ndat, double pps, uint32_t nbin, const dspx_fold_chunk* chunks,
                          uint32_t nchunk, const dspx_fold_opts* opts, dspx_fold** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && out && chunks, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipSetDevice(ctx->device));
    auto* f = new dspx_fold;
    f->hold.reset(ctx);
    try {
      fold_plan(f, ctx, ndat, pps, nbin, chunks, nchunk, opts);
    } catch (...) {
      delete f;
      throw;
    }
    *out = f;
  });
}

int dspx_fold_replan(dspx_fold* f, uint64_t ndat, double pps, uint32_t nbin, const dspx_fold_chunk* chunks,
                     uint32_t nchunk, const dspx_fold_opts* opts)
{
  dspx_ctx* ctx = f ? f->plan.ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(f && chunks, DSPX_E_PARAM, "null argument");
    fold_plan(f, ctx, ndat, pps, nbin, chunks, nchunk, opts);
  });
}

void dspx_fold_destroy(dspx_fold* f)
{
  if (!f) return;
  (void) hipSetDevice(f->plan.ctx->device);
  (void) hipStreamSynchronize(f->plan.ctx->stream);
  delete f;
}

int dspx_fold_ndat_folded(const dspx_fold* f, uint64_t* n)
{
  if (!f || !n) return DSPX_E_PARAM;
  *n = f->zeroed ? 0 : f->plan.nfold;
  return DSPX_OK;
}

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
    DSPX_REQUIRE(!f->zeroed, DSPX_E_STATE, "zeroed data: hits are counted by dspx_fold_execute_zeroed");
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

int dspx_fold_execute_zeroed(dspx_fold* f, const float* d_in, uint64_t in_chan_stride, uint64_t in_pol_stride,
                             uint32_t nchan, uint32_t npol, uint32_t ndim, float* d_prof, uint32_t* d_hits,
                             uint32_t hits_nchan, unsigned long long* d_ndat_folded)
{
  dspx_ctx* ctx = f ? f->plan.ctx : nullptr;
  return guard(ctx, [&] {
    DSPX_REQUIRE(f && d_in && d_prof && d_hits, DSPX_E_PARAM, "null argument");
    f->plan.fold_fpt(d_in, in_chan_stride, in_pol_stride, nchan, npol, ndim, d_prof, d_hits, hits_nchan,
                     d_ndat_folded);
  });
}

int dspx_fold_static(dspx_ctx* ctx, uint64_t ndat, double pps, uint32_t nbin, const dspx_fold_chunk* chunks,
                     uint32_t nchunk, const float* d_in, uint64_t in_chan_stride, uint64_t in_pol_stride,
                     uint32_t nchan, uint32_t npol, uint32_t ndim, float* d_prof, uint32_t* d_hits)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && chunks, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(!d_in == !d_prof, DSPX_E_PARAM, "input and profiles go together");
    fold_static(ctx, ndat, pps, nbin, chunks, nchunk, d_in, in_chan_stride, in_pol_stride, nchan, npol, ndim,
                d_prof, d_hits);
  });
}

}  // extern "C"
