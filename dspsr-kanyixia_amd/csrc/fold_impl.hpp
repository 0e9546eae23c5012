// fold_impl.hpp -- the device bin plan shared by the Fold engine and the
// fused pipeline.
#pragma once
#include <vector>

#include "dspx_internal.hpp"
#include "segments.hpp"

namespace dspx {

// WeightedTimeSeries weights of a block (Fold.C:700-733, 792-810): samples of
// a zero weight are not folded unless the data are zeroed instead.
struct FoldWeights {
  std::vector<uint16_t> w;
  uint32_t ndatperweight = 0;
  uint32_t weight_idat = 0;
};

struct FoldPlan {
  dspx_ctx* ctx = nullptr;
  uint64_t ndat = 0;
  double pps = 0;
  uint32_t nbin = 0;
  uint32_t nsub = 0;
  uint32_t nseg = 0;
  uint64_t nfold = 0;  // samples the plan folds (Fold::fold's ndat_folded without zeroed data)
  DevBuf seg;   // FoldSegment[nseg]
  DevBuf plan;  // uint32[ndat]: (isub << 16) | bin, ~0 = not folded (with a sample plan)
  bool have_sample_plan = false;
  HostStage stage;
  std::vector<dspx_fold_chunk> chunk_table;

  // Segments of the chunks (uploaded without waiting for the stream) and,
  // with sample_plan, the per-sample plan evaluated from them.
  void build(dspx_ctx* c, uint64_t ndat, double pps, uint32_t nbin, const dspx_fold_chunk* chunks,
             uint32_t nchunk, bool sample_plan = true, const FoldWeights* weights = nullptr);
  void hits(uint32_t* d_hits);
  // zhits (zeroed data, Fold.C:911-930): hits[isub][hits_nchan][nbin] counted
  // where value 0 of pol 0 is non-zero, the count added into *nfolded
  void fold_fpt(const float* d_in, uint64_t chan_stride, uint64_t pol_stride, uint32_t nchan,
                uint32_t npol, uint32_t ndim, float* d_prof, uint32_t* zhits = nullptr,
                uint32_t hits_nchan = 1, unsigned long long* nfolded = nullptr);
  // first and last sub-integration touched by samples [i0, i1)
  void sub_range(uint64_t i0, uint64_t i1, uint32_t& first, uint32_t& last) const
  {
    first = 0xffffffffu;
    last = 0;
    for (auto& c : chunk_table) {
      if (c.ndat == 0) continue;
      uint64_t a = c.idat_start, b = c.idat_start + c.ndat;
      if (b <= i0 || a >= i1) continue;
      if (c.isub < first) first = c.isub;
      if (c.isub > last) last = c.isub;
    }
  }
};

// Bin-plan entry of output sample i in segment sg (bit-identical wherever the
// plan is evaluated): (isub << 16) | bin, or ~0 for samples not folded.
__device__ __forceinline__ uint32_t seg_key(const FoldSegment& sg, uint64_t i, double nbin_d)
{
  if (sg.isub == 0xffffffffu) return 0xffffffffu;
  uint64_t m = sg.M + (i - sg.j0) * sg.D;
  double x = ldexp(double(m), sg.e - 52);
  uint32_t bin = uint32_t(x * nbin_d);
  return (sg.isub << 16) | bin;
}

// Index of the segment holding output sample i (segments sorted by j0, the
// first at j0 = 0).
__device__ __forceinline__ uint32_t seg_find(const FoldSegment* __restrict__ seg, uint32_t nseg, uint64_t i)
{
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (seg[mid].j0 <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Where the fold kernels read each sample's key: the per-sample plan, or
// the segments themselves (no per-sample buffer).
struct PlanKeys {
  const uint32_t* plan;
  __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return plan[i]; }
};
struct SegKeys {
  const FoldSegment* seg;
  uint32_t nseg;
  double nbin_d;
  __device__ __forceinline__ uint32_t operator()(uint64_t i) const
  {
    return seg_key(seg[seg_find(seg, nseg, i)], i, nbin_d);
  }
};

}  // namespace dspx
