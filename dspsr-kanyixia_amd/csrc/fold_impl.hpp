// fold_impl.hpp -- the device bin plan shared by the Fold engine and the
// fused pipeline.
#pragma once
#include <vector>

#include "dspx_internal.hpp"

namespace dspx {

struct FoldPlan {
  dspx_ctx* ctx = nullptr;
  uint64_t ndat = 0;
  double pps = 0;
  uint32_t nbin = 0;
  uint32_t nsub = 0;
  uint32_t nseg = 0;
  DevBuf seg;   // FoldSegment[nseg]
  DevBuf plan;  // uint32[ndat]: (isub << 16) | bin, ~0 = not folded
  std::vector<dspx_fold_chunk> chunk_table;
  std::vector<uint32_t> sub_of_sample_first;

  void build(dspx_ctx* c, uint64_t ndat, double pps, uint32_t nbin, const dspx_fold_chunk* chunks,
             uint32_t nchunk);
  void hits(uint32_t* d_hits);
  void fold_fpt(const float* d_in, uint64_t chan_stride, uint64_t pol_stride, uint32_t nchan,
                uint32_t npol, uint32_t ndim, float* d_prof);
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

}  // namespace dspx
