// binplan.cpp -- host entry points of the exact, parallel evaluation of the
// reference fold recurrence (the segment walk is in segments.hpp).
#include <cmath>
#include <vector>

#include "dspx_internal.hpp"
#include "segments.hpp"

#pragma clang fp contract(off)

namespace dspx {

void fold_segments(double phi, double pps, uint64_t j_begin, uint64_t n, uint32_t isub,
                   std::vector<FoldSegment>& out)
{
  fold_segments_to(phi, pps, j_begin, n, isub, [&](const FoldSegment& s, uint64_t) { out.push_back(s); });
}

}  // namespace dspx

extern "C" int dspx_fold_binplan_host(double phi, double phase_per_sample, uint32_t nbin, uint64_t ndat,
                                      uint32_t* bins)
{
  if (!bins || nbin == 0 || !(phase_per_sample > 0.0 && phase_per_sample < 0.5)) return DSPX_E_PARAM;
  std::vector<dspx::FoldSegment> segs;
  dspx::fold_segments(phi, phase_per_sample, 0, ndat, 0, segs);
  for (size_t s = 0; s < segs.size(); s++) {
    uint64_t end = s + 1 < segs.size() ? segs[s + 1].j0 : ndat;
    for (uint64_t j = segs[s].j0; j < end; j++) {
      double x = std::ldexp(double(segs[s].M + (j - segs[s].j0) * segs[s].D), segs[s].e - 52);
      bins[j] = uint32_t(x * double(nbin));
    }
  }
  return DSPX_OK;
}

extern "C" int dspx_fold_nsegments_host(double phi, double phase_per_sample, uint64_t ndat, uint64_t* nseg)
{
  if (!nseg || !(phase_per_sample > 0.0 && phase_per_sample < 0.5)) return DSPX_E_PARAM;
  std::vector<dspx::FoldSegment> segs;
  dspx::fold_segments(phi, phase_per_sample, 0, ndat, 0, segs);
  *nseg = segs.size();
  return DSPX_OK;
}
