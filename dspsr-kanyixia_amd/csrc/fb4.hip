// fb4.hip -- stage B of the convolving filterbank's four-step form
// (k_stageB_fb in stages.hpp): each row's forward transform, the response
// product and the sub-bands' backward transforms, for windows of 2^13 to 2^22
// points.  Kept apart from conv.hip so that its instantiations compile in
// parallel with it.
//
// Reference: FilterbankEngineCPU::perform (Signal/General/
// FilterbankEngineCPU.C:184-370).
#include "fb4.hpp"

namespace dspx {

void fb4_stage_b(uint32_t N2, uint32_t F2, unsigned nb, hipStream_t st, const ConvShape& s, const ChirpArgs& c,
                 const float2* twr, const float2* twf, uint32_t ppg, uint32_t nsub)
{
  switch (N2 * 65536u + F2) {
#define CASE_FB(n2, f2) \
  case n2 * 65536u + f2: k_stageB_fb<n2, f2><<<nb, n2 / 16, 0, st>>>(s, c, twr, twf, ppg, nsub); break;
    CASE_FB(512, 8) CASE_FB(512, 16) CASE_FB(512, 32) CASE_FB(512, 64) CASE_FB(512, 128) CASE_FB(512, 256)
    CASE_FB(1024, 8) CASE_FB(1024, 16) CASE_FB(1024, 32) CASE_FB(1024, 64) CASE_FB(1024, 128) CASE_FB(1024, 256)
    CASE_FB(1024, 512)
    CASE_FB(2048, 8) CASE_FB(2048, 16) CASE_FB(2048, 32) CASE_FB(2048, 64) CASE_FB(2048, 128) CASE_FB(2048, 256)
    CASE_FB(2048, 512) CASE_FB(2048, 1024)
    CASE_FB(4096, 8) CASE_FB(4096, 16) CASE_FB(4096, 32) CASE_FB(4096, 64) CASE_FB(4096, 128) CASE_FB(4096, 256)
    CASE_FB(4096, 512) CASE_FB(4096, 1024) CASE_FB(4096, 2048)
#undef CASE_FB
    default: fail(DSPX_E_INTERNAL, "four-step filterbank rows N2=%u F2=%u", N2, F2);
  }
  DSPX_HIP(hipGetLastError());
}

}  // namespace dspx
