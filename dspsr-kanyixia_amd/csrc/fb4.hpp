// fb4.hpp -- launcher of the convolving filterbank's four-step stage B
// (k_stageB_fb, stages.hpp), instantiated in fb4.hip.
#pragma once
#include "stages.hpp"

namespace dspx {

// Whether k_stageB_fb is built for N2-point rows in F2-point segments:
// N2 = 512 .. 4096, F2 = 8 .. N2/2.
inline bool fb4_rows_ok(uint32_t N2, uint32_t F2)
{
  return N2 >= 512 && N2 <= 4096 && !(N2 & (N2 - 1)) && F2 >= 8 && F2 <= N2 / 2 && !(F2 & (F2 - 1));
}

// Launches k_stageB_fb<N2, F2> with nb workgroups of N2/16 threads.
void fb4_stage_b(uint32_t N2, uint32_t F2, unsigned nb, hipStream_t st, const ConvShape& s, const ChirpArgs& c,
                 const float2* twr, const float2* twf, uint32_t ppg, uint32_t nsub);

}  // namespace dspx
