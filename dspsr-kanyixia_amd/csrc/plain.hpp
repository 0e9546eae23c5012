// plain.hpp -- the plain FP32 GPU form of the convolution: the reference's
// CUDA::ConvolutionEngine structure (ConvolutionCUDA.cu:567-770 /
// ConvolutionCUDASpectral.cu:536-607) on MI355X: overlapping windows copied
// out, a batched library FFT, a chirp multiply kernel, the inverse library
// FFT and an overlap-discard copy -- every step a full pass through HBM.
// It is the baseline the fused three-stage form is measured against.
#pragma once
#include <hipfft/hipfft.h>

#include "dspx_internal.hpp"

namespace dspx {

struct PlainConv {
  dspx_ctx* ctx = nullptr;
  uint32_t nchan = 0, npol = 0, N = 0, nfilt_pos = 0, nfilt_neg = 0, step = 0;
  uint32_t chirp_mode = 0;
  hipfftHandle plan = 0;
  bool have_plan = false;
  uint32_t plan_batch = 0;
  DevBuf resp;     // [nchan][N] FFT-bin order
  DevBuf scratch;  // [batch][N]
  bool have_resp = false;

  void init(dspx_ctx* c, uint32_t nchan, uint32_t npol, uint32_t N, uint32_t pos, uint32_t neg,
            uint32_t mode, const dspx_dedisp_desc& d);
  void set_table(const float* host);
  void perform(const float* in, uint64_t ics, uint64_t ips, float* out, uint64_t ocs, uint64_t ops,
               uint32_t npart);
  ~PlainConv();
};

}  // namespace dspx
