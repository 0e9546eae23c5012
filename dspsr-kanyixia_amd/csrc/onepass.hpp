// onepass.hpp -- the LoadToFold chain in one kernel for short transforms.
//
// For an N-point overlap-save transform that fits one LDS row (N = 2^7 ..
// 2^13), every window is read once as raw 8/16-bit samples and leaves the
// chip only as fold histograms (or, for the filterbank output, as its kept
// samples): unpack -> forward N-point Stockham FFT -> dedispersion chirp
// (FP64 phase, on chip) -> inverse FFT -> overlap discard -> detection ->
// phase-bin accumulation in LDS.  The HBM traffic per input sample is the
// raw bytes of the windows, r*B_in (r = N / step), against r*(B_in + 32) for
// the three-stage form of the longer transforms.
//
// A workgroup owns one channel and a range of its windows (ppg of them, a
// multiple of WF), WF windows in flight, one per group of NTW = N/16
// threads, each with its own LDS exchange row; the histogram of the
// workgroup's sub-integration slots follows the rows.  The channel's chirp
// is evaluated once per workgroup, into LDS after the rows (N <= 4096) or
// into registers (the bins each thread holds).
//
// Reference: Convolution::transformation (Signal/General/Convolution.C:
// 646-813), Detection (Detection.C:225-573), Fold::fold (Fold.C:790-949),
// chained as LoadToFold1.C:109-258 chains them.
#pragma once
#include "stages.hpp"

namespace dspx {

template <int N> struct OnePass {
  // complex values per thread per pol: 16, so that a transform of up to 1024
  // points is one wavefront's (or part of one), whose exchanges need no
  // workgroup barrier, and the fold takes 16 consecutive samples per lane
  static constexpr int E = 16;
  static constexpr int NTW = N / E;               // threads per window
  static constexpr int NT = NTW < 256 ? 256 : NTW;  // threads per workgroup
  static constexpr int WF = NT / NTW;             // windows in flight
  static constexpr int LROW = N + N / 16;         // exchange row (RowLayout padding)
  // exchange rows, then the channel's chirp (N values, CH_LDS)
  static constexpr size_t lds_bytes = (size_t(WF) * LROW + (N <= 4096 ? N : 0)) * sizeof(float2);
  static constexpr int R0 = RadixPlan<N, E>::r[0];                      // first forward = last inverse pass
  static constexpr int RL = RadixPlan<N, E>::r[RadixPlan<N, E>::n - 1];  // last forward = first inverse pass
  static constexpr bool WAVE = NTW <= 64;         // windows within one wavefront: no workgroup barriers
  // the channel's chirp in LDS after the rows (N <= 4096), else in registers
  static constexpr bool CH_LDS = N <= 4096;
};

constexpr int kOnePassMinLog2 = 7, kOnePassMaxLog2 = 13;
// The longest transform the fused form (DSPX_FORM_FUSED) runs in one pass: at
// 8192 points (one window per 512-thread workgroup, the chirp in registers,
// scratch spills) the three stages were faster on each beam measured (an
// 8-bit 512-channel beam 10.1 against 11.7 ms per block, a 16-bit one 24.4
// against 27.2), at 4096 points the one-pass kernel (7.1 against 11.0 ms).
constexpr int kOnePassFusedMaxLog2 = 12;

// LDS bytes of the exchange rows for an n-point transform (host side).
inline uint32_t onepass_ntw(uint32_t n) { return n / 16; }
inline uint32_t onepass_threads(uint32_t n) { return onepass_ntw(n) < 256 ? 256 : onepass_ntw(n); }
inline uint32_t onepass_wf(uint32_t n) { return onepass_threads(n) / onepass_ntw(n); }
inline size_t onepass_lds_bytes(uint32_t n)
{
  return (size_t(onepass_wf(n)) * (n + n / 16) + (n <= 4096 ? n : 0)) * sizeof(float2);
}

struct OnePassArgs {
  uint32_t nchan, npol, npart;
  uint32_t step, nfilt_pos;
  uint32_t ppg;   // windows per workgroup, a multiple of WF
  uint32_t ngrp;  // workgroups per channel
};

// grid.x = nchan * ngrp (group fastest), block = OnePass<N>::NT, dynamic LDS
// = lds_bytes (+ the fold histogram).  Fold partials go to
// fa.partial[chan][grp][nslot][hrow]; fa.keys is the key plan [npart][N]
// (slots relative to the first sub-integration of the window's workgroup).
// ZH: the zap mask and per-channel hits of zeroed data (fold_lane).
template <int N, int MODE, int NV, class LD, bool ZH = false>
__global__ __launch_bounds__(OnePass<N>::NT) void k_onepass(LD ld, OnePassArgs a, ChirpArgs c, FilterbankArgs fb,
                                                            FoldArgs fa, const float2* __restrict__ tw)
{
  using S = OnePass<N>;
  constexpr int E = S::E, NTW = S::NTW, NT = S::NT, WF = S::WF, R0 = S::R0, RL = S::RL;
  using LAY = RowLayout<N, NTW>;
  extern __shared__ float2 smem[];
  const int slot = WF > 1 ? int(threadIdx.x) / NTW : 0;
  const int tid0 = WF > 1 ? int(threadIdx.x) % NTW : int(threadIdx.x);
  float2* const lds0 = smem + slot * S::LROW;
  float2* chl = smem + WF * S::LROW;
  float* hist = reinterpret_cast<float*>(chl + (S::CH_LDS ? N : 0));

  const uint32_t grp = blockIdx.x % a.ngrp;
  const uint32_t chan = blockIdx.x / a.ngrp;
  const uint32_t w_begin = grp * a.ppg;
  const uint32_t w_end = min(a.npart, w_begin + a.ppg);

  const uint32_t hist_len = MODE == kStageCFold ? fa.nslot * fa.hrow : 0;
  if (MODE == kStageCFold)
    for (uint32_t i = threadIdx.x; i < hist_len; i += NT) hist[i] = 0.f;
  // fold groups: QF consecutive samples per thread (runs of equal bins
  // accumulated in registers, so that short periods with bins a few samples
  // wide add to the LDS histogram once per run rather than once per sample)
  constexpr int QF = E;
  constexpr int QG = E / QF;

  // the channel's chirp, evaluated once per workgroup: into LDS (read after
  // the first exchange's barrier), or the bins X[j + r*N/RL] (j = tid +
  // q*NTW) this thread holds after the forward transform into registers
  float2 ch[S::CH_LDS ? 1 : E];
  if constexpr (S::CH_LDS) {
    for (uint32_t k = threadIdx.x; k < uint32_t(N); k += NT) chl[k] = chirp_onchip(c, chan, k, N);
  } else {
#pragma unroll
    for (int q = 0; q < E / RL; q++)
#pragma unroll
      for (int r = 0; r < RL; r++) ch[q * RL + r] = chirp_onchip(c, chan, uint32_t(tid0 + q * NTW + r * (N / RL)), N);
  }
  // the histogram's zeros and the chirp table are seen by every wavefront
  // (the loop below has no workgroup barrier where windows are a
  // wavefront's)
  __syncthreads();
  int tid = tid0;
  float2* lds = lds0;
  // two passes of one radix (256 points): the second pass's twiddles, the
  // same for every window (conjugated in the inverse), held in registers
  // rather than rebuilt from the table per window and pol (a 512-channel
  // beam: 16.6 -> 16.2 ms per block with 16-bit samples, 4.5 -> 4.3 ms with
  // 2048 bins)
  constexpr bool TW2 = RadixPlan<N, E>::n == 2 && RadixPlan<N, E>::r[0] == RadixPlan<N, E>::r[1];
  float2 wq[TW2 ? E : 1];
  if constexpr (TW2) pass_twiddles<N, 1, NTW, E, E>(tw, tid0, wq);

  // every slot runs the same number of iterations (the exchanges are
  // workgroup barriers); a slot past the group's last window transforms its
  // group's first window again and neither stores nor folds it
  const uint32_t niter = (w_end - w_begin + WF - 1) / WF;
  // the raw words of a window (both pols packed), in the forward plan's
  // first-pass order, loaded at the start of its iteration: loading the next
  // window's during this one's transforms held 16-32 more registers, which
  // cost the 16-bit kernels occupancy (one wave per SIMD from 512 points,
  // with spills to AGPRs: a 2048-point beam 5.33 ms per block against 3.38
  // without) and gained nothing for 8-bit samples
  using Word = typename LD::Word;
  auto load_words = [&](uint32_t it_, Word* wd) {
    const uint32_t w_ = w_begin + it_ * WF + uint32_t(slot);
    const uint64_t t_ = uint64_t(w_ < w_end ? w_ : w_begin) * a.step;
#pragma unroll
    for (int q = 0; q < E / R0; q++)
#pragma unroll
      for (int r = 0; r < R0; r++) wd[q * R0 + r] = ld.word(chan, t_, uint32_t(tid0 + q * NTW + r * (N / R0)));
  };
  for (uint32_t it = 0; it < niter; it++) {
    const uint32_t w = w_begin + it * WF + uint32_t(slot);
    const bool valid = w < w_end;
    Word wcur[E];
    load_words(it, wcur);
    // the lane's index arithmetic (LDS addresses of every pass) is redone
    // per window rather than held in registers across the loop
    tid = opaque_i(tid0);
    lds = smem + slot * S::LROW;
    float2 y0[E], v[E];
    // one polarisation: transform, chirp, transform back, and (folding)
    // regroup into dst; written out for pol 0 (into y0) and pol 1 (into v)
    // rather than looped over, so that neither the word's half nor the
    // destination array is selected per value (a select per float of
    // both arrays in the rolled loop)
    auto one_pol = [&](const uint32_t pol, float2* dst) __attribute__((always_inline)) {
#pragma unroll
      for (int e = 0; e < E; e++) v[e] = ld.pol(wcur[e], pol);
      // other lengths: the pass twiddles are reloaded per window (opaque)
      // rather than held in registers across the loop (registers go to
      // occupancy)
      if constexpr (TW2)
        fft_regs_2w<LAY, N, NTW, -1>(lds, v, wq, tw, tid);
      else
        fft_regs<LAY, N, 1, NTW, -1>(lds, v, opaque(tw), tid);
      // X[j + r*N/RL] (j = tid + q*NTW) times the chirp of its bin
#pragma unroll
      for (int q = 0; q < E / RL; q++)
#pragma unroll
        for (int r = 0; r < RL; r++)
          v[q * RL + r] = cmul(v[q * RL + r], S::CH_LDS ? chl[tid + q * NTW + r * (N / RL)] : ch[q * RL + r]);
      if constexpr (TW2)
        fft_regs_2w<LAY, N, NTW, +1>(lds, v, wq, tw, tid);
      else
        fft_regs<LAY, N, 1, NTW, +1>(lds, v, opaque(tw), tid);
      // v[q*R0 + r] = y[tid + q*NTW + r*N/R0]
      if constexpr (MODE == kStageCFilterbank) {
        if (valid) {
          float* o = fb.out + uint64_t(chan) * fb.chan_stride + uint64_t(pol) * fb.pol_stride;
#pragma unroll
          for (int q = 0; q < E / R0; q++)
#pragma unroll
            for (int r = 0; r < R0; r++) {
              const uint32_t t = uint32_t(tid + q * NTW + r * (N / R0));
              if (t < a.nfilt_pos || t >= a.nfilt_pos + a.step) continue;
              *reinterpret_cast<float2*>(o + 2 * (uint64_t(w) * a.step + (t - a.nfilt_pos))) = v[q * R0 + r];
            }
        }
      } else {
        // regroup through LDS: each thread then holds QG groups of QF
        // consecutive samples, (g*NTW + tid)*QF + k
        xsync<NTW>();
#pragma unroll
        for (int q = 0; q < E / R0; q++)
#pragma unroll
          for (int r = 0; r < R0; r++) lds[LAY::idx(tid + q * NTW + r * (N / R0), 0)] = v[q * R0 + r];
        xsync<NTW>();
#pragma unroll
        for (int g = 0; g < QG; g++)
#pragma unroll
          for (int k = 0; k < QF; k++) dst[g * QF + k] = lds[LAY::idx((g * NTW + tid) * QF + k, 0)];
      }
    };
    one_pol(0u, y0);
    if (a.npol == 2) one_pol(1u, v);
    if constexpr (MODE == kStageCFold) {
      const uint32_t* kw = fa.keys + uint64_t(valid ? w : w_begin) * N;
#pragma unroll
      for (int g = 0; g < QG; g++) {
        uint32_t kq[QF];
        load_keys<QF>(kw + (g * NTW + tid) * QF, kq);
        if (!valid)
#pragma unroll
          for (int k = 0; k < QF; k++) kq[k] = 0xffffffffu;
        float2 b[QF];
#pragma unroll
        for (int k = 0; k < QF; k++) b[k] = a.npol == 2 ? v[g * QF + k] : make_float2(0.f, 0.f);
        // output index of the group's first sample
        const int32_t i0 = int32_t(w * a.step + uint32_t(g * NTW + tid) * QF) - int32_t(a.nfilt_pos);
        fold_lane<NV, QF, ZH>(hist, fa, a.npol, kq, y0 + g * QF, b, chan, i0);
      }
    }
  }
  if constexpr (MODE == kStageCFold) {
    __syncthreads();
    float* part = fa.partial + (uint64_t(chan) * a.ngrp + grp) * hist_len;
    for (uint32_t i = threadIdx.x; i < hist_len; i += NT) part[i] = hist[i];
  }
}

}  // namespace dspx
