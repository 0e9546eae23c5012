// stageb_mfma.hpp -- stage B's 4096-point row transforms on the matrix cores.
//
// Each radix-16 Stockham pass is a product of the 16-point DFT matrix with 16
// butterflies at a time, as v_mfma_f32_16x16x32_f16: A (16 x 32) is the DFT
// matrix embedded in reals (row m: output m; columns 2r, 2r+1: the real and
// imaginary part of input r), B (32 x 16) holds the 16 butterflies' inputs,
// D (16 x 16) their outputs, one MFMA for the real parts and one for the
// imaginary parts.  Both operands are split into FP16 hi and lo parts (x =
// hi + lo, lo = x - hi rounded to FP16) and the three products hi*hi, hi*lo,
// lo*hi are accumulated in FP32, which keeps about 22 bits of each operand.
//
// Lane layout of the 16x16x32 FP16 MFMA (checked on the device): lane l
// holds A[l%16][8(l/16) .. 8(l/16)+7], B[8(l/16) .. +7][l%16] and
// D[4(l/16) .. +3][l%16].  So lane (n = l%16, g = l/16) holds inputs
// r = 4g .. 4g+3 of butterfly n and receives its outputs m = 4g .. 4g+3: the
// ownership is the same before and after a pass, and a forward transform's
// last pass leaves X[j + 256m], which the inverse's first pass reads.
//
// Range: FP16 holds magnitudes below 65504.  The row is scaled on load by
// 2^-p (p from the largest value stage A can produce, so that every complex
// magnitude is at most 46341), each pass's matrix carries 1/16 (a pass then
// never grows the largest magnitude), and the chirp carries 2^floor(m/2) in
// place of most of its 1/sqrt(N); the conjugate twiddle of the store undoes
// all of it exactly (powers of two).
#pragma once
#include "stages.hpp"

namespace dspx {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MfmaScale {
  float in;     // 2^-p on load
  float chirp;  // 2^floor(m/2) on the chirp
  float out;    // 2^(p + 24 - floor(m/2)) on store
};

// The 16-point DFT matrix (DIR = -1 forward, +1 inverse) times 1/16 as this
// lane's A fragments: real-part rows (are) and imaginary-part rows (aim),
// hi and lo, for output row m = lane % 16 and inputs r = 4g' .. 4g'+3
// (g' = lane / 16), columns interleaved (re, im) per input.
template <int DIR>
__device__ __forceinline__ void dft16_fragments(int lane, f16x8& are_hi, f16x8& are_lo, f16x8& aim_hi,
                                                f16x8& aim_lo)
{
  const int m = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = 4 * g + i;
    const int e = (m * r) & 15;
    const float c = kCos16[e] * 0.0625f, s = DIR * kSin16[e] * 0.0625f;  // W^{mr} / 16
    // Yr = c*xr - s*xi ; Yi = s*xr + c*xi
    const float v[4] = {c, -s, s, c};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const _Float16 h = (_Float16) v[q];
      const _Float16 l = (_Float16) (v[q] - (float) h);
      if (q < 2) {
        are_hi[2 * i + q] = h;
        are_lo[2 * i + q] = l;
      } else {
        aim_hi[2 * i + q - 2] = h;
        aim_lo[2 * i + q - 2] = l;
      }
    }
  }
}

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// FP16 hi and lo parts of two reals: hi = fp16(a, b); lo = fp16(a - hi.x,
// b - hi.y), each difference formed and rounded once by a mixed-precision
// FMA (v_fma_mixlo/mixhi_f16 write one half of the packed result).
__device__ __forceinline__ void split2(float a, float b, f16x2& hi, f16x2& lo)
{
  hi = f16x2{(_Float16) a, (_Float16) b};
  uint32_t h = __builtin_bit_cast(uint32_t, hi);
  uint32_t l = 0;
  asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "+v"(l) : "v"(h), "v"(a));
  asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(l) : "v"(h), "v"(b));
  lo = __builtin_bit_cast(f16x2, l);
}

// One tile's 16-point DFTs: x holds this lane's 4 inputs (r = 4g + i) of
// butterfly n, and receives its 4 outputs (m = 4g + i).  CONJ: the inverse
// transform, as conj(DFT(conj(x))) with the forward matrix.
template <bool CONJ>
__device__ __forceinline__ void dft16_mfma(float2* x, const f16x8& are_hi, const f16x8& are_lo, const f16x8& aim_hi,
                                           const f16x8& aim_lo)
{
  f16x8 bh, bl;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    f16x2 h, l;
    split2(x[i].x, CONJ ? -x[i].y : x[i].y, h, l);
    bh[2 * i] = h.x;
    bh[2 * i + 1] = h.y;
    bl[2 * i] = l.x;
    bl[2 * i + 1] = l.y;
  }
  f32x4 dr = {0.f, 0.f, 0.f, 0.f}, di = {0.f, 0.f, 0.f, 0.f};
  dr = __builtin_amdgcn_mfma_f32_16x16x32_f16(are_hi, bh, dr, 0, 0, 0);
  di = __builtin_amdgcn_mfma_f32_16x16x32_f16(aim_hi, bh, di, 0, 0, 0);
  dr = __builtin_amdgcn_mfma_f32_16x16x32_f16(are_hi, bl, dr, 0, 0, 0);
  di = __builtin_amdgcn_mfma_f32_16x16x32_f16(aim_hi, bl, di, 0, 0, 0);
  dr = __builtin_amdgcn_mfma_f32_16x16x32_f16(are_lo, bh, dr, 0, 0, 0);
  di = __builtin_amdgcn_mfma_f32_16x16x32_f16(aim_lo, bh, di, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; i++) x[i] = make_float2(dr[i], CONJ ? -di[i] : di[i]);
}

// grid.x = N1 * npg * nchan, block = 256 (4 wavefronts; wave w owns
// butterflies j = 16t + n of tiles t = 4w .. 4w+3).  As k_stageB<4096, 16, 1>.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_stageB_mfma(ConvShape s, ChirpArgs c, const float2* __restrict__ twr,
                                                     uint32_t ppg, MfmaScale sc)
{
  constexpr int N2 = 4096;
  using LAY = RowLayout<N2, 256>;
  __shared__ float2 lds[N2 + N2 / 16];
  __shared__ float2 hlds[N2];
  const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
  const int n = lane & 15, g = lane >> 4;
  const uint32_t npg = (s.npart + ppg - 1) / ppg;
  uint32_t bid = blockIdx.x;
  const uint32_t k1 = bid % s.N1;
  bid /= s.N1;
  const uint32_t pg = bid % npg;
  const uint32_t chan = bid / npg;
  const uint32_t p_begin = pg * ppg;
  const uint32_t p_end = min(s.npart, p_begin + ppg);
  const uint32_t N = s.N1 * N2;
  const uint64_t plane = uint64_t(s.N1) * N2;

  // the row's chirp (bins k2 = 0..4095 of row k1), scaled, in LDS
  for (int k2 = threadIdx.x; k2 < N2; k2 += 256) {
    const float2 h = chirp_onchip(c, chan, k1 + s.N1 * k2, N);
    hlds[k2] = make_float2(h.x * sc.chirp, h.y * sc.chirp);
  }
  f16x8 fre_hi, fre_lo, fim_hi, fim_lo;
  dft16_fragments<-1>(lane, fre_hi, fre_lo, fim_hi, fim_lo);
  // this lane's elements: tile tt (t = 4*wave + tt), input/output i, at row
  // index j + 256*(4g + i) with j = 16t + n
  auto elem = [&](int tt, int i) { return uint32_t(16 * (4 * wave + tt) + n + 256 * (4 * g + i)); };
  // inter-stage twiddles W_N^{idx*k1} of the load (times 2^-p) and their
  // conjugates' factors on the store (times the inverse scale)
  float2 w[16];
#pragma unroll
  for (int tt = 0; tt < 4; tt++)
#pragma unroll
    for (int i = 0; i < 4; i++) w[4 * tt + i] = tw_N(s.tw, elem(tt, i) * k1);
  // pass-2 twiddles W_256^{r*n} (the same in every tile), r = 4g + i
  float2 t2[4];
#pragma unroll
  for (int i = 0; i < 4; i++) t2[i] = twr[((4 * g + i) * n * 16) & (N2 - 1)];
  __syncthreads();

  const uint32_t nrow = (p_end - p_begin) * s.npol;
  auto row_ptr = [&](uint32_t i) {
    uint32_t p = p_begin + i / s.npol, pol = i % s.npol;
    return s.work + ((uint64_t(chan) * s.npart + p) * s.npol + pol) * plane + uint64_t(k1) * N2;
  };
  // the LDS position of pass output m of butterfly j for a pass of span ns
  auto out_idx = [&](uint32_t j, int m, uint32_t ns) { return (j / ns) * ns * 16 + (j % ns) + uint32_t(m) * ns; };

  for (uint32_t row = 0; row < nrow; row++) {
    float2* rp = row_ptr(row);
    float2 v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const float2 x = rp[elem(e / 4, e % 4)];
      const float2 ww = w[e];
      v[e] = make_float2((x.x * ww.x - x.y * ww.y) * sc.in, (x.x * ww.y + x.y * ww.x) * sc.in);
    }
    // forward: passes of span 1, 16, 256; inverse: the same spans
#pragma unroll
    for (int dir = 0; dir < 2; dir++) {
#pragma unroll
      for (int pass = 0; pass < 3; pass++) {
        const uint32_t ns = pass == 0 ? 1u : pass == 1 ? 16u : 256u;
        if (pass > 0) {
          // inputs x[j + 256r] of this pass from the previous pass's outputs
#pragma unroll
          for (int e = 0; e < 16; e++) v[e] = lds[LAY::idx(elem(e / 4, e % 4), 0)];
          // twiddles W_{16 ns}^{r (j % ns)}, conjugated for the inverse
#pragma unroll
          for (int e = 0; e < 16; e++) {
            const int i = e % 4;
            float2 tw;
            if (pass == 1) {
              tw = t2[i];
            } else {
              const uint32_t j = 16 * (4 * wave + e / 4) + n;
              tw = twr[((4 * g + i) * j) & (N2 - 1)];
            }
            v[e] = dir == 0 ? cmul(v[e], tw) : cmulc(v[e], tw);
          }
        }
#pragma unroll
        for (int tt = 0; tt < 4; tt++) {
          if (dir == 0)
            dft16_mfma<false>(v + 4 * tt, fre_hi, fre_lo, fim_hi, fim_lo);
          else
            dft16_mfma<true>(v + 4 * tt, fre_hi, fre_lo, fim_hi, fim_lo);
        }
        if (pass < 2) {
          __syncthreads();  // the previous reads of the row are done
#pragma unroll
          for (int e = 0; e < 16; e++) {
            const uint32_t j = 16 * (4 * wave + e / 4) + n;
            lds[LAY::idx(out_idx(j, 4 * g + e % 4, ns), 0)] = v[e];
          }
          __syncthreads();
        }
      }
      if (dir == 0) {
        // X[j + 256m] times the chirp of bin k2 = j + 256m
#pragma unroll
        for (int e = 0; e < 16; e++) v[e] = cmul(v[e], hlds[elem(e / 4, e % 4)]);
      }
    }
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const float2 ww = w[e];
      const float2 y = make_float2(v[e].x * sc.out, v[e].y * sc.out);
      rp[elem(e / 4, e % 4)] = cmulc(y, ww);
    }
  }
}

}  // namespace dspx
