// stages.hpp -- the three fused kernels of the overlap-save convolution.
//
// An nfft-point transform is split N = N1 * N2 (four-step):
//   input  x[N2*n1 + n2]          (n1 < N1, n2 < N2)
//   stage A: per column n2, forward N1-point FFT over n1 -> work[k1][n2]
//            (unpack fused in)
//   stage B: per row k1, times W_N^{n2*k1}, forward N2-point FFT over n2
//            (bin k = k1 + N1*k2), times the chirp H[k], inverse N2-point FFT,
//            times W_N^{-n2*k1} (chirp and twiddles generated once per row and
//            reused for every part and polarisation)
//   stage C: per column n2, inverse N1-point FFT over k1 -> y[N2*n1 + n2],
//            overlap discard, then either the filterbank store or detection
//            plus phase-bin accumulation in LDS.
// Each stage moves 8 bytes per complex sample per direction through HBM once.
#pragma once
#include "dspx_internal.hpp"
#include "fft_core.hpp"

namespace dspx {

constexpr int kColTile = 4096;  // complex values per pol in a column tile
constexpr int kColThreads = 256;

// ------------------------------------------------------------ input loaders

// 8-bit complex samples; both pols of (t, chan) in one word at
// raw + chan*cstride + t*tstride (TFP: cstride = npol*2, tstride = nchan*npol*2;
// channel-major after the transpose pass: cstride = ndat*npol*2, tstride = npol*2).
struct LoadI8 {
  const uint8_t* raw;
  uint64_t tstride;  // bytes between time samples
  uint64_t cstride;  // bytes between channels
  uint32_t npol;
  uint32_t offset_binary;
  float scale;
  // offset binary b is two's complement b ^ 0x80 (b - 127.5 = int8(b ^ 0x80)
  // + 0.5 exactly), so one exclusive or of the whole word with a uniform
  // mask serves both encodings
  __device__ __forceinline__ uint32_t flip() const { return offset_binary ? 0x80808080u : 0u; }
  __device__ __forceinline__ float cvt(uint32_t byte) const { return (float(int8_t(uint8_t(byte))) + 0.5f) * scale; }
  // packed raw word of both pols of one sample
  using Word = uint32_t;
  __device__ __forceinline__ Word word(uint32_t chan, uint64_t t) const
  {
    const uint8_t* q = raw + t * tstride + uint64_t(chan) * cstride;
    return npol == 2 ? *reinterpret_cast<const uint32_t*>(q) : uint32_t(*reinterpret_cast<const uint16_t*>(q));
  }
  // the word of sample t0 + j, where j * tstride < 2^32: the per-sample part
  // of the address is a 32-bit product (the window's base once per window)
  __device__ __forceinline__ Word word(uint32_t chan, uint64_t t0, uint32_t j) const
  {
    const uint8_t* q = raw + t0 * tstride + uint64_t(chan) * cstride + j * uint32_t(tstride);
    return npol == 2 ? *reinterpret_cast<const uint32_t*>(q) : uint32_t(*reinterpret_cast<const uint16_t*>(q));
  }
  __device__ __forceinline__ float2 pol(Word w, uint32_t ipol) const
  {
    w = (w ^ flip()) >> (16 * ipol);
    return make_float2(cvt(w & 0xff), cvt((w >> 8) & 0xff));
  }
  __device__ __forceinline__ void operator()(uint32_t chan, uint64_t t, float2& a, float2& b) const
  {
    Word w = word(chan, t);
    a = pol(w, 0);
    b = npol == 2 ? pol(w, 1) : make_float2(0.f, 0.f);
  }
};

// 16-bit complex samples: float(v) + 0.5 (offset binary: v ^ 0x8000); strides
// in int16 elements, laid out as for LoadI8.
struct LoadI16 {
  const int16_t* raw;
  uint64_t tstride;
  uint64_t cstride;
  uint32_t npol;
  uint32_t offset_binary;
  __device__ __forceinline__ uint32_t flip() const { return offset_binary ? 0x80008000u : 0u; }
  __device__ __forceinline__ float cvt(int16_t v) const { return float(v) + 0.5f; }
  using Word = uint2;
  __device__ __forceinline__ Word word(uint32_t chan, uint64_t t) const
  {
    const int16_t* q = raw + t * tstride + uint64_t(chan) * cstride;
    return npol == 2 ? *reinterpret_cast<const uint2*>(q) : make_uint2(*reinterpret_cast<const uint32_t*>(q), 0u);
  }
  __device__ __forceinline__ Word word(uint32_t chan, uint64_t t0, uint32_t j) const
  {
    const int16_t* q = raw + t0 * tstride + uint64_t(chan) * cstride + j * uint32_t(tstride);
    return npol == 2 ? *reinterpret_cast<const uint2*>(q) : make_uint2(*reinterpret_cast<const uint32_t*>(q), 0u);
  }
  __device__ __forceinline__ float2 pol(Word w, uint32_t ipol) const
  {
    const uint32_t v = (ipol ? w.y : w.x) ^ flip();
    return make_float2(cvt(int16_t(v & 0xffff)), cvt(int16_t(v >> 16)));
  }
  __device__ __forceinline__ void operator()(uint32_t chan, uint64_t t, float2& a, float2& b) const
  {
    Word w = word(chan, t);
    a = pol(w, 0);
    b = npol == 2 ? pol(w, 1) : make_float2(0.f, 0.f);
  }
};

// FPT float complex (the Convolution engine's unpacked TimeSeries).
struct LoadF32 {
  const float* in;
  uint64_t chan_stride;  // floats
  uint64_t pol_stride;   // floats
  uint32_t npol;
  // the "word" is the sample's offset; each pol is loaded when converted
  using Word = uint64_t;
  __device__ __forceinline__ Word word(uint32_t chan, uint64_t t) const { return uint64_t(chan) * chan_stride + 2 * t; }
  __device__ __forceinline__ Word word(uint32_t chan, uint64_t t0, uint32_t j) const { return word(chan, t0 + j); }
  __device__ __forceinline__ float2 pol(Word w, uint32_t ipol) const
  {
    return *reinterpret_cast<const float2*>(in + w + (ipol ? pol_stride : 0));
  }
  __device__ __forceinline__ void operator()(uint32_t chan, uint64_t t, float2& a, float2& b) const
  {
    Word w = word(chan, t);
    a = pol(w, 0);
    b = npol == 2 ? pol(w, 1) : make_float2(0.f, 0.f);
  }
};

// --------------------------------------------------------------- parameters

struct ConvShape {
  float2* work;        // [nchan][npart][npol][N1][N2]
  uint32_t N1, N2, log2N;
  uint32_t nchan, npol, npart;
  uint32_t step, nfilt_pos;
  Tw2 tw;              // W_N^m
};

struct ChirpArgs {
  uint32_t mode;       // DSPX_CHIRP_*
  float snorm;         // float(1/sqrt(N))
  double binwidth;     // chanwidth / N
  double half_cw;      // 0.5 * chanwidth
  const double* coeff; // per channel
  const double* cf;    // per channel centre frequency (MHz)
  const double* fdelay; // per channel fractional inter-channel delay (us), 0 without -K
  const float2* table; // DSPX_CHIRP_TABLE: [nchan][N1][N2] (transposed bins)
  // real-sampled input: bins in frequency order (no dual-sideband swap) and
  // the first channel's DC zero (Response.C:335-345, 418-420)
  uint32_t nyquist = 0;
};

// The SpectralKurtosis zap mask a fused fold applies (LoadToFold1.C:
// 265-377): output sample idat of channel chan is zeroed, so neither folded
// nor counted, where its block of M samples is flagged, mask[idat / M][chan],
// or lies past the nblk blocks the estimator formed (the cleaned series of
// SpectralKurtosis::mask is nblk*M samples, SpectralKurtosis.C:1465-1566).
struct ZapArgs {
  const uint8_t* mask;  // [nblk][nchan]
  uint32_t M, nblk, nchan;
  uint64_t magic;       // ceil(2^64 / M): idat / M = umul64hi(idat, magic) for idat < 2^32
};

struct FoldArgs {
  const uint32_t* keys;      // [npart][nfft]: (slot << 16) | bin of each transform output, ~0 skip
  float* partial;            // [nchan][npart][G][nslot][hrow]
  uint32_t nbin, nval, ndim, nslot, G;
  uint32_t state, npol_out;
  // profile offset of detected value i of bin b: voff[i] + b * ndim
  uint32_t voff[4];
  // histogram floats per slot: nval*nbin, then with per-channel hits (zeroed
  // data) nbin hit counts
  uint32_t hrow;
  ZapArgs zap;
  // bins narrower than 64 samples (the launch's choice of ownership at 32
  // and 64 points, FoldOwn)
  uint32_t narrow = 1;
};

struct FilterbankArgs {
  float* out;            // FPT complex
  uint64_t chan_stride;  // floats
  uint64_t pol_stride;   // floats
  // the convolving filterbank's four-step form (k_stageB_fb): stage B left
  // sub-band s's backward transform at row positions s*f2 + ob, so output
  // sample ob + f2*n1 of sub-band s goes to output channel chan*nsub + s,
  // kept where it lies in [nfilt_pos, nfilt_pos + nkeep); f2 = 0: the
  // convolution's discard
  uint32_t f2 = 0, lf2 = 0, nsub = 1, nkeep = 0;
};

__device__ __forceinline__ float2 tw_N(const Tw2& t, uint32_t m)
{
  return cmul(t.lo[m & ((1u << t.h) - 1u)], t.hi[m >> t.h]);
}

// ------------------------------------------------------------------ stage A
// grid.x = ncb * npart * nchan, block = 256.  Tile: N1 rows x C columns.
template <int N1, class LD>
__global__ __launch_bounds__(kColThreads) void k_stageA(LD ld, ConvShape s, const float2* __restrict__ twc)
{
  constexpr int C = kColTile / N1;
  constexpr int E = kColTile / kColThreads;  // 16
  constexpr int R0 = FirstRadix<N1, 16, -1>::v;
  constexpr int RL = LastRadix<N1, 16, -1>::v;
  using LAY = ColLayout<N1, C, C, kColThreads>;
  __shared__ float2 lds[kColTile];

  const int tid = threadIdx.x;
  const uint32_t ncb = s.N2 / C;
  uint32_t bid = blockIdx.x;
  const uint32_t cb = bid % ncb;
  bid /= ncb;
  const uint32_t p = bid % s.npart;
  const uint32_t chan = bid / s.npart;
  const uint32_t c0 = cb * C;
  const uint64_t t0 = uint64_t(p) * s.step + c0;

  // raw words of both pols stay packed until their pol is transformed
  typename LD::Word w[E];
#pragma unroll
  for (int q = 0; q < E / R0; q++) {
    int g = tid + q * kColThreads;
    int b = g % C, j = g / C;
#pragma unroll
    for (int r = 0; r < R0; r++) w[q * R0 + r] = ld.word(chan, t0 + uint64_t(s.N2) * uint32_t(j + r * (N1 / R0)) + b);
  }

  const uint64_t plane = uint64_t(s.N1) * s.N2;
  float2* wbase = s.work + (uint64_t(chan) * s.npart + p) * s.npol * plane;
  for (uint32_t pol = 0; pol < s.npol; pol++) {
    float2 v[E];
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = ld.pol(w[i], pol);
    fft_regs<LAY, N1, C, kColThreads, -1>(lds, v, twc, tid);
    float2* out = wbase + pol * plane;
#pragma unroll
    for (int q = 0; q < E / RL; q++) {
      int g = tid + q * kColThreads;
      int b = g % C, j = g / C;
#pragma unroll
      for (int r = 0; r < RL; r++) {
        uint32_t k1 = j + r * (N1 / RL);
        uint32_t n2 = c0 + b;
        out[uint64_t(k1) * s.N2 + n2] = v[q * RL + r];
      }
    }
  }
}

// ------------------------------------------------------------------ stage B

// Chirp of FFT bin k of channel chan.  The phase is evaluated in FP64 with
// the reference's operation order (Dedispersion.C:166-186); in parity mode it
// is rounded to float as the reference stores it (Dedispersion.C:104-111).
// The (possibly huge) angle is reduced modulo 2 pi in FP64 (two-term
// Cody-Waite with FMA), and only the reduced angle, |r| <= pi, goes through
// single-precision sincos.
__device__ __forceinline__ float2 chirp_onchip(const ChirpArgs& c, uint32_t chan, uint32_t k, uint32_t N)
{
#pragma clang fp contract(off)
  if (c.nyquist && chan == 0 && k == 0) return make_float2(0.f, 0.f);
  uint32_t ipt = c.nyquist ? k : (k + (N >> 1)) & (N - 1);
  double freq = double(ipt) * c.binwidth - c.half_cw;
  double cf = c.cf[chan];
  // coeff*sqr(freq)/(chan_freq+freq) + delay_phase, delay_phase = -2 pi freq delay
  double delay_phase = -6.283185307179586 * freq * c.fdelay[chan];
  double phase = c.coeff[chan] * (freq * freq) / (cf + freq) + delay_phase;
  double th = c.mode == DSPX_CHIRP_PARITY ? double(float(phase)) : phase;
  double n = rint(th * 0.15915494309189535);          // 1 / (2 pi)
  double r = fma(-n, 6.283185307179586, th);           // 2 pi, high part
  r = fma(-n, 2.4492935982947064e-16, r);              // 2 pi - high part
  float sn, cs;
  sincosf(float(r), &sn, &cs);
  return make_float2(c.snorm * cs, c.snorm * sn);
}

// Rows in flight per stage B workgroup for N2-point rows with E values per
// thread: the fastest of 1, 2, 4, ... measured on a 64-channel beam
// (256 points: 3.4x faster than one row per 16-thread workgroup; 512: 1.9x
// with E = 8; 2048: 1.25x; 1024 and 4096 points: one row).
constexpr int stage_b_rows(int n2, int e)
{
  return e == 16 ? (n2 == 256 ? 16 : n2 == 512 || n2 == 2048 ? 2 : 1) : (n2 == 512 ? 4 : 1);
}

// grid.x = N1 * npg * nchan, block = RW * N2/E.  A workgroup owns row k1 of
// channel chan for the parts [pg*ppg, min(npart, (pg+1)*ppg)): the chirp of
// the row is generated once and applied to every part and polarisation.
// Rows shorter than 1024 points fill less than a wavefront with N2/E
// threads, so RW of the workgroup's rows are in flight at once, one per
// group of N2/E threads, each with its own LDS exchange buffer; the chirp
// and the inter-stage twiddles (functions of k1 only) are shared.
template <int N2, int E, int RW>
__global__ __launch_bounds__(RW * N2 / E) void k_stageB(ConvShape s, ChirpArgs c, const float2* __restrict__ twr,
                                                        uint32_t ppg)
{
  constexpr int NT = N2 / E;
  constexpr int LROW = N2 + N2 / 16;
  using LAY = RowLayout<N2, NT>;
  __shared__ float2 lds_all[RW * LROW];
  constexpr int RF = FirstRadix<N2, E, -1>::v;   // forward first pass
  constexpr int RI = FirstRadix<N2, E, +1>::v;   // inverse first pass (= forward last)
  constexpr int RL = LastRadix<N2, E, +1>::v;    // inverse last pass

  const int tid = threadIdx.x % NT;
  const int sub = RW > 1 ? int(threadIdx.x) / NT : 0;
  float2* lds = lds_all + sub * LROW;
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

  // Chirp: the bins this thread holds after the forward transform
  // (k2 = tid + r * N2/RI, one butterfly per thread), evaluated once per
  // workgroup into LDS and read back per row.
  //
  // Inter-stage twiddles W_N^{n2*k1} of the elements a thread loads and
  // stores (n2 = j + r*N2/RF; the forward plan's first pass and the inverse
  // plan's last pass have the same radix) factor as
  //   W_N^{j*k1} * W_N^{r*(N2/RF)*k1}:
  // a per-thread base times RF workgroup-uniform steps held in scalar
  // registers, recomposed per row (4 multiply-adds per element) so that the
  // registers go to prefetching the next row instead.
  static_assert(E / RI == 1, "row plan must end with a pass of radix E");
  static_assert(RF == RL, "first forward and last inverse radix must agree");
  __shared__ float2 hlds[N2];
#pragma unroll 1
  for (int k2 = threadIdx.x; k2 < N2; k2 += RW * NT)
    hlds[k2] = c.mode == DSPX_CHIRP_TABLE ? c.table[(uint64_t(chan) * s.N1 + k1) * N2 + k2]
                                          : chirp_onchip(c, chan, k1 + s.N1 * k2, N);
  constexpr int Q = E / RF;
  float2 base[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) base[q] = tw_N(s.tw, uint32_t(tid + q * NT) * k1);
  float2 step[RF];
#pragma unroll
  for (int r = 0; r < RF; r++) step[r] = uniform(tw_N(s.tw, uint32_t(r * (N2 / RF)) * k1));
  __syncthreads();

  // Row i = it * RW + sub.  Every group runs the same number of iterations
  // (the exchanges inside the transforms are workgroup barriers); a group
  // past the last row transforms stale registers and stores nothing.
  const uint32_t nrow = (p_end - p_begin) * s.npol;
  const uint32_t niter = (nrow + RW - 1) / RW;
  auto row_ptr = [&](uint32_t i) {
    uint32_t p = p_begin + i / s.npol, pol = i % s.npol;
    return s.work + ((uint64_t(chan) * s.npart + p) * s.npol + pol) * plane + uint64_t(k1) * N2;
  };
  float2 pf[E] = {};
  if (uint32_t(sub) < nrow) {
    const float2* row = row_ptr(sub);
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
      for (int r = 0; r < RF; r++) pf[q * RF + r] = row[tid + q * NT + r * (N2 / RF)];
  }
  // 4096-point rows: the workgroup's occupancy is set by its LDS (two per
  // CU), so registers are spent on keeping the row loop's twiddles (the
  // inter-stage twiddles here, the pass twiddles by the compiler's hoisting).
  // Shorter rows recompute them per row, keeping registers for occupancy.
  constexpr bool HOIST = N2 >= 4096;
  float2 w[E];
  if constexpr (HOIST) {
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
      for (int r = 0; r < RF; r++) w[q * RF + r] = cmul(base[q], step[r]);
  }
  for (uint32_t it = 0; it < niter; it++) {
    const uint32_t i = it * RW + sub;
    float2 v[E];
    if constexpr (HOIST) {
#pragma unroll
      for (int e = 0; e < E; e++) v[e] = cmul(pf[e], w[e]);
    } else {
#pragma unroll
      for (int q = 0; q < Q; q++) {
        float2 b = opaque_v(base[q]);
#pragma unroll
        for (int r = 0; r < RF; r++) v[q * RF + r] = cmul(pf[q * RF + r], cmul(b, step[r]));
      }
    }
    if (i + RW < nrow) {
      const float2* nxt = row_ptr(i + RW);
#pragma unroll
      for (int q = 0; q < Q; q++)
#pragma unroll
        for (int r = 0; r < RF; r++) pf[q * RF + r] = nxt[tid + q * NT + r * (N2 / RF)];
    }
    const float2* tw = HOIST ? twr : opaque(twr);
    fft_regs<LAY, N2, 1, NT, -1>(lds, v, tw, tid);
#pragma unroll
    for (int r = 0; r < RI; r++) v[r] = cmul(v[r], hlds[tid + r * (N2 / RI)]);
    fft_regs<LAY, N2, 1, NT, +1>(lds, v, HOIST ? tw : opaque(tw), tid);
    if (i < nrow) {
      float2* row = row_ptr(i);
#pragma unroll
      for (int q = 0; q < Q; q++) {
        float2 b = HOIST ? base[q] : opaque_v(base[q]);
#pragma unroll
        for (int r = 0; r < RL; r++)
          row[tid + q * NT + r * (N2 / RL)] = cmulc(v[q * RL + r], HOIST ? w[q * RL + r] : cmul(b, step[r]));
      }
    }
  }
}

// --------------------------------- stage B of the convolving filterbank

// The convolving filterbank (Filterbank.C:75-570, FilterbankEngineCPU.C:
// 184-370) transforms an N = nsub*fres point window forward, multiplies the
// response, and transforms each sub-band's fres bins back.  In the four-step
// split (N = N1*N2, fres = N1*F2) the forward transform leaves bin
// k1 + N1*k2 at row k1, position k2, and sub-band s owns bins s*fres + m
// with m = k1 + N1*j (j < F2): the row's segment k2 = s*F2 + j.  Its
// fres-point backward transform splits the same way,
//   Y_s[ob + F2*a] = sum_k1 e^{+2 pi i k1 a / N1} conj(W_fres^{k1 ob})
//                    sum_j Z[k1][s*F2 + j] e^{+2 pi i j ob / F2},
// so stage B runs each row's NSUB = N2/F2 backward F2-point transforms
// (batched columns in LDS), multiplies by conj(W_N^{nsub k1 ob}), and stores
// Y at row position s*F2 + ob; stage C's backward column transform over k1
// then yields output sample ob + F2*a of sub-band s (FilterbankArgs.f2).
// The response is the transposed table [nchan][N1][N2] (ConvCore::set_table).
// grid.x = N1 * npg * nchan, block = N2/16; one row in flight.
template <int N2, int F2>
__global__ __launch_bounds__(N2 / 16) void k_stageB_fb(ConvShape s, ChirpArgs c, const float2* __restrict__ twr,
                                                       const float2* __restrict__ twf, uint32_t ppg, uint32_t nsub)
{
  constexpr int E = 16, NT = N2 / E, NSUB = N2 / F2;
  constexpr int LROW = N2 + N2 / 16, LSEG = F2 * (NSUB + 1);
  using LAY = RowLayout<N2, NT>;
  using SEG = ColLayout<F2, NSUB, NSUB + 1, NT>;
  __shared__ float2 lds[LROW > LSEG ? LROW : LSEG];
  __shared__ float2 hlds[N2];
  constexpr int RF = FirstRadix<N2, E, -1>::v;   // forward first pass
  constexpr int RI = FirstRadix<N2, E, +1>::v;   // forward last pass
  constexpr int RS0 = FirstRadix<F2, E, +1>::v;  // segment transform, first pass
  constexpr int RSL = LastRadix<F2, E, +1>::v;   // segment transform, last pass
  static_assert(E / RI == 1, "row plan must end with a pass of radix E");
  constexpr int Q = E / RF;

  const int tid = threadIdx.x;
  const uint32_t npg = (s.npart + ppg - 1) / ppg;
  uint32_t bid = blockIdx.x;
  const uint32_t k1 = bid % s.N1;
  bid /= s.N1;
  const uint32_t pg = bid % npg;
  const uint32_t chan = bid / npg;
  const uint32_t p_begin = pg * ppg;
  const uint32_t p_end = min(s.npart, p_begin + ppg);
  const uint64_t plane = uint64_t(s.N1) * N2;

  // a null table: the unit response (the filterbank's zero-DM output)
  for (int k2 = tid; k2 < N2; k2 += NT)
    hlds[k2] = c.table ? c.table[(uint64_t(chan) * s.N1 + k1) * N2 + k2] : make_float2(1.f, 0.f);
  // inter-stage twiddles W_N^{n2 k1} of the loaded elements (n2 = tid + q*NT
  // + r*N2/RF), and the output twiddles conj(W_N^{nsub k1 ob}) of the stored
  // ones (n2 = tid + e*NT, ob = n2 % F2), both fixed for the workgroup
  float2 w[E], wo[E];
#pragma unroll
  for (int q = 0; q < Q; q++)
#pragma unroll
    for (int r = 0; r < RF; r++) w[q * RF + r] = tw_N(s.tw, uint32_t(tid + q * NT + r * (N2 / RF)) * k1);
#pragma unroll
  for (int e = 0; e < E; e++) wo[e] = tw_N(s.tw, nsub * k1 * (uint32_t(tid + e * NT) & uint32_t(F2 - 1)));
  __syncthreads();

  const uint32_t nrow = (p_end - p_begin) * s.npol;
  for (uint32_t i = 0; i < nrow; i++) {
    const uint32_t p = p_begin + i / s.npol, pol = i % s.npol;
    float2* row = s.work + ((uint64_t(chan) * s.npart + p) * s.npol + pol) * plane + uint64_t(k1) * N2;
    float2 v[E];
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
      for (int r = 0; r < RF; r++) v[q * RF + r] = cmul(row[tid + q * NT + r * (N2 / RF)], w[q * RF + r]);
    fft_regs<LAY, N2, 1, NT, -1>(lds, v, twr, tid);
    // bin k2 = tid + r*N2/RI times the response, into the segments' layout
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RI; r++) {
      const int k2 = tid + r * (N2 / RI);
      lds[SEG::idx(k2 % F2, k2 / F2)] = cmul(v[r], hlds[k2]);
    }
    __syncthreads();
    lds_load_pass<SEG, F2, NSUB, NT, RS0>(lds, v, tid);
    fft_regs<SEG, F2, NSUB, NT, +1>(lds, v, twf, tid);
    // v[q*RSL + r] = Y_b[j + r*F2/RSL] (g = tid + q*NT, b = g % NSUB, j = g / NSUB)
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E / RSL; q++) {
      const int g = tid + q * NT, b = g % NSUB, j = g / NSUB;
#pragma unroll
      for (int r = 0; r < RSL; r++) lds[LAY::idx(b * F2 + j + r * (F2 / RSL), 0)] = v[q * RSL + r];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int n2 = tid + e * NT;
      row[n2] = cmulc(lds[LAY::idx(n2, 0)], wo[e]);
    }
  }
}

// ---------------------------------- stage B of real-sampled (Nyquist) input

// Real-sampled input (Convolution.C:385-413, 700-790): a window of 2N real
// samples x is taken as N complex samples z[n] = x[2n] + i x[2n+1] (stage A
// unchanged), whose transform Z gives the real transform's bins 0..N-1 as
//   X[k] = E[k] + W_2N^k O[k],  E = (Z[k] + Z*[N-k]) / 2,
//                               O = (Z[k] - Z*[N-k]) / 2i
// (the reference's frc1d forward transform); X times the response, then the
// N-point backward transform (bcc1d) as for complex input.  In the four-step
// split bin k = k1 + N1*k2 sits at row k1, position k2, and its mirror N - k
// at row N1 - k1, position N2 - 1 - k2 (row 0: position (N2 - k2) mod N2),
// so a workgroup takes rows a = ip and b = N1 - ip together (rows 0 and
// N1/2 alone): both forward transforms, both rows' bins into LDS, each
// bin's mirror from there, the response products, both backward
// transforms.  splitw: W_2N^{k1 + N1*k2} at [k1][k2].  grid.x = (N1/2 + 1) *
// npg * nchan, block = N2/16.
__device__ __forceinline__ float2 real_split(float2 z, float2 zm, float2 w)
{
  const float2 e = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y + zm.y));
  const float2 d = make_float2(0.5f * (z.x - zm.x), 0.5f * (z.y - zm.y));
  return cadd(e, cmul(w, make_float2(d.y, -d.x)));
}

template <int N2>
__global__ __launch_bounds__(N2 / 16) void k_stageB_real(ConvShape s, ChirpArgs c, const float2* __restrict__ twr,
                                                         const float2* __restrict__ splitw, uint32_t ppg)
{
  constexpr int E = 16, NT = N2 / E;
  constexpr int LROW = N2 + N2 / 16;
  using LAY = RowLayout<N2, NT>;
  __shared__ float2 lds[2 * LROW + 2 * N2];
  float2* bufA = lds;
  float2* bufB = lds + LROW;
  float2* hA = lds + 2 * LROW;
  float2* hB = hA + N2;
  constexpr int RF = FirstRadix<N2, E, -1>::v;  // forward first pass = backward last
  constexpr int RI = FirstRadix<N2, E, +1>::v;  // forward last pass
  static_assert(E / RI == 1, "row plan must end with a pass of radix E");
  constexpr int Q = E / RF;

  const int tid = threadIdx.x;
  const uint32_t npair = s.N1 / 2 + 1;
  const uint32_t npg = (s.npart + ppg - 1) / ppg;
  uint32_t bid = blockIdx.x;
  const uint32_t ip = bid % npair;
  bid /= npair;
  const uint32_t pg = bid % npg;
  const uint32_t chan = bid / npg;
  const uint32_t p_begin = pg * ppg;
  const uint32_t p_end = min(s.npart, p_begin + ppg);
  const uint32_t N = s.N1 * N2;
  const uint64_t plane = uint64_t(s.N1) * N2;
  const uint32_t ra = ip, rb = s.N1 - ip;
  const bool pair = ip != 0 && ip != s.N1 / 2;

  for (int k2 = tid; k2 < N2; k2 += NT) {
    hA[k2] = c.mode == DSPX_CHIRP_TABLE ? c.table[(uint64_t(chan) * s.N1 + ra) * N2 + k2]
                                        : chirp_onchip(c, chan, ra + s.N1 * k2, N);
    if (pair)
      hB[k2] = c.mode == DSPX_CHIRP_TABLE ? c.table[(uint64_t(chan) * s.N1 + rb) * N2 + k2]
                                          : chirp_onchip(c, chan, rb + s.N1 * k2, N);
  }
  float2 wa[E], wb[E], sa[RI], sb[RI];
#pragma unroll
  for (int q = 0; q < Q; q++)
#pragma unroll
    for (int r = 0; r < RF; r++) {
      const uint32_t n2 = uint32_t(tid + q * NT + r * (N2 / RF));
      wa[q * RF + r] = tw_N(s.tw, n2 * ra);
      wb[q * RF + r] = tw_N(s.tw, (n2 * rb) & (N - 1));
    }
#pragma unroll
  for (int r = 0; r < RI; r++) {
    const int k2 = tid + r * (N2 / RI);
    sa[r] = splitw[uint64_t(ra) * N2 + k2];
    sb[r] = pair ? splitw[uint64_t(rb) * N2 + k2] : make_float2(0.f, 0.f);
  }
  __syncthreads();

  const uint32_t nrow = (p_end - p_begin) * s.npol;
  for (uint32_t i = 0; i < nrow; i++) {
    const uint32_t p = p_begin + i / s.npol, pol = i % s.npol;
    float2* base = s.work + ((uint64_t(chan) * s.npart + p) * s.npol + pol) * plane;
    float2* rowa = base + uint64_t(ra) * N2;
    float2* rowb = base + uint64_t(rb) * N2;
    float2 va[E], vb[E];
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
      for (int r = 0; r < RF; r++) {
        const int n2 = tid + q * NT + r * (N2 / RF);
        va[q * RF + r] = cmul(rowa[n2], wa[q * RF + r]);
        vb[q * RF + r] = pair ? cmul(rowb[n2], wb[q * RF + r]) : make_float2(0.f, 0.f);
      }
    fft_regs<LAY, N2, 1, NT, -1>(bufA, va, twr, tid);
    if (pair) fft_regs<LAY, N2, 1, NT, -1>(bufB, vb, twr, tid);
    // the bins of both rows into LDS, then each bin's mirror from there
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RI; r++) {
      const int k2 = tid + r * (N2 / RI);
      bufA[LAY::idx(k2, 0)] = va[r];
      if (pair) bufB[LAY::idx(k2, 0)] = vb[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RI; r++) {
      const int k2 = tid + r * (N2 / RI);
      const int mk = N2 - 1 - k2;
      float2 ma;
      if (pair)
        ma = bufB[LAY::idx(mk, 0)];
      else
        ma = bufA[LAY::idx(ra == 0 ? (N2 - k2) & (N2 - 1) : mk, 0)];
      va[r] = cmul(real_split(va[r], make_float2(ma.x, -ma.y), sa[r]), hA[k2]);
      if (pair) {
        const float2 mb = bufA[LAY::idx(mk, 0)];
        vb[r] = cmul(real_split(vb[r], make_float2(mb.x, -mb.y), sb[r]), hB[k2]);
      }
    }
    __syncthreads();  // the mirrors are read before the backward transforms reuse the buffers
    fft_regs<LAY, N2, 1, NT, +1>(bufA, va, twr, tid);
    if (pair) fft_regs<LAY, N2, 1, NT, +1>(bufB, vb, twr, tid);
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
      for (int r = 0; r < RF; r++) {
        const int n2 = tid + q * NT + r * (N2 / RF);
        rowa[n2] = cmulc(va[q * RF + r], wa[q * RF + r]);
        if (pair) rowb[n2] = cmulc(vb[q * RF + r], wb[q * RF + r]);
      }
  }
}

// ------------------------------------------------------------------ stage C

enum { kStageCFilterbank = 0, kStageCFold = 1 };

// Detected values of one sample (Detection.C:252-318; cross_detect.ic,
// stokes_detect.ic), evaluated without contraction like the reference.
__device__ __forceinline__ int detect_values(uint32_t state, uint32_t npol, float2 a, float2 b, float* d)
{
#pragma clang fp contract(off)
  float pp = a.x * a.x;
  pp += a.y * a.y;
  if (state == DSPX_INTENSITY) {
    if (npol == 2) {
      pp += b.x * b.x;
      pp += b.y * b.y;
    }
    d[0] = pp;
    return 1;
  }
  float qq = b.x * b.x;
  qq += b.y * b.y;
  if (state == DSPX_PPQQ) {
    d[0] = pp;
    d[1] = qq;
    return 2;
  }
  float pq_r = a.x * b.x + a.y * b.y;
  float pq_i = a.x * b.y - a.y * b.x;
  if (state == DSPX_STOKES) {
    float p2 = a.x * a.x + a.y * a.y;
    float q2 = b.x * b.x + b.y * b.y;
    d[0] = p2 + q2;
    d[1] = p2 - q2;
    d[2] = 2.0f * pq_r;
    d[3] = 2.0f * pq_i;
  } else {
    d[0] = a.x * a.x + a.y * a.y;
    d[1] = b.x * b.x + b.y * b.y;
    d[2] = pq_r;
    d[3] = pq_i;
  }
  return 4;
}

// Products of one sample accumulated with fused multiply-adds, for the
// fold: NV = 1: c0 += |p|^2 (+ |q|^2 with two pols); NV = 2: c0 += |p|^2,
// c1 += |q|^2; NV = 4: also c2 += Re(p q*) = p.re q.re + p.im q.im and
// c3 += Im(p q*) = p.re q.im - p.im q.re (cross_detect.ic, stokes_detect.ic).
// The detected values of a sum of samples are formed from the sums by
// detect_finish: the fold adds them anyway, in an order of its own.
template <int NV>
__device__ __forceinline__ void detect_acc(uint32_t npol, float2 a, float2 b, float* c)
{
  c[0] = fmaf(a.x, a.x, c[0]);
  c[0] = fmaf(a.y, a.y, c[0]);
  if constexpr (NV == 1) {
    if (npol == 2) {
      c[0] = fmaf(b.x, b.x, c[0]);
      c[0] = fmaf(b.y, b.y, c[0]);
    }
  } else {
    c[1] = fmaf(b.x, b.x, c[1]);
    c[1] = fmaf(b.y, b.y, c[1]);
    if constexpr (NV == 4) {
      c[2] = fmaf(a.x, b.x, c[2]);
      c[2] = fmaf(a.y, b.y, c[2]);
      c[3] = fmaf(a.x, b.y, c[3]);
      c[3] = fmaf(-a.y, b.x, c[3]);
    }
  }
}

// Stokes (I = PP + QQ, Q = PP - QQ, U = 2 Re, V = 2 Im) or the coherence
// products themselves, from accumulated products.
template <int NV>
__device__ __forceinline__ void detect_finish(bool stokes, const float* c, float* d)
{
  if constexpr (NV == 4) {
    d[0] = stokes ? c[0] + c[1] : c[0];
    d[1] = stokes ? c[0] - c[1] : c[1];
    d[2] = stokes ? 2.0f * c[2] : c[2];
    d[3] = stokes ? 2.0f * c[3] : c[3];
  } else {
#pragma unroll
    for (int i = 0; i < NV; i++) d[i] = c[i];
  }
}

// Detection specialised on the number of detected values: NV = 1 Intensity,
// 2 PPQQ, 4 Coherence (stokes = false) or Stokes.
template <int NV>
__device__ __forceinline__ void detect_nv(bool stokes, uint32_t npol, float2 a, float2 b, float* d)
{
#pragma clang fp contract(off)
  float pp = a.x * a.x;
  pp += a.y * a.y;
  if constexpr (NV == 1) {
    if (npol == 2) {
      pp += b.x * b.x;
      pp += b.y * b.y;
    }
    d[0] = pp;
  } else {
    float qq = b.x * b.x;
    qq += b.y * b.y;
    if constexpr (NV == 2) {
      d[0] = pp;
      d[1] = qq;
    } else {
      float pq_r = a.x * b.x + a.y * b.y;
      float pq_i = a.x * b.y - a.y * b.x;
      float p2 = a.x * a.x + a.y * a.y;
      float q2 = b.x * b.x + b.y * b.y;
      d[0] = stokes ? p2 + q2 : p2;
      d[1] = stokes ? p2 - q2 : q2;
      d[2] = stokes ? 2.0f * pq_r : pq_r;
      d[3] = stokes ? 2.0f * pq_i : pq_i;
    }
  }
}

// QN consecutive keys starting at a multiple of QN (vector loads)
template <int QN>
__device__ __forceinline__ void load_keys(const uint32_t* __restrict__ k, uint32_t* out)
{
  if constexpr (QN == 1) {
    out[0] = k[0];
  } else if constexpr (QN == 2) {
    const uint2 v = *reinterpret_cast<const uint2*>(k);
    out[0] = v.x;
    out[1] = v.y;
  } else {
    static_assert(QN % 4 == 0, "key vectors of 1, 2 or 4k words");
#pragma unroll
    for (int i = 0; i < QN; i += 4) {
      const uint4 v = *reinterpret_cast<const uint4*>(k + i);
      out[i] = v.x;
      out[i + 1] = v.y;
      out[i + 2] = v.z;
      out[i + 3] = v.w;
    }
  }
}

// One lane's QN consecutive output samples of a fold row (a[q], b[q]: the
// two pols; kq[q]: their histogram keys, ~0 where not folded) added into the
// LDS histogram: samples of the lane's first key accumulated in registers
// and added once per 16-lane row when the row's first keys agree, else once
// per lane; samples after a bin edge accumulated per run of equal keys and
// added once per run (bins a few samples wide, short periods with many
// bins, put an edge in most lanes).
// ZH, zeroed data with per-channel hits (Fold.C:911-930): samples the zap
// mask flags are dropped, and each sample whose first detected value is
// non-zero counts one hit, accumulated as one more value and added at
// hist[slot][nval*nbin + bin].  i0: output sample index of q = 0 (negative
// within the discarded overlap, where every key is ~0).
template <int NV, int QN, bool ZH>
__device__ __forceinline__ void fold_lane(float* hist, const FoldArgs& fa, uint32_t npol, uint32_t* kq,
                                          const float2* a, const float2* b, uint32_t chan, int32_t i0)
{
  constexpr int NA = ZH ? NV + 1 : NV;
  constexpr uint32_t NONE = 0xffffffffu;
  const bool stokes = fa.state == DSPX_STOKES;
  if constexpr (ZH) {
    // block k and offset r of the first sample at or past output 0, then
    // advanced sample by sample (M >= 1: at most one block edge per step)
    const ZapArgs& z = fa.zap;
    const uint32_t base = i0 < 0 ? 0u : uint32_t(i0);
    uint32_t k = uint32_t(__umul64hi(uint64_t(base), z.magic));
    uint32_t r = base - k * z.M;
    bool zap = k >= z.nblk || z.mask[uint64_t(k) * z.nchan + chan] != 0;
#pragma unroll
    for (int q = 0; q < QN; q++) {
      if (q > 0 && i0 + q > int32_t(base)) {
        if (++r == z.M) {
          r = 0;
          k++;
          zap = k >= z.nblk || z.mask[uint64_t(k) * z.nchan + chan] != 0;
        }
      }
      if (zap) kq[q] = NONE;
    }
  }
  auto acc_sample = [&](int q, float* c) {
    detect_acc<NV>(npol, a[q], b[q], c);
    if constexpr (ZH) {
      // the first detected value (PP, PP + QQ or Stokes I) is a sum of
      // squares: zero exactly when each square is
      float x0 = a[q].x * a[q].x + a[q].y * a[q].y;
      if ((NV == 1 && npol == 2) || (NV == 4 && stokes)) x0 += b[q].x * b[q].x + b[q].y * b[q].y;
      c[NV] += x0 != 0.f ? 1.f : 0.f;
    }
  };
  auto finish = [&](const float* c, float* d) {
    detect_finish<NV>(stokes, c, d);
    if constexpr (ZH) d[NV] = c[NV];
  };
  // the slot, the bin and the LDS histogram's row and bin strides are all
  // below 2^24 (the histogram fits in LDS), so the offsets are 24-bit
  // products (full-rate multiplies instead of quarter-rate 32-bit ones)
  auto add = [&](uint32_t key, const float* d) {
    float* hs = hist + __umul24(key >> 16, fa.hrow);
    const uint32_t bin = key & 0xffffu;
#pragma unroll
    for (int i = 0; i < NV; i++) atomicAdd(hs + __umul24(bin, fa.ndim) + fa.voff[i], d[i]);
    if constexpr (ZH) atomicAdd(hs + NV * fa.nbin + bin, d[NV]);
  };
  float cacc[NA], racc[NA];
#pragma unroll
  for (int i = 0; i < NA; i++) cacc[i] = racc[i] = 0.f;
  const uint32_t lane_key = kq[0];
  uint32_t run_key = NONE;
  auto flush_run = [&]() {
    if (run_key != NONE) {
      float d[NA];
      finish(racc, d);
      add(run_key, d);
    }
  };
#pragma unroll
  for (int q = 0; q < QN; q++) {
    const uint32_t key = kq[q];
    if (key == lane_key) {
      acc_sample(q, cacc);
    } else if (key != NONE) {
      if (key != run_key) {
        flush_run();
        run_key = key;
#pragma unroll
        for (int i = 0; i < NA; i++) racc[i] = 0.f;
      }
      acc_sample(q, racc);
    }
  }
  flush_run();
  float acc[NA];
  finish(cacc, acc);
  // is the whole 16-lane row in one bin?
  const int lane = int(threadIdx.x & 63u);
  const uint32_t k15 = __shfl(lane_key, lane | 15);
  const uint64_t same = __ballot(lane_key == k15);
  const bool row_uniform = ((same >> (lane & 48)) & 0xffffull) == 0xffffull;
  float rs[NA];
#pragma unroll
  for (int i = 0; i < NA; i++) rs[i] = row_sum_lane15(acc[i]);
  if (row_uniform) {
    if ((lane & 15) == 15 && k15 != NONE) add(k15, rs);
  } else if (lane_key != NONE) {
    add(lane_key, acc);
  }
}

// The fold's ownership of an L-point inverse column transform's outputs:
// each thread holds QN consecutive columns of R rows.  A multi-pass
// transform's last pass delivers E / (last radix) of them; a single-pass
// transform (L = 16) would leave one column per thread.  The outputs are
// regrouped through LDS into QN = 16 (or the tile's column count, if fewer)
// wherever the last pass leaves fewer (at 32 and 64 points only for narrow
// bins, NB): a lane's 16 consecutive samples are
// accumulated per run of equal bins (bins a few samples wide, short periods,
// add to the histogram once per run), and a 16-lane row spans 256 samples
// for the row sums of wide bins.
template <int L, bool NB = true> struct FoldOwn {
  static constexpr int E = kColTile / kColThreads;
  static constexpr int RL = LastRadix<L, 16, +1>::v;
  static constexpr int C = kColTile / L;             // columns of a tile
  static constexpr int QMAX = C < 16 ? C : 16;
  // with wide bins 32- and 64-point columns keep the last pass's 8 and 4
  // samples per lane: the regrouping costs them one more exchange (the
  // 128-channel workload's stage C 527 -> 564 us with it)
  static constexpr bool REGROUP = RadixPlan<L, 16>::n == 1 || (E / RL < QMAX && (L >= 128 || NB));
  static constexpr int QN = REGROUP ? QMAX : E / RL;
  static constexpr int R = E / QN;
};

// Inverse column transform ending in the fold's ownership (FoldOwn): on
// entry v holds the first pass's inputs in the standard ownership; on exit
// v[q*R + r] holds y(row j + r*L/R, column (tid % (C/QN))*QN + q) with
// j = tid / (C/QN).
template <class LAY, int L, int C, bool NB>
__device__ __forceinline__ void last_pass_cols(float2* lds, float2* v, const float2* __restrict__ tw, int tid)
{
  using OWN = FoldOwn<L, NB>;
  constexpr int NPASS = RadixPlan<L, 16>::n;
  constexpr int RL = OWN::RL;
  constexpr int QN = OWN::QN;
  const int b0 = (tid % (C / QN)) * QN;
  const int j = tid / (C / QN);
  if constexpr (OWN::REGROUP) {
    if constexpr (NPASS == 1) {
      // single pass in the standard ownership (one column per thread, the
      // caller has synchronised), then one exchange
      stockham_compute<L, C, kColThreads, RL, 1, +1>(v, tw, tid);
      lds_store_pass<LAY, L, C, kColThreads, RL, 1>(lds, v, tid);
    } else {
      // every pass in the standard ownership, then one more exchange
      run_passes<LAY, L, C, kColThreads, +1, 0, NPASS>(lds, v, tw, tid);
      __syncthreads();
      lds_store_pass<LAY, L, C, kColThreads, RL, InvNs<L, 16, NPASS - 1>::v>(lds, v, tid);
    }
    __syncthreads();
    constexpr int R = OWN::R;
#pragma unroll
    for (int q = 0; q < QN; q++)
#pragma unroll
      for (int r = 0; r < R; r++) v[q * R + r] = lds[LAY::idx(j + r * (L / R), b0 + q)];
  } else {
    run_passes<LAY, L, C, kColThreads, +1, 0, NPASS - 1>(lds, v, tw, tid);
    constexpr int RP = InvRadix<L, 16, NPASS - 2>::v;
    constexpr int NSP = InvNs<L, 16, NPASS - 2>::v;
    constexpr int NS = InvNs<L, 16, NPASS - 1>::v;
    __syncthreads();
    lds_store_pass<LAY, L, C, kColThreads, RP, NSP>(lds, v, tid);
    __syncthreads();
    const int k = j % NS;
#pragma unroll
    for (int q = 0; q < QN; q++) {
#pragma unroll
      for (int r = 0; r < RL; r++) v[q * RL + r] = lds[LAY::idx(j + r * (L / RL), b0 + q)];
#pragma unroll
      for (int r = 1; r < RL; r++) v[q * RL + r] = cmulc(v[q * RL + r], tw[r * k * (L / (NS * RL))]);
      dft<RL, +1>(v + q * RL);
    }
  }
}

// grid.x = G * npart * nchan (fold) or ncb * npart * nchan (filterbank).
// NV: detected values per sample in fold mode (1, 2 or 4).  NB: narrow
// bins (the fold's ownership, FoldOwn).  PFT: the fold
// loads each tile's inputs while the previous tile is transformed (64 more
// registers; the launch chooses it where occupancy is set by the LDS).
template <int N1, int MODE, int NV = 1, bool PFT = false, bool ZH = false, bool NB = true>
__global__ __launch_bounds__(kColThreads) void k_stageC(ConvShape s, FilterbankArgs fb, FoldArgs fa,
                                                       const float2* __restrict__ twc)
{
  constexpr int C = kColTile / N1;
  constexpr int E = kColTile / kColThreads;
  constexpr int RI = FirstRadix<N1, 16, +1>::v;
  constexpr int RL = LastRadix<N1, 16, +1>::v;
  // swizzled: the fold's regrouping reads each lane's 16 consecutive
  // columns without bank conflicts
  using LAY = ColLayoutSwz<N1, C, C, kColThreads>;
  extern __shared__ float2 smem[];
  float2* lds = smem;
  float* hist = reinterpret_cast<float*>(smem + kColTile);

  const int tid = threadIdx.x;
  const uint32_t ncb = s.N2 / C;
  const uint32_t G = MODE == kStageCFold ? fa.G : ncb;
  uint32_t bid = blockIdx.x;
  const uint32_t g = bid % G;
  bid /= G;
  const uint32_t p = bid % s.npart;
  const uint32_t chan = bid / s.npart;
  uint32_t cb_begin, cb_end;
  if (MODE == kStageCFold) {
    cb_begin = uint32_t(uint64_t(ncb) * g / G);
    cb_end = uint32_t(uint64_t(ncb) * (g + 1) / G);
  } else {
    cb_begin = g;
    cb_end = g + 1;
  }

  const uint32_t hist_len = MODE == kStageCFold ? fa.nslot * fa.hrow : 0;
  if (MODE == kStageCFold) {
    for (uint32_t i = tid; i < hist_len; i += kColThreads) hist[i] = 0.f;
  }
  // this part's keys, indexed like the transform outputs
  const uint32_t* kp = MODE == kStageCFold ? fa.keys + uint64_t(p) * s.N1 * s.N2 : nullptr;

  const uint64_t plane = uint64_t(s.N1) * s.N2;
  const float2* wbase = s.work + (uint64_t(chan) * s.npart + p) * s.npol * plane;

  // The tile's first-pass inputs.  Each address is a workgroup-uniform row
  // base (scalar registers: tile column, row r of the butterfly) plus a
  // per-thread 32-bit byte offset that does not change from tile to tile,
  // so the loads need no per-element 64-bit address arithmetic.
  const char* wbytes = reinterpret_cast<const char*>(wbase);
  const uint64_t plane_bytes = plane * sizeof(float2);
  uint32_t voff[E / RI];
#pragma unroll
  for (int q = 0; q < E / RI; q++) {
    const int gg = tid + q * kColThreads;
    voff[q] = (uint32_t(gg / C) * s.N2 + uint32_t(gg % C)) * uint32_t(sizeof(float2));
  }
  auto load_tile = [&](uint32_t cb, float2* xa, float2* xb) {
    const char* t0 = wbytes + uint64_t(cb) * C * sizeof(float2);
#pragma unroll
    for (int q = 0; q < E / RI; q++) {
#pragma unroll
      for (int r = 0; r < RI; r++) {
        const char* rowp = t0 + uint64_t(r * (N1 / RI)) * s.N2 * sizeof(float2);
        xa[q * RI + r] = *reinterpret_cast<const float2*>(rowp + voff[q]);
        xb[q * RI + r] =
            s.npol == 2 ? *reinterpret_cast<const float2*>(rowp + plane_bytes + voff[q]) : make_float2(0.f, 0.f);
      }
    }
  };
  // One tile: column transform, then the filterbank store or the fold.
  // va, vb hold the tile's inputs and are overwritten.
  auto tile = [&](uint32_t cb, float2* va, float2* vb) {
    const uint32_t c0 = cb * C;
    const float2* tw = opaque(twc);
    if constexpr (MODE == kStageCFilterbank) {
      __syncthreads();  // previous tile's LDS reads are complete
      fft_regs<LAY, N1, C, kColThreads, +1>(lds, va, tw, tid);
      if (s.npol == 2) {
        __syncthreads();
        fft_regs<LAY, N1, C, kColThreads, +1>(lds, vb, opaque(tw), tid);
      }
#pragma unroll
      for (int q = 0; q < E / RL; q++) {
        int gg = tid + q * kColThreads;
        int b = gg % C, j = gg / C;
#pragma unroll
        for (int r = 0; r < RL; r++) {
          uint32_t n1 = j + r * (N1 / RL);
          float* o;
          if (fb.f2) {
            const uint32_t n2 = c0 + b, sb = n2 >> fb.lf2;
            const uint32_t t = (n2 & (fb.f2 - 1)) + (n1 << fb.lf2);
            if (t < s.nfilt_pos || t >= s.nfilt_pos + fb.nkeep) continue;
            const uint64_t idat = uint64_t(p) * fb.nkeep + (t - s.nfilt_pos);
            o = fb.out + (uint64_t(chan) * fb.nsub + sb) * fb.chan_stride + 2 * idat;
          } else {
            uint32_t t = s.N2 * n1 + c0 + b;
            if (t < s.nfilt_pos || t >= s.nfilt_pos + s.step) continue;
            uint64_t idat = uint64_t(p) * s.step + (t - s.nfilt_pos);
            o = fb.out + uint64_t(chan) * fb.chan_stride + 2 * idat;
          }
          *reinterpret_cast<float2*>(o) = va[q * RL + r];
          if (s.npol == 2) *reinterpret_cast<float2*>(o + fb.pol_stride) = vb[q * RL + r];
        }
      }
    } else {
      // all passes but the last in the standard ownership; the last pass
      // gives each thread QN consecutive columns (consecutive samples) of the
      // rows j + r*N1/RL, so a 16-lane DPP row spans whole tile rows
      __syncthreads();
      last_pass_cols<LAY, N1, C, NB>(lds, va, tw, tid);
      if (s.npol == 2) {
        __syncthreads();
        last_pass_cols<LAY, N1, C, NB>(lds, vb, opaque(tw), tid);
      }
      constexpr int QN = FoldOwn<N1, NB>::QN;
      constexpr int RF = FoldOwn<N1, NB>::R;
      const uint32_t bq0 = (tid % (C / QN)) * QN;
      const uint32_t jr = tid / (C / QN);
      // key address: uniform row base plus a per-thread 32-bit byte offset
      const uint32_t kvo = (s.N2 * jr + bq0) * uint32_t(sizeof(uint32_t));
#pragma unroll
      for (int r = 0; r < RF; r++) {
        uint32_t kq[QN];
        const char* krow = reinterpret_cast<const char*>(kp + c0 + uint64_t(r * (N1 / RF)) * s.N2);
        load_keys<QN>(reinterpret_cast<const uint32_t*>(krow + kvo), kq);
        float2 xa[QN], xb[QN];
#pragma unroll
        for (int q = 0; q < QN; q++) {
          xa[q] = va[q * RF + r];
          xb[q] = vb[q * RF + r];
        }
        // output index of the lane's first sample (row n1, column c0 + bq0)
        const int32_t i0 = int32_t(p * s.step + s.N2 * (jr + r * (N1 / RF)) + c0 + bq0) - int32_t(s.nfilt_pos);
        fold_lane<NV, QN, ZH>(hist, fa, s.npol, kq, xa, xb, chan, i0);
      }
    }
  };
  // With PF the next tile's inputs are loaded while the current one is
  // transformed (two register sets used in turn, without the copy, measured
  // 5% slower on the 128-channel fold: 0.545 against 0.517 ms).
  constexpr bool PF = MODE == kStageCFold && PFT;
  if constexpr (PF) {
    float2 pa[E], pb[E];
    if (cb_begin < cb_end) load_tile(cb_begin, pa, pb);
    for (uint32_t cb = cb_begin; cb < cb_end; cb++) {
      float2 va[E], vb[E];
#pragma unroll
      for (int e = 0; e < E; e++) {
        va[e] = pa[e];
        vb[e] = pb[e];
      }
      if (cb + 1 < cb_end) load_tile(cb + 1, pa, pb);
      tile(cb, va, vb);
    }
  } else {
    for (uint32_t cb = cb_begin; cb < cb_end; cb++) {
      float2 va[E], vb[E];
      load_tile(cb, va, vb);
      tile(cb, va, vb);
    }
  }

  if (MODE == kStageCFold) {
    __syncthreads();
    float* part = fa.partial + ((uint64_t(chan) * s.npart + p) * fa.G + g) * hist_len;
    for (uint32_t i = tid; i < hist_len; i += kColThreads) part[i] = hist[i];
  }
}

}  // namespace dspx
