// sk.hip -- spectral kurtosis engine: block sums, the three detectors of
// SpectralKurtosis::detect and the mask (blocks of M samples starting every
// M / noverlap samples; the averaging detectors restricted to the included
// channels), for one resolution or several (the coarser ones' sums
// time-scrunched from the finer ones', their zaps mapped onto the finest
// resolution's mask rows).
//
// Reference: SpectralKurtosis.C:674-798 (compute), 830-909 (detect),
// 1105-1222 (detect_skfb), 1312-1462 (detect_fscr), 1014-1103 (detect_tscr),
// 1465-1566 (mask).  Arithmetic is kept in the reference's types and order
// where a decision depends on it: float sq = re*re + im*im and sq*sq, both
// summed in double; the float estimator of detect_skfb; double averages of
// detect_fscr; sequential double sums of detect_tscr.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "dspx_internal.hpp"
#include "stages.hpp"

struct dspx_sk {
  dspx::CtxRef hold;
  dspx_ctx* ctx = nullptr;
  dspx_sk_desc d{};
  dspx::DevBuf lim;  // lower[M+1], upper[M+1], symmetric[M+1]
  // detect_tscr's thresholds of the totals above M: totals (ascending),
  // then lower and upper, ntscr of each
  dspx::DevBuf tscr;
  uint32_t ntscr = 0;
  dspx::DevBuf status;  // detect_tscr: [0] totals without thresholds, [1] the smallest of them
  dspx::DevBuf tot;     // per-channel totals (dspx_sk_tscr_totals)
  dspx::DevBuf chans;   // channels the averaging detectors include (dspx_sk_set_channels), 1 byte each
  bool have_chans = false;
  uint32_t nov = 1, off = 0;  // noverlap; blocks start every off = M / noverlap samples
  double pcsd = 0;
  bool have_limits = false;
  // resolutions after the first (dspx_sk_add_resolution), increasing M
  struct Res {
    uint32_t M = 0, nov = 1, off = 0;
    dspx::DevBuf lim;    // lower[M+1], upper[M+1], symmetric[M+1]
    dspx::DevBuf chans;  // include flags, if have_chans
    bool have_chans = false;
  };
  std::vector<std::unique_ptr<Res>> res;
  dspx::DevBuf rsum[2], rcnt[2];  // sums and counts of the coarser resolutions, used in turn
  dspx::DevBuf rtest, rrow;       // fscr of a coarser resolution: per-block test, per-row zap
  // the sums dspx_sk_detect_res left for detect_tscr, and the last
  // resolution's shape and thresholds
  const float* last_sums = nullptr;
  const uint32_t* last_counts = nullptr;
  uint64_t last_npart = 0, last_rows = 0;
  uint32_t last_M = 0, last_nov = 1;
  const float* last_lim = nullptr;
  bool have_last = false;
};

namespace dspx {

namespace {

constexpr int kSkThreads = 256;

// Fixed-order tree sum over the block (deterministic, like a sequential sum
// to within double rounding; the float results match the reference's
// sequential double sums except within 1e-16 of a float rounding boundary).
__device__ __forceinline__ void block_sum(double* a, double* b, uint32_t* c, int tid)
{
  for (int s = kSkThreads / 2; s > 0; s >>= 1) {
    __syncthreads();
    if (tid < s) {
      a[tid] += a[tid + s];
      b[tid] += b[tid + s];
      c[tid] += c[tid + s];
    }
  }
  __syncthreads();
}

// grid (npart, nchan): S1 = sum sq, S2 = sum sq^2 over block ipart of each pol
// (SpectralKurtosis.C:752-766), counts of pol 0 (:781-787).
__global__ __launch_bounds__(kSkThreads) void k_sk_sums_fpt(const float* __restrict__ in, uint64_t cs, uint64_t ps,
                                                           uint32_t nchan, uint32_t npol, uint32_t M, uint32_t off,
                                                           int zeroed,
                                                           float* __restrict__ sums, uint32_t* __restrict__ counts)
{
#pragma clang fp contract(off)
  __shared__ double r1[kSkThreads], r2[kSkThreads];
  __shared__ uint32_t rc[kSkThreads];
  const int tid = threadIdx.x;
  const uint64_t part = blockIdx.x;
  const uint32_t chan = blockIdx.y;
  for (uint32_t pol = 0; pol < npol; pol++) {
    const float2* x = reinterpret_cast<const float2*>(in + uint64_t(chan) * cs + uint64_t(pol) * ps) + part * off;
    double s1 = 0, s2 = 0;
    uint32_t c = 0;
    for (uint32_t i = tid; i < M; i += kSkThreads) {
      float2 v = x[i];
      float sq = (v.x * v.x) + (v.y * v.y);
      if (!zeroed || sq != 0) {
        s1 += sq;
        s2 += sq * sq;
        c++;
      }
    }
    r1[tid] = s1;
    r2[tid] = s2;
    rc[tid] = c;
    block_sum(r1, r2, rc, tid);
    if (tid == 0) {
      float* o = sums + ((part * nchan + chan) * npol + pol) * 2;
      o[0] = float(r1[0]);
      o[1] = float(r2[0]);
      if (pol == 0) counts[part * nchan + chan] = rc[0];
    }
  }
}

// Raw TFP input (the unpacked zero-DM samples): grid (npart, ceil(nchan/32));
// lanes span 32 consecutive channels of a time sample (one TFP row segment),
// 8 row groups stride the block; the row groups are summed in order.
template <class LD>
__global__ __launch_bounds__(kSkThreads) void k_sk_sums_raw(LD ld, uint64_t t0, uint32_t nchan, uint32_t M,
                                                           uint32_t off, float* __restrict__ sums,
                                                           uint32_t* __restrict__ counts)
{
#pragma clang fp contract(off)
  constexpr int CH = 32, RG = kSkThreads / CH;
  __shared__ double r1[RG][2][CH], r2[RG][2][CH];
  const int tid = threadIdx.x;
  const int cl = tid % CH, rg = tid / CH;
  const uint64_t part = blockIdx.x;
  const uint32_t chan = blockIdx.y * CH + cl;
  const uint32_t npol = ld.npol;
  double s1[2] = {0, 0}, s2[2] = {0, 0};
  if (chan < nchan) {
    const uint64_t tb = t0 + part * off;
    for (uint32_t i = rg; i < M; i += RG) {
      typename LD::Word w = ld.word(chan, tb + i);
#pragma unroll
      for (uint32_t pol = 0; pol < 2; pol++) {
        if (pol < npol) {
          float2 v = ld.pol(w, pol);
          float sq = (v.x * v.x) + (v.y * v.y);
          s1[pol] += sq;
          s2[pol] += sq * sq;
        }
      }
    }
  }
#pragma unroll
  for (int pol = 0; pol < 2; pol++) {
    r1[rg][pol][cl] = s1[pol];
    r2[rg][pol][cl] = s2[pol];
  }
  __syncthreads();
  if (rg == 0 && chan < nchan) {
    for (uint32_t pol = 0; pol < npol; pol++) {
      double a = 0, b = 0;
      for (int g = 0; g < RG; g++) {
        a += r1[g][pol][cl];
        b += r2[g][pol][cl];
      }
      float* o = sums + ((part * nchan + chan) * npol + pol) * 2;
      o[0] = float(a);
      o[1] = float(b);
    }
    counts[part * nchan + chan] = M;  // raw samples are never zeroed data
  }
}

// The same sums for 8-bit samples of two pols with the channel count a
// multiple of 4: a thread takes four neighbouring channels (one 16-byte
// load of their words), a workgroup 128 channels.  The row groups and the
// order of the additions are those of k_sk_sums_raw, so the sums agree.
__global__ __launch_bounds__(kSkThreads) void k_sk_sums_raw8x4(LoadI8 ld, uint64_t t0, uint32_t nchan, uint32_t M,
                                                              uint32_t off, float* __restrict__ sums,
                                                              uint32_t* __restrict__ counts)
{
#pragma clang fp contract(off)
  constexpr int CL = 32, CW = 4, RG = kSkThreads / CL;
  __shared__ double r1[RG][2][CL * CW], r2[RG][2][CL * CW];
  const int tid = threadIdx.x;
  const int cl = tid % CL, rg = tid / CL;
  const uint64_t part = blockIdx.x;
  const uint32_t c0 = blockIdx.y * (CL * CW) + cl * CW;
  double s1[CW][2] = {}, s2[CW][2] = {};
  if (c0 < nchan) {
    const uint64_t tb = t0 + part * off;
    for (uint32_t i = rg; i < M; i += RG) {
      const uint4 w = *reinterpret_cast<const uint4*>(ld.raw + (tb + i) * ld.tstride + uint64_t(c0) * 4);
      const uint32_t wd[CW] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int c = 0; c < CW; c++)
#pragma unroll
        for (int pol = 0; pol < 2; pol++) {
          const float2 v = ld.pol(wd[c], pol);
          const float sq = (v.x * v.x) + (v.y * v.y);
          s1[c][pol] += sq;
          s2[c][pol] += sq * sq;
        }
    }
  }
#pragma unroll
  for (int c = 0; c < CW; c++)
#pragma unroll
    for (int pol = 0; pol < 2; pol++) {
      r1[rg][pol][cl * CW + c] = s1[c][pol];
      r2[rg][pol][cl * CW + c] = s2[c][pol];
    }
  __syncthreads();
  // 128 threads finish one channel each
  if (tid < CL * CW) {
    const uint32_t chan = blockIdx.y * (CL * CW) + tid;
    if (chan < nchan) {
      for (int pol = 0; pol < 2; pol++) {
        double a = 0, b = 0;
        for (int g = 0; g < RG; g++) {
          a += r1[g][pol][tid];
          b += r2[g][pol][tid];
        }
        float* o = sums + ((part * nchan + chan) * 2 + pol) * 2;
        o[0] = float(a);
        o[1] = float(b);
      }
      counts[part * nchan + chan] = M;
    }
  }
}

__global__ void k_fill_u8(uint8_t* p, uint64_t n, uint8_t v)
{
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Rows of the first resolution's mask that a block of a coarser resolution
// flags (detect_skfb / detect_fscr, SpectralKurtosis.C:1121-1131, 1355-1364):
// block b flags rows b * fo + k * fs, k < nflag; (1, 1, 1) for the first.
struct SkFlagMap {
  uint64_t fo;
  uint32_t fs, nflag;
};

// detect_skfb (SpectralKurtosis.C:1148-1210): one thread per (block, channel).
// Counts above M (never produced by the compute kernels) are not tested.
__global__ void k_sk_skfb(float* __restrict__ sums, uint32_t* __restrict__ counts, uint64_t n, uint32_t npol,
                          uint32_t M, uint32_t Nd, int omit, const float* __restrict__ lower, const float* __restrict__ upper,
                          uint8_t* __restrict__ mask, uint32_t nchan, SkFlagMap fm)
{
#pragma clang fp contract(off)
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned Mprime = counts[i];
  if (Mprime == 0 || Mprime > M) return;
  float* sd = sums + i * npol * 2;
  bool zap = true;
  if (Mprime > 2) {
    const float M_fac = float(Mprime * Nd + 1) / (Mprime - 1);
    zap = false;
    for (uint32_t pol = 0; pol < npol; pol++) {
      float S1 = sd[2 * pol], S2 = sd[2 * pol + 1];
      float V = M_fac * (Mprime * (S2 / (S1 * S1)) - 1);
      if (V > upper[Mprime] || V < lower[Mprime]) zap = true;
    }
  }
  if (zap) {
    if (omit) {
      counts[i] = 0;
      for (uint32_t pol = 0; pol < npol; pol++) sd[2 * pol] = sd[2 * pol + 1] = 0.f;
    }
    const uint64_t b = i / nchan, ch = i % nchan;
    for (uint32_t k = 0; k < fm.nflag; k++) mask[(b * fm.fo + uint64_t(k) * fm.fs) * nchan + ch] = 1;
  }
}

// detect_fscr (SpectralKurtosis.C:1369-1451): one workgroup per block; the
// channel sums of each pol are reduced in a fixed order, over the included
// channels (chans, or every channel when null; Resolution::get_channels).
// A coarser resolution's block reads its first mask row (part * fo) and,
// with test set, records its outcome there instead of zapping (k_sk_fscr_scan
// then applies the outcomes in block order).
__global__ __launch_bounds__(kSkThreads) void k_sk_fscr(const float* __restrict__ sums,
                                                       const uint32_t* __restrict__ counts, uint32_t nchan,
                                                       uint32_t npol, uint32_t M, uint32_t Nd, int omit,
                                                       const float* __restrict__ symmetric,
                                                       const uint8_t* __restrict__ chans,
                                                       uint8_t* __restrict__ mask, uint64_t fo,
                                                       uint8_t* __restrict__ test)
{
#pragma clang fp contract(off)
  __shared__ double ra[kSkThreads], rm[kSkThreads], rq[kSkThreads], rn[kSkThreads];
  __shared__ int zap;
  const int tid = threadIdx.x;
  const uint64_t part = blockIdx.x;
  uint8_t* mrow = mask + part * fo * nchan;
  if (tid == 0) zap = 0;
  for (uint32_t pol = 0; pol < npol; pol++) {
    double sk_avg = 0, M_tot = 0, Msq_tot = 0, chan_count = 0;
    for (uint32_t ch = tid; ch < nchan; ch += kSkThreads) {
      if ((!chans || chans[ch]) && (!omit || mrow[ch] == 0)) {
        const float* sd = sums + ((part * nchan + ch) * npol + pol) * 2;
        float S1 = sd[0], S2 = sd[1];
        unsigned Mprime = counts[part * nchan + ch];
        const float M_fac = float(Mprime * Nd + 1) / (Mprime - 1);
        float V = M_fac * (Mprime * (S2 / (S1 * S1)) - 1);
        sk_avg += V * Mprime;
        M_tot += Mprime;
        Msq_tot += Mprime * Mprime;
        chan_count += 1;
      }
    }
    ra[tid] = sk_avg;
    rm[tid] = M_tot;
    rq[tid] = Msq_tot;
    rn[tid] = chan_count;
    for (int s = kSkThreads / 2; s > 0; s >>= 1) {
      __syncthreads();
      if (tid < s) {
        ra[tid] += ra[tid + s];
        rm[tid] += rm[tid + s];
        rq[tid] += rq[tid + s];
        rn[tid] += rn[tid + s];
      }
    }
    __syncthreads();
    if (tid == 0 && rn[0] > 0) {
      double avg = ra[0] / rm[0];
      double M_mean = rm[0] / rn[0];
      double Msq_mean = rq[0] / rn[0];
      double Deff = Msq_mean / (M_mean * M_mean);
      double eff = rn[0] / Deff;
      float threshold = symmetric[min(unsigned(M_mean + 0.5), M)] / sqrt(eff);
      float up = 1.0 + threshold;
      float lo = 1.0 - threshold;
      if (avg > up || avg < lo) zap = 1;
    }
    __syncthreads();
  }
  if (test) {
    if (tid == 0) test[part] = uint8_t(zap);
    return;
  }
  if (zap)
    for (uint32_t ch = tid; ch < nchan; ch += kSkThreads) mrow[ch] = 1;
}

// detect_fscr of a coarser resolution, in block order: a block whose first
// row a preceding block's zap has filled has no channel left to average
// (omit_outliers) and is not zapped itself (SpectralKurtosis.C:1376-1385,
// 1418-1432); one thread, rows[] (the first resolution's rows, cleared by
// the caller) marks the zapped rows.  Without omit_outliers every tested
// block stands.
__global__ void k_sk_fscr_scan(const uint8_t* __restrict__ test, uint64_t npart, SkFlagMap fm, int omit,
                               uint8_t* __restrict__ rows)
{
  for (uint64_t b = 0; b < npart; b++) {
    if (!test[b] || (omit && rows[b * fm.fo])) continue;
    for (uint32_t k = 0; k < fm.nflag; k++) rows[b * fm.fo + uint64_t(k) * fm.fs] = 1;
  }
}

// mask[row][ch] |= rows[row]: grid over nrow * nchan bytes
__global__ void k_sk_rows_or(const uint8_t* __restrict__ rows, uint64_t nrow, uint32_t nchan,
                             uint8_t* __restrict__ mask)
{
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nrow * nchan && rows[i / nchan]) mask[i] = 1;
}

// tscrunch_sums (SpectralKurtosis.C:917-972) out of place: block b of the
// coarser resolution sums blocks (b + i) * offset, i < nsum, of the finer, in
// that order (float sums and integer counts); one thread per (block, chan).
__global__ void k_sk_tscrunch(const float* __restrict__ in_s, const uint32_t* __restrict__ in_c, uint64_t npart,
                              uint32_t nchan, uint32_t npol, uint32_t nsum, uint32_t offset,
                              float* __restrict__ out_s, uint32_t* __restrict__ out_c)
{
#pragma clang fp contract(off)
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= npart * nchan) return;
  const uint64_t b = i / nchan, ch = i % nchan;
  const uint32_t nf = 2 * npol;
  float acc[4];
  uint32_t c = 0;
  for (uint32_t k = 0; k < nf; k++) acc[k] = 0.f;
  for (uint32_t j = 0; j < nsum; j++) {
    const uint64_t src = (b + j) * offset * nchan + ch;
    for (uint32_t k = 0; k < nf; k++) acc[k] = j ? acc[k] + in_s[src * nf + k] : in_s[src * nf + k];
    c += in_c[src];
  }
  for (uint32_t k = 0; k < nf; k++) out_s[i * nf + k] = acc[k];
  out_c[i] = c;
}

// The per-channel totals of the counts (detect_tscr's M_t before its test,
// SpectralKurtosis.C:1036-1040: every noverlap-th block), one workgroup per
// channel.
__global__ __launch_bounds__(kSkThreads) void k_sk_totals(const uint32_t* __restrict__ counts, uint64_t npart,
                                                         uint32_t nchan, uint32_t nov,
                                                         unsigned long long* __restrict__ totals)
{
  __shared__ unsigned long long rc[kSkThreads];
  const uint32_t ch = blockIdx.x;
  const int tid = threadIdx.x;
  unsigned long long c = 0;
  for (uint64_t p = uint64_t(tid) * nov; p < npart; p += uint64_t(kSkThreads) * nov) c += counts[p * nchan + ch];
  rc[tid] = c;
  for (int st = kSkThreads / 2; st > 0; st >>= 1) {
    __syncthreads();
    if (tid < st) rc[tid] += rc[tid + st];
  }
  if (tid == 0) totals[ch] = rc[0];
}

// detect_tscr (SpectralKurtosis.C:1033-1095): one workgroup per channel.
// The reference adds the blocks' sums in block order; here they are added in
// a fixed tree, in double precision as there, which can move the statistic
// only in the last bits of a double.  Every total is tested with
// SKLimits(total) (get_limits, SpectralKurtosis.C:989-1012, 1045): totals up
// to M from the M+1 tables, larger ones from the table of totals (ttot,
// ascending), else, from 32768 on, the closed form of SKLimits.C:152-160.
// A total in (M, 32768) absent from the table is not tested and is counted
// in status[0] (status[1]: the smallest such total), so that the caller
// fails rather than return a mask that misses it.
__global__ __launch_bounds__(kSkThreads) void k_sk_tscr(const float* __restrict__ sums,
                                                       const uint32_t* __restrict__ counts, uint64_t npart,
                                                       uint32_t nchan, uint32_t npol, uint32_t M, uint32_t Nd,
                                                       const float* __restrict__ lower,
                                                       const float* __restrict__ upper, double pcsd,
                                                       const unsigned long long* __restrict__ ttot,
                                                       const float* __restrict__ tlo,
                                                       const float* __restrict__ tup, uint32_t ntscr,
                                                       unsigned long long* __restrict__ status,
                                                       const uint8_t* __restrict__ chans, uint32_t nov,
                                                       uint8_t* __restrict__ mask)
{
#pragma clang fp contract(off)
  // one workgroup per channel: the count and the double sums over the
  // blocks, each thread taking every kSkThreads-th block, then a fixed-order
  // tree in LDS
  __shared__ double r1[2][kSkThreads], r2[2][kSkThreads];
  __shared__ uint64_t rc[kSkThreads];
  __shared__ int zap;
  const uint32_t ch = blockIdx.x;
  if (chans && !chans[ch]) return;  // not an included channel (SpectralKurtosis.C:1032-1033)
  const int tid = threadIdx.x;
  uint64_t count = 0;
  double s1[2] = {0.0, 0.0}, s2[2] = {0.0, 0.0};
  // the non-overlapping blocks: every noverlap-th (:1037, 1067)
  for (uint64_t p = uint64_t(tid) * nov; p < npart; p += uint64_t(kSkThreads) * nov) {
    count += counts[p * nchan + ch];
    for (uint32_t pol = 0; pol < npol; pol++) {
      const uint64_t idx = (p * nchan + ch) * npol + pol;
      s1[pol] += sums[idx * 2];
      s2[pol] += sums[idx * 2 + 1];
    }
  }
  rc[tid] = count;
  for (int pol = 0; pol < 2; pol++) {
    r1[pol][tid] = s1[pol];
    r2[pol][tid] = s2[pol];
  }
  for (int st = kSkThreads / 2; st > 0; st >>= 1) {
    __syncthreads();
    if (tid < st) {
      rc[tid] += rc[tid + st];
      for (int pol = 0; pol < 2; pol++) {
        r1[pol][tid] += r1[pol][tid + st];
        r2[pol][tid] += r2[pol][tid + st];
      }
    }
  }
  if (tid == 0) {
    zap = 0;
    const uint64_t total = rc[0];
    float lo = 0.f, up = 0.f;
    bool test = true;
    if (total == 0) {
      test = false;
    } else if (total <= M) {
      lo = lower[total];
      up = upper[total];
    } else {
      // binary search of the table of totals
      uint32_t a = 0, b = ntscr;
      while (a < b) {
        const uint32_t m = (a + b) / 2;
        if (ttot[m] < total)
          a = m + 1;
        else
          b = m;
      }
      if (a < ntscr && ttot[a] == total) {
        lo = tlo[a];
        up = tup[a];
      } else if (total >= 32768) {
        double factor = sqrt(4.0 / double(total)) * pcsd;
        lo = float(1.0f - factor);
        up = float(1.0f + factor);
      } else {
        test = false;
        atomicAdd(status, 1ull);
        atomicMin(status + 1, (unsigned long long) total);
      }
    }
    if (test) {
      const float M_t = float(total);
      const float M_fac = (M_t * Nd + 1) / (M_t - 1);
      for (uint32_t pol = 0; pol < npol; pol++) {
        const double S1 = r1[pol][0], S2 = r2[pol][0];
        const double SK = M_fac * (M_t * (S2 / (S1 * S1)) - 1);
        if (SK > up || SK < lo) zap = 1;
      }
    }
  }
  __syncthreads();
  if (zap)
    for (uint64_t p = tid; p < npart; p += kSkThreads) mask[p * nchan + ch] = 1;
}

// SpectralKurtosis::mask (SpectralKurtosis.C:1465-1566), one resolution:
// output sample t (< n) is input sample t + idat_start (idat_start = M - off)
// of block min(t / off, npm - 1) + nskip's channel, zeroed in every pol when
// that block is flagged.  (The reference writes the blocks in order, each
// M samples from ipart * off, so the last block over a sample decides it.)
// With noverlap 1: off = M, idat_start = nskip = 0, out[t] = in[t] masked by
// block t / M.  grid (ceil(n / 512), nchan * npol): two complex samples per
// thread, one 16-byte move where both rows allow it and off is even (the
// pair then lies in one block), else two 8-byte moves.
struct SkMaskShape {
  uint32_t nchan, npol, off, nskip;
  uint64_t npm, idat_start, n;
};
__global__ void k_sk_mask(const uint8_t* __restrict__ mask, const float* __restrict__ in, uint64_t ics, uint64_t ips,
                          float* __restrict__ out, uint64_t ocs, uint64_t ops, SkMaskShape m)
{
  const uint64_t t = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 2;
  if (t >= m.n) return;
  const uint32_t chan = blockIdx.y / m.npol, pol = blockIdx.y % m.npol;
  const float2* x = reinterpret_cast<const float2*>(in + chan * ics + pol * ips) + m.idat_start;
  float2* y = reinterpret_cast<float2*>(out + chan * ocs + pol * ops);
  auto flagged = [&](uint64_t k) { return mask[(min(k / m.off, m.npm - 1) + m.nskip) * m.nchan + chan] != 0; };
  const uintptr_t al = reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y);
  if (t + 2 <= m.n && (al & 15) == 0 && (m.off & 1) == 0) {
    float4 v = *reinterpret_cast<const float4*>(x + t);
    if (flagged(t)) v = make_float4(0.f, 0.f, 0.f, 0.f);
    *reinterpret_cast<float4*>(y + t) = v;
    return;
  }
  for (uint64_t k = t; k < t + 2 && k < m.n; k++) {
    float2 v = x[k];
    if (flagged(k)) v = make_float2(0.f, 0.f);
    y[k] = v;
  }
}

__global__ void k_sk_status_init(unsigned long long* status)
{
  status[0] = 0;
  status[1] = ~0ull;
}

// detect_tscr over the blocks' sums, OR-ed into the mask; waits for the
// stream and fails when a channel's total has no thresholds.
// M, nov, lim: the resolution whose blocks the sums are (the last one).
void run_tscr(dspx_sk* s, const float* d_sums, const uint32_t* d_counts, uint64_t npart, uint8_t* d_mask,
              uint32_t M, uint32_t nov, const float* lim)
{
  const dspx_sk_desc& d = s->d;
  hipStream_t st = s->ctx->stream;
  s->status.reserve(2 * sizeof(unsigned long long));
  auto* status = s->status.as<unsigned long long>();
  const float* lower = lim;
  const float* upper = lower + M + 1;
  const auto* ttot = s->tscr.as<unsigned long long>();
  const float* tlo = reinterpret_cast<const float*>(ttot + s->ntscr);
  k_sk_status_init<<<1, 1, 0, st>>>(status);
  k_sk_tscr<<<d.nchan, kSkThreads, 0, st>>>(d_sums, d_counts, npart, d.nchan, d.npol, M, d.Nd, lower, upper,
                                            s->pcsd, ttot, tlo, tlo + s->ntscr, s->ntscr, status,
                                            s->have_chans ? s->chans.as<uint8_t>() : nullptr, nov, d_mask);
  DSPX_HIP(hipGetLastError());
  unsigned long long h[2];
  DSPX_HIP(hipMemcpyAsync(h, status, sizeof(h), hipMemcpyDeviceToHost, st));
  DSPX_HIP(hipStreamSynchronize(st));
  if (h[0])
    fail(DSPX_E_STATE, "detect_tscr: %llu channel totals (the smallest %llu) have no thresholds "
         "(dspx_sk_set_tscr_limits)", h[0], h[1]);
}

}  // namespace

}  // namespace dspx

using namespace dspx;

extern "C" {

int dspx_sk_create(dspx_ctx* ctx, const dspx_sk_desc* desc, dspx_sk** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && desc && out, DSPX_E_PARAM, "null argument");
    const dspx_sk_desc& d = *desc;
    DSPX_REQUIRE(d.nchan >= 1 && d.nchan <= 65535, DSPX_E_PARAM, "nchan=%u", d.nchan);
    DSPX_REQUIRE(d.npol == 1 || d.npol == 2, DSPX_E_PARAM, "npol=%u", d.npol);
    DSPX_REQUIRE(d.M >= 2, DSPX_E_PARAM, "M=%u < 2 (SKLimits::calc_limits)", d.M);
    DSPX_REQUIRE(d.Nd >= 1, DSPX_E_PARAM, "Nd=%u", d.Nd);
    const uint32_t nov = d.noverlap ? d.noverlap : 1;
    DSPX_REQUIRE(d.M % nov == 0, DSPX_E_STATE, "noverlap=%u does not divide M=%u", nov, d.M);
    auto* s = new dspx_sk;
    s->hold.reset(ctx);
    s->ctx = ctx;
    s->d = d;
    s->nov = nov;
    s->off = d.M / nov;
    *out = s;
  });
}

void dspx_sk_destroy(dspx_sk* s)
{
  if (!s) return;
  (void) hipSetDevice(s->ctx->device);
  (void) hipStreamSynchronize(s->ctx->stream);
  delete s;
}

int dspx_sk_set_limits(dspx_sk* s, const float* lower, const float* upper, const float* symmetric,
                       double pc_std_devs)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && lower && upper && symmetric, DSPX_E_PARAM, "null argument");
    const size_t n = size_t(s->d.M) + 1;
    std::vector<float> h(3 * n);
    std::copy(lower, lower + n, h.begin());
    std::copy(upper, upper + n, h.begin() + n);
    std::copy(symmetric, symmetric + n, h.begin() + 2 * n);
    s->lim.reserve(h.size() * sizeof(float));
    DSPX_HIP(hipMemcpyAsync(s->lim.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, s->ctx->stream));
    DSPX_HIP(hipStreamSynchronize(s->ctx->stream));
    s->pcsd = pc_std_devs;
    s->have_limits = true;
  });
}

int dspx_sk_compute(dspx_sk* s, const float* d_in, uint64_t cs, uint64_t ps, uint64_t npart, int zeroed_data,
                    float* d_sums, uint32_t* d_counts)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_in && d_sums && d_counts, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(npart < (1u << 31), DSPX_E_UNSUPPORTED, "npart=%llu", (unsigned long long) npart);
    if (!npart) return;
    dim3 grid(unsigned(npart), s->d.nchan);
    k_sk_sums_fpt<<<grid, kSkThreads, 0, s->ctx->stream>>>(d_in, cs, ps, s->d.nchan, s->d.npol, s->d.M, s->off,
                                                            zeroed_data, d_sums, d_counts);
    DSPX_HIP(hipGetLastError());
  });
}

int dspx_sk_compute_raw(dspx_sk* s, const void* d_raw, uint32_t nbit, uint32_t offset_binary, float scale,
                        uint64_t t0, uint64_t npart, float* d_sums, uint32_t* d_counts)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_raw && d_sums && d_counts, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(npart < (1u << 31), DSPX_E_UNSUPPORTED, "npart=%llu", (unsigned long long) npart);
    if (!npart) return;
    const uint32_t nchan = s->d.nchan, npol = s->d.npol;
    dim3 grid(unsigned(npart), (nchan + 31) / 32);
    if (nbit == 8) {
      LoadI8 ld{static_cast<const uint8_t*>(d_raw), uint64_t(nchan) * npol * 2, uint64_t(npol) * 2, npol,
                offset_binary, scale};
      if (npol == 2 && nchan % 4 == 0 && reinterpret_cast<uintptr_t>(d_raw) % 16 == 0) {
        // rows of 4*nchan bytes: 16-byte aligned with the base
        dim3 g4(unsigned(npart), (nchan + 127) / 128);
        k_sk_sums_raw8x4<<<g4, kSkThreads, 0, s->ctx->stream>>>(ld, t0, nchan, s->d.M, s->off, d_sums, d_counts);
      } else {
        k_sk_sums_raw<<<grid, kSkThreads, 0, s->ctx->stream>>>(ld, t0, nchan, s->d.M, s->off, d_sums, d_counts);
      }
    } else if (nbit == 16) {
      LoadI16 ld{static_cast<const int16_t*>(d_raw), uint64_t(nchan) * npol * 2, uint64_t(npol) * 2, npol,
                 offset_binary};
      k_sk_sums_raw<<<grid, kSkThreads, 0, s->ctx->stream>>>(ld, t0, nchan, s->d.M, s->off, d_sums, d_counts);
    } else {
      fail(DSPX_E_PARAM, "nbit=%u (8 or 16 supported)", nbit);
    }
    DSPX_HIP(hipGetLastError());
  });
}

int dspx_sk_detect(dspx_sk* s, float* d_sums, uint32_t* d_counts, uint64_t npart, uint32_t flags, uint8_t* d_mask)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_sums && d_counts && d_mask, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(s->have_limits, DSPX_E_STATE, "thresholds not set (dspx_sk_set_limits)");
    DSPX_REQUIRE((flags & ~7u) == 0, DSPX_E_PARAM, "flags=%u", flags);
    DSPX_REQUIRE(npart < (1u << 31), DSPX_E_UNSUPPORTED, "npart=%llu", (unsigned long long) npart);
    const dspx_sk_desc& d = s->d;
    hipStream_t st = s->ctx->stream;
    const uint64_t n = npart * d.nchan;
    if (!n) return;
    const float* lower = s->lim.as<float>();
    const float* upper = lower + d.M + 1;
    const float* sym = upper + d.M + 1;
    // reset_mask (a fill kernel rather than a memset node, so that the
    // sequence can be captured into a graph)
    k_fill_u8<<<unsigned((n + 255) / 256), 256, 0, st>>>(d_mask, n, 0);
    if (flags & DSPX_SK_SKFB)
      k_sk_skfb<<<unsigned((n + 255) / 256), 256, 0, st>>>(d_sums, d_counts, n, d.npol, d.M, d.Nd,
                                                          d.omit_outliers, lower, upper, d_mask, d.nchan,
                                                          SkFlagMap{1, 1, 1});
    if (flags & DSPX_SK_FSCR)
      k_sk_fscr<<<unsigned(npart), kSkThreads, 0, st>>>(d_sums, d_counts, d.nchan, d.npol, d.M, d.Nd,
                                                        d.omit_outliers, sym,
                                                        s->have_chans ? s->chans.as<uint8_t>() : nullptr, d_mask,
                                                        1, nullptr);
    DSPX_HIP(hipGetLastError());
    if (flags & DSPX_SK_TSCR) run_tscr(s, d_sums, d_counts, npart, d_mask, d.M, s->nov, lower);
  });
}

int dspx_sk_set_tscr_limits(dspx_sk* s, uint32_t n, const uint64_t* totals, const float* lower, const float* upper)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && (n == 0 || (totals && lower && upper)), DSPX_E_PARAM, "null argument");
    for (uint32_t i = 0; i < n; i++) {
      DSPX_REQUIRE(totals[i] > s->d.M, DSPX_E_PARAM, "total %llu <= M=%u (the M+1 tables hold it)",
                   (unsigned long long) totals[i], s->d.M);
      DSPX_REQUIRE(i == 0 || totals[i] > totals[i - 1], DSPX_E_PARAM, "totals not ascending at %u", i);
    }
    const size_t bytes = size_t(n) * (sizeof(unsigned long long) + 2 * sizeof(float));
    std::vector<unsigned char> h(std::max<size_t>(bytes, 1));
    std::memcpy(h.data(), totals, size_t(n) * sizeof(uint64_t));
    std::memcpy(h.data() + size_t(n) * 8, lower, size_t(n) * sizeof(float));
    std::memcpy(h.data() + size_t(n) * 12, upper, size_t(n) * sizeof(float));
    s->tscr.reserve(h.size());
    DSPX_HIP(hipMemcpyAsync(s->tscr.p, h.data(), h.size(), hipMemcpyHostToDevice, s->ctx->stream));
    DSPX_HIP(hipStreamSynchronize(s->ctx->stream));
    s->ntscr = n;
  });
}

int dspx_sk_set_channels(dspx_sk* s, const uint8_t* include)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s, DSPX_E_PARAM, "null argument");
    if (!include) {
      s->have_chans = false;
      return;
    }
    s->chans.reserve(s->d.nchan);
    DSPX_HIP(hipMemcpyAsync(s->chans.p, include, s->d.nchan, hipMemcpyHostToDevice, s->ctx->stream));
    DSPX_HIP(hipStreamSynchronize(s->ctx->stream));
    s->have_chans = true;
  });
}

int dspx_sk_blocks(const dspx_sk* s, uint64_t ndat, uint64_t* npart, uint64_t* ndat_out)
{
  if (!s || !npart || !ndat_out) return DSPX_E_PARAM;
  // Resolution::prepare (SpectralKurtosis.C:263-281)
  const uint64_t off = s->off, idat_start = s->d.M - off;
  *npart = 0;
  *ndat_out = 0;
  if (ndat < idat_start) return DSPX_OK;
  const uint64_t nblock = (ndat - idat_start) / off;
  if (nblock * off < idat_start) return DSPX_OK;
  *npart = nblock;
  *ndat_out = nblock * off - idat_start;
  return DSPX_OK;
}

int dspx_sk_tscr_totals(dspx_sk* s, const uint32_t* d_counts, uint64_t npart, uint64_t* h_totals)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_counts && h_totals, DSPX_E_PARAM, "null argument");
    const uint32_t nchan = s->d.nchan;
    hipStream_t st = s->ctx->stream;
    s->tot.reserve(size_t(nchan) * sizeof(unsigned long long));
    k_sk_totals<<<nchan, kSkThreads, 0, st>>>(d_counts, npart, nchan, s->nov, s->tot.as<unsigned long long>());
    DSPX_HIP(hipGetLastError());
    DSPX_HIP(hipMemcpyAsync(h_totals, s->tot.p, size_t(nchan) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    DSPX_HIP(hipStreamSynchronize(st));
  });
}

int dspx_sk_detect_tscr(dspx_sk* s, const float* d_sums, const uint32_t* d_counts, uint64_t npart, uint8_t* d_mask)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_sums && d_counts && d_mask, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(s->have_limits, DSPX_E_STATE, "thresholds not set (dspx_sk_set_limits)");
    DSPX_REQUIRE(npart < (1u << 31), DSPX_E_UNSUPPORTED, "npart=%llu", (unsigned long long) npart);
    if (!npart) return;
    run_tscr(s, d_sums, d_counts, npart, d_mask, s->d.M, s->nov, s->lim.as<float>());
  });
}

int dspx_sk_mask(dspx_sk* s, const uint8_t* d_mask, const float* d_in, uint64_t ics, uint64_t ips, float* d_out,
                 uint64_t ocs, uint64_t ops, uint64_t npart)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_mask && d_in && d_out, DSPX_E_PARAM, "null argument");
    if (!npart) return;
    // Resolution::prepare (SpectralKurtosis.C:263-281) and mask (:1480-1510)
    // with one resolution: the output holds (npart + 1) * off - M samples
    const uint64_t M = s->d.M, off = s->off;
    DSPX_REQUIRE((npart + 1) * off >= 2 * M, DSPX_E_PARAM, "%llu blocks of %llu with noverlap %u: no output",
                 (unsigned long long) npart, (unsigned long long) M, s->nov);
    SkMaskShape m{s->d.nchan, s->d.npol, uint32_t(off), s->nov - 1, 0, M - off, (npart + 1) * off - M};
    m.npm = (m.n - M) / off + 1;
    DSPX_REQUIRE(m.npm + m.nskip <= npart, DSPX_E_INTERNAL, "mask blocks");
    DSPX_REQUIRE(d_in != d_out || s->nov == 1, DSPX_E_PARAM, "in place with noverlap %u", s->nov);
    const uint64_t n = m.n;
    DSPX_REQUIRE((n + 511) / 512 < (1ull << 31), DSPX_E_UNSUPPORTED, "block too long");
    dim3 grid(unsigned((n + 511) / 512), s->d.nchan * s->d.npol);
    DSPX_REQUIRE(grid.y <= 65535, DSPX_E_UNSUPPORTED, "nchan*npol=%u", grid.y);
    k_sk_mask<<<grid, 256, 0, s->ctx->stream>>>(d_mask, d_in, ics, ips, d_out, ocs, ops, m);
    DSPX_HIP(hipGetLastError());
  });
}


int dspx_sk_add_resolution(dspx_sk* s, uint32_t M, uint32_t noverlap, const float* lower, const float* upper,
                           const float* symmetric, const uint8_t* include)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && lower && upper && symmetric, DSPX_E_PARAM, "null argument");
    const uint32_t nov = noverlap ? noverlap : 1;
    DSPX_REQUIRE(M >= 2 && M % nov == 0, DSPX_E_STATE, "noverlap=%u does not divide M=%u", nov, M);
    const uint32_t pM = s->res.empty() ? s->d.M : s->res.back()->M;
    const uint32_t poff = s->res.empty() ? s->off : s->res.back()->off;
    // Resolution::compatible (SpectralKurtosis.C:283-294), in increasing M
    DSPX_REQUIRE(M > pM && M % pM == 0, DSPX_E_STATE, "M=%u not a larger multiple of the previous M=%u", M, pM);
    DSPX_REQUIRE((M / nov) % poff == 0, DSPX_E_STATE, "overlap offset %u not divisible by the previous %u",
                 M / nov, poff);
    std::unique_ptr<dspx_sk::Res> r(new dspx_sk::Res);
    r->M = M;
    r->nov = nov;
    r->off = M / nov;
    const size_t n = size_t(M) + 1;
    std::vector<float> h(3 * n);
    std::copy(lower, lower + n, h.begin());
    std::copy(upper, upper + n, h.begin() + n);
    std::copy(symmetric, symmetric + n, h.begin() + 2 * n);
    r->lim.reserve(h.size() * sizeof(float));
    DSPX_HIP(hipMemcpy(r->lim.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    if (include) {
      r->chans.reserve(s->d.nchan);
      DSPX_HIP(hipMemcpy(r->chans.p, include, s->d.nchan, hipMemcpyHostToDevice));
      r->have_chans = true;
    }
    s->res.push_back(std::move(r));
    s->have_last = false;
  });
}

int dspx_sk_clear_resolutions(dspx_sk* s)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipStreamSynchronize(s->ctx->stream));
    s->res.clear();
    s->have_last = false;
  });
}

namespace {

// Resolution::prepare (SpectralKurtosis.C:263-281) of blocks of M every off.
uint64_t sk_nblock(uint64_t ndat, uint64_t M, uint64_t off, uint64_t* ndat_out)
{
  const uint64_t idat_start = M - off;
  *ndat_out = 0;
  if (ndat < idat_start) return 0;
  const uint64_t nblock = (ndat - idat_start) / off;
  if (nblock * off < idat_start) return 0;
  *ndat_out = nblock * off - idat_start;
  return nblock;
}

}  // namespace

int dspx_sk_blocks_res(const dspx_sk* s, uint64_t ndat, uint64_t* npart, uint64_t* ndat_out)
{
  if (!s || !npart || !ndat_out) return DSPX_E_PARAM;
  uint64_t nout0;
  *npart = sk_nblock(ndat, s->d.M, s->off, &nout0);
  if (s->res.empty()) {
    *ndat_out = nout0;
  } else {
    const auto& r = *s->res.back();
    sk_nblock(ndat, r.M, r.off, ndat_out);
  }
  return DSPX_OK;
}

int dspx_sk_detect_res(dspx_sk* s, float* d_sums, uint32_t* d_counts, uint64_t ndat, uint32_t flags, uint8_t* d_mask)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_sums && d_counts && d_mask, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(s->have_limits, DSPX_E_STATE, "thresholds not set (dspx_sk_set_limits)");
    DSPX_REQUIRE((flags & ~7u) == 0, DSPX_E_PARAM, "flags=%u", flags);
    const dspx_sk_desc& d = s->d;
    hipStream_t st = s->ctx->stream;
    uint64_t nout;
    const uint64_t np0 = sk_nblock(ndat, d.M, s->off, &nout);
    DSPX_REQUIRE(np0 < (1u << 31), DSPX_E_UNSUPPORTED, "npart=%llu", (unsigned long long) np0);
    s->have_last = false;
    const uint64_t n0 = np0 * d.nchan;
    if (!n0) return;
    // the first resolution, as dspx_sk_detect without its tscr
    k_fill_u8<<<unsigned((n0 + 255) / 256), 256, 0, st>>>(d_mask, n0, 0);
    const float* lim0 = s->lim.as<float>();
    if (flags & DSPX_SK_SKFB)
      k_sk_skfb<<<unsigned((n0 + 255) / 256), 256, 0, st>>>(d_sums, d_counts, n0, d.npol, d.M, d.Nd,
                                                           d.omit_outliers, lim0, lim0 + d.M + 1, d_mask, d.nchan,
                                                           SkFlagMap{1, 1, 1});
    if (flags & DSPX_SK_FSCR)
      k_sk_fscr<<<unsigned(np0), kSkThreads, 0, st>>>(d_sums, d_counts, d.nchan, d.npol, d.M, d.Nd,
                                                      d.omit_outliers, lim0 + 2 * (d.M + 1),
                                                      s->have_chans ? s->chans.as<uint8_t>() : nullptr, d_mask, 1,
                                                      nullptr);
    DSPX_HIP(hipGetLastError());
    const float* cs = d_sums;
    const uint32_t* cc = d_counts;
    uint64_t have = np0, np = np0;
    uint32_t pM = d.M, poff = s->off, M = d.M, nov = s->nov;
    const float* lim = lim0;
    const size_t fpd = size_t(d.nchan) * d.npol * 2;
    for (size_t ir = 0; ir < s->res.size(); ir++) {
      const auto& r = *s->res[ir];
      uint64_t nr_out;
      np = sk_nblock(ndat, r.M, r.off, &nr_out);
      const uint32_t nsum = r.M / pM, offset = r.off / poff;
      // tscrunch_sums asserts that the last block it reads exists
      DSPX_REQUIRE(np == 0 || (np - 1) * offset + uint64_t(nsum - 1) * offset < have, DSPX_E_STATE,
                   "resolution M=%u noverlap=%u reads block %llu of %llu (tscrunch_sums)", r.M, r.nov,
                   (unsigned long long) ((np - 1) * offset + uint64_t(nsum - 1) * offset),
                   (unsigned long long) have);
      if (!np) break;
      DevBuf& os = s->rsum[ir & 1];
      DevBuf& oc = s->rcnt[ir & 1];
      os.reserve(np * fpd * sizeof(float));
      oc.reserve(np * d.nchan * sizeof(uint32_t));
      const uint64_t nn = np * d.nchan;
      k_sk_tscrunch<<<unsigned((nn + 255) / 256), 256, 0, st>>>(cs, cc, np, d.nchan, d.npol, nsum, offset,
                                                               os.as<float>(), oc.as<uint32_t>());
      cs = os.as<float>();
      cc = oc.as<uint32_t>();
      const uint32_t nflag = r.M / d.M;
      const SkFlagMap fm{uint64_t(nflag) * s->nov / r.nov, s->nov, nflag};
      const float* rl = r.lim.as<float>();
      if (flags & DSPX_SK_SKFB)
        k_sk_skfb<<<unsigned((nn + 255) / 256), 256, 0, st>>>(os.as<float>(), oc.as<uint32_t>(), nn, d.npol, r.M,
                                                             d.Nd, d.omit_outliers, rl, rl + r.M + 1, d_mask,
                                                             d.nchan, fm);
      if (flags & DSPX_SK_FSCR) {
        s->rtest.reserve(np);
        s->rrow.reserve(np0);
        k_sk_fscr<<<unsigned(np), kSkThreads, 0, st>>>(cs, cc, d.nchan, d.npol, r.M, d.Nd, d.omit_outliers,
                                                       rl + 2 * (r.M + 1),
                                                       r.have_chans ? r.chans.as<uint8_t>() : nullptr, d_mask, fm.fo,
                                                       s->rtest.as<uint8_t>());
        k_fill_u8<<<unsigned((np0 + 255) / 256), 256, 0, st>>>(s->rrow.as<uint8_t>(), np0, 0);
        k_sk_fscr_scan<<<1, 1, 0, st>>>(s->rtest.as<uint8_t>(), np, fm, d.omit_outliers, s->rrow.as<uint8_t>());
        k_sk_rows_or<<<unsigned((n0 + 255) / 256), 256, 0, st>>>(s->rrow.as<uint8_t>(), np0, d.nchan, d_mask);
      }
      DSPX_HIP(hipGetLastError());
      have = np;
      pM = r.M;
      poff = r.off;
      M = r.M;
      nov = r.nov;
      lim = rl;
    }
    s->last_sums = cs;
    s->last_counts = cc;
    s->last_npart = np;
    s->last_rows = np0;
    s->last_M = M;
    s->last_nov = nov;
    s->last_lim = lim;
    s->have_last = true;
    if ((flags & DSPX_SK_TSCR) && np) run_tscr(s, cs, cc, np, d_mask, M, nov, lim);
  });
}

int dspx_sk_res_totals(dspx_sk* s, uint64_t* h_totals)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && h_totals, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(s->have_last, DSPX_E_STATE, "no dspx_sk_detect_res to follow");
    const uint32_t nchan = s->d.nchan;
    hipStream_t st = s->ctx->stream;
    s->tot.reserve(size_t(nchan) * sizeof(unsigned long long));
    if (s->last_npart)
      k_sk_totals<<<nchan, kSkThreads, 0, st>>>(s->last_counts, s->last_npart, nchan, s->last_nov,
                                                s->tot.as<unsigned long long>());
    else
      k_fill_u8<<<unsigned((8 * nchan + 255) / 256), 256, 0, st>>>(s->tot.as<uint8_t>(), 8ull * nchan, 0);
    DSPX_HIP(hipGetLastError());
    DSPX_HIP(hipMemcpyAsync(h_totals, s->tot.p, size_t(nchan) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    DSPX_HIP(hipStreamSynchronize(st));
  });
}

int dspx_sk_res_tscr(dspx_sk* s, uint8_t* d_mask)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_mask, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(s->have_last, DSPX_E_STATE, "no dspx_sk_detect_res to follow");
    if (!s->last_npart) return;
    run_tscr(s, s->last_sums, s->last_counts, s->last_npart, d_mask, s->last_M, s->last_nov, s->last_lim);
  });
}

int dspx_sk_mask_res(dspx_sk* s, const uint8_t* d_mask, const float* d_in, uint64_t ics, uint64_t ips,
                     float* d_out, uint64_t ocs, uint64_t ops, uint64_t ndat)
{
  return guard(s ? s->ctx : nullptr, [&] {
    DSPX_REQUIRE(s && d_mask && d_in && d_out, DSPX_E_PARAM, "null argument");
    const uint64_t M = s->d.M, off = s->off;
    const uint64_t Ml = s->res.empty() ? M : s->res.back()->M;
    const uint64_t offl = s->res.empty() ? off : s->res.back()->off;
    uint64_t nout0, nout;
    const uint64_t np0 = sk_nblock(ndat, M, off, &nout0);
    sk_nblock(ndat, Ml, offl, &nout);
    if (nout < M) return;  // no whole block of the first resolution
    // SpectralKurtosis::mask (:1480-1510): nskip rows skipped, npm blocks
    SkMaskShape m{s->d.nchan, s->d.npol, uint32_t(off), uint32_t((Ml - offl) / off), 0, Ml - offl, nout};
    DSPX_REQUIRE((Ml - offl) % off == 0 && (nout - M) % off == 0, DSPX_E_STATE, "mask blocks not aligned");
    m.npm = (nout - M) / off + 1;
    DSPX_REQUIRE(m.npm + m.nskip <= np0, DSPX_E_INTERNAL, "mask blocks");
    DSPX_REQUIRE(d_in != d_out || (Ml == offl && M == off), DSPX_E_PARAM, "in place with overlapping blocks");
    DSPX_REQUIRE((nout + 511) / 512 < (1ull << 31), DSPX_E_UNSUPPORTED, "block too long");
    dim3 grid(unsigned((nout + 511) / 512), s->d.nchan * s->d.npol);
    DSPX_REQUIRE(grid.y <= 65535, DSPX_E_UNSUPPORTED, "nchan*npol=%u", grid.y);
    k_sk_mask<<<grid, 256, 0, s->ctx->stream>>>(d_mask, d_in, ics, ips, d_out, ocs, ops, m);
    DSPX_HIP(hipGetLastError());
  });
}

}  // extern "C"
