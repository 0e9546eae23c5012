// rfinet.hip -- RFI proposal network: per-block features from the spectral
// kurtosis sums, and a 4-layer 1-D convolutional classifier across frequency
// channels (forward, backward, loss and Adam) trained on the SK detector's
// masks.  No counterpart in the reference: it proposes blocks ahead of the
// SK detector (SpectralKurtosis.C:830-909), which keeps the final decision.
//
// Layout: activations [nblk][C][nchan] float (the channel axis is the
// convolution's axis, contiguous); weights W[co][ci][k] then bias[co] per
// layer, an FP32 master copy and the BF16-rounded copy the forward and
// backward passes use (values widened to FP32, FP32 accumulation).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "dspx_internal.hpp"

namespace dspx {
namespace {

constexpr int kF = 4;    // input features
constexpr int kW = 32;   // hidden width
constexpr int kK = 19;   // taps
constexpr int kP = kK / 2;
constexpr int kL = 4;    // layers
constexpr int kTileX = 256;

constexpr int cin_of(int l) { return l == 0 ? kF : kW; }
constexpr int cout_of(int l) { return l == kL - 1 ? 1 : kW; }
constexpr int nparam_of(int l) { return cout_of(l) * cin_of(l) * kK + cout_of(l); }
constexpr int off_of(int l) { return l == 0 ? 0 : off_of(l - 1) + nparam_of(l - 1); }
constexpr int kNParam = off_of(kL);

__device__ __forceinline__ float bf16_round(float x)
{
  // round to nearest even on the upper 16 bits (finite inputs)
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}

// ---------------------------------------------------------------- features
// Per (block, channel): log total power, SK estimator of each pol (the
// float estimator of detect_skfb, SpectralKurtosis.C:1166-1175).
__global__ void k_rfi_stats(const float* __restrict__ sums, const uint32_t* __restrict__ counts, uint64_t n,
                            uint32_t npol, uint32_t Nd, float* __restrict__ lp, float* __restrict__ v1,
                            float* __restrict__ vmax)
{
#pragma clang fp contract(off)
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned Mp = counts[i];
  const float* sd = sums + i * npol * 2;
  float s1 = 0.f, vs = 0.f, vm = 0.f;
  if (Mp > 2) {
    const float M_fac = float(Mp * Nd + 1) / (Mp - 1);
    for (uint32_t p = 0; p < npol; p++) {
      float S1 = sd[2 * p], S2 = sd[2 * p + 1];
      float V = S1 > 0.f ? M_fac * (Mp * (S2 / (S1 * S1)) - 1) : 1.f;
      s1 += S1;
      vs += V - 1.f;
      vm = fmaxf(vm, fabsf(V - 1.f));
    }
    lp[i] = logf(s1 / float(Mp) + 1e-20f);
  } else {
    lp[i] = 0.f;
  }
  v1[i] = vs / float(npol);
  vmax[i] = vm;
}

// per-channel mean of the log power over the blocks: a workgroup per 32
// channels, whose 8 row groups each add every 8th block (rows of 32
// channels read together), then the 8 partials added in a fixed order
__global__ __launch_bounds__(256) void k_rfi_chan_mean(const float* __restrict__ lp, uint64_t nblk, uint32_t nchan,
                                                       float* __restrict__ mean)
{
  constexpr int CH = 32, RG = 8;
  __shared__ double part[RG][CH];
  const int cl = threadIdx.x % CH, rg = threadIdx.x / CH;
  const uint32_t c = blockIdx.x * CH + cl;
  double s = 0;
  if (c < nchan)
    for (uint64_t b = rg; b < nblk; b += RG) s += lp[b * nchan + c];
  part[rg][cl] = s;
  __syncthreads();
  if (rg == 0 && c < nchan) {
    double t = 0;
    for (int g = 0; g < RG; g++) t += part[g][cl];
    mean[c] = float(t / double(nblk));
  }
}

// feat[b][f][c]: f0 = log power - channel mean, f1 = mean SK deviation,
// f2 = largest |SK - 1| of the pols, f3 = change of log power since the
// previous block (0 for the first)
__global__ void k_rfi_features(const float* __restrict__ lp, const float* __restrict__ v1,
                               const float* __restrict__ vmax, const float* __restrict__ mean, uint64_t nblk,
                               uint32_t nchan, float* __restrict__ feat)
{
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nblk * nchan) return;
  uint64_t b = i / nchan;
  uint32_t c = uint32_t(i % nchan);
  float* f = feat + b * kF * nchan + c;
  f[0] = lp[i] - mean[c];
  f[nchan] = v1[i];
  f[2 * nchan] = vmax[i];
  f[3 * nchan] = b ? lp[i] - lp[i - nchan] : 0.f;
}

// ------------------------------------------------------------------ layers
// Y[b][co][x] = bias[co] + sum_ci sum_k W[co][ci][k] * A[b][ci][x + k - P]
// (zero padding), optionally ReLU.  grid (ceil(nchan/256), nblk); the input
// tile of all CIN rows sits in LDS, the weights are workgroup-uniform.
// With MASK, the output is multiplied by (Z[b][co][x] > 0): the gradient
// through the ReLU of the layer whose output Z is.
template <int CIN, int COUT, bool RELU, bool MASK>
__global__ __launch_bounds__(kTileX) void k_conv(const float* __restrict__ A, const float* __restrict__ W,
                                                  uint32_t nchan, float* __restrict__ Y, const float* __restrict__ Z)
{
  __shared__ float t[CIN][kTileX + kK - 1];
  const int tid = threadIdx.x;
  const uint32_t x0 = blockIdx.x * kTileX;
  const uint64_t b = blockIdx.y;
  const float* a = A + b * CIN * nchan;
  for (int ci = 0; ci < CIN; ci++)
    for (int j = tid; j < kTileX + kK - 1; j += kTileX) {
      int64_t x = int64_t(x0) + j - kP;
      t[ci][j] = (x >= 0 && x < nchan) ? a[uint64_t(ci) * nchan + x] : 0.f;
    }
  __syncthreads();
  const uint32_t x = x0 + tid;
  if (x >= nchan) return;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; co++) acc[co] = W[COUT * CIN * kK + co];
  for (int ci = 0; ci < CIN; ci++) {
#pragma unroll
    for (int k = 0; k < kK; k++) {
      float v = t[ci][tid + k];
#pragma unroll
      for (int co = 0; co < COUT; co++) acc[co] += W[(co * CIN + ci) * kK + k] * v;
    }
  }
  float* y = Y + b * COUT * nchan + x;
  const float* z = MASK ? Z + b * COUT * nchan + x : nullptr;
#pragma unroll
  for (int co = 0; co < COUT; co++) {
    float v = acc[co];
    if (RELU) v = fmaxf(v, 0.f);
    if (MASK) v = z[uint64_t(co) * nchan] > 0.f ? v : 0.f;
    y[uint64_t(co) * nchan] = v;
  }
}

// ------------------------------------------------- 32 -> 32 layers on MFMA
// The same convolution as an implicit GEMM on the matrix cores:
// Y[co][x] = bias[co] + sum_k W'[co][k] * B[k][x], k = tap * 32 + ci,
// B[k][x] = A[ci][x + tap - P], with 16x16x32 BF16 products accumulated in
// FP32.  The weights are BF16 already (the network's weights are kept
// rounded); each activation is split into BF16 hi = bf16(a) and lo =
// bf16(a - hi), and both halves are multiplied, so the products carry
// about 16 mantissa bits of the activation (relative error near 2^-17)
// instead of 8.
//
// Lane layout of v_mfma_f32_16x16x32_bf16 (lane l, r = l % 16, q = l / 16):
// A operand A[r][8q + j], B operand B[8q + j][r], j < 8 (8 BF16 each);
// result D[4q + j][r].  A workgroup of 4 waves computes 32 output channels
// (two 16-row tiles) at kMT positions (kMT / 64 16-column tiles per wave); k runs
// in 19 steps of 32 (one tap, input channels 8q + j).
constexpr int kMT = 128;             // positions per workgroup (a multiple of 64)
constexpr int kMXP = kMT + kK - 1;   // activation tile width
constexpr int kMRS = kW + 4;         // LDS row stride (BF16 units): rows 72 bytes apart
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef short s16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// eight BF16 (bit patterns) from two 8-byte loads
__device__ __forceinline__ bf16x8_t bf16x8_of(bf16x4_t lo, bf16x4_t hi)
{
  const s16x8_t v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8_t, v);
}

__device__ __forceinline__ uint16_t bf16_bits_rne(float f)
{
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

// BF16 weights of a 32 -> 32 layer as [tap][co][ci] (exact: the upper half
// of each BF16-rounded weight)
__global__ void k_conv_w_bf16(const float* __restrict__ W, uint16_t* __restrict__ Wb)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= uint32_t(kK * kW * kW)) return;
  const uint32_t ci = i % kW, co = (i / kW) % kW, tap = i / (kW * kW);
  Wb[i] = uint16_t(__float_as_uint(W[(co * kW + ci) * kK + tap]) >> 16);
}

// grid (ceil(nchan / kMT), nblk), block 256.  The weight fragments are read
// from global memory (every workgroup reads the same 39 KB, which stay in
// the caches); each wave takes NT 16-position column tiles.
template <bool RELU>
__global__ __launch_bounds__(256) void k_conv_mfma(const float* __restrict__ A, const uint16_t* __restrict__ Wb,
                                                   const float* __restrict__ bias, uint32_t nchan,
                                                   float* __restrict__ Y)
{
  constexpr int NT = kMT / 64;   // column tiles per wave
  __shared__ __align__(16) uint16_t ahi[kMXP * kMRS], alo[kMXP * kMRS];
  const int tid = threadIdx.x;
  const uint32_t x0 = blockIdx.x * kMT;
  const uint64_t b = blockIdx.y;
  // activations [ci][x0 - P + j] -> hi/lo at [j][ci]
  const float* a = A + b * kW * nchan;
  for (int i = tid; i < kMXP * kW; i += 256) {
    const int ci = i / kMXP, j = i % kMXP;
    const int64_t x = int64_t(x0) + j - kP;
    const float v = (x >= 0 && x < int64_t(nchan)) ? a[uint64_t(ci) * nchan + uint64_t(x)] : 0.f;
    const uint16_t h = bf16_bits_rne(v);
    ahi[j * kMRS + ci] = h;
    alo[j * kMRS + ci] = bf16_bits_rne(v - __uint_as_float(uint32_t(h) << 16));
  }
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6;
  const int r = lane & 15, q = lane >> 4;
  f32x4_t acc[2][NT];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[m][t][j] = bias[16 * m + 4 * q + j];
  const bf16x4_t* wg = reinterpret_cast<const bf16x4_t*>(Wb);
  const int ci = 8 * q;
#pragma unroll 2
  for (int tap = 0; tap < kK; tap++) {
    bf16x8_t aw[2];
#pragma unroll
    for (int m = 0; m < 2; m++) {
      const int w0 = ((tap * kW + 16 * m + r) * kW + ci) / 4;
      aw[m] = bf16x8_of(wg[w0], wg[w0 + 1]);
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int jj = 16 * (wv * NT + t) + r + tap;
      const bf16x4_t* ph = reinterpret_cast<const bf16x4_t*>(&ahi[jj * kMRS + ci]);
      const bf16x4_t* pl = reinterpret_cast<const bf16x4_t*>(&alo[jj * kMRS + ci]);
      const bf16x8_t bh = bf16x8_of(ph[0], ph[1]);
      const bf16x8_t bl = bf16x8_of(pl[0], pl[1]);
#pragma unroll
      for (int m = 0; m < 2; m++) {
        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[m], bh, acc[m][t], 0, 0, 0);
        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[m], bl, acc[m][t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const uint32_t x = x0 + 16 * (wv * NT + t) + r;
    if (x >= nchan) continue;
    float* y = Y + b * kW * nchan + x;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        float v = acc[m][t][j];
        if (RELU) v = fmaxf(v, 0.f);
        y[uint64_t(16 * m + 4 * q + j) * nchan] = v;
      }
  }
}

// Weights of the layer transposed and flipped for the data gradient:
// WT[ci][co][k] = W[co][ci][K-1-k], bias 0.
template <int CIN, int COUT>
__global__ void k_transpose_w(const float* __restrict__ W, float* __restrict__ WT)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < CIN * COUT * kK) {
    int k = i % kK, co = (i / kK) % COUT, ci = i / (kK * COUT);
    WT[i] = W[(co * CIN + ci) * kK + (kK - 1 - k)];
  } else if (i < CIN * COUT * kK + CIN) {
    WT[i] = 0.f;
  }
}

// Weight gradient partials: dW[co][ci][k] = sum_{b,x} G[b][co][x] A[b][ci][x+k-P],
// dbias[co] = sum G.  grid (nwg): workgroup g takes blocks g, g+nwg, ...;
// thread j owns the (co, ci) pairs j, j+256, ... with a K-tap sliding window.
// Writes part[g][nparam of the layer].
template <int CIN, int COUT>
__global__ __launch_bounds__(kTileX) void k_conv_dw(const float* __restrict__ A, const float* __restrict__ G,
                                                     uint64_t nblk, uint32_t nchan, float* __restrict__ part)
{
  constexpr int NPAIR = CIN * COUT;
  constexpr int PPT = (NPAIR + kTileX - 1) / kTileX;
  constexpr int TX = 128;
  __shared__ float ta[CIN][TX + kK - 1];
  __shared__ float tg[COUT][TX];
  const int tid = threadIdx.x;
  float acc[PPT][kK];
  float bacc = 0.f;
#pragma unroll
  for (int p = 0; p < PPT; p++)
#pragma unroll
    for (int k = 0; k < kK; k++) acc[p][k] = 0.f;
  for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const float* a = A + b * CIN * nchan;
    const float* g = G + b * COUT * nchan;
    for (uint32_t x0 = 0; x0 < nchan; x0 += TX) {
      __syncthreads();
      for (int i = tid; i < CIN * (TX + kK - 1); i += kTileX) {
        int ci = i / (TX + kK - 1), j = i % (TX + kK - 1);
        int64_t x = int64_t(x0) + j - kP;
        ta[ci][j] = (x >= 0 && x < nchan) ? a[uint64_t(ci) * nchan + x] : 0.f;
      }
      for (int i = tid; i < COUT * TX; i += kTileX) {
        int co = i / TX, j = i % TX;
        tg[co][j] = x0 + j < nchan ? g[uint64_t(co) * nchan + x0 + j] : 0.f;
      }
      __syncthreads();
#pragma unroll
      for (int p = 0; p < PPT; p++) {
        const int pair = tid + p * kTileX;
        if (pair < NPAIR) {
          const int co = pair / CIN, ci = pair % CIN;
          float win[kK];
#pragma unroll
          for (int k = 0; k < kK; k++) win[k] = ta[ci][k];
          for (int j = 0; j < TX; j++) {
            float gv = tg[co][j];
#pragma unroll
            for (int k = 0; k < kK; k++) acc[p][k] += gv * win[k];
#pragma unroll
            for (int k = 0; k < kK - 1; k++) win[k] = win[k + 1];
            win[kK - 1] = ta[ci][j + kK < TX + kK - 1 ? j + kK : TX + kK - 2];
          }
        }
      }
      if (tid < COUT)
        for (int j = 0; j < TX; j++) bacc += tg[tid][j];
    }
  }
  float* o = part + uint64_t(blockIdx.x) * (COUT * CIN * kK + COUT);
#pragma unroll
  for (int p = 0; p < PPT; p++) {
    const int pair = tid + p * kTileX;
    if (pair < NPAIR) {
      const int co = pair / CIN, ci = pair % CIN;
#pragma unroll
      for (int k = 0; k < kK; k++) o[(co * CIN + ci) * kK + k] = acc[p][k];
    }
  }
  if (tid < COUT) o[COUT * CIN * kK + tid] = bacc;
}

// grad[i] = sum over the nwg partials, in order
__global__ void k_sum_parts(const float* __restrict__ part, uint32_t nwg, uint32_t n, float* __restrict__ grad)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (uint32_t g = 0; g < nwg; g++) s += part[uint64_t(g) * n + i];
  grad[i] = s;
}

// Class-weighted binary cross-entropy with logits, mean over the n
// positions: loss_i = w_i (softplus(z) - y z), w = pos_weight for y = 1.
// dz_i = w_i (sigmoid(z) - y) / n; per-workgroup loss partials.
__global__ __launch_bounds__(256) void k_bce(const float* __restrict__ z, const uint8_t* __restrict__ y, uint64_t n,
                                             float pos_weight, float* __restrict__ dz, float* __restrict__ lpart)
{
  __shared__ float r[256];
  const int tid = threadIdx.x;
  float s = 0.f;
  for (uint64_t i = uint64_t(blockIdx.x) * 256 + tid; i < n; i += uint64_t(gridDim.x) * 256) {
    float zi = z[i];
    float yi = y[i] ? 1.f : 0.f;
    float w = y[i] ? pos_weight : 1.f;
    float sp = fmaxf(zi, 0.f) + log1pf(expf(-fabsf(zi)));
    s += w * (sp - yi * zi);
    float sg = 1.f / (1.f + expf(-zi));
    dz[i] = w * (sg - yi) / float(n);
  }
  r[tid] = s;
  for (int st = 128; st > 0; st >>= 1) {
    __syncthreads();
    if (tid < st) r[tid] += r[tid + st];
  }
  if (tid == 0) lpart[blockIdx.x] = r[0];
}

__global__ void k_loss_total(const float* __restrict__ lpart, uint32_t n, uint64_t npos, float* __restrict__ loss)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0;
    for (uint32_t i = 0; i < n; i++) s += lpart[i];
    loss[0] = float(s / double(npos));
  }
}

// Adam on the FP32 master weights; the BF16-rounded copy follows.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ wq, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, uint32_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wq[i] = bf16_round(wi);
}

__global__ void k_round_bf16(const float* __restrict__ w, float* __restrict__ wq, uint32_t n)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wq[i] = bf16_round(w[i]);
}

__global__ void k_propose(const float* __restrict__ z, uint64_t n, float thr, uint8_t* __restrict__ mask)
{
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) mask[i] = z[i] > thr ? 1 : 0;
}

// splitmix64 for the initial weights
uint64_t splitmix(uint64_t& s)
{
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace
}  // namespace dspx

using namespace dspx;

struct dspx_rfinet {
  CtxRef hold;
  dspx_ctx* ctx = nullptr;
  DevBuf w, wq, m, v, grad, wt;  // wt: transposed weights of one layer
  DevBuf wbf;                    // BF16 [tap][co][ci] weights of the two 32 -> 32 layers
  DevBuf act;                    // A1..A3, logits of the last forward
  DevBuf gbuf[2];                // gradient ping-pong
  DevBuf part, lpart, loss;
  uint64_t cap_pos = 0;          // positions the activation buffers hold
};

namespace {

void rfinet_reserve(dspx_rfinet* N, uint64_t npos)
{
  if (npos <= N->cap_pos) return;
  N->act.reserve(npos * (3 * kW + 1) * sizeof(float));
  for (auto& g : N->gbuf) g.reserve(npos * kW * sizeof(float));
  N->cap_pos = npos;
}

float* act_of(dspx_rfinet* N, int l, uint64_t npos)
{
  // output of layer l (0..3): A1, A2, A3 of kW rows, logits of 1 row
  return N->act.as<float>() + uint64_t(l) * kW * npos;
}

template <int CIN, int COUT, bool RELU, bool MASK>
void launch_conv(hipStream_t st, const float* A, const float* W, uint64_t nblk, uint32_t nchan, float* Y,
                 const float* Z = nullptr)
{
  dim3 grid((nchan + kTileX - 1) / kTileX, unsigned(nblk));
  k_conv<CIN, COUT, RELU, MASK><<<grid, kTileX, 0, st>>>(A, W, nchan, Y, Z);
  DSPX_HIP(hipGetLastError());
}

void forward(dspx_rfinet* N, const float* feat, uint64_t nblk, uint32_t nchan, float* logits_out)
{
  hipStream_t st = N->ctx->stream;
  const uint64_t npos = nblk * nchan;
  const float* wq = N->wq.as<float>();
  float* a1 = act_of(N, 0, npos);
  float* a2 = act_of(N, 1, npos);
  float* a3 = act_of(N, 2, npos);
  float* z = logits_out ? logits_out : act_of(N, 3, npos);
  launch_conv<kF, kW, true, false>(st, feat, wq + off_of(0), nblk, nchan, a1);
  static const bool valu = [] {
    const char* e = std::getenv("DSPX_RFINET_VALU");
    return e && std::atoi(e) != 0;
  }();
  if (valu) {
    launch_conv<kW, kW, true, false>(st, a1, wq + off_of(1), nblk, nchan, a2);
    launch_conv<kW, kW, true, false>(st, a2, wq + off_of(2), nblk, nchan, a3);
  } else {
    // the two 32 -> 32 layers on the matrix cores (DSPX_RFINET_VALU=1: on the VALU)
    uint16_t* wb = N->wbf.as<uint16_t>();
    constexpr int nw = kK * kW * kW;
    for (int l = 1; l <= 2; l++) {
      k_conv_w_bf16<<<(nw + 255) / 256, 256, 0, st>>>(wq + off_of(l), wb + (l - 1) * nw);
      DSPX_HIP(hipGetLastError());
    }
    DSPX_REQUIRE(nblk <= 65535, DSPX_E_UNSUPPORTED, "nblk=%llu", (unsigned long long) nblk);
    dim3 grid((nchan + kMT - 1) / kMT, unsigned(nblk));
    k_conv_mfma<true><<<grid, 256, 0, st>>>(a1, wb, wq + off_of(1) + kW * kW * kK, nchan, a2);
    k_conv_mfma<true><<<grid, 256, 0, st>>>(a2, wb + nw, wq + off_of(2) + kW * kW * kK, nchan, a3);
    DSPX_HIP(hipGetLastError());
  }
  launch_conv<kW, 1, false, false>(st, a3, wq + off_of(3), nblk, nchan, z);
}

template <int CIN, int COUT>
void weight_grad(dspx_rfinet* N, const float* A, const float* G, uint64_t nblk, uint32_t nchan, float* grad)
{
  hipStream_t st = N->ctx->stream;
  const uint32_t nwg = uint32_t(std::min<uint64_t>(nblk, 512));
  const uint32_t np = COUT * CIN * kK + COUT;
  N->part.reserve(size_t(nwg) * np * sizeof(float));
  k_conv_dw<CIN, COUT><<<nwg, kTileX, 0, st>>>(A, G, nblk, nchan, N->part.as<float>());
  DSPX_HIP(hipGetLastError());
  k_sum_parts<<<(np + 255) / 256, 256, 0, st>>>(N->part.as<float>(), nwg, np, grad);
  DSPX_HIP(hipGetLastError());
}

template <int CIN, int COUT>
const float* transposed(dspx_rfinet* N, int l)
{
  hipStream_t st = N->ctx->stream;
  const int n = CIN * COUT * kK + CIN;
  k_transpose_w<CIN, COUT><<<(n + 255) / 256, 256, 0, st>>>(N->wq.as<float>() + off_of(l), N->wt.as<float>());
  DSPX_HIP(hipGetLastError());
  return N->wt.as<float>();
}

}  // namespace

extern "C" {

uint32_t dspx_rfinet_nparam(void) { return uint32_t(kNParam); }

int dspx_rfinet_create(dspx_ctx* ctx, uint64_t seed, dspx_rfinet** out)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && out, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<dspx_rfinet> N(new dspx_rfinet);
    N->hold.reset(ctx);
    N->ctx = ctx;
    const size_t bytes = size_t(kNParam) * sizeof(float);
    for (DevBuf* b : {&N->w, &N->wq, &N->m, &N->v, &N->grad}) b->reserve(bytes);
    N->wt.reserve(size_t(kW * kW * kK + kW) * sizeof(float));
    N->wbf.reserve(size_t(2 * kK * kW * kW) * sizeof(uint16_t));
    // He-uniform weights (bound sqrt(6 / fan_in)), zero biases
    std::vector<float> h(kNParam, 0.f);
    uint64_t s = seed;
    const int cins[kL] = {cin_of(0), cin_of(1), cin_of(2), cin_of(3)};
    const int couts[kL] = {cout_of(0), cout_of(1), cout_of(2), cout_of(3)};
    const int offs[kL] = {off_of(0), off_of(1), off_of(2), off_of(3)};
    for (int l = 0; l < kL; l++) {
      const double bound = std::sqrt(6.0 / double(cins[l] * kK));
      for (int i = 0; i < couts[l] * cins[l] * kK; i++) {
        double u = double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
        h[offs[l] + i] = float((2 * u - 1) * bound);
      }
    }
    std::vector<float> z(kNParam, 0.f);
    DSPX_HIP(hipMemcpy(N->w.p, h.data(), bytes, hipMemcpyHostToDevice));
    DSPX_HIP(hipMemcpy(N->m.p, z.data(), bytes, hipMemcpyHostToDevice));
    DSPX_HIP(hipMemcpy(N->v.p, z.data(), bytes, hipMemcpyHostToDevice));
    k_round_bf16<<<(kNParam + 255) / 256, 256, 0, ctx->stream>>>(N->w.as<float>(), N->wq.as<float>(), kNParam);
    DSPX_HIP(hipGetLastError());
    DSPX_HIP(hipStreamSynchronize(ctx->stream));
    *out = N.release();
  });
}

void dspx_rfinet_destroy(dspx_rfinet* N)
{
  if (!N) return;
  (void) hipSetDevice(N->ctx->device);
  (void) hipStreamSynchronize(N->ctx->stream);
  delete N;
}

int dspx_rfinet_get_params(dspx_rfinet* N, float* host, int rounded)
{
  return guard(N ? N->ctx : nullptr, [&] {
    DSPX_REQUIRE(N && host, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipStreamSynchronize(N->ctx->stream));
    DSPX_HIP(hipMemcpy(host, rounded ? N->wq.p : N->w.p, size_t(kNParam) * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int dspx_rfinet_set_params(dspx_rfinet* N, const float* host)
{
  return guard(N ? N->ctx : nullptr, [&] {
    DSPX_REQUIRE(N && host, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipStreamSynchronize(N->ctx->stream));
    DSPX_HIP(hipMemcpy(N->w.p, host, size_t(kNParam) * sizeof(float), hipMemcpyHostToDevice));
    k_round_bf16<<<(kNParam + 255) / 256, 256, 0, N->ctx->stream>>>(N->w.as<float>(), N->wq.as<float>(), kNParam);
    DSPX_HIP(hipGetLastError());
    DSPX_HIP(hipStreamSynchronize(N->ctx->stream));
  });
}

int dspx_rfi_features(dspx_ctx* ctx, const float* d_sums, const uint32_t* d_counts, uint64_t nblk, uint32_t nchan,
                      uint32_t npol, uint32_t Nd, float* d_scratch, float* d_feat)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && d_sums && d_counts && d_scratch && d_feat, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(npol == 1 || npol == 2, DSPX_E_PARAM, "npol=%u", npol);
    const uint64_t n = nblk * nchan;
    if (!n) return;
    float* lp = d_scratch;
    float* v1 = lp + n;
    float* vm = v1 + n;
    float* mean = vm + n;
    hipStream_t st = ctx->stream;
    k_rfi_stats<<<unsigned((n + 255) / 256), 256, 0, st>>>(d_sums, d_counts, n, npol, Nd, lp, v1, vm);
    k_rfi_chan_mean<<<(nchan + 31) / 32, 256, 0, st>>>(lp, nblk, nchan, mean);
    k_rfi_features<<<unsigned((n + 255) / 256), 256, 0, st>>>(lp, v1, vm, mean, nblk, nchan, d_feat);
    DSPX_HIP(hipGetLastError());
  });
}

int dspx_rfinet_forward(dspx_rfinet* N, const float* d_feat, uint64_t nblk, uint32_t nchan, float* d_logits)
{
  return guard(N ? N->ctx : nullptr, [&] {
    DSPX_REQUIRE(N && d_feat && d_logits, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(nblk < (1u << 31), DSPX_E_UNSUPPORTED, "nblk=%llu", (unsigned long long) nblk);
    if (!nblk || !nchan) return;
    rfinet_reserve(N, nblk * nchan);
    forward(N, d_feat, nblk, nchan, d_logits);
  });
}

int dspx_rfinet_backward(dspx_rfinet* N, const float* d_feat, const uint8_t* d_label, uint64_t nblk, uint32_t nchan,
                         float pos_weight, float* d_loss)
{
  return guard(N ? N->ctx : nullptr, [&] {
    DSPX_REQUIRE(N && d_feat && d_label, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(nblk < (1u << 31), DSPX_E_UNSUPPORTED, "nblk=%llu", (unsigned long long) nblk);
    const uint64_t npos = nblk * nchan;
    if (!npos) return;
    hipStream_t st = N->ctx->stream;
    rfinet_reserve(N, npos);
    forward(N, d_feat, nblk, nchan, nullptr);
    float* a1 = act_of(N, 0, npos);
    float* a2 = act_of(N, 1, npos);
    float* a3 = act_of(N, 2, npos);
    float* z = act_of(N, 3, npos);
    float* g4 = N->gbuf[0].as<float>();  // dL/dz (one row)
    const uint32_t nl = uint32_t(std::min<uint64_t>((npos + 255) / 256, 1024));
    N->lpart.reserve(nl * sizeof(float));
    N->loss.reserve(sizeof(float));
    k_bce<<<nl, 256, 0, st>>>(z, d_label, npos, pos_weight, g4, N->lpart.as<float>());
    k_loss_total<<<1, 1, 0, st>>>(N->lpart.as<float>(), nl, npos, d_loss ? d_loss : N->loss.as<float>());
    DSPX_HIP(hipGetLastError());
    float* grad = N->grad.as<float>();
    float* g3 = N->gbuf[1].as<float>();
    // layer 4 (32 -> 1)
    weight_grad<kW, 1>(N, a3, g4, nblk, nchan, grad + off_of(3));
    launch_conv<1, kW, false, true>(st, g4, transposed<kW, 1>(N, 3), nblk, nchan, g3, a3);
    // layer 3 (32 -> 32)
    weight_grad<kW, kW>(N, a2, g3, nblk, nchan, grad + off_of(2));
    float* g2 = N->gbuf[0].as<float>();
    launch_conv<kW, kW, false, true>(st, g3, transposed<kW, kW>(N, 2), nblk, nchan, g2, a2);
    // layer 2 (32 -> 32)
    weight_grad<kW, kW>(N, a1, g2, nblk, nchan, grad + off_of(1));
    float* g1 = N->gbuf[1].as<float>();
    launch_conv<kW, kW, false, true>(st, g2, transposed<kW, kW>(N, 1), nblk, nchan, g1, a1);
    // layer 1 (4 -> 32)
    weight_grad<kF, kW>(N, d_feat, g1, nblk, nchan, grad + off_of(0));
  });
}

int dspx_rfinet_grad(dspx_rfinet* N, float* host)
{
  return guard(N ? N->ctx : nullptr, [&] {
    DSPX_REQUIRE(N && host, DSPX_E_PARAM, "null argument");
    DSPX_HIP(hipStreamSynchronize(N->ctx->stream));
    DSPX_HIP(hipMemcpy(host, N->grad.p, size_t(kNParam) * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int dspx_rfinet_adam(dspx_rfinet* N, float lr, float beta1, float beta2, float eps, uint32_t t)
{
  return guard(N ? N->ctx : nullptr, [&] {
    DSPX_REQUIRE(N && t >= 1, DSPX_E_PARAM, "invalid argument");
    const float c1 = 1.f - std::pow(beta1, float(t)), c2 = 1.f - std::pow(beta2, float(t));
    k_adam<<<(kNParam + 255) / 256, 256, 0, N->ctx->stream>>>(N->w.as<float>(), N->wq.as<float>(),
                                                               N->grad.as<float>(), N->m.as<float>(),
                                                               N->v.as<float>(), kNParam, lr, beta1, beta2, eps, c1, c2);
    DSPX_HIP(hipGetLastError());
  });
}

int dspx_rfinet_propose(dspx_ctx* ctx, const float* d_logits, uint64_t n, float threshold, uint8_t* d_mask)
{
  return guard(ctx, [&] {
    DSPX_REQUIRE(ctx && d_logits && d_mask, DSPX_E_PARAM, "null argument");
    if (!n) return;
    k_propose<<<unsigned((n + 255) / 256), 256, 0, ctx->stream>>>(d_logits, n, threshold, d_mask);
    DSPX_HIP(hipGetLastError());
  });
}

}  // extern "C"
