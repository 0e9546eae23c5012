// fft_core.hpp -- register/LDS Stockham FFT building blocks for gfx950.
//
// A workgroup of NT threads transforms B independent sequences of length L
// held in LDS ("batched columns", element (n, b) at n*RS + b) or one row
// (B = 1, element n at n + n/16: one pad slot every 16 complex values keeps
// the radix-16 scatter of the first Stockham pass conflict-free across the
// 64 LDS banks).  Every thread owns E = B*L/NT complex values in registers;
// a radix-R pass gives each thread E/R butterflies g = tid + q*NT.
//
// Stockham autosort pass (size Ns already transformed, radix R):
//   j-th butterfly reads  x[j + r*L/R],            r = 0..R-1
//   multiplies by         W_{Ns*R}^{r*(j%Ns)}
//   applies a DFT_R and writes y[(j/Ns)*Ns*R + j%Ns + r*Ns].
// The output is in natural order.  The last pass (Ns = L/R) leaves the
// thread holding y[j + r*L/R] -- the very elements the first pass of a
// transform with the same radix reads, so a forward transform, a pointwise
// product and an inverse transform chain without an LDS round trip.
//
// Twiddles come from a table of exp(-2 pi i m / L) computed in double on the
// host and rounded once to float; DIR = +1 uses the conjugate.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dspx {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by DIR*i  (forward DIR=-1: -i ; inverse DIR=+1: +i)
template <int DIR> __device__ __forceinline__ float2 mul_dir_i(float2 a)
{
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// cos/sin(2 pi m / 16), m = 0..15, correctly rounded to float
__device__ constexpr float kCos16[16] = {
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
__device__ constexpr float kSin16[16] = {
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f};

// x * W_16^m with W_16 = exp(DIR * 2 pi i / 16)
template <int DIR> __device__ __forceinline__ float2 mul_w16(float2 x, int m)
{
  m &= 15;
  if (m == 0) return x;
  if (m == 4) return mul_dir_i<DIR>(x);
  if (m == 8) return make_float2(-x.x, -x.y);
  if (m == 12) return mul_dir_i<-DIR>(x);
  float2 w = make_float2(kCos16[m], DIR * kSin16[m]);
  return cmul(x, w);
}

template <int DIR> __device__ __forceinline__ void dft2(float2& a, float2& b)
{
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR> __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3)
{
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_dir_i<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// In-register DFT of length R (2, 4, 8, 16) on v[0..R-1], natural order in
// and out.  R = P*Q with P = 4 first: y[n2][k1] = DFT_P_n1 x[Q n1 + n2],
// y *= W_R^{n2 k1}, X[k1 + P k2] = DFT_Q_n2 y[n2][k1].
template <int R, int DIR> __device__ __forceinline__ void dft(float2* v)
{
  if constexpr (R == 2) {
    dft2<DIR>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<DIR>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    // Q = 2: x[2 n1 + n2]
    dft4<DIR>(v[0], v[2], v[4], v[6]);
    dft4<DIR>(v[1], v[3], v[5], v[7]);
    // y[1][k1] *= W8^{k1} = W16^{2 k1}
    v[3] = mul_w16<DIR>(v[3], 2);
    v[5] = mul_w16<DIR>(v[5], 4);
    v[7] = mul_w16<DIR>(v[7], 6);
    float2 o[8];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
      float2 a = v[2 * k1], b = v[2 * k1 + 1];
      o[k1] = cadd(a, b);
      o[k1 + 4] = csub(a, b);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = o[i];
  } else if constexpr (R == 16) {
    // Q = 4: y[n2][k1] from x[4 n1 + n2]
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) dft4<DIR>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    // now v[4*k1 + n2] = y[n2][k1]; twiddle W16^{n2 k1}
#pragma unroll
    for (int k1 = 1; k1 < 4; k1++)
#pragma unroll
      for (int n2 = 1; n2 < 4; n2++) v[4 * k1 + n2] = mul_w16<DIR>(v[4 * k1 + n2], n2 * k1);
    float2 o[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
      float2 a0 = v[4 * k1], a1 = v[4 * k1 + 1], a2 = v[4 * k1 + 2], a3 = v[4 * k1 + 3];
      dft4<DIR>(a0, a1, a2, a3);
      o[k1] = a0;
      o[k1 + 4] = a1;
      o[k1 + 8] = a2;
      o[k1 + 12] = a3;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = o[i];
  }
}

// ---------------------------------------------------------------------------
// Batched-column layout: element (n, b) at lds[n * RS + b].
// Butterfly g -> b = g % B, j = g / B (consecutive lanes -> consecutive b).
// ---------------------------------------------------------------------------
template <int L, int B, int RS, int NT> struct ColLayout {
  static constexpr int E = L * B / NT;
  __device__ static __forceinline__ int idx(int n, int b) { return n * RS + b; }
};

// The same with the low four column bits (the low log2 B, for B < 16)
// XOR-ed by a key of the row and of column bits 5 and up: a permutation of
// each aligned group of 16 columns (the key is constant over the group), so
// that 16 consecutive columns of a row still fall on 16 distinct bank pairs,
// while a read of column q of each lane's group of 16 consecutive columns
// (the fold's regrouping: lane l reads row j(l), column 16 l' + q, with
// B/16 lanes l' per row) puts the 16 lanes of a DPP row on 16 distinct bank
// pairs: lanes of one row differ in (l' & 1, l' >> 1), those of the next
// rows are offset by M = B/32 in the key (float2 = two of the 64 banks).
template <int L, int B, int RS, int NT> struct ColLayoutSwz {
  static constexpr int E = L * B / NT;
  static constexpr int M = B >= 64 ? B / 32 : 1;
  static constexpr int KM = (B < 16 ? B : 16) - 1;
  __device__ static __forceinline__ int idx(int n, int b) { return n * RS + (b ^ (((b >> 5) + n * M) & KM)); }
};

// Row layout (B = 1): element n at n + (n >> 4).
template <int L, int NT> struct RowLayout {
  static constexpr int E = L / NT;
  __device__ static __forceinline__ int idx(int n, int /*b*/) { return n + (n >> 4); }
};

// The exchange barrier of a transform done by NT threads: a workgroup
// barrier, or, when one wavefront (or part of one) owns the transform, a
// wavefront-scope fence and scheduling barrier only (a wavefront's LDS
// accesses complete in order, so its lanes see each other's writes without
// waiting for the rest of the workgroup).
template <int NT> __device__ __forceinline__ void xsync()
{
  if constexpr (NT <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Compute one Stockham pass on registers v (which hold, for each of the
// thread's E/R butterflies q, the R inputs x[j + r*L/R]).  After the call
// v[q*R + r] holds the output destined for index (j/Ns)*Ns*R + j%Ns + r*Ns.
// Twiddle memory (float2 units from the base of the table the context
// builds): exp(-2 pi i m / L) for L = 2..16384 at offset L, m < L; from
// kTwPowBase, power records of 4 float2 {W_S^k, W_S^2k, W_S^4k, W_S^8k}
// (W_S = exp(-2 pi i / S)) for spans S = 2..16384 at record S/2 + k, k < S/2.
// Kernels are handed the pointer to the exp table of their length L.
constexpr int kTwMaxL = 16384;
constexpr int kTwPowBase = 2 * kTwMaxL;
constexpr int kTwTotal = kTwPowBase + 4 * kTwMaxL;

template <int L, int B, int NT, int R, int NS, int DIR>
__device__ __forceinline__ void stockham_compute(float2* v, const float2* __restrict__ tw, int tid)
{
  constexpr int E = L * B / NT;
  static_assert(E % R == 0, "radix must divide the per-thread count");
#pragma unroll
  for (int q = 0; q < E / R; q++) {
    int g = tid + q * NT;
    int j = g / B;
    if constexpr (NS > 1) {
      int k = j % NS;
      float2 w[R];
      if constexpr (R >= 8 && B == 1 && NS >= 8) {
        // Row transforms (one butterfly index per lane, so lanes gather
        // different twiddles): power record {W^k, W^2k, W^4k, W^8k} of span
        // S = NS*R, two contiguous 16-byte loads, the other powers by at most
        // three products each, instead of R-1 scattered table loads.  Column
        // transforms (B > 1) share k across the lanes of a column group and
        // keep the direct loads.
        constexpr int S = NS * R;
        const float4* rec = reinterpret_cast<const float4*>(tw - L + kTwPowBase) + 2 * (S / 2 + k);
        float4 a = rec[0], b = rec[1];
        w[1] = make_float2(a.x, a.y);
        w[2] = make_float2(a.z, a.w);
        w[4] = make_float2(b.x, b.y);
        w[3] = cmul(w[2], w[1]);
        w[5] = cmul(w[4], w[1]);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
        if constexpr (R == 16) {
          w[8] = make_float2(b.z, b.w);
#pragma unroll
          for (int r = 9; r < 16; r++) w[r] = cmul(w[8], w[r - 8]);
        }
      } else {
#pragma unroll
        for (int r = 1; r < R; r++) w[r] = tw[r * k * (L / (NS * R))];
      }
#pragma unroll
      for (int r = 1; r < R; r++) v[q * R + r] = DIR < 0 ? cmul(v[q * R + r], w[r]) : cmulc(v[q * R + r], w[r]);
    }
    dft<R, DIR>(v + q * R);
  }
}

// The twiddles stockham_compute applies in a pass (R, NS), for each of the
// thread's E/R butterflies q: w[q*R + r], r >= 1 (w[q*R] unused), computed
// from the same table entries and products, so that a pass applied with
// them (stockham_apply) gives the same values.
template <int L, int B, int NT, int R, int NS>
__device__ __forceinline__ void pass_twiddles(const float2* __restrict__ tw, int tid, float2* wq)
{
  constexpr int E = L * B / NT;
  static_assert(NS > 1, "the first pass has no twiddles");
#pragma unroll
  for (int q = 0; q < E / R; q++) {
    int g = tid + q * NT;
    int k = (g / B) % NS;
    float2* w = wq + q * R;
    w[0] = make_float2(1.f, 0.f);
    if constexpr (R >= 8 && B == 1 && NS >= 8) {
      constexpr int S = NS * R;
      const float4* rec = reinterpret_cast<const float4*>(tw - L + kTwPowBase) + 2 * (S / 2 + k);
      float4 a = rec[0], b = rec[1];
      w[1] = make_float2(a.x, a.y);
      w[2] = make_float2(a.z, a.w);
      w[4] = make_float2(b.x, b.y);
      w[3] = cmul(w[2], w[1]);
      w[5] = cmul(w[4], w[1]);
      w[6] = cmul(w[4], w[2]);
      w[7] = cmul(w[4], w[3]);
      if constexpr (R == 16) {
        w[8] = make_float2(b.z, b.w);
#pragma unroll
        for (int r = 9; r < 16; r++) w[r] = cmul(w[8], w[r - 8]);
      }
    } else {
#pragma unroll
      for (int r = 1; r < R; r++) w[r] = tw[r * k * (L / (NS * R))];
    }
  }
}

// A Stockham pass (as stockham_compute) with its twiddles given (from
// pass_twiddles; DIR > 0 applies their conjugates).
template <int L, int B, int NT, int R, int DIR>
__device__ __forceinline__ void stockham_apply(float2* v, const float2* wq)
{
  constexpr int E = L * B / NT;
#pragma unroll
  for (int q = 0; q < E / R; q++) {
#pragma unroll
    for (int r = 1; r < R; r++)
      v[q * R + r] = DIR < 0 ? cmul(v[q * R + r], wq[q * R + r]) : cmulc(v[q * R + r], wq[q * R + r]);
    dft<R, DIR>(v + q * R);
  }
}

template <class LAY, int L, int B, int NT, int R>
__device__ __forceinline__ void lds_load_pass(const float2* lds, float2* v, int tid)
{
  constexpr int E = L * B / NT;
#pragma unroll
  for (int q = 0; q < E / R; q++) {
    int g = tid + q * NT;
    int b = g % B, j = g / B;
#pragma unroll
    for (int r = 0; r < R; r++) v[q * R + r] = lds[LAY::idx(j + r * (L / R), b)];
  }
}

template <class LAY, int L, int B, int NT, int R, int NS>
__device__ __forceinline__ void lds_store_pass(float2* lds, const float2* v, int tid)
{
  constexpr int E = L * B / NT;
#pragma unroll
  for (int q = 0; q < E / R; q++) {
    int g = tid + q * NT;
    int b = g % B, j = g / B;
    int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
    for (int r = 0; r < R; r++) lds[LAY::idx(base + r * NS, b)] = v[q * R + r];
  }
}

// Radix plans, keyed by the transform length L and the complex values each
// thread owns (E = 16 for the column tiles, 8 for the row transforms).  The
// inverse uses the reversed plan, so a forward transform's last pass and the
// inverse's first pass share thread ownership.
template <int L, int E> struct RadixPlan;
#define DSPX_PLAN(L, E, N, R0, R1, R2, R3) \
  template <> struct RadixPlan<L, E> { static constexpr int n = N; static constexpr int r[4] = {R0, R1, R2, R3}; };
DSPX_PLAN(2, 16, 1, 2, 0, 0, 0)
DSPX_PLAN(4, 16, 1, 4, 0, 0, 0)
DSPX_PLAN(8, 16, 1, 8, 0, 0, 0)
DSPX_PLAN(16, 16, 1, 16, 0, 0, 0)
DSPX_PLAN(32, 16, 2, 2, 16, 0, 0)
DSPX_PLAN(64, 16, 2, 4, 16, 0, 0)
DSPX_PLAN(128, 16, 2, 8, 16, 0, 0)
DSPX_PLAN(256, 16, 2, 16, 16, 0, 0)
DSPX_PLAN(512, 16, 3, 2, 16, 16, 0)
DSPX_PLAN(1024, 16, 3, 4, 16, 16, 0)
DSPX_PLAN(2048, 16, 3, 8, 16, 16, 0)
DSPX_PLAN(4096, 16, 3, 16, 16, 16, 0)
DSPX_PLAN(8192, 16, 4, 2, 16, 16, 16)
DSPX_PLAN(16384, 16, 4, 4, 16, 16, 16)
DSPX_PLAN(128, 8, 3, 2, 8, 8, 0)
DSPX_PLAN(256, 8, 3, 4, 8, 8, 0)
DSPX_PLAN(512, 8, 3, 8, 8, 8, 0)
DSPX_PLAN(1024, 8, 4, 2, 8, 8, 8)
DSPX_PLAN(2048, 8, 4, 4, 8, 8, 8)
DSPX_PLAN(4096, 8, 4, 8, 8, 8, 8)
#undef DSPX_PLAN

// Ns before pass p of the forward plan (product of earlier radices).
template <int L, int E, int P> struct FwdNs {
  static constexpr int v = FwdNs<L, E, P - 1>::v * RadixPlan<L, E>::r[P - 1];
};
template <int L, int E> struct FwdNs<L, E, 0> { static constexpr int v = 1; };

// Inverse plan uses the reversed radix order: pass p has radix r[n-1-p].
template <int L, int E, int P> struct InvRadix { static constexpr int v = RadixPlan<L, E>::r[RadixPlan<L, E>::n - 1 - P]; };
template <int L, int E, int P> struct InvNs {
  static constexpr int v = InvNs<L, E, P - 1>::v * InvRadix<L, E, P - 1>::v;
};
template <int L, int E> struct InvNs<L, E, 0> { static constexpr int v = 1; };

// Run passes [P0, P1) of the forward (DIR=-1, forward plan) or inverse
// (DIR=+1, reversed plan) transform.  On entry v holds the inputs of pass P0
// in pass-P0 ownership; on exit v holds the outputs of pass P1-1 in its
// output ownership.  Passes after the first go through LDS.
template <class LAY, int L, int B, int NT, int DIR, int P, int PEND>
__device__ __forceinline__ void run_passes(float2* lds, float2* v, const float2* __restrict__ tw, int tid)
{
  constexpr int E = L * B / NT;
  if constexpr (P < PEND) {
    constexpr int R = DIR < 0 ? RadixPlan<L, E>::r[P] : InvRadix<L, E, P>::v;
    constexpr int NS = DIR < 0 ? FwdNs<L, E, P>::v : InvNs<L, E, P>::v;
    if constexpr (P > 0) {
      // previous pass left outputs in registers: scatter them, then gather
      constexpr int RP = DIR < 0 ? RadixPlan<L, E>::r[P - 1] : InvRadix<L, E, P - 1>::v;
      constexpr int NSP = DIR < 0 ? FwdNs<L, E, P - 1>::v : InvNs<L, E, P - 1>::v;
      xsync<NT>();
      lds_store_pass<LAY, L, B, NT, RP, NSP>(lds, v, tid);
      xsync<NT>();
      lds_load_pass<LAY, L, B, NT, R>(lds, v, tid);
    }
    stockham_compute<L, B, NT, R, NS, DIR>(v, tw, tid);
    run_passes<LAY, L, B, NT, DIR, P + 1, PEND>(lds, v, tw, tid);
  }
}

// Full transform: v holds pass-0 inputs (x[j + r*L/R0] for the thread's
// butterflies) and ends holding outputs X[j + r*L/Rlast] of the last pass.
template <class LAY, int L, int B, int NT, int DIR>
__device__ __forceinline__ void fft_regs(float2* lds, float2* v, const float2* __restrict__ tw, int tid)
{
  run_passes<LAY, L, B, NT, DIR, 0, RadixPlan<L, L * B / NT>::n>(lds, v, tw, tid);
}

// A transform of two passes of one radix R (B = 1) whose second pass's
// twiddles are given in registers (wq, from pass_twiddles<L, 1, NT, R, R>):
// the forward and the inverse plan then share them, conjugated.  The same
// operations, hence the same values, as fft_regs.
template <class LAY, int L, int NT, int DIR>
__device__ __forceinline__ void fft_regs_2w(float2* lds, float2* v, const float2* wq, const float2* __restrict__ tw,
                                            int tid)
{
  constexpr int E = L / NT;
  constexpr int R = RadixPlan<L, E>::r[0];
  static_assert(RadixPlan<L, E>::n == 2 && RadixPlan<L, E>::r[1] == R, "two passes of one radix");
  stockham_compute<L, 1, NT, R, 1, DIR>(v, tw, tid);  // NS = 1: no twiddles
  xsync<NT>();
  lds_store_pass<LAY, L, 1, NT, R, 1>(lds, v, tid);
  xsync<NT>();
  lds_load_pass<LAY, L, 1, NT, R>(lds, v, tid);
  stockham_apply<L, 1, NT, R, DIR>(v, wq);
}

// Returns p unchanged, opaquely: loads through the result cannot be hoisted
// out of the loop that calls this.  Keeps per-pass twiddle loads inside the
// loop over rows instead of holding every pass's twiddles in registers.  p
// must point to global memory (the twiddle tables): the value passes through
// the asm as a global-address-space pointer, so that the loads stay global
// loads (through a generic pointer they would be flat loads, which also
// count against the LDS wait counter and hold up the LDS exchanges).
template <class T> __device__ __forceinline__ const T* opaque(const T* p)
{
  auto q = (const __attribute__((address_space(1))) T*) p;
  asm volatile("" : "+s"(q));
  return (const T*) q;
}

// Returns x unchanged, opaquely (a vector-register value): arithmetic on the
// result cannot be hoisted out of the loop that calls this.
__device__ __forceinline__ float2 opaque_v(float2 x)
{
  asm volatile("" : "+v"(x.x), "+v"(x.y));
  return x;
}

// Returns a per-lane integer unchanged, opaquely: index arithmetic on the
// result is redone where it is used instead of being hoisted out of the loop
// that calls this (and held in registers for the loop's duration).
__device__ __forceinline__ int opaque_i(int x)
{
  asm volatile("" : "+v"(x));
  return x;
}

// A workgroup-uniform value moved to scalar registers.
__device__ __forceinline__ float2 uniform(float2 x)
{
  return make_float2(__builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x.x))),
                     __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x.y))));
}

// Radix of the first / last pass of each direction.
template <int L, int E, int DIR> struct FirstRadix {
  static constexpr int v = DIR < 0 ? RadixPlan<L, E>::r[0] : RadixPlan<L, E>::r[RadixPlan<L, E>::n - 1];
};
template <int L, int E, int DIR> struct LastRadix {
  static constexpr int v = DIR < 0 ? RadixPlan<L, E>::r[RadixPlan<L, E>::n - 1] : RadixPlan<L, E>::r[0];
};

// Wavefront sum delivered in lane 63, by DPP row shifts and row broadcasts
// (no LDS traffic): partial sums of 4 lanes, then of each 16-lane row in its
// lane 15, then rows 0+1 into lane 31 and all four rows into lane 63.
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ float dpp_shift(float v)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK,
                                                               BANK_MASK, true));
}
__device__ __forceinline__ float wave_sum_lane63(float v)
{
  float s = v + dpp_shift<0x111, 0xf, 0xf>(v);  // row_shr:1
  s += dpp_shift<0x112, 0xf, 0xf>(v);           // row_shr:2
  s += dpp_shift<0x113, 0xf, 0xf>(v);           // row_shr:3
  s += dpp_shift<0x114, 0xf, 0xe>(s);           // row_shr:4, lanes 4..15 of each row
  s += dpp_shift<0x118, 0xf, 0xc>(s);           // row_shr:8, lanes 8..15
  s += dpp_shift<0x142, 0xa, 0xf>(s);           // row_bcast:15 into rows 1 and 3
  s += dpp_shift<0x143, 0xc, 0xf>(s);           // row_bcast:31 into rows 2 and 3
  return s;
}

// Sum over the 16 lanes of each DPP row, delivered in the row's lane 15.
__device__ __forceinline__ float row_sum_lane15(float v)
{
  float s = v + dpp_shift<0x111, 0xf, 0xf>(v);  // row_shr:1
  s += dpp_shift<0x112, 0xf, 0xf>(v);           // row_shr:2
  s += dpp_shift<0x113, 0xf, 0xf>(v);           // row_shr:3
  s += dpp_shift<0x114, 0xf, 0xe>(s);           // row_shr:4, lanes 4..15
  s += dpp_shift<0x118, 0xf, 0xc>(s);           // row_shr:8, lanes 8..15
  return s;
}

}  // namespace dspx
