// dspx_internal.hpp -- shared host/device declarations of libdspx.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "dspx.h"
#include "refcount.hpp"

struct dspx_ctx {
  // the caller's handle and every engine made on the context (refcount.hpp)
  dspx::RefCount refs;
  int device = 0;
  int ncu = 0;  // compute units of the device
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  // twiddle tables exp(-2 pi i m / L) for L = 2..8192, table L at offset L
  float2* d_tw = nullptr;
  // events for per-stage timing (dspx_conv_stage_ms)
  hipEvent_t ev[8] = {};
  int nev = 0;
};

namespace dspx {

// Frees a context whose last reference was dropped (ctx.hip).
void ctx_free(dspx_ctx* c);
// An engine's reference on its context: declared as the engine's first
// member, so it is dropped after the engine's own buffers are released.
using CtxRef = Ref<dspx_ctx, ctx_free>;

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const char* fmt, ...);

// A failed call's error is reported through the status and also cleared
// from the thread's last-error slot, so that the caller's next
// hipGetLastError (a launch check of its own) does not see it.
#define DSPX_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      (void) hipGetLastError();                                                          \
      ::dspx::fail(DSPX_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                   __FILE__, __LINE__);                                                  \
    }                                                                                    \
  } while (0)

#define DSPX_REQUIRE(cond, code, ...)            \
  do {                                           \
    if (!(cond)) ::dspx::fail(code, __VA_ARGS__); \
  } while (0)

// Wraps the body of every C-ABI entry point: exceptions never cross the ABI.
template <class F> int guard(dspx_ctx* ctx, F&& f)
{
  try {
    f();
    if (ctx) ctx->last_error.clear();
    return DSPX_OK;
  } catch (const Error& e) {
    if (ctx) ctx->last_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return DSPX_E_INTERNAL;
  }
}

inline const float2* twiddles(const dspx_ctx* ctx, int L) { return ctx->d_tw + L; }

// Lets a kernel request, as dynamic LDS, the whole LDS of a workgroup
// (160 KB on gfx950) less its static LDS; requests above 64 KB need this.
inline void allow_full_lds(const void* kernel)
{
  hipFuncAttributes a{};
  DSPX_HIP(hipFuncGetAttributes(&a, kernel));
  DSPX_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(160 * 1024 - a.sharedSizeBytes)));
}

// Two-level table of W_N^m = exp(-2 pi i m / N): lo[m & (2^h - 1)] * hi[m >> h].
struct Tw2 {
  const float2* lo;
  const float2* hi;
  int h;
};

// Device buffer that grows on demand (engine-owned scratch).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void reserve(size_t n)
  {
    if (n <= bytes) return;
    if (p) (void) hipFree(p);
    p = nullptr;
    bytes = 0;
    DSPX_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  void release()
  {
    if (p) (void) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Pinned host staging for small asynchronous uploads made every block (fold
// plans): two slots used in turn, a slot reused only after the copy that last
// read it has completed, so that the caller never waits for the stream.
struct HostStage {
  void* slot[2] = {};
  size_t cap[2] = {};
  hipEvent_t done[2] = {};
  bool pending[2] = {};
  int next = 0;
  // a pinned buffer of n bytes for the caller to fill before upload(s, ...)
  void* acquire(size_t n, int& s)
  {
    s = next;
    next ^= 1;
    if (pending[s]) {
      DSPX_HIP(hipEventSynchronize(done[s]));
      pending[s] = false;
    }
    if (cap[s] < n) {
      if (slot[s]) (void) hipHostFree(slot[s]);
      slot[s] = nullptr;
      cap[s] = 0;
      DSPX_HIP(hipHostMalloc(&slot[s], n, hipHostMallocDefault));
      cap[s] = n;
    }
    if (!done[s]) DSPX_HIP(hipEventCreateWithFlags(&done[s], hipEventDisableTiming));
    return slot[s];
  }
  void upload(int s, void* dst, size_t n, hipStream_t st)
  {
    DSPX_HIP(hipMemcpyAsync(dst, slot[s], n, hipMemcpyHostToDevice, st));
    DSPX_HIP(hipEventRecord(done[s], st));
    pending[s] = true;
  }
  ~HostStage()
  {
    for (int s = 0; s < 2; s++) {
      if (pending[s]) (void) hipEventSynchronize(done[s]);
      if (done[s]) (void) hipEventDestroy(done[s]);
      if (slot[s]) (void) hipHostFree(slot[s]);
    }
  }
};

// ---- engines shared between entry points (ctx.hip) ----
void unpack(dspx_ctx* ctx, const void* raw, uint32_t nbit, uint32_t offset_binary, uint32_t nchan,
            uint32_t npol, uint64_t ndat, float scale, float* out, uint64_t chan_stride,
            uint64_t pol_stride);
void detect(dspx_ctx* ctx, uint32_t state, uint32_t ndim, const float* in, uint64_t ics, uint64_t ips,
            uint32_t nchan, uint32_t npol, uint64_t ndat, float* out, uint64_t ocs, uint64_t ops);

// ---- fold without allocation (fold.hip): the per-device table address ----
void fold_static_setup(int device);

// ---- host-side algorithms (geometry.cpp) ----
int geometry(const dspx_dedisp_desc* d, dspx_geometry* g);
double channel_centre(double cfreq, double bw, unsigned nchan, int dc_centred, unsigned ichan);
double chirp_coeff(double bw, double dm, double chan_freq);
void sample_delay(double cfreq, double bw, unsigned nchan, double dm, double freq, int64_t& idelay, double& frac);
double interchan_fractional_delay(double cfreq, double bw, unsigned nchan, double dm, double chan_freq);
int sample_delays(const dspx_dedisp_desc* d, int64_t* delay, double* frac, int64_t* applied, int64_t* zero_delay,
                  uint64_t* total_delay);

// Bin-plan segments (binplan.cpp): x_j = (M + (j - j0) * D) * 2^(e - 52)
struct FoldSegment {
  uint64_t j0;  // absolute output sample index of the first element
  uint64_t M;
  uint64_t D;
  int32_t e;
  uint32_t isub;  // 0xffffffff: samples not folded
};
void fold_segments(double phi, double pps, uint64_t j_begin, uint64_t n, uint32_t isub,
                   std::vector<FoldSegment>& out);

// The convolving filterbank's four-step form (conv.hip): windows of 2^13 to
// 2^22 points whose sub-bands hold at least 8 * N1 bins; fb4_create returns
// null for the other shapes.
struct Fb4;
Fb4* fb4_create(dspx_ctx* ctx, const dspx_fb_desc& d);
void fb4_destroy(Fb4* f);
void fb4_set_response(Fb4* f, const float* host);
// d_zdm: the zero-DM output (sub-bands without the response), or null
void fb4_perform(Fb4* f, const float* d_in, uint64_t ics, uint64_t ips, uint64_t npart, float* d_out, uint64_t ocs,
                 uint64_t ops, float* d_zdm);

}  // namespace dspx
