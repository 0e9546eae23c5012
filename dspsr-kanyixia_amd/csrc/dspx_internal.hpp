// dspx_internal.hpp -- shared host/device declarations of libdspx.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "dspx.h"

struct dspx_ctx {
  int device = 0;
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

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const char* fmt, ...);

#define DSPX_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      ::dspx::fail(DSPX_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                   __FILE__, __LINE__);                                                  \
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

// ---- engines shared between entry points (ctx.hip) ----
void unpack(dspx_ctx* ctx, const void* raw, uint32_t nbit, uint32_t offset_binary, uint32_t nchan,
            uint32_t npol, uint64_t ndat, float scale, float* out, uint64_t chan_stride,
            uint64_t pol_stride);
void detect(dspx_ctx* ctx, uint32_t state, uint32_t ndim, const float* in, uint64_t ics, uint64_t ips,
            uint32_t nchan, uint32_t npol, uint64_t ndat, float* out, uint64_t ocs, uint64_t ops);

// ---- host-side algorithms (geometry.cpp) ----
int geometry(const dspx_dedisp_desc* d, dspx_geometry* g);
double channel_centre(double cfreq, double bw, unsigned nchan, int dc_centred, unsigned ichan);
double chirp_coeff(double bw, double dm, double chan_freq);

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

}  // namespace dspx
