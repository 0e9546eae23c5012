// binplan.cpp -- exact, parallel evaluation of the reference fold recurrence.
//
// Fold::fold (Signal/Pulsar/Fold.C:790-838) walks every sample serially:
//     phi -= floor(phi);  ibin = unsigned(phi * nbin);  phi += pps;
// in IEEE double.  Within one binade [2^e, 2^(e+1)) every double is an
// integer multiple M of u = 2^(e-52), and fl(M*u + pps) = (M + D)*u with
// D = round(pps/u), as long as the sum stays inside the binade (and pps/u is
// not an exact half-integer tie).  The sequence is therefore piecewise
// arithmetic: x_j = (M + (j - j0)*D) * u exactly, over runs that end at a
// binade boundary or at the wrap through 1.  The host walks run to run
// (O(number of binades crossed)), stepping across every run boundary with the
// reference's own double arithmetic, and the device evaluates each sample of
// a run independently -- the same doubles, hence the same bins, as the
// serial loop.
#include <cmath>

#include "dspx_internal.hpp"

#pragma clang fp contract(off)

namespace dspx {

static const double kTwo52 = 4503599627370496.0;   // 2^52
static const uint64_t kTop = (uint64_t(1) << 53);  // 2^53

// x = M * 2^(e-52) with M in [2^52, 2^53), or M = 0 for x = 0.
static void decompose(double x, uint64_t& M, int& e)
{
  if (x == 0.0) {
    M = 0;
    e = 0;
    return;
  }
  int ex;
  double f = std::frexp(x, &ex);  // x = f * 2^ex, f in [0.5, 1)
  M = uint64_t(std::ldexp(f, 53));
  e = ex - 1;
}

// One step of the reference recurrence from a wrapped phase x.
static double next_phase(double x, double pps)
{
  double phi = x + pps;
  return phi - std::floor(phi);
}

void fold_segments(double phi, double pps, uint64_t j_begin, uint64_t n, uint32_t isub,
                   std::vector<FoldSegment>& out)
{
  double x = phi - std::floor(phi);
  uint64_t j = 0;
  while (j < n) {
    uint64_t M;
    int e;
    decompose(x, M, e);
    uint64_t D = 0;
    bool serial = (M == 0);
    if (!serial) {
      double q = std::ldexp(pps, 52 - e);  // pps / u, exact
      if (!(q < kTwo52)) {
        serial = true;
      } else {
        double fl = std::floor(q);
        double frac = q - fl;  // exact
        if (frac == 0.5)
          serial = true;  // round-half-even would depend on M's parity
        else
          D = uint64_t(fl) + (frac > 0.5 ? 1 : 0);
      }
    }
    uint64_t len = 1;
    if (!serial) {
      // further steps that keep M + s*D strictly inside the binade
      uint64_t s = (D == 0) ? (n - j) : (kTop - 1 - M) / D;
      len = (s + 1 < n - j) ? s + 1 : n - j;
    }
    FoldSegment seg;
    seg.j0 = j_begin + j;
    seg.M = M;
    seg.D = serial ? 0 : D;
    seg.e = e;
    seg.isub = isub;
    out.push_back(seg);
    double x_last = std::ldexp(double(M + (len - 1) * seg.D), e - 52);
    j += len;
    x = next_phase(x_last, pps);
  }
}

}  // namespace dspx

extern "C" int dspx_fold_binplan_host(double phi, double phase_per_sample, uint32_t nbin, uint64_t ndat,
                                      uint32_t* bins)
{
  if (!bins || nbin == 0 || !(phase_per_sample > 0.0 && phase_per_sample < 0.5)) return DSPX_E_PARAM;
  std::vector<dspx::FoldSegment> segs;
  dspx::fold_segments(phi, phase_per_sample, 0, ndat, 0, segs);
  for (size_t s = 0; s < segs.size(); s++) {
    uint64_t end = s + 1 < segs.size() ? segs[s + 1].j0 : ndat;
    for (uint64_t j = segs[s].j0; j < end; j++) {
      double x = std::ldexp(double(segs[s].M + (j - segs[s].j0) * segs[s].D), segs[s].e - 52);
      bins[j] = uint32_t(x * double(nbin));
    }
  }
  return DSPX_OK;
}

extern "C" int dspx_fold_nsegments_host(double phi, double phase_per_sample, uint64_t ndat, uint64_t* nseg)
{
  if (!nseg || !(phase_per_sample > 0.0 && phase_per_sample < 0.5)) return DSPX_E_PARAM;
  std::vector<dspx::FoldSegment> segs;
  dspx::fold_segments(phi, phase_per_sample, 0, ndat, 0, segs);
  *nseg = segs.size();
  return DSPX_OK;
}
