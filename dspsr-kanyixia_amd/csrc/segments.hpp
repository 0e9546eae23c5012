// segments.hpp -- host walk of the reference fold recurrence as arithmetic
// runs, and the fold plan of a block built from those runs.
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
// serial loop.  Nothing here allocates: segments go to a caller's sink.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "dspx_internal.hpp"

namespace dspx {

namespace seg_detail {

// x = M * 2^(e-52) with M in [2^52, 2^53), or M = 0 for x = 0.
inline void decompose(double x, uint64_t& M, int& e)
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

}  // namespace seg_detail

// The runs of samples [j_begin, j_begin + n) folded from phase phi, each
// passed to push(segment, length).
template <class PUSH>
void fold_segments_to(double phi, double pps, uint64_t j_begin, uint64_t n, uint32_t isub, PUSH&& push)
{
#pragma clang fp contract(off)
  const double two52 = 4503599627370496.0;   // 2^52
  const uint64_t top = (uint64_t(1) << 53);  // 2^53
  double x = phi - std::floor(phi);
  uint64_t j = 0;
  while (j < n) {
    uint64_t M;
    int e;
    seg_detail::decompose(x, M, e);
    uint64_t D = 0;
    bool serial = (M == 0);
    if (!serial) {
      double q = std::ldexp(pps, 52 - e);  // pps / u, exact
      if (!(q < two52)) {
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
      uint64_t s = (D == 0) ? (n - j) : (top - 1 - M) / D;
      len = (s + 1 < n - j) ? s + 1 : n - j;
    }
    FoldSegment seg;
    seg.j0 = j_begin + j;
    seg.M = M;
    seg.D = serial ? 0 : D;
    seg.e = e;
    seg.isub = isub;
    push(seg, len);
    double x_last = std::ldexp(double(M + (len - 1) * seg.D), e - 52);
    j += len;
    // one step of the reference recurrence from the run's last phase
    double ph = x_last + pps;
    x = ph - std::floor(ph);
  }
}

// Sample weights of a WeightedTimeSeries (Fold.C:700-733, 792-810): sample
// idat of the block belongs to weight (idat + weight_idat) / ndatperweight.
struct WeightView {
  const uint16_t* w = nullptr;
  uint64_t nweights = 0;
  uint32_t ndatperweight = 0;
  uint32_t weight_idat = 0;
  bool active() const { return w && ndatperweight; }
  uint64_t index(uint64_t idat) const { return (idat + weight_idat) / ndatperweight; }
  // first sample of the next weight after the one holding idat
  uint64_t next(uint64_t idat) const { return (index(idat) + 1) * ndatperweight - weight_idat; }
  bool bad(uint64_t idat) const { return w[index(idat)] == 0; }
};

// The fold plan of a block of ndat samples: the chunks' runs (chunks sorted by
// idat_start, not overlapping), gaps and bad-weight samples marked not folded
// (isub = ~0), sorted and covering [0, ndat).  push(segment) receives each
// segment in order; nsub is one past the largest isub and nfold the number
// of samples folded.
template <class PUSH>
void plan_segments(uint64_t ndat, double pps, const dspx_fold_chunk* chunks, uint32_t nchunk,
                   const WeightView& wv, PUSH&& push, uint32_t& nsub, uint64_t& nfold)
{
  nsub = 0;
  nfold = 0;
  uint64_t cursor = 0;
  bool last_unfolded = false;
  bool any = false;
  auto unfolded = [&](uint64_t j0) {
    if (any && last_unfolded) return;  // extends the previous unfolded run
    push(FoldSegment{j0, 0, 0, 0, 0xffffffffu});
    any = true;
    last_unfolded = true;
  };
  auto folded = [&](const FoldSegment& s) {
    push(s);
    any = true;
    last_unfolded = false;
  };
  for (uint32_t c = 0; c < nchunk; c++) {
    const dspx_fold_chunk& ch = chunks[c];
    DSPX_REQUIRE(ch.idat_start >= cursor, DSPX_E_PARAM, "fold chunks overlap or are not sorted at sample %llu",
                 (unsigned long long) ch.idat_start);
    DSPX_REQUIRE(ch.idat_start + ch.ndat <= ndat, DSPX_E_PARAM, "fold chunk [%llu, +%llu) beyond ndat=%llu",
                 (unsigned long long) ch.idat_start, (unsigned long long) ch.ndat, (unsigned long long) ndat);
    DSPX_REQUIRE(ch.isub < 0xffff, DSPX_E_PARAM, "isub=%u too large", ch.isub);
    if (ch.ndat && wv.active()) {
      // Fold.C:721-733 (first weight) and the assert of :798 (last)
      DSPX_REQUIRE(wv.index(ch.idat_start) < wv.nweights, DSPX_E_STATE,
                   "iweight=%llu >= nweight=%llu", (unsigned long long) wv.index(ch.idat_start),
                   (unsigned long long) wv.nweights);
      DSPX_REQUIRE(wv.index(ch.idat_start + ch.ndat - 1) < wv.nweights, DSPX_E_PARAM,
                   "the weights end before sample %llu", (unsigned long long) (ch.idat_start + ch.ndat - 1));
    }
    if (ch.idat_start > cursor) unfolded(cursor);
    if (ch.ndat) {
      fold_segments_to(ch.phi, pps, ch.idat_start, ch.ndat, ch.isub, [&](const FoldSegment& s, uint64_t len) {
        if (!wv.active()) {
          folded(s);
          nfold += len;
          return;
        }
        // pieces of equal weight status
        const uint64_t end = s.j0 + len;
        uint64_t x = s.j0;
        while (x < end) {
          const bool bad = wv.bad(x);
          uint64_t y = x;
          while (y < end && wv.bad(y) == bad) y = std::min(end, wv.next(y));
          if (bad) {
            unfolded(x);
          } else {
            FoldSegment p = s;
            p.j0 = x;
            p.M = s.M + (x - s.j0) * s.D;
            folded(p);
            nfold += y - x;
          }
          x = y;
        }
      });
    }
    cursor = ch.idat_start + ch.ndat;
    if (ch.isub + 1 > nsub) nsub = ch.isub + 1;
  }
  if (cursor < ndat || !any) unfolded(cursor);
}

}  // namespace dspx
