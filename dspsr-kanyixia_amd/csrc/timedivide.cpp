// timedivide.cpp -- sub-integration boundaries across blocks.
//
// Restates dsp::TimeDivide (Signal/Pulsar/TimeDivide.C:36-67 set_start_time,
// :124-338 set_bounds, :346-506 set_boundaries) for divisions of a fixed
// length in seconds, or in turns of a constant folding period (the branch
// Subint::prepare takes without a predictor, Signal/Pulsar/dsp/Subint.h:
// 217-240), and the loop of Subint::transformation (Subint.h:253-330) that
// calls set_bounds until the block is used up.  Epochs are held as the
// reference's MJD holds them (integer day, integer second of the day,
// fractional second; sums and differences formed per component and then
// settled), so that offsets between nearby epochs keep the fractional
// second's precision.  The MJD class itself belongs to a dependency that the
// reference tree does not carry; this follows its published representation.
#include <cmath>
#include <vector>

#include "dspx_internal.hpp"

namespace dspx {
namespace {

struct Epoch {
  int64_t day = 0, sec = 0;
  double frac = 0.0;
  static Epoch make(int64_t d, int64_t s, double f)
  {
    Epoch e{d, s, f};
    e.settle();
    return e;
  }
  void settle()
  {
    const double fs = std::floor(frac);
    sec += int64_t(fs);
    frac -= fs;
    int64_t ds = sec / 86400;
    sec -= ds * 86400;
    if (sec < 0) {
      sec += 86400;
      ds -= 1;
    }
    day += ds;
  }
  double in_seconds() const { return double(day) * 86400.0 + double(sec) + frac; }
  bool zero() const { return day == 0 && sec == 0 && frac == 0.0; }
};

Epoch operator+(const Epoch& a, double s) { return Epoch::make(a.day, a.sec, a.frac + s); }
Epoch operator-(const Epoch& a, const Epoch& b) { return Epoch::make(a.day - b.day, a.sec - b.sec, a.frac - b.frac); }
bool operator<(const Epoch& a, const Epoch& b) { return (a - b).in_seconds() < 0.0; }
bool operator>(const Epoch& a, const Epoch& b) { return (a - b).in_seconds() > 0.0; }
const Epoch& emax(const Epoch& a, const Epoch& b) { return a < b ? b : a; }
const Epoch& emin(const Epoch& a, const Epoch& b) { return b < a ? b : a; }

Epoch from(const dspx_mjd& m) { return Epoch::make(m.day, m.sec, m.frac); }

}  // namespace
}  // namespace dspx

using namespace dspx;

struct dspx_timediv {
  Epoch start_time;
  double seconds = 0, turns = 0, period = 0;
  // boundaries of the current division, the end of the last bounds, and the
  // input of the last set_bounds (the Observation the rounding refers to)
  Epoch lower, upper, current_end;
  bool is_valid = false;
  uint64_t division = 0, division_ndat = 0;
  Epoch obs_start;
  double obs_rate = 0;

  double interval() const { return seconds > 0 ? seconds : period * turns; }

  // TimeDivide::set_boundaries (const MJD&, const MJD&), TimeDivide.C:467-506:
  // both edges rounded to whole samples of the current input
  void set_edges(const Epoch& mjd1, const Epoch& mjd2)
  {
    const int64_t samples = std::lrint((mjd1 - obs_start).in_seconds() * obs_rate);
    lower = obs_start + double(samples) / obs_rate;
    const double secs = (mjd2 - lower).in_seconds();
    DSPX_REQUIRE(secs > 0, DSPX_E_INTERNAL, "division of %g s", secs);
    division_ndat = uint64_t(std::lrint(secs * obs_rate));
    upper = lower + double(division_ndat) / obs_rate;
  }

  // TimeDivide::set_boundaries (const MJD&), TimeDivide.C:346-433 (the
  // division in seconds or in turns of a constant period)
  void set_division(const Epoch& input_start)
  {
    const Epoch divide_start = emax(start_time, input_start);
    const double s = (divide_start - start_time).in_seconds();
    const double spi = interval();
    division = uint64_t(s / spi);  // truncates, as the reference's cast
    set_edges(start_time + double(division) * spi, start_time + double(division + 1) * spi);
  }

  // TimeDivide::set_bounds (TimeDivide.C:124-338)
  void set_bounds(const Epoch& input_start, double rate, uint64_t input_ndat, dspx_timediv_bounds& b)
  {
    obs_start = input_start;
    obs_rate = rate;
    const Epoch input_end = input_start + double(input_ndat) / rate;
    Epoch divide_start = is_valid ? emax(current_end, input_start) : input_start;
    b = dspx_timediv_bounds{};
    if (input_end < lower || divide_start + 0.5 / rate > upper) {
      b.new_division = 1;
      set_division(divide_start + 0.55 / rate);
    }
    divide_start = emax(lower, divide_start);
    const double start_sample = std::rint((divide_start - input_start).in_seconds() * rate);
    DSPX_REQUIRE(start_sample >= 0, DSPX_E_INTERNAL, "start sample %g", start_sample);
    const uint64_t idat_start = uint64_t(start_sample);
    b.division = division;
    b.division_ndat = division_ndat;
    if (idat_start >= input_ndat) {
      is_valid = false;
      return;
    }
    const Epoch divide_end = emin(input_end, upper);
    const double end_sample = std::rint((divide_end - input_start).in_seconds() * rate);
    uint64_t idat_end = uint64_t(end_sample);
    DSPX_REQUIRE(end_sample >= 0 && idat_end > idat_start, DSPX_E_STATE, "idat_end=%llu <= idat_start=%llu",
                 (unsigned long long) idat_end, (unsigned long long) idat_start);
    if (idat_end > input_ndat)
      idat_end = input_ndat;  // rounding in rint
    else if (idat_end < input_ndat)
      b.in_next = 1;
    b.idat_start = idat_start;
    b.ndat = idat_end - idat_start;
    if ((upper - divide_end).in_seconds() * rate < 0.5) b.end_reached = 1;
    b.is_valid = 1;
    is_valid = true;
    current_end = input_start + double(idat_end) / rate;
  }
};

extern "C" {

int dspx_timediv_create(const dspx_mjd* start_time, double seconds, double turns, double period,
                        const dspx_mjd* reference_epoch, dspx_timediv** out)
{
  return guard(nullptr, [&] {
    DSPX_REQUIRE(start_time && out, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(std::isfinite(start_time->frac), DSPX_E_PARAM, "start time");
    DSPX_REQUIRE((seconds > 0) != (turns > 0), DSPX_E_PARAM, "one of seconds=%g, turns=%g", seconds, turns);
    DSPX_REQUIRE(turns == 0 || period > 0, DSPX_E_PARAM, "turns need a folding period (period=%g)", period);
    auto* t = new dspx_timediv;
    t->seconds = seconds;
    t->turns = turns;
    t->period = period;
    // set_start_time (TimeDivide.C:36-67): the reference epoch, else with a
    // whole number of seconds per division the start is rounded down to a
    // multiple of it within the day
    Epoch st = from(*start_time);
    if (reference_epoch && !from(*reference_epoch).zero()) {
      st = from(*reference_epoch);
    } else if (seconds > 0 && seconds == double(unsigned(seconds))) {
      const int64_t is = int64_t(unsigned(seconds));
      st = Epoch::make(st.day, (st.sec / is) * is, 0.0);
    }
    t->start_time = st;
    *out = t;
  });
}

void dspx_timediv_destroy(dspx_timediv* t) { delete t; }

int dspx_timediv_set_bounds(dspx_timediv* t, const dspx_mjd* input_start, double rate, uint64_t input_ndat,
                            dspx_timediv_bounds* out)
{
  return guard(nullptr, [&] {
    DSPX_REQUIRE(t && input_start && out, DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(rate > 0 && std::isfinite(rate), DSPX_E_PARAM, "rate=%g", rate);
    t->set_bounds(from(*input_start), rate, input_ndat, *out);
  });
}

int dspx_timediv_block(dspx_timediv* t, const dspx_mjd* input_start, double rate, uint64_t input_ndat,
                       dspx_timediv_bounds* out, uint32_t max, uint32_t* n)
{
  return guard(nullptr, [&] {
    DSPX_REQUIRE(t && input_start && n && (out || !max), DSPX_E_PARAM, "null argument");
    DSPX_REQUIRE(rate > 0 && std::isfinite(rate), DSPX_E_PARAM, "rate=%g", rate);
    const Epoch st = from(*input_start);
    uint32_t k = 0;
    dspx_timediv_bounds b;
    do {
      t->set_bounds(st, rate, input_ndat, b);
      if (!b.is_valid) continue;
      DSPX_REQUIRE(k < max, DSPX_E_PARAM, "more than %u divisions in the block", max);
      out[k++] = b;
    } while (b.in_next);
    *n = k;
  });
}

}  // extern "C"
