// geometry.cpp -- host-side convolution geometry and chirp coefficients.
//
// Restates, for the library's own use, the reference rules that fix the
// FFT length and the overlap of the overlap-save convolution:
//   PlasmaResponse::smearing_time/smearing_samples/prepare
//     (Signal/General/PlasmaResponse.C:91-250),
//   optimal_fft_length (Signal/General/optimize_fft.c:64-135),
//   Convolution::prepare (Signal/General/Convolution.C:316-455),
// and the per-channel terms of Dedispersion::build_compute
//   (Signal/General/Dedispersion.C:166-186) and
//   Observation::get_centre_frequency (Kernel/Classes/Observation.C:521-535).
// Every expression keeps the reference's operation order and is compiled
// without floating-point contraction, so the doubles it produces are the
// ones the reference computes.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dspx_internal.hpp"

#pragma clang fp contract(off)

namespace dspx {

static const double kDmDispersion = 2.41e-4;        // Dedispersion.C:32
static const float kSmearingBuffer = 0.1f;          // PlasmaResponse.C:20
static const unsigned kSmearThreshold = 16u << 20;  // PlasmaResponse.C:89

static double delay_time(double dm, double freq)
{
  double dispersion = dm / kDmDispersion;
  return dispersion * 1.0 / (freq * freq);
}

static double smearing_time(double cfreq, double bw, unsigned nchan, double dm, int half,
                            unsigned nunsupported)
{
  double abs_bw = std::fabs(bw);
  double ch_abs_bw = abs_bw / double(nchan);
  double lower_ch_cfreq = cfreq - (abs_bw - ch_abs_bw) / 2.0;
  for (unsigned i = 0; i < nunsupported && i < nchan; i++) lower_ch_cfreq += ch_abs_bw;
  if (half) {
    ch_abs_bw /= 2.0;
    lower_ch_cfreq += double(half) * ch_abs_bw;
  }
  double f1 = lower_ch_cfreq - std::fabs(0.5 * ch_abs_bw);
  double f2 = lower_ch_cfreq + std::fabs(0.5 * ch_abs_bw);
  return delay_time(dm, f1) - delay_time(dm, f2);
}

static unsigned smearing_samples(double cfreq, double bw, unsigned nchan, double dm, int half,
                                 unsigned nunsupported)
{
  double tsmear = smearing_time(cfreq, bw, nchan, dm, half, nunsupported);
  double sampling_rate = std::fabs(bw) / double(nchan) * 1e6;
  tsmear *= (1.0 + kSmearingBuffer);
  return unsigned(std::ceil(tsmear * sampling_rate));
}

static int64_t optimal_fft_length(uint64_t nbad, uint64_t nfft_max, unsigned factor)
{
  if (!nbad) return -1;
  uint64_t nfft_min = factor;
  while (nbad > nfft_min) nfft_min *= 2;
  if (nfft_max && nfft_max < nfft_min) return -1;
  uint64_t nfft = nfft_min;
  double timescale = double(nfft) * std::log(double(nfft)) / double(nfft - nbad);
  while (nfft_max == 0 || nfft * 2 < nfft_max) {
    double prev = timescale;
    nfft *= 2;
    timescale = double(nfft) * std::log(double(nfft)) / double(nfft - nbad);
    if (timescale > prev) {
      nfft /= 2;
      break;
    }
  }
  return int64_t(nfft);
}

int geometry(const dspx_dedisp_desc* d, dspx_geometry* g)
{
  if (!d || !g) return DSPX_E_PARAM;
  if (d->nchan == 0 || !(d->bandwidth != 0.0) || !(d->centre_frequency > 0.0) ||
      !(d->dispersion_measure >= 0.0))
    return DSPX_E_PARAM;
  unsigned threshold = kSmearThreshold / d->nchan;
  unsigned ichan = 0, neg = 0;
  while ((neg = smearing_samples(d->centre_frequency, d->bandwidth, d->nchan,
                                 d->dispersion_measure, -1, ichan)) > threshold) {
    ichan++;
    if (ichan == d->nchan) return DSPX_E_STATE;  // PlasmaResponse.C:101-105
  }
  unsigned pos = smearing_samples(d->centre_frequency, d->bandwidth, d->nchan,
                                  d->dispersion_measure, 1, ichan);
  uint64_t nov = uint64_t(neg) + pos;
  int64_t nfft = d->nfft ? int64_t(d->nfft) : optimal_fft_length(nov, 0, 1);
  if (nfft < 2 || uint64_t(nfft) <= nov || nfft > (int64_t(1) << 31))
    return DSPX_E_STATE;  // Convolution.C:426-428
  g->impulse_neg = neg;
  g->impulse_pos = pos;
  g->nfft = uint32_t(nfft);
  g->nsamp_overlap = uint32_t(nov);
  g->nsamp_step = uint32_t(nfft - int64_t(nov));
  g->nunsupported = ichan;
  return DSPX_OK;
}

double channel_centre(double cfreq, double bw, unsigned nchan, int dc_centred, unsigned ichan)
{
  double base;
  if (dc_centred)
    base = cfreq - 0.5 * bw;
  else
    base = cfreq - 0.5 * bw + 0.5 * bw / double(nchan);
  return base + double(ichan) * bw / double(nchan);
}

// coeff = -sign * 2 * pi * (1e6 * DM / 2.41e-4) / f_c^2   (Dedispersion.C:168-173)
double chirp_coeff(double bw, double dm, double chan_freq)
{
  double sign = bw / std::fabs(bw);
  double dispersion_per_MHz = 1e6 * dm / kDmDispersion;
  return -sign * 2 * M_PI * dispersion_per_MHz / (chan_freq * chan_freq);
}

// Dedispersion::SampleDelay::get_sample_delay (DedispersionSampleDelay.C:
// 82-97) with the channel sample rate |bw|/nchan MHz that Dedispersion::build
// configures it with (Dedispersion.C:86-97): the dispersion delay of
// frequency freq relative to the band centre, in samples, split into the
// nearest integer and the remainder.
void sample_delay(double cfreq, double bw, unsigned nchan, double dm, double freq, int64_t& idelay, double& frac)
{
  double dispersion = dm / kDmDispersion;
  double chanwidth = std::fabs(bw) / double(nchan);
  double rate = chanwidth * 1e6;
  double time_delay = dispersion * (1.0 / (cfreq * cfreq) - 1.0 / (freq * freq));
  double sd = time_delay * rate;
  idelay = int64_t(std::floor(sd + 0.5));
  frac = sd - double(idelay);
}

// Dedispersion::build_setup (Dedispersion.C:142-164): the fractional sample
// delay of channel frequency chan_freq as a time in microseconds (Doppler 1).
double interchan_fractional_delay(double cfreq, double bw, unsigned nchan, double dm, double chan_freq)
{
  int64_t idelay;
  double frac;
  sample_delay(cfreq, bw, nchan, dm, chan_freq, idelay, frac);
  double chanwidth = bw / double(nchan);
  return frac / chanwidth;
}

// SampleDelay::build with one span (SampleDelay.C:56-160) and the applied
// offsets of SampleDelayEngine::set_delays (SampleDelayCUDA.cu:92-107).
int sample_delays(const dspx_dedisp_desc* d, int64_t* delay, double* frac, int64_t* applied, int64_t* zero_delay,
                  uint64_t* total_delay)
{
  if (!d || d->nchan == 0 || !(d->bandwidth != 0.0) || !(d->centre_frequency > 0.0)) return DSPX_E_PARAM;
  const unsigned nchan = d->nchan;
  std::vector<int64_t> dl(nchan);
  std::vector<double> fr(nchan), freqs(nchan);
  for (unsigned c = 0; c < nchan; c++) {
    freqs[c] = channel_centre(d->centre_frequency, d->bandwidth, nchan, int(d->dc_centred), c);
    sample_delay(d->centre_frequency, d->bandwidth, nchan, d->dispersion_measure, freqs[c], dl[c], fr[c]);
  }
  // get_delay_range(0, nchan - 1) (DedispersionSampleDelay.C:122-135)
  const double dispersion = d->dispersion_measure / kDmDispersion;
  const double rate = (std::fabs(d->bandwidth) / double(nchan)) * 1e6;
  const double fmid = (freqs[0] + freqs[nchan - 1]) / 2;
  const double dmid = dispersion * (1.0 / (d->centre_frequency * d->centre_frequency) - 1.0 / (fmid * fmid));
  const int64_t centre = int64_t(std::floor(dmid * rate + 0.5));
  int64_t zd = 0;
  for (unsigned c = 0; c < nchan; c++) zd = std::max(zd, dl[c] - centre);
  uint64_t total = 0;
  for (unsigned c = 0; c < nchan; c++) {
    const int64_t zc = centre + zd;  // zero_delays[c]
    if (zc - dl[c] > int64_t(total)) total = uint64_t(zc - dl[c]);
    const int64_t a = zc ? zc - dl[c] : dl[c];
    if (a < 0) return DSPX_E_STATE;  // SampleDelayCUDA.cu:100-104
    // every read stays inside the input (ndat_out = ndat - total)
    if (uint64_t(a) > total) total = uint64_t(a);
    if (applied) applied[c] = a;
    if (delay) delay[c] = dl[c];
    if (frac) frac[c] = fr[c];
  }
  if (zero_delay) *zero_delay = zd;
  if (total_delay) *total_delay = total;
  return DSPX_OK;
}

}  // namespace dspx
