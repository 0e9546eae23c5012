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
#include <cmath>

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

}  // namespace dspx
