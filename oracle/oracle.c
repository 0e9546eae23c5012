/*
 * oracle.c -- CPU restatement of the dspsr coherent-dedispersion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing that is measured or shipped.  The product
 * path (libdspx.so) does not link, load or call anything in oracle/.
 *
 * Every function restates the reference algorithm it cites (paths relative to
 * the dspsr source tree).  Compile with -ffp-contract=off so that every
 * floating-point operation is the separate IEEE operation the reference
 * source writes (no fused multiply-add contraction).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   ox_optimal_fft_length   -- against the reference's own optimize_fft.c,
 *                              compiled from its sources into oracle/_ref.
 *   ox_detect_* (Coherence/Stokes) -- against the reference's cross_detect.c /
 *                              stokes_detect.c compiled into oracle/_ref, and the
 *                              closed forms of test/Detection/test_Detection.cpp.
 *   ox_fft_*                -- brute-force DFT (N <= 4096) and numpy FP64 FFT.
 *   ox_convolve_*           -- the ConvolutionTest tukey known-answer test
 *                              (Signal/General/tests/ConvolutionTest.cpp:148-308)
 *                              and FP64 circular convolution by direct sum.
 *   ox_dedisp_response_*    -- physics: an impulse dispersed with the cold-plasma
 *                              group delay (closed form) is restored to one sample.
 *                              The frequency grid itself lives in a header that
 *                              is absent from the reference mirror: parity of the
 *                              grid is unpinned by source (pinned by the KAT).
 *   ox_smearing_samples / ox_conv_geometry -- closed-form delay differences
 *                              recomputed in the tests at several (f, bw, DM).
 *   ox_fold                 -- brute-force restatement in the tests and the
 *                              constant-period closed form for bin indices.
 *   ox_unpack8/16           -- closed forms (v+0.5)*scale, float(v)+0.5.
 *   ox_sk_*                 -- closed-form SK estimator on constructed sums; the
 *                              thresholds come from a committed table
 *                              (tests/golden), parity unpinned beyond that table.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Geometry                                                                 */
/* ------------------------------------------------------------------------ */

/* Dedispersion.C:32 */
static const double ox_dm_dispersion = 2.41e-4;
/* PlasmaResponse.C:20 (a float: 1.0 + smearing_buffer is 1.100000001490116) */
static const float ox_smearing_buffer = 0.1f;
/* PlasmaResponse.C:89 */
static const unsigned ox_smearing_samples_threshold = 16u * 1024u * 1024u;

/* Dedispersion.C:136-140 -- dispersion delay in seconds at freq (MHz) */
double ox_delay_time(double dm, double freq)
{
  double dispersion = dm / ox_dm_dispersion;
  return dispersion * 1.0 / (freq * freq);
}

/* PlasmaResponse.C:132-140 */
static double ox_smearing_time_band(double dm, double cfreq, double bw)
{
  double f1 = cfreq - fabs(0.5 * bw);
  double f2 = cfreq + fabs(0.5 * bw);
  return ox_delay_time(dm, f1) - ox_delay_time(dm, f2);
}

/* PlasmaResponse.C:161-208; nunsupported = number of leading channels
   flagged unsupported by PlasmaResponse::prepare */
double ox_smearing_time(double cfreq, double bw, unsigned nchan, double dm,
                        int half, unsigned nunsupported)
{
  double abs_bw = fabs(bw);
  double ch_abs_bw = abs_bw / (double) nchan;
  double lower_ch_cfreq = cfreq - (abs_bw - ch_abs_bw) / 2.0;
  unsigned ichan;
  for (ichan = 0; ichan < nunsupported && ichan < nchan; ichan++)
    lower_ch_cfreq += ch_abs_bw;
  if (half)
  {
    ch_abs_bw /= 2.0;
    lower_ch_cfreq += (double) half * ch_abs_bw;
  }
  return ox_smearing_time_band(dm, lower_ch_cfreq, ch_abs_bw);
}

/* PlasmaResponse.C:210-250 (psrdisp_compatible = false) */
unsigned ox_smearing_samples(double cfreq, double bw, unsigned nchan, double dm,
                             int half, unsigned nunsupported)
{
  double tsmear = ox_smearing_time(cfreq, bw, nchan, dm, half, nunsupported);
  double sampling_rate = fabs(bw) / (double) nchan * 1e6; /* :265-269 */
  tsmear *= (1.0 + ox_smearing_buffer);
  return (unsigned) ceil(tsmear * sampling_rate);
}

/* PlasmaResponse.C:91-119.  Returns 0, or -1 when every channel exceeds the
   threshold (the reference throws InvalidState). */
int ox_plasma_prepare(double cfreq, double bw, unsigned nchan, double dm,
                      unsigned* impulse_neg, unsigned* impulse_pos,
                      unsigned* nunsupported)
{
  unsigned threshold = ox_smearing_samples_threshold / nchan;
  unsigned ichan = 0;
  unsigned neg;
  while ((neg = ox_smearing_samples(cfreq, bw, nchan, dm, -1, ichan)) > threshold)
  {
    ichan++;
    if (ichan == nchan)
      return -1;
  }
  *impulse_neg = neg;
  *impulse_pos = ox_smearing_samples(cfreq, bw, nchan, dm, 1, ichan);
  *nunsupported = ichan;
  return 0;
}

/* optimize_fft.c:64-135 */
int64_t ox_optimal_fft_length(uint64_t nbadperfft, uint64_t nfft_max, unsigned factor)
{
  uint64_t nfft_min, nfft;
  double order_fft, timescale, prev_timescale;
  if (!nbadperfft)
    return -1;
  nfft_min = factor;
  while (nbadperfft > nfft_min)
    nfft_min *= 2;
  if (nfft_max && nfft_max < nfft_min)
    return -1;
  nfft = nfft_min;
  order_fft = (double) nfft * log((double) nfft);
  timescale = order_fft / (double) (nfft - nbadperfft);
  while (nfft_max == 0 || nfft * 2 < nfft_max)
  {
    prev_timescale = timescale;
    nfft *= 2;
    order_fft = (double) nfft * log((double) nfft);
    timescale = order_fft / (double) (nfft - nbadperfft);
    if (timescale > prev_timescale)
    {
      nfft /= 2;
      break;
    }
  }
  return (int64_t) nfft;
}

/* Convolution geometry: PlasmaResponse::prepare -> Response::set_optimal_ndat
   (Response.C:553-588) -> Convolution::prepare/prepare_output
   (Convolution.C:316-495), Analytic (complex) input.
   out[0..5] = impulse_neg, impulse_pos, nfft, nsamp_overlap, nsamp_step,
   nunsupported.  nfft_user != 0 mimics a user-set frequency resolution. */
int ox_conv_geometry(double cfreq, double bw, unsigned nchan, double dm,
                     unsigned nfft_user, uint64_t* out)
{
  unsigned ineg = 0, ipos = 0, nuns = 0;
  int64_t nfft;
  if (ox_plasma_prepare(cfreq, bw, nchan, dm, &ineg, &ipos, &nuns) != 0)
    return -1;
  if (nfft_user)
    nfft = nfft_user;
  else
    nfft = ox_optimal_fft_length((uint64_t) ineg + ipos, 0, 1);
  if (nfft < 2 || (uint64_t) nfft < (uint64_t) ineg + ipos)
    return -2;
  out[0] = ineg;
  out[1] = ipos;
  out[2] = (uint64_t) nfft;
  out[3] = (uint64_t) ineg + ipos;
  out[4] = (uint64_t) nfft - ((uint64_t) ineg + ipos);
  out[5] = nuns;
  return 0;
}

/* Convolution.C:485-487 */
uint64_t ox_conv_npart(uint64_t ndat, uint64_t nfft, uint64_t noverlap)
{
  if (ndat < nfft)
    return 0;
  return (ndat - noverlap) / (nfft - noverlap);
}

/* ------------------------------------------------------------------------ */
/* Unpack                                                                    */
/* ------------------------------------------------------------------------ */

/* GenericEightBitUnpackerCUDA.cu:23-45 with offset +0.5 (two's complement,
   :81-85) or -127.5 (offset binary, :87-92).  raw is TFP
   [ndat][nchan][npol][2]; out is FPT [nchan][npol][ndat][2]. */
void ox_unpack8(const uint8_t* raw, unsigned nchan, unsigned npol, uint64_t ndat,
                int offset_binary, float scale, float* out)
{
  uint64_t idat;
  unsigned ichan, ipol, idim;
  float offset = offset_binary ? -127.5f : 0.5f;
  for (ichan = 0; ichan < nchan; ichan++)
    for (ipol = 0; ipol < npol; ipol++)
    {
      float* o = out + ((uint64_t) ichan * npol + ipol) * ndat * 2;
      for (idat = 0; idat < ndat; idat++)
        for (idim = 0; idim < 2; idim++)
        {
          uint8_t b = raw[((idat * nchan + ichan) * npol + ipol) * 2 + idim];
          float v = offset_binary ? (float) b : (float) (int8_t) b;
          o[idat * 2 + idim] = (v + offset) * scale;
        }
    }
}

/* SixteenBitUnpacker.C:56-78: float(v)+0.5, offset binary XORs 0x8000. */
void ox_unpack16(const int16_t* raw, unsigned nchan, unsigned npol, uint64_t ndat,
                 int offset_binary, float* out)
{
  uint64_t idat;
  unsigned ichan, ipol, idim;
  for (ichan = 0; ichan < nchan; ichan++)
    for (ipol = 0; ipol < npol; ipol++)
    {
      float* o = out + ((uint64_t) ichan * npol + ipol) * ndat * 2;
      for (idat = 0; idat < ndat; idat++)
        for (idim = 0; idim < 2; idim++)
        {
          int16_t v = raw[((idat * nchan + ichan) * npol + ipol) * 2 + idim];
          if (offset_binary)
            v = (int16_t) (v ^ 0x8000);
          o[idat * 2 + idim] = (float) v + 0.5f;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* FFT: iterative radix-2, unnormalised, twiddles from FP64 (FFTW-like       */
/* accuracy).  sign = -1 forward (FFTW_FORWARD), +1 backward.               */
/* ------------------------------------------------------------------------ */

#define OX_DEFINE_FFT(T, NAME)                                                 \
  void NAME(uint64_t n, T* x, int sign)                                        \
  {                                                                            \
    uint64_t i, j, len, k, half, step;                                         \
    T* tw;                                                                     \
    if (n < 2) return;                                                         \
    tw = (T*) malloc(sizeof(T) * n);                                           \
    for (k = 0; k < n / 2; k++)                                                \
    {                                                                          \
      double a = 2.0 * M_PI * (double) k / (double) n;                         \
      tw[2 * k] = (T) cos(a);                                                  \
      tw[2 * k + 1] = (T) ((double) sign * sin(a));                            \
    }                                                                          \
    for (i = 1, j = 0; i < n; i++)                                             \
    {                                                                          \
      uint64_t bit = n >> 1;                                                   \
      for (; j & bit; bit >>= 1) j ^= bit;                                     \
      j ^= bit;                                                                \
      if (i < j)                                                               \
      {                                                                        \
        T tr = x[2 * i], ti = x[2 * i + 1];                                    \
        x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];                      \
        x[2 * j] = tr; x[2 * j + 1] = ti;                                      \
      }                                                                        \
    }                                                                          \
    for (len = 2; len <= n; len <<= 1)                                         \
    {                                                                          \
      half = len >> 1;                                                         \
      step = n / len;                                                          \
      for (i = 0; i < n; i += len)                                             \
        for (k = 0; k < half; k++)                                             \
        {                                                                      \
          T wr = tw[2 * k * step], wi = tw[2 * k * step + 1];                  \
          T* a = x + 2 * (i + k);                                              \
          T* b = x + 2 * (i + k + half);                                       \
          T vr = b[0] * wr - b[1] * wi;                                        \
          T vi = b[0] * wi + b[1] * wr;                                        \
          T ur = a[0], ui = a[1];                                              \
          a[0] = ur + vr; a[1] = ui + vi;                                      \
          b[0] = ur - vr; b[1] = ui - vi;                                      \
        }                                                                      \
    }                                                                          \
    free(tw);                                                                  \
  }

OX_DEFINE_FFT(float, ox_fft_f32)
OX_DEFINE_FFT(double, ox_fft_f64)

/* ------------------------------------------------------------------------ */
/* Dedispersion response                                                     */
/* ------------------------------------------------------------------------ */

/* Observation.C:521-535 -- centre frequency of channel ichan (unswapped). */
double ox_channel_centre(double cfreq, double bw, unsigned nchan, int dc_centred,
                         unsigned ichan)
{
  double base;
  if (dc_centred)
    base = cfreq - 0.5 * bw;
  else
    base = cfreq - 0.5 * bw + 0.5 * bw / (double) nchan;
  return base + (double) ichan * bw / (double) nchan;
}

/* Dedispersion.C:166-186 (build_compute).  freq is the offset (MHz) from
   the channel centre chan_freq; delay_us the channel's fractional
   inter-channel delay (Dedispersion.C:142-164; 0 without -K). */
double ox_chirp_phase_delay(double bw, double dm, double chan_freq, double freq, double delay_us)
{
  double sign = bw / fabs(bw);
  double dispersion_per_MHz = 1e6 * dm / ox_dm_dispersion;
  double coeff = -sign * 2 * M_PI * dispersion_per_MHz / (chan_freq * chan_freq);
  double delay_phase = -2.0 * M_PI * freq * delay_us;
  return coeff * (freq * freq) / (chan_freq + freq) + delay_phase;
}

double ox_chirp_phase(double bw, double dm, double chan_freq, double freq)
{
  return ox_chirp_phase_delay(bw, dm, chan_freq, freq, 0.0);
}

/* Dedispersion::SampleDelay::get_sample_delay (DedispersionSampleDelay.C:
   82-97), configured by Dedispersion::build with the band's centre frequency,
   DM and the channel rate |bw|/nchan MHz (Dedispersion.C:86-97): the delay of
   frequency freq relative to the band centre in samples, as the nearest
   integer and the remainder. */
void ox_sample_delay(double cfreq, double bw, unsigned nchan, double dm, double freq, int64_t* idelay,
                     double* frac)
{
  double dispersion = dm / ox_dm_dispersion;
  double chanwidth = fabs(bw) / (double) nchan;
  double rate = chanwidth * 1e6;
  double time_delay = dispersion * (1.0 / (cfreq * cfreq) - 1.0 / (freq * freq));
  double sd = time_delay * rate;
  *idelay = (int64_t) floor(sd + 0.5);
  *frac = sd - (double) *idelay;
}

/* Dedispersion::build_setup (Dedispersion.C:142-164): the fractional delay
   as a time in microseconds (Doppler 1). */
double ox_interchan_delay(double cfreq, double bw, unsigned nchan, double dm, double chan_freq)
{
  int64_t idelay;
  double frac, chanwidth = bw / (double) nchan;
  ox_sample_delay(cfreq, bw, nchan, dm, chan_freq, &idelay, &frac);
  return frac / chanwidth;
}

/* SampleDelay::build with one span (SampleDelay.C:56-160) and the offsets
   SampleDelayEngine::set_delays applies (SampleDelayCUDA.cu:92-107):
   applied[c] = zero_delays[c] - delay[c] (delay[c] where zero_delays[c] is
   0); total = the largest applied offset, so that output sample i of every
   channel reads input sample i + applied[c] < ndat.  Returns -1 for a
   negative offset (the reference throws). */
int ox_sample_delays(double cfreq, double bw, unsigned nchan, int dc_centred, double dm, int64_t* delay,
                     double* frac, int64_t* applied, int64_t* zero_delay, uint64_t* total)
{
  unsigned c;
  double dispersion = dm / ox_dm_dispersion;
  double rate = (fabs(bw) / (double) nchan) * 1e6;
  double f0 = ox_channel_centre(cfreq, bw, nchan, dc_centred, 0);
  double f1 = ox_channel_centre(cfreq, bw, nchan, dc_centred, nchan - 1);
  double fmid = (f0 + f1) / 2;
  double dmid = dispersion * (1.0 / (cfreq * cfreq) - 1.0 / (fmid * fmid));
  int64_t centre = (int64_t) floor(dmid * rate + 0.5), zd = 0;
  uint64_t tot = 0;
  for (c = 0; c < nchan; c++)
  {
    ox_sample_delay(cfreq, bw, nchan, dm, ox_channel_centre(cfreq, bw, nchan, dc_centred, c), &delay[c], &frac[c]);
    if (delay[c] - centre > zd)
      zd = delay[c] - centre;
  }
  for (c = 0; c < nchan; c++)
  {
    int64_t zc = centre + zd;
    if (zc - delay[c] > (int64_t) tot)
      tot = (uint64_t) (zc - delay[c]);
    applied[c] = zc ? zc - delay[c] : delay[c];
    if (applied[c] < 0)
      return -1;
    if ((uint64_t) applied[c] > tot)
      tot = (uint64_t) applied[c];
  }
  *zero_delay = zd;
  *total = tot;
  return 0;
}

/* Offset frequency of point ipt of an npt-point channel response, lowest
   frequency first.  (The loop that evaluates this lives in a header absent
   from the reference mirror; restated as (ipt - npt/2) * chanwidth / npt.) */
double ox_chirp_freq(double chanwidth, unsigned npt, unsigned ipt)
{
  double binwidth = chanwidth / (double) npt;
  return (double) ipt * binwidth - 0.5 * chanwidth;
}

/* Full response as the Convolution applies it, in FFT-bin order:
   Dedispersion::build (Dedispersion.C:84-129: FP64 phase stored as float,
   then polar(1.0f, phase)), times the normaliser ScalarFilter 1/sqrt(nfft)
   (Convolution.C:401-414, ResponseProduct.C:142-248, Response.C:675-726),
   then swapped so that FFT bin k holds point (k + nfft/2) % nfft
   (Response.C:325-400, dual-sideband Analytic input).
   mode 0 = reference parity (float phase); mode 1 = FP64 phase (truth
   phase, evaluated in double and rounded to float at the end).
   resp: [nchan][nfft][2] float. */
void ox_dedisp_response_f32(double cfreq, double bw, unsigned nchan, int dc_centred,
                            double dm, unsigned nfft, int mode, int interchan, float* resp)
{
  unsigned ichan, k;
  double chanwidth = bw / (double) nchan;
  float s = (float) (1.0 / sqrt((double) nfft));
  for (ichan = 0; ichan < nchan; ichan++)
  {
    double cf = ox_channel_centre(cfreq, bw, nchan, dc_centred, ichan);
    double dly = interchan ? ox_interchan_delay(cfreq, bw, nchan, dm, cf) : 0.0;
    for (k = 0; k < nfft; k++)
    {
      unsigned ipt = (k + nfft / 2) % nfft;
      double ph = ox_chirp_phase_delay(bw, dm, cf, ox_chirp_freq(chanwidth, nfft, ipt), dly);
      float c, sn;
      if (mode == 0)
      {
        float ph32 = (float) ph;
        c = cosf(ph32);
        sn = sinf(ph32);
      }
      else
      {
        c = (float) cos(ph);
        sn = (float) sin(ph);
      }
      /* product with the scalar filter (s + 0i) * (c + i sn) */
      resp[2 * ((uint64_t) ichan * nfft + k)] = s * c - 0.0f * sn;
      resp[2 * ((uint64_t) ichan * nfft + k) + 1] = 0.0f * c + s * sn;
    }
  }
}

/* FP64 truth response: exp(i * phase) / sqrt(nfft), all in double. */
void ox_dedisp_response_f64(double cfreq, double bw, unsigned nchan, int dc_centred,
                            double dm, unsigned nfft, int interchan, double* resp)
{
  unsigned ichan, k;
  double chanwidth = bw / (double) nchan;
  double s = 1.0 / sqrt((double) nfft);
  for (ichan = 0; ichan < nchan; ichan++)
  {
    double cf = ox_channel_centre(cfreq, bw, nchan, dc_centred, ichan);
    double dly = interchan ? ox_interchan_delay(cfreq, bw, nchan, dm, cf) : 0.0;
    for (k = 0; k < nfft; k++)
    {
      unsigned ipt = (k + nfft / 2) % nfft;
      double ph = ox_chirp_phase_delay(bw, dm, cf, ox_chirp_freq(chanwidth, nfft, ipt), dly);
      resp[2 * ((uint64_t) ichan * nfft + k)] = s * cos(ph);
      resp[2 * ((uint64_t) ichan * nfft + k) + 1] = s * sin(ph);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Overlap-save convolution: Convolution.C:700-803 (Analytic input).         */
/* in:  [nchan][npol][ndat][2]; resp: [nchan][nfft][2] (npol 1 response      */
/* applies to every pol, Response.C:665-667); out: [nchan][npol][npart*step][2].*/
/* zdm (optional) receives the zero-DM copy (Convolution.C:792-799).         */
/* ------------------------------------------------------------------------ */

#define OX_DEFINE_CONV(T, NAME, FFT)                                           \
  uint64_t NAME(const T* in, unsigned nchan, unsigned npol, uint64_t ndat,     \
                const T* resp, unsigned nfft, unsigned nfilt_pos,              \
                unsigned nfilt_neg, T* out, T* zdm)                            \
  {                                                                            \
    uint64_t nov = (uint64_t) nfilt_pos + nfilt_neg;                           \
    uint64_t step = nfft - nov;                                                \
    uint64_t npart = (ndat < nfft) ? 0 : (ndat - nov) / step;                  \
    uint64_t outndat = npart * step;                                           \
    unsigned ichan, ipol;                                                      \
    uint64_t ipart, k;                                                         \
    T* buf = (T*) malloc(sizeof(T) * 2 * nfft);                                \
    for (ichan = 0; ichan < nchan; ichan++)                                    \
      for (ipol = 0; ipol < npol; ipol++)                                      \
      {                                                                        \
        const T* x = in + ((uint64_t) ichan * npol + ipol) * ndat * 2;         \
        T* y = out + ((uint64_t) ichan * npol + ipol) * outndat * 2;           \
        const T* h = resp + (uint64_t) ichan * nfft * 2;                       \
        for (ipart = 0; ipart < npart; ipart++)                                \
        {                                                                      \
          memcpy(buf, x + ipart * step * 2, sizeof(T) * 2 * nfft);             \
          FFT(nfft, buf, -1);                                                  \
          for (k = 0; k < nfft; k++)                                           \
          {                                                                    \
            T d_r = buf[2 * k], d_i = buf[2 * k + 1];                          \
            T f_r = h[2 * k], f_i = h[2 * k + 1];                              \
            buf[2 * k] = f_r * d_r - f_i * d_i;                                \
            buf[2 * k + 1] = f_i * d_r + f_r * d_i;                            \
          }                                                                    \
          FFT(nfft, buf, +1);                                                  \
          memcpy(y + ipart * step * 2, buf + 2 * (uint64_t) nfilt_pos,         \
                 sizeof(T) * 2 * step);                                        \
          if (zdm)                                                             \
            memcpy(zdm + ((uint64_t) ichan * npol + ipol) * outndat * 2        \
                       + ipart * step * 2,                                     \
                   x + ipart * step * 2 + 2 * (uint64_t) nfilt_pos,            \
                   sizeof(T) * 2 * step);                                      \
        }                                                                      \
      }                                                                        \
    free(buf);                                                                 \
    return npart;                                                              \
  }

OX_DEFINE_CONV(float, ox_convolve_f32, ox_fft_f32)
OX_DEFINE_CONV(double, ox_convolve_f64, ox_fft_f64)

/* Overlap-save convolution of real-sampled (Nyquist) input: Convolution.C:
   385-413 (nsamp_fft = 2*nfft real samples, nsamp_overlap = 2*nfilt_tot,
   normaliser 1/sqrt(nfft)), 597-608 (output ndat = npart*step/2, Analytic),
   700-790 (frc1d forward, response over the nfft bins, bcc1d backward of
   nfft points, nfilt_pos complex samples dropped).  The forward transform
   of the 2*nfft real window is its 2*nfft-point complex DFT (imaginary part
   zero), of which bins 0..nfft-1 are used (the Nyquist bin is not).
     in  [nchan][npol][ndat] real; resp [nchan][nfft][2] in frequency order
     (real input is not dual-sideband: no swap, Observation.C:95-102,
     Response.C:335-345), DC of the first channel zero (Response.C:418-420);
     out [nchan][npol][npart*step][2], step = nfft - nfilt_pos - nfilt_neg,
     npart = (ndat - 2*nfilt) / (2*step).  Returns npart. */
#define OX_DEFINE_CONV_REAL(T, NAME, FFT)                                      \
  uint64_t NAME(const T* in, unsigned nchan, unsigned npol, uint64_t ndat,     \
                const T* resp, unsigned nfft, unsigned nfilt_pos,              \
                unsigned nfilt_neg, T* out)                                    \
  {                                                                            \
    uint64_t nov = (uint64_t) nfilt_pos + nfilt_neg;                           \
    uint64_t step = nfft - nov;                                                \
    uint64_t npart = (ndat < 2 * (uint64_t) nfft) ? 0                          \
                                                  : (ndat - 2 * nov) / (2 * step); \
    uint64_t outndat = npart * step;                                           \
    unsigned ichan, ipol;                                                      \
    uint64_t ipart, k;                                                         \
    T* buf = (T*) malloc(sizeof(T) * 4 * nfft);                                \
    for (ichan = 0; ichan < nchan; ichan++)                                    \
      for (ipol = 0; ipol < npol; ipol++)                                      \
      {                                                                        \
        const T* x = in + ((uint64_t) ichan * npol + ipol) * ndat;             \
        T* y = out + ((uint64_t) ichan * npol + ipol) * outndat * 2;           \
        const T* h = resp + (uint64_t) ichan * nfft * 2;                       \
        for (ipart = 0; ipart < npart; ipart++)                                \
        {                                                                      \
          for (k = 0; k < 2 * (uint64_t) nfft; k++)                            \
          {                                                                    \
            buf[2 * k] = x[ipart * 2 * step + k];                              \
            buf[2 * k + 1] = 0;                                                \
          }                                                                    \
          FFT(2 * nfft, buf, -1);                                              \
          for (k = 0; k < nfft; k++)                                           \
          {                                                                    \
            T d_r = buf[2 * k], d_i = buf[2 * k + 1];                          \
            T f_r = h[2 * k], f_i = h[2 * k + 1];                              \
            buf[2 * k] = f_r * d_r - f_i * d_i;                                \
            buf[2 * k + 1] = f_i * d_r + f_r * d_i;                            \
          }                                                                    \
          FFT(nfft, buf, +1);                                                  \
          memcpy(y + ipart * step * 2, buf + 2 * (uint64_t) nfilt_pos,         \
                 sizeof(T) * 2 * step);                                        \
        }                                                                      \
      }                                                                        \
    free(buf);                                                                 \
    return npart;                                                              \
  }

OX_DEFINE_CONV_REAL(float, ox_convolve_real_f32, ox_fft_f32)
OX_DEFINE_CONV_REAL(double, ox_convolve_real_f64, ox_fft_f64)

/* ------------------------------------------------------------------------ */
/* Detection                                                                 */
/* ------------------------------------------------------------------------ */

/* Detection.C:252-318, FPT, Analytic input.  Intensity: out [nchan][ndat],
   summed in the order ((p.re^2 + p.im^2) + q.re^2) + q.im^2.
   PPQQ: out [nchan][npol][ndat]. */
void ox_detect_square_law(const float* in, unsigned nchan, unsigned npol,
                          uint64_t ndat, int ppqq, float* out)
{
  unsigned ichan, ipol;
  uint64_t idat;
  for (ichan = 0; ichan < nchan; ichan++)
  {
    unsigned opol = 0;
    for (ipol = 0; ipol < npol; ipol++)
    {
      const float* ip = in + ((uint64_t) ichan * npol + ipol) * ndat * 2;
      unsigned onpol = ppqq ? npol : 1;
      float* op = out + ((uint64_t) ichan * onpol + opol) * ndat;
      int sum = (!ppqq && npol == 2 && ipol == 1);
      for (idat = 0; idat < ndat; idat++)
      {
        float re = ip[2 * idat], im = ip[2 * idat + 1];
        if (sum)
          op[idat] += re * re;
        else
          op[idat] = re * re;
        op[idat] += im * im;
      }
      if (ppqq)
        opol++;
    }
  }
}

/* Detection.C:471-630 out-of-place polarimetry with cross_detect.ic:30-43
   (stokes = 0, Coherence) or stokes_detect.ic:30-44 (stokes = 1).
   in: [nchan][2][ndat][2]; out: [nchan][4/ndim][ndat*ndim] with the result
   pointers of Detection::get_result_pointers (Detection.C:578-624). */
void ox_detect_polarimetry(const float* in, unsigned nchan, uint64_t ndat,
                           unsigned ndim, int stokes, float* out)
{
  unsigned ichan;
  uint64_t j;
  unsigned onpol = 4 / ndim;
  for (ichan = 0; ichan < nchan; ichan++)
  {
    const float* p = in + (uint64_t) ichan * 2 * ndat * 2;
    const float* q = p + ndat * 2;
    float* base = out + (uint64_t) ichan * onpol * ndat * ndim;
    float* r[4];
    if (ndim == 1)
    {
      r[0] = base; r[1] = base + ndat; r[2] = base + 2 * ndat; r[3] = base + 3 * ndat;
    }
    else if (ndim == 2)
    {
      r[0] = base; r[1] = base + 1; r[2] = base + 2 * ndat; r[3] = r[2] + 1;
    }
    else
    {
      r[0] = base; r[1] = base + 1; r[2] = base + 2; r[3] = base + 3;
    }
    for (j = 0; j < ndat; j++)
    {
      float p_r = p[2 * j], p_i = p[2 * j + 1];
      float q_r = q[2 * j], q_i = q[2 * j + 1];
      if (stokes)
      {
        float pp = p_r * p_r + p_i * p_i;
        float qq = q_r * q_r + q_i * q_i;
        r[0][j * ndim] = pp + qq;
        r[1][j * ndim] = pp - qq;
        r[2][j * ndim] = (float) (2.0 * (double) (p_r * q_r + p_i * q_i));
        r[3][j * ndim] = (float) (2.0 * (double) (p_r * q_i - p_i * q_r));
      }
      else
      {
        r[0][j * ndim] = p_r * p_r + p_i * p_i;
        r[1][j * ndim] = q_r * q_r + q_i * q_i;
        r[2][j * ndim] = p_r * q_r + p_i * q_i;
        r[3][j * ndim] = p_r * q_i - p_i * q_r;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Fold: Fold.C:662-949 (no engine, FPT order).                              */
/* in: [nchan][npol][ndat*ndim]; prof: [nchan][npol][nbin*ndim] (accumulated);*/
/* hits: [hits_nchan][nbin] (accumulated).  phi is get_phi(start_time) of the */
/* first folded sample's midpoint (Fold.C:686-693) and pps is                */
/* sampling_interval / pfold (Fold.C:762-764), both computed by the caller.  */
/* weights may be NULL (no WeightedTimeSeries).  binplan (optional) receives */
/* the per-sample bin (nbin marks a sample in a bad weight).  Returns        */
/* ndat_folded (Fold.C:831-836, 911-930).                                    */
/* ------------------------------------------------------------------------ */
uint64_t ox_fold(const float* in, unsigned nchan, unsigned npol, unsigned ndim,
                 uint64_t ndat, uint64_t idat_start, uint64_t ndat_fold,
                 double phi, double phase_per_sample, unsigned nbin,
                 float* prof, uint32_t* hits, unsigned hits_nchan,
                 int zeroed_samples, const uint16_t* weights, uint64_t nweights,
                 unsigned ndatperweight, unsigned weight_idat,
                 uint32_t* binplan_out)
{
  uint64_t idat, idat_end = idat_start + ndat_fold;
  uint64_t iweight = 0, idat_nextweight = 0, ndat_folded = 0;
  int bad_data = 0;
  double double_nbin = (double) nbin;
  unsigned ichan, ipol, idim;
  uint32_t* binplan;
  uint32_t* h = hits;
  if (idat_end > ndat)
    return (uint64_t) -1;
  binplan = binplan_out ? binplan_out
                        : (uint32_t*) malloc(sizeof(uint32_t) * (ndat_fold ? ndat_fold : 1));
  if (ndatperweight)
  {
    iweight = (idat_start + weight_idat) / ndatperweight;
    idat_nextweight = (iweight + 1) * ndatperweight - weight_idat;
    if (iweight >= nweights)
    {
      if (!binplan_out) free(binplan);
      return (uint64_t) -2;
    }
    if (!zeroed_samples && weights[iweight] == 0)
      bad_data = 1;
  }
  for (idat = idat_start; idat < idat_end; idat++)
  {
    unsigned ibin;
    if (ndatperweight && idat >= idat_nextweight)
    {
      iweight++;
      bad_data = (!zeroed_samples && weights[iweight] == 0);
      idat_nextweight += ndatperweight;
    }
    phi -= floor(phi);
    ibin = (unsigned) (phi * double_nbin);
    phi += phase_per_sample;
    if (bad_data)
      binplan[idat - idat_start] = nbin;
    else
    {
      binplan[idat - idat_start] = ibin;
      if (!zeroed_samples)
      {
        hits[ibin]++;
        ndat_folded++;
      }
    }
  }
  for (ichan = 0; ichan < nchan; ichan++)
  {
    for (ipol = 0; ipol < npol; ipol++)
    {
      const float* timep = in + ((uint64_t) ichan * npol + ipol) * ndat * ndim
                           + idat_start * ndim;
      float* phasep = prof + ((uint64_t) ichan * npol + ipol) * nbin * ndim;
      for (idat = 0; idat < ndat_fold; idat++)
      {
        if (binplan[idat] != nbin)
        {
          float* phdimp = phasep + (uint64_t) binplan[idat] * ndim;
          for (idim = 0; idim < ndim; idim++)
            phdimp[idim] += timep[idim];
          if (zeroed_samples && ipol == 0 && timep[0] != 0)
          {
            h[binplan[idat]]++;
            ndat_folded++;
          }
        }
        timep += ndim;
      }
    }
    if (zeroed_samples && hits_nchan == nchan)
      h += nbin;
  }
  if (!binplan_out)
    free(binplan);
  return ndat_folded;
}

/* ------------------------------------------------------------------------ */
/* Spectral kurtosis (SpectralKurtosis.C, CPU path).                          */
/* ------------------------------------------------------------------------ */

/* Resolution::prepare (SpectralKurtosis.C:263-281): blocks of M samples
   every M / noverlap samples; npart of them fit in ndat, and the masked
   output holds *ndat_out samples. */
uint64_t ox_sk_blocks(uint64_t ndat, unsigned M, unsigned noverlap, uint64_t* ndat_out)
{
  uint64_t idat_off = M / noverlap, idat_start = M - idat_off, nblock, idat_end;
  *ndat_out = 0;
  if (ndat < idat_start)
    return 0;
  nblock = (ndat - idat_start) / idat_off;
  idat_end = nblock * idat_off;
  if (idat_end < idat_start)
    return 0;
  *ndat_out = idat_end - idat_start;
  return nblock;
}

/* SpectralKurtosis.C:674-798 (compute, FPT input, no engine).  Per block of
   M samples (block ipart from ipart * M / noverlap): sqld = re*re + im*im
   (float), S1 += sqld, S2 += sqld*sqld (float product added to a double
   accumulator), skipping zero samples when the input carries zeroed data.
   The sums are stored as float, TFP [npart][nchan][npol][2], and the count
   of the first pol per [npart][nchan].  Analytic input
   [nchan][npol][ndat][2].  Returns npart. */
uint64_t ox_sk_compute(const float* in, unsigned nchan, unsigned npol, uint64_t ndat,
                       unsigned M, unsigned noverlap, int zeroed_data, float* sums, uint32_t* counts)
{
  uint64_t ndat_out, npart = ox_sk_blocks(ndat, M, noverlap, &ndat_out), ipart, i;
  const uint64_t off = M / noverlap;
  unsigned ichan, ipol;
  for (ipart = 0; ipart < npart; ipart++)
    for (ichan = 0; ichan < nchan; ichan++)
      for (ipol = 0; ipol < npol; ipol++)
      {
        const float* x = in + ((uint64_t) ichan * npol + ipol) * ndat * 2 + ipart * off * 2;
        double S1_sum = 0, S2_sum = 0;
        unsigned count = 0;
        for (i = 0; i < M; i++)
        {
          float re = x[2 * i], im = x[2 * i + 1];
          float sqld = (re * re) + (im * im);
          if (!zeroed_data || sqld != 0)
          {
            S1_sum += sqld;
            S2_sum += sqld * sqld;
            count++;
          }
        }
        sums[((ipart * nchan + ichan) * npol + ipol) * 2] = (float) S1_sum;
        sums[((ipart * nchan + ichan) * npol + ipol) * 2 + 1] = (float) S2_sum;
        if (ipol == 0)
          counts[ipart * nchan + ichan] = count;
      }
  return npart;
}

/* SpectralKurtosis.C:1140-1200 (detect_skfb, single resolution): per block
   and channel, V = M_fac * (M' * (S2 / (S1*S1)) - 1) in float with
   M_fac = (float)(M'*Nd + 1) / (M' - 1); the block is zapped when any pol
   falls outside [lower, upper] or when M' <= 2.  Blocks with M' == 0 are
   skipped.  lower/upper are indexed by M' (tables of length M+1, computed by
   the caller from SKLimits).  mask: [npart][nchan] bytes, set to 1 when
   zapped (not cleared).  Returns the number of zapped blocks. */
uint64_t ox_sk_detect_skfb(const float* sums, const uint32_t* counts, uint64_t npart,
                           unsigned nchan, unsigned npol, unsigned Nd,
                           const float* lower, const float* upper, uint8_t* mask)
{
  uint64_t ipart, nzap = 0;
  unsigned ichan, ipol;
  for (ipart = 0; ipart < npart; ipart++)
    for (ichan = 0; ichan < nchan; ichan++)
    {
      unsigned Mp = counts[ipart * nchan + ichan];
      int zap = 1;
      if (Mp == 0)
        continue;
      if (Mp > 2)
      {
        const float M_fac = (float) (Mp * Nd + 1) / (Mp - 1);
        zap = 0;
        for (ipol = 0; ipol < npol; ipol++)
        {
          const float* sd = sums + ((ipart * nchan + ichan) * npol + ipol) * 2;
          float S1_sum = sd[0], S2_sum = sd[1];
          float V = M_fac * (Mp * (S2_sum / (S1_sum * S1_sum)) - 1);
          if (V > upper[Mp] || V < lower[Mp])
            zap = 1;
        }
      }
      if (zap)
      {
        nzap++;
        mask[ipart * nchan + ichan] = 1;
      }
    }
  return nzap;
}

/* SKLimits.C:152-160: thresholds of a count M >= 32768,
   1 -/+ sqrt(4/M) * per_comparison_std_devs (pcsd), stored as float. */
static void ox_sk_limits_large(uint64_t M, double pcsd, float* lower, float* upper)
{
  double one_std_dev = sqrt(4.0 / (double) M);
  double factor = one_std_dev * pcsd;
  *lower = (float) (1.0f - factor);
  *upper = (float) (1.0f + factor);
}

/* SpectralKurtosis::detect (SpectralKurtosis.C:830-909) with one resolution
   (noverlap 1, every channel included): the mask is reset, then
     flags & 1: detect_skfb  (SpectralKurtosis.C:1105-1222)
     flags & 2: detect_fscr  (SpectralKurtosis.C:1312-1462)
     flags & 4: detect_tscr  (SpectralKurtosis.C:1014-1103)
   lower/upper/symmetric: SKLimits thresholds by count (tables of M+1).
   detect_tscr tests every total with SKLimits(total) (get_limits, :1045):
   totals up to M from the tables, larger ones from tscr_totals (ascending,
   ntscr of them, thresholds tscr_lower/tscr_upper), else, from 32768 on, the
   closed form of SKLimits.C:152-160 with pcsd (per-comparison sigma).  A
   total in (M, 32768) without thresholds is an error: returns UINT64_MAX.
   With omit_outliers, skfb zeroes the sums and counts of the blocks it zaps
   (:1189-1199) and fscr skips them (:1382).  Unsigned arithmetic as in the
   reference (Mprime - 1 wraps for Mprime = 0).  Returns the number of mask
   bytes set. */
uint64_t ox_sk_detect(float* sums, uint32_t* counts, uint64_t npart, unsigned nchan, unsigned npol,
                      unsigned M, unsigned Nd, int omit_outliers, unsigned flags, const float* lower,
                      const float* upper, const float* symmetric, double pcsd, const uint64_t* tscr_totals,
                      const float* tscr_lower, const float* tscr_upper, unsigned ntscr, unsigned noverlap,
                      const uint8_t* channels, uint8_t* mask)
{
  uint64_t ipart, i, nset = 0;
  unsigned ichan, ipol;
  memset(mask, 0, npart * nchan);
  if (flags & 1)
    for (ipart = 0; ipart < npart; ipart++)
      for (ichan = 0; ichan < nchan; ichan++)
      {
        uint32_t* wt = counts + ipart * nchan + ichan;
        float* sd = sums + (ipart * nchan + ichan) * npol * 2;
        unsigned Mprime = *wt;
        int zap = 1;
        if (Mprime == 0)
          continue;
        if (Mprime > 2)
        {
          const float M_fac = (float) (Mprime * Nd + 1) / (Mprime - 1);
          zap = 0;
          for (ipol = 0; ipol < npol; ipol++)
          {
            float S1_sum = sd[2 * ipol], S2_sum = sd[2 * ipol + 1];
            float V = M_fac * (Mprime * (S2_sum / (S1_sum * S1_sum)) - 1);
            if (V > upper[Mprime] || V < lower[Mprime])
              zap = 1;
          }
        }
        if (zap)
        {
          if (omit_outliers)
          {
            *wt = 0;
            for (ipol = 0; ipol < npol; ipol++)
              sd[2 * ipol] = sd[2 * ipol + 1] = 0.0f;
          }
          mask[ipart * nchan + ichan] = 1;
        }
      }
  if (flags & 2)
    for (ipart = 0; ipart < npart; ipart++)
    {
      int zap_ipart = 0;
      for (ipol = 0; ipol < npol; ipol++)
      {
        double sk_avg = 0, M_tot = 0, Msq_tot = 0, chan_count = 0;
        for (ichan = 0; ichan < nchan; ichan++)
          if ((!channels || channels[ichan]) && (!omit_outliers || mask[ipart * nchan + ichan] == 0))
          {
            const float* sd = sums + ((ipart * nchan + ichan) * npol + ipol) * 2;
            float S1_sum = sd[0], S2_sum = sd[1];
            unsigned Mprime = counts[ipart * nchan + ichan];
            const float M_fac = (float) (Mprime * Nd + 1) / (Mprime - 1);
            float V = M_fac * (Mprime * (S2_sum / (S1_sum * S1_sum)) - 1);
            sk_avg += V * Mprime;
            M_tot += Mprime;
            Msq_tot += Mprime * Mprime;
            chan_count++;
          }
        if (chan_count > 0)
        {
          double M_mean, Msq_mean, Deff, effective_count;
          float threshold, up, lo;
          sk_avg /= M_tot;
          M_mean = M_tot / chan_count;
          Msq_mean = Msq_tot / chan_count;
          Deff = Msq_mean / (M_mean * M_mean);
          effective_count = chan_count / Deff;
          threshold = symmetric[(unsigned) (M_mean + 0.5)] / sqrt(effective_count);
          up = 1.0 + threshold;
          lo = 1.0 - threshold;
          if (sk_avg > up || sk_avg < lo)
            zap_ipart = 1;
        }
      }
      if (zap_ipart)
        for (ichan = 0; ichan < nchan; ichan++)
          mask[ipart * nchan + ichan] = 1;
    }
  if (flags & 4)
    for (ichan = 0; ichan < nchan; ichan++)
    {
      uint64_t count = 0;
      float lo, up, M_t, M_fac;
      int zap_chan = 0;
      if (channels && !channels[ichan])
        continue;
      for (ipart = 0; ipart < npart; ipart += noverlap)
        count += counts[ipart * nchan + ichan];
      if (count == 0)
        continue;
      if (count <= M)
      {
        lo = lower[count];
        up = upper[count];
      }
      else
      {
        unsigned k;
        for (k = 0; k < ntscr && tscr_totals[k] != count; k++)
          ;
        if (k < ntscr)
        {
          lo = tscr_lower[k];
          up = tscr_upper[k];
        }
        else if (count >= 32768)
          ox_sk_limits_large(count, pcsd, &lo, &up);
        else
          return UINT64_MAX;
      }
      M_t = (float) count;
      M_fac = (M_t * Nd + 1) / (M_t - 1);
      for (ipol = 0; ipol < npol; ipol++)
      {
        double S1 = 0.0, S2 = 0.0, SK;
        for (ipart = 0; ipart < npart; ipart += noverlap)
        {
          uint64_t idx = (ipart * nchan + ichan) * npol + ipol;
          S1 += sums[idx * 2];
          S2 += sums[idx * 2 + 1];
        }
        SK = M_fac * (M_t * (S2 / (S1 * S1)) - 1);
        if (SK > up || SK < lo)
          zap_chan = 1;
      }
      if (zap_chan)
        for (ipart = 0; ipart < npart; ipart++)
          mask[ipart * nchan + ichan] = 1;
    }
  for (i = 0; i < npart * nchan; i++)
    nset += mask[i];
  return nset;
}

/* SpectralKurtosis::detect with several resolutions (SpectralKurtosis.C:
   360-380 prepare/compatible, 849-905 detect, 917-972 tscrunch_sums, 1105-1222
   detect_skfb, 1312-1462 detect_fscr, 1014-1103 detect_tscr), sorted by M:
   resolution r has blocks of Ms[r] samples every off_r = Ms[r] / novs[r]
   samples, npart_r of them in ndat samples (Resolution::prepare), thresholds
   tables[r] = lower, upper, symmetric (Ms[r]+1 each, consecutive; the tables
   of the resolutions follow one another), and include flags channels[r][nchan].
   The sums and counts of resolution 0 (npart_0 blocks) are tscrunched in
   place from resolution r-1 to r before its tests:
     out[ipart] = sum_{isum < nsum} in[(ipart + isum) * offset],
     nsum = Ms[r] / Ms[r-1], offset = off_r / off_{r-1}
   (the reference's stride for isum is the blocks' offset; it asserts that
   the last block read lies inside the input, an error here).  The mask has
   resolution 0's rows; the tests of resolution r > 0 flag rows ipart *
   flag_offset + iflag * flag_step (iflag < nflag; nflag = Ms[r] / Ms[0],
   flag_step = novs[0], flag_offset = nflag * flag_step / novs[r]), and its
   fscr with omit_outliers reads row ipart * flag_offset.  detect_tscr takes
   the last resolution's sums (every novs[last]-th block of npart_last), its
   tables and tscr_* for larger totals, resolution 0's channels, and flags the
   first npart_last rows of the mask.  Returns the number of mask bytes set;
   UINT64_MAX for a tscr total without thresholds, UINT64_MAX - 1 for
   resolutions that are not compatible. */
uint64_t ox_sk_detect_res(float* sums, uint32_t* counts, uint64_t ndat, unsigned nchan, unsigned npol,
                          unsigned nres, const unsigned* Ms, const unsigned* novs, const float* tables,
                          const uint8_t* channels, unsigned Nd, int omit_outliers, unsigned flags, double pcsd,
                          const uint64_t* tscr_totals, const float* tscr_lower, const float* tscr_upper,
                          unsigned ntscr, uint8_t* mask)
{
  uint64_t npart[16], off[16], nd_out, i, nset = 0, have;
  const float* tab[16];
  unsigned r, ichan, ipol;
  const uint64_t fpd = (uint64_t) nchan * npol * 2;
  if (nres < 1 || nres > 16)
    return UINT64_MAX - 1;
  tab[0] = tables;
  for (r = 0; r < nres; r++)
  {
    if (Ms[r] % novs[r])
      return UINT64_MAX - 1;
    off[r] = Ms[r] / novs[r];
    npart[r] = ox_sk_blocks(ndat, Ms[r], novs[r], &nd_out);
    if (r > 0)
    {
      if (Ms[r] % Ms[r - 1] || off[r] % off[r - 1])
        return UINT64_MAX - 1;
      tab[r] = tab[r - 1] + 3 * (uint64_t) (Ms[r - 1] + 1);
    }
  }
  memset(mask, 0, npart[0] * nchan);
  have = npart[0];
  for (r = 0; r < nres; r++)
  {
    const float *lower = tab[r], *upper = tab[r] + Ms[r] + 1, *symmetric = tab[r] + 2 * (Ms[r] + 1);
    const uint8_t* chans = channels + (uint64_t) r * nchan;
    uint64_t nflag = 1, flag_step = 1, flag_offset = 1, ipart, iflag;
    if (r > 0)
    {
      /* tscrunch_sums (SpectralKurtosis.C:917-972), in place, in order */
      const uint64_t nsum = Ms[r] / Ms[r - 1], offset = off[r] / off[r - 1];
      uint64_t isum, k;
      if (npart[r] && (npart[r] - 1) * offset + (nsum - 1) * offset >= have)
        return UINT64_MAX - 1;
      for (ipart = 0; ipart < npart[r]; ipart++)
      {
        for (k = 0; k < fpd; k++)
          sums[ipart * fpd + k] = sums[ipart * offset * fpd + k];
        for (k = 0; k < nchan; k++)
          counts[ipart * nchan + k] = counts[ipart * offset * nchan + k];
        for (isum = 1; isum < nsum; isum++)
        {
          for (k = 0; k < fpd; k++)
            sums[ipart * fpd + k] += sums[(ipart + isum) * offset * fpd + k];
          for (k = 0; k < nchan; k++)
            counts[ipart * nchan + k] += counts[(ipart + isum) * offset * nchan + k];
        }
      }
      have = npart[r];
      nflag = Ms[r] / Ms[0];
      flag_step = novs[0];
      flag_offset = nflag * flag_step / novs[r];
    }
    if (flags & 1)
      for (ipart = 0; ipart < npart[r]; ipart++)
        for (ichan = 0; ichan < nchan; ichan++)
        {
          uint32_t* wt = counts + ipart * nchan + ichan;
          float* sd = sums + (ipart * nchan + ichan) * npol * 2;
          unsigned Mprime = *wt;
          int zap = 1;
          if (Mprime == 0)
            continue;
          if (Mprime > 2)
          {
            const float M_fac = (float) (Mprime * Nd + 1) / (Mprime - 1);
            zap = 0;
            for (ipol = 0; ipol < npol; ipol++)
            {
              float S1_sum = sd[2 * ipol], S2_sum = sd[2 * ipol + 1];
              float V = M_fac * (Mprime * (S2_sum / (S1_sum * S1_sum)) - 1);
              if (V > upper[Mprime] || V < lower[Mprime])
                zap = 1;
            }
          }
          if (zap)
          {
            if (omit_outliers)
            {
              *wt = 0;
              for (ipol = 0; ipol < npol; ipol++)
                sd[2 * ipol] = sd[2 * ipol + 1] = 0.0f;
            }
            for (iflag = 0; iflag < nflag; iflag++)
              mask[(ipart * flag_offset + iflag * flag_step) * nchan + ichan] = 1;
          }
        }
    if (flags & 2)
      for (ipart = 0; ipart < npart[r]; ipart++)
      {
        const uint8_t* row = mask + ipart * flag_offset * nchan;
        int zap_ipart = 0;
        for (ipol = 0; ipol < npol; ipol++)
        {
          double sk_avg = 0, M_tot = 0, Msq_tot = 0, chan_count = 0;
          for (ichan = 0; ichan < nchan; ichan++)
            if (chans[ichan] && (!omit_outliers || row[ichan] == 0))
            {
              const float* sd = sums + ((ipart * nchan + ichan) * npol + ipol) * 2;
              float S1_sum = sd[0], S2_sum = sd[1];
              unsigned Mprime = counts[ipart * nchan + ichan];
              const float M_fac = (float) (Mprime * Nd + 1) / (Mprime - 1);
              float V = M_fac * (Mprime * (S2_sum / (S1_sum * S1_sum)) - 1);
              sk_avg += V * Mprime;
              M_tot += Mprime;
              Msq_tot += Mprime * Mprime;
              chan_count++;
            }
          if (chan_count > 0)
          {
            double M_mean, Msq_mean, Deff, effective_count;
            float threshold, up, lo;
            sk_avg /= M_tot;
            M_mean = M_tot / chan_count;
            Msq_mean = Msq_tot / chan_count;
            Deff = Msq_mean / (M_mean * M_mean);
            effective_count = chan_count / Deff;
            threshold = symmetric[(unsigned) (M_mean + 0.5)] / sqrt(effective_count);
            up = 1.0 + threshold;
            lo = 1.0 - threshold;
            if (sk_avg > up || sk_avg < lo)
              zap_ipart = 1;
          }
        }
        if (zap_ipart)
          for (iflag = 0; iflag < nflag; iflag++)
            memset(mask + (ipart * flag_offset + iflag * flag_step) * nchan, 1, nchan);
      }
  }
  if (flags & 4)
  {
    const unsigned L = nres - 1, M = Ms[L], nov = novs[L];
    const float *lower = tab[L], *upper = tab[L] + M + 1;
    const uint64_t np = npart[L];
    uint64_t ipart;
    for (ichan = 0; ichan < nchan; ichan++)
    {
      uint64_t count = 0;
      float lo, up, M_t, M_fac;
      int zap_chan = 0;
      if (!channels[ichan])
        continue;
      for (ipart = 0; ipart < np; ipart += nov)
        count += counts[ipart * nchan + ichan];
      if (count == 0)
        continue;
      if (count <= M)
      {
        lo = lower[count];
        up = upper[count];
      }
      else
      {
        unsigned k;
        for (k = 0; k < ntscr && tscr_totals[k] != count; k++)
          ;
        if (k < ntscr)
        {
          lo = tscr_lower[k];
          up = tscr_upper[k];
        }
        else if (count >= 32768)
          ox_sk_limits_large(count, pcsd, &lo, &up);
        else
          return UINT64_MAX;
      }
      M_t = (float) count;
      M_fac = (M_t * Nd + 1) / (M_t - 1);
      for (ipol = 0; ipol < npol; ipol++)
      {
        double S1 = 0.0, S2 = 0.0, SK;
        for (ipart = 0; ipart < np; ipart += nov)
        {
          uint64_t idx = (ipart * nchan + ichan) * npol + ipol;
          S1 += sums[idx * 2];
          S2 += sums[idx * 2 + 1];
        }
        SK = M_fac * (M_t * (S2 / (S1 * S1)) - 1);
        if (SK > up || SK < lo)
          zap_chan = 1;
      }
      if (zap_chan)
        for (ipart = 0; ipart < np; ipart++)
          mask[ipart * nchan + ichan] = 1;
    }
  }
  for (i = 0; i < npart[0] * nchan; i++)
    nset += mask[i];
  return nset;
}

/* SpectralKurtosis::mask (SpectralKurtosis.C:1480-1554) with several
   resolutions: blocks of the finest (M0 every off0 samples), the output
   starting idat_start = M_last - off_last samples in and holding ndat_out =
   the last resolution's output samples; nskip = idat_start / off0 mask rows
   are skipped and npm = (ndat_out - M0) / off0 + 1 blocks written in order,
   block ipart zeroed in every pol where mask[ipart + nskip][ichan] is set.
   out rows hold ndat samples.  Returns ndat_out. */
uint64_t ox_sk_mask_res(const float* in, unsigned nchan, unsigned npol, uint64_t ndat, unsigned M0, unsigned nov0,
                        unsigned Ml, unsigned novl, const uint8_t* mask, float* out)
{
  const uint64_t off = M0 / nov0, offl = Ml / novl, idat_start = Ml - offl, nskip = idat_start / off;
  uint64_t ndat_out, npm, ipart, j;
  unsigned ichan, ipol;
  ox_sk_blocks(ndat, Ml, novl, &ndat_out);
  if (ndat_out < M0)
    return 0;
  npm = (ndat_out - M0) / off + 1;
  for (ichan = 0; ichan < nchan; ichan++)
    for (ipol = 0; ipol < npol; ipol++)
    {
      const float* x = in + ((uint64_t) ichan * npol + ipol) * ndat * 2 + idat_start * 2;
      float* y = out + ((uint64_t) ichan * npol + ipol) * ndat * 2;
      for (ipart = 0; ipart < npm; ipart++)
        for (j = 0; j < 2 * (uint64_t) M0; j++)
          y[ipart * off * 2 + j] = mask[(ipart + nskip) * nchan + ichan] ? 0.0f : x[ipart * off * 2 + j];
    }
  return ndat_out;
}

/* SpectralKurtosis::mask (SpectralKurtosis.C:1480-1554), one resolution,
   FPT complex input of ndat samples per row, mask of the npart blocks
   [npart][nchan]: the blocks written in order, block ipart (of npm) the M
   samples from ipart * off of the output from input sample idat_start +
   ipart * off, zeroed in every pol when mask[ipart + nskip][ichan] is set
   (off = M / noverlap, idat_start = M - off, nskip = idat_start / off, npm =
   (ndat_out - M) / off + 1).  out rows hold ndat samples (the output's
   ndat_out first).  Returns ndat_out. */
uint64_t ox_sk_mask(const float* in, unsigned nchan, unsigned npol, uint64_t ndat, unsigned M, unsigned noverlap,
                    const uint8_t* mask, uint64_t npart, float* out)
{
  const uint64_t off = M / noverlap, idat_start = M - off, nskip = idat_start / off;
  const uint64_t ndat_out = (npart + 1) * off - M, npm = (ndat_out - M) / off + 1;
  unsigned ichan, ipol;
  uint64_t ipart, j;
  for (ichan = 0; ichan < nchan; ichan++)
    for (ipol = 0; ipol < npol; ipol++)
    {
      const float* x = in + ((uint64_t) ichan * npol + ipol) * ndat * 2 + idat_start * 2;
      float* y = out + ((uint64_t) ichan * npol + ipol) * ndat * 2;
      for (ipart = 0; ipart < npm; ipart++)
        for (j = 0; j < 2 * (uint64_t) M; j++)
          y[ipart * off * 2 + j] = mask[(ipart + nskip) * nchan + ichan] ? 0.0f : x[ipart * off * 2 + j];
    }
  return ndat_out;
}

/* FilterbankEngineCPU::perform (FilterbankEngineCPU.C:184-370), complex
   input, no apodization, passband or zero-DM output: for each input channel,
   FFT part and pol, the nsub*fres-point forward transform of the window at
   ipart*in_step, times the response (resp[ichan_in][nsub*fres] complex, in
   the order the reference's Response holds it after its swaps), then for
   each of the nsub sub-bands the fres-point backward transform of its
   contiguous slice, of which samples [nfilt_pos, nfilt_pos + nkeep) are
   kept (nkeep = fres - nfilt_pos - nfilt_neg).
     in  [nchan][npol][ndat] complex, out [nchan*nsub][npol][npart*nkeep].
   in_step = nsub*nkeep (Filterbank.C:154-192), npart = (ndat - nsub*nfilt)
   / in_step.  Unnormalized transforms; the response carries the scale.
   Returns npart. */
#define OX_DEFINE_FB(T, NAME, FFT)                                             \
  uint64_t NAME(const T* in, unsigned nchan, unsigned npol, uint64_t ndat,     \
                unsigned nsub, unsigned fres, unsigned nfilt_pos,              \
                unsigned nfilt_neg, const T* resp, T* out)                     \
  {                                                                            \
    const uint64_t nfft = (uint64_t) nsub * fres;                              \
    const unsigned nkeep = fres - nfilt_pos - nfilt_neg;                       \
    const uint64_t novl = (uint64_t) (nfilt_pos + nfilt_neg) * nsub;           \
    const uint64_t in_step = nfft - novl;                                      \
    const uint64_t npart = ndat > novl ? (ndat - novl) / in_step : 0;          \
    const uint64_t nout = npart * nkeep;                                       \
    T* buf = (T*) malloc(sizeof(T) * 2 * nfft);                                \
    unsigned ichan, ipol, isub, m;                                             \
    uint64_t ipart, k;                                                         \
    for (ichan = 0; ichan < nchan; ichan++)                                    \
      for (ipart = 0; ipart < npart; ipart++)                                  \
        for (ipol = 0; ipol < npol; ipol++)                                    \
        {                                                                      \
          const T* x = in + (((uint64_t) ichan * npol + ipol) * ndat           \
                             + ipart * in_step) * 2;                           \
          const T* r = resp + (uint64_t) ichan * nfft * 2;                     \
          if (ipart * in_step + nfft > ndat) continue;                         \
          memcpy(buf, x, sizeof(T) * 2 * nfft);                                \
          FFT(nfft, buf, -1);                                                  \
          for (k = 0; k < nfft; k++)                                           \
          {                                                                    \
            T dr = buf[2 * k], di = buf[2 * k + 1];                            \
            T fr_ = r[2 * k], fi = r[2 * k + 1];                               \
            buf[2 * k] = fr_ * dr - fi * di;                                   \
            buf[2 * k + 1] = fi * dr + fr_ * di;                               \
          }                                                                    \
          for (isub = 0; isub < nsub; isub++)                                  \
          {                                                                    \
            T* s = buf + (uint64_t) isub * fres * 2;                           \
            T* o = out + ((((uint64_t) ichan * nsub + isub) * npol + ipol)     \
                          * nout + ipart * nkeep) * 2;                         \
            FFT(fres, s, +1);                                                  \
            for (m = 0; m < nkeep; m++)                                        \
            {                                                                  \
              o[2 * m] = s[2 * (nfilt_pos + m)];                               \
              o[2 * m + 1] = s[2 * (nfilt_pos + m) + 1];                       \
            }                                                                  \
          }                                                                    \
        }                                                                      \
    free(buf);                                                                 \
    return npart;                                                              \
  }

OX_DEFINE_FB(float, ox_filterbank_f32, ox_fft_f32)
OX_DEFINE_FB(double, ox_filterbank_f64, ox_fft_f64)
