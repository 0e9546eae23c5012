// Operations.C -- data model and Transformation logic of the hot path
// (engine branches only; see dsp/Operations.h).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>

#include "dsp/Operations.h"
#include "dspx.h"

namespace dsp {

// ------------------------------------------------------------------ Memory

static Memory g_host{false};
static Memory g_device{true};
Memory* Memory::host() { return &g_host; }
Memory* Memory::device() { return &g_device; }

void* Memory::allocate(size_t bytes)
{
  if (!bytes) return nullptr;
  void* p = nullptr;
  if (on_device) {
    if (hipMalloc(&p, bytes) != hipSuccess) throw Error(FailedCall, "dsp::Memory::allocate", "hipMalloc(%zu)", bytes);
  } else {
    p = std::malloc(bytes);
    if (!p) throw Error(FailedCall, "dsp::Memory::allocate", "malloc(%zu)", bytes);
  }
  return p;
}

void Memory::free(void* p)
{
  if (!p) return;
  if (on_device)
    (void) hipFree(p);
  else
    std::free(p);
}

void Memory::zero(void* p, size_t bytes, void* stream)
{
  if (!p || !bytes) return;
  if (on_device) {
    if (hipMemsetAsync(p, 0, bytes, static_cast<hipStream_t>(stream)) != hipSuccess)
      throw Error(FailedCall, "dsp::Memory::zero", "hipMemsetAsync");
  } else {
    std::memset(p, 0, bytes);
  }
}

void TimeSeries::release()
{
  memory->free(buffer);
  buffer = nullptr;
  capacity = 0;
}

void TimeSeries::resize(uint64_t n)
{
  uint64_t need = n * ndim;
  if (need > capacity || uint64_t(npol) * nchan * need > uint64_t(npol) * nchan * capacity) {
    release();
    buffer = static_cast<float*>(memory->allocate(size_t(nchan) * npol * need * sizeof(float)));
    capacity = need;
  }
  pol_span = need;
  ndat = n;
}

void TimeSeries::zero() { memory->zero(buffer, size_t(nchan) * npol * pol_span * sizeof(float)); }

BitSeries::~BitSeries() { memory->free(data); }

void BitSeries::resize(uint64_t n, unsigned nb)
{
  memory->free(data);
  ndat = n;
  nbit = nb;
  data = static_cast<unsigned char*>(memory->allocate(size_t(n) * nchan * npol * 2 * nbit / 8));
}

// ----------------------------------------------------------- Dedispersion

void Dedispersion::prepare(const Observation* input, unsigned channels)
{
  unsigned n = channels ? channels : input->get_nchan();
  if (n != nchan || input->get_centre_frequency() != centre_frequency || input->get_bandwidth() != bandwidth ||
      input->get_dc_centred() != dc_centred)
    changed = true;
  nchan = n;
  centre_frequency = input->get_centre_frequency();
  bandwidth = input->get_bandwidth();
  dc_centred = input->get_dc_centred();
  dspx_dedisp_desc d{centre_frequency, bandwidth, dispersion_measure, nchan, dc_centred ? 1u : 0u, user_ndat, 0};
  dspx_geometry g;
  int rc = dspx_dedisp_geometry(&d, &g);
  if (rc == DSPX_E_STATE)
    throw Error(InvalidState, "dsp::PlasmaResponse::prepare", "smearing exceeds the threshold in every channel");
  if (rc != DSPX_OK) throw Error(InvalidParam, "dsp::Dedispersion::prepare", "invalid band or DM");
  impulse_neg = g.impulse_neg;
  impulse_pos = g.impulse_pos;
  ndat = g.nfft;
}

void TableResponse::set_table(unsigned nchan_, unsigned ndat_, unsigned impulse_pos_, unsigned impulse_neg_,
                              std::vector<float> t)
{
  if (t.size() != size_t(nchan_) * ndat_ * 2)
    throw Error(InvalidParam, "dsp::TableResponse::set_table", "table of %zu floats for nchan=%u ndat=%u", t.size(),
                nchan_, ndat_);
  nchan = nchan_;
  ndat = ndat_;
  impulse_pos = impulse_pos_;
  impulse_neg = impulse_neg_;
  data = std::move(t);
  changed = true;
}

// ------------------------------------------------------------- Filterbank

void Filterbank::prepare()
{
  if (prepared) return;
  if (!engine) throw Error(InvalidState, "dsp::Filterbank::prepare", "no engine (the CPU branch is the oracle)");
  if (input->get_state() != Signal::Analytic)
    throw Error(InvalidState, "dsp::Filterbank::prepare", "only Analytic (complex) input is on this path");
  const unsigned in_nchan = input->get_nchan();
  if (nchan < in_nchan)
    throw Error(InvalidState, "dsp::Filterbank::prepare", "output nchan=%d < input nchan=%d", nchan, in_nchan);
  if (nchan % in_nchan)
    throw Error(InvalidState, "dsp::Filterbank::prepare", "output nchan=%d not a multiple of input nchan=%d", nchan,
                in_nchan);
  nchan_subband = nchan / in_nchan;
  nfilt_pos = nfilt_neg = 0;
  if (response) {
    response->prepare(input, nchan);
    if (response->get_nchan() != nchan)
      throw Error(InvalidState, "dsp::Filterbank::prepare", "response nchan=%d != output nchan=%d",
                  response->get_nchan(), nchan);
    nfilt_pos = response->get_impulse_pos();
    nfilt_neg = response->get_impulse_neg();
    freq_res = response->get_ndat();
  }
  if (freq_res == 0) throw Error(InvalidState, "dsp::Filterbank::prepare", "Response.ndat = 0");
  if (nfilt_pos + nfilt_neg >= freq_res)
    throw Error(InvalidState, "dsp::Filterbank::prepare", "nfilt=%u >= freq_res=%u", nfilt_pos + nfilt_neg,
                freq_res);
  nsamp_fft = nchan_subband * freq_res;
  nsamp_overlap = (nfilt_pos + nfilt_neg) * nchan_subband;
  nsamp_step = nsamp_fft - nsamp_overlap;
  build_table();
  engine->setup(this);
  prepared = true;
}

void Filterbank::build_table()
{
  // ScalarFilter 1/sqrt(nsamp_fft*freq_res) appended to the response for the
  // unnormalised transforms (Filterbank.C:176-185, Convolution.C:300-313)
  const float norm = float(1.0 / std::sqrt(double(nsamp_fft) * double(freq_res)));
  const size_t n = size_t(nchan) * freq_res;  // complex values
  table.assign(2 * n, 0.f);
  if (response) {
    const auto& d = response->get_data();
    if (d.size() != 2 * n)
      throw Error(InvalidState, "dsp::Filterbank::prepare", "response has no host table (%zu floats, need %zu)",
                  d.size(), 2 * n);
    for (size_t i = 0; i < 2 * n; i++) table[i] = d[i] * norm;
    response->set_changed(false);
  } else {
    for (size_t i = 0; i < n; i++) table[2 * i] = norm;
  }
  table_changed = true;
}

void Filterbank::operate()
{
  if (!prepared) prepare();
  if (response && response->has_changed()) build_table();
  const uint64_t ndat = input->get_ndat();
  npart = ndat > nsamp_overlap ? (ndat - nsamp_overlap) / nsamp_step : 0;  // Filterbank.C:418-424
  const unsigned nkeep = freq_res - nfilt_pos - nfilt_neg;
  output->copy_configuration(input);
  output->set_nchan(nchan);
  output->set_state(Signal::Analytic);
  output->set_ndim(2);
  output->set_rate(input->get_rate() / nchan_subband);  // Filterbank.C:356
  output->resize(npart * nkeep);
  output->change_start_time(nfilt_pos);  // Filterbank.C:387
  if (zero_DM) {  // Filterbank.C:388-396
    if (!zero_DM_output) throw Error(InvalidState, "dsp::Filterbank::operate", "zero_DM without output");
    zero_DM_output->copy_configuration(output);
    zero_DM_output->resize(output->get_ndat());
  }
  if (!npart) return;
  if (zero_DM)
    engine->perform(input, output, zero_DM_output, npart, uint64_t(nsamp_step) * 2, uint64_t(nkeep) * 2);
  else
    engine->perform(input, output, npart, uint64_t(nsamp_step) * 2, uint64_t(nkeep) * 2);
  engine->finish();
}

// ----------------------------------------------------- SpectralKurtosis

void SpectralKurtosis::set_limits(std::vector<float> lo, std::vector<float> hi, std::vector<float> sym, double pcsd)
{
  if (lo.size() != size_t(M) + 1 || hi.size() != size_t(M) + 1 || sym.size() != size_t(M) + 1)
    throw Error(InvalidParam, "dsp::SpectralKurtosis::set_limits", "tables need M+1=%u entries", M + 1);
  lower = std::move(lo);
  upper = std::move(hi);
  symmetric = std::move(sym);
  pc_std_devs = pcsd;
  prepared = false;
}

std::vector<uint8_t> SpectralKurtosis::get_channels(unsigned nchan) const
{
  std::vector<uint8_t> c(nchan, include.empty() ? 1 : 0);
  auto set = [&](const std::vector<std::pair<unsigned, unsigned>>& r, uint8_t v) {
    for (const auto& p : r) {
      if (p.second >= nchan || p.first > p.second)
        throw Error(InvalidParam, "dsp::SpectralKurtosis::Resolution::get_channels", "range [%u, %u] of %u channels",
                    p.first, p.second, nchan);
      for (unsigned i = p.first; i <= p.second; i++) c[i] = v;
    }
  };
  set(include, 1);
  set(exclude, 0);
  return c;
}

SpectralKurtosis::~SpectralKurtosis()
{
  Memory::device()->free(sums);
  Memory::device()->free(counts);
  Memory::device()->free(zapmask);
}

void SpectralKurtosis::operate()
{
  if (!engine) throw Error(InvalidState, "dsp::SpectralKurtosis::prepare", "no engine");
  if (input->get_state() != Signal::Analytic)
    throw Error(InvalidState, "dsp::SpectralKurtosis::prepare", "only Analytic (complex) input is on this path");
  if (lower.empty()) throw Error(InvalidState, "dsp::SpectralKurtosis::prepare", "thresholds not set");
  if (!prepared) {
    engine->setup(this);
    prepared = true;
  }
  // Resolution::prepare (SpectralKurtosis.C:263-281)
  if (M % noverlap)
    throw Error(InvalidState, "dsp::SpectralKurtosis::Resolution::prepare", "noverlap=%u does not divide M=%u",
                noverlap, M);
  const uint64_t off = M / noverlap, idat_start = M - off, ndat = input->get_ndat();
  npart = ndat < idat_start ? 0 : (ndat - idat_start) / off;
  if (npart * off < idat_start) npart = 0;
  const uint64_t ndat_out = npart ? npart * off - idat_start : 0;
  if (npart > capacity) {
    Memory* dev = Memory::device();
    dev->free(sums);
    dev->free(counts);
    dev->free(zapmask);
    const size_t nc = size_t(npart) * input->get_nchan();
    sums = static_cast<float*>(dev->allocate(nc * input->get_npol() * 2 * sizeof(float)));
    counts = static_cast<uint32_t*>(dev->allocate(nc * sizeof(uint32_t)));
    zapmask = static_cast<uint8_t*>(dev->allocate(nc));
    capacity = npart;
  }
  output->copy_configuration(input);
  output->resize(ndat_out);  // output_ndat = Nblock * off - idat_start
  output->set_zeroed_data(true);  // zapped samples are zeroed (SpectralKurtosis.C:1468)
  if (!npart) return;
  engine->compute(input, sums, counts, npart);
  engine->detect(sums, counts, npart, detectors & ~unsigned(DSPX_SK_TSCR), zapmask);
  if (detectors & DSPX_SK_TSCR) {
    // detect_tscr tests every total with SKLimits(total) (SpectralKurtosis.C:1045)
    for (uint64_t t : engine->tscr_totals(counts, npart))
      if (t > M && t < 32768 && !tscr_limits.count(t))
        throw Error(InvalidState, "dsp::SpectralKurtosis::detect_tscr", "no thresholds for the total count %llu",
                    (unsigned long long) t);
    engine->detect_tscr(sums, counts, npart, zapmask, tscr_limits);
  }
  engine->mask(zapmask, input, output, npart);
}

// ------------------------------------------------------------ SampleDelay

void SampleDelay::build()
{
  const unsigned nchan = input->get_nchan();
  dspx_dedisp_desc d{input->get_centre_frequency(), input->get_bandwidth(), input->get_dispersion_measure(),
                     nchan, input->get_dc_centred() ? 1u : 0u, 0, 1};
  std::vector<int64_t> a(nchan);
  int rc = dspx_sample_delays(&d, nullptr, nullptr, a.data(), &zero_delay, &total_delay);
  if (rc == DSPX_E_STATE)
    throw Error(InvalidState, "dsp::SampleDelay::build", "a channel's delay is negative");
  if (rc != DSPX_OK) throw Error(InvalidParam, "dsp::Dedispersion::SampleDelay::match", "invalid input");
  const unsigned npol = input->get_npol();
  applied.assign(size_t(nchan) * npol, 0);
  for (unsigned c = 0; c < nchan; c++)
    for (unsigned p = 0; p < npol; p++) applied[size_t(c) * npol + p] = a[c];
  engine->set_delays(npol, nchan, applied);
  built_cfreq = input->get_centre_frequency();
  built_bw = input->get_bandwidth();
  built_dm = input->get_dispersion_measure();
  built_rate = input->get_rate();
  built_nchan = nchan;
  built_npol = npol;
}

void SampleDelay::operate()
{
  if (!engine) throw Error(InvalidState, "dsp::SampleDelay::transformation", "no engine");
  if (built_dm != input->get_dispersion_measure() || built_cfreq != input->get_centre_frequency() ||
      built_bw != input->get_bandwidth() || built_rate != input->get_rate() || built_nchan != input->get_nchan() ||
      built_npol != input->get_npol())
    build();  // Dedispersion::SampleDelay::match (DedispersionSampleDelay.C:25-80)
  const uint64_t ndat = input->get_ndat();
  const uint64_t output_ndat = ndat < total_delay ? 0 : ndat - total_delay;  // SampleDelay.C:168-182
  output->copy_configuration(input);
  output->resize(output_ndat);
  output->change_start_time(zero_delay);  // SampleDelay.C:221
  if (output_ndat) engine->retard(input, output, output_ndat);
}

// ------------------------------------------------------------ Convolution

void Convolution::prepare()
{
  if (prepared) return;
  if (!response) throw Error(InvalidState, "dsp::Convolution::prepare", "no frequency response");
  if (input->get_detected()) throw Error(InvalidState, "dsp::Convolution::prepare", "input data are detected");
  const Signal::State state = input->get_state();
  if (state != Signal::Analytic && state != Signal::Nyquist)
    throw Error(InvalidState, "dsp::Convolution::prepare", "Cannot transform Signal::State");  // Convolution.C:385-400
  if (state == Signal::Nyquist && zero_DM)
    throw Error(InvalidState, "dsp::Convolution::prepare", "no zero-DM output of real-sampled input on this path");
  if (!engine) throw Error(InvalidState, "dsp::Convolution::prepare", "no engine (the CPU branch is the oracle)");
  response->prepare(input);
  if (response->get_ndat() < 2)
    throw Error(InvalidState, "dsp::Convolution::prepare", "invalid response size = %u", response->get_ndat());
  if (response->get_nchan() != input->get_nchan())
    throw Error(InvalidState, "dsp::Convolution::prepare", "invalid response nsub=%d != nchan=%d",
                response->get_nchan(), input->get_nchan());
  nfilt_pos = response->get_impulse_pos();
  nfilt_neg = response->get_impulse_neg();
  // real-sampled input: 2 real samples per complex point of the response
  // (Convolution.C:385-395)
  const unsigned nreal = state == Signal::Nyquist ? 2 : 1;
  nsamp_fft = nreal * response->get_ndat();
  nsamp_overlap = nreal * (nfilt_pos + nfilt_neg);
  if (nsamp_fft < nsamp_overlap)
    throw Error(InvalidState, "dsp::Convolution::prepare", "error nfft=%d < nfilt=%d", nsamp_fft, nsamp_overlap);
  prepare_output();
  engine->prepare(this);
  prepared = true;
}

void Convolution::prepare_output()
{
  nsamp_step = nsamp_fft - nsamp_overlap;
  uint64_t ndat = input->get_ndat();
  npart = 0;
  if (ndat >= nsamp_fft) npart = (ndat - nsamp_overlap) / nsamp_step;  // Convolution.C:485-487
  output->copy_configuration(input);
  output->set_state(Signal::Analytic);
  output->set_ndim(2);
  if (input->get_state() == Signal::Nyquist) output->set_rate(0.5 * input->get_rate());  // Convolution.C:515-517
}

void Convolution::operate()
{
  if (!prepared) prepare();
  if (response->has_changed()) {  // Response::match -> Engine::copy_response
    engine->copy_response(response);
    response->set_changed(false);
  }
  prepare_output();
  uint64_t output_ndat = npart * nsamp_step;  // Convolution::reserve (:601-611)
  if (input->get_state() == Signal::Nyquist) output_ndat /= 2;
  output->resize(output_ndat);
  output->set_start_time(input->get_start_time());
  output->change_start_time(nfilt_pos);
  if (zero_DM) {
    if (!zero_DM_output) throw Error(InvalidState, "dsp::Convolution::operate", "zero_DM without output");
    zero_DM_output->copy_configuration(output);
    zero_DM_output->resize(output_ndat);
    engine->perform(input, output, zero_DM_output, unsigned(npart));
  } else {
    engine->perform(input, output, unsigned(npart));
  }
}

// -------------------------------------------------------------- Detection

void Detection::operate()
{
  if (!engine) throw Error(InvalidState, "dsp::Detection::operate", "no engine");
  if (input->get_ndat() == 0) return;
  unsigned npol = input->get_npol();
  output->copy_configuration(input);
  output->set_state(state);
  if (state == Signal::Intensity) {
    output->set_npol(1);
    output->set_ndim(1);
  } else if (state == Signal::PPQQ) {
    output->set_npol(npol);
    output->set_ndim(1);
  } else if (state == Signal::Coherence || state == Signal::Stokes) {
    if (npol != 2)
      throw Error(InvalidState, "dsp::Detection::checks", "invalid npol=%d for %s formation", npol,
                  state == Signal::Stokes ? "Stokes" : "Coherence");
    if (!(ndim == 1 || ndim == 2 || ndim == 4))
      throw Error(InvalidState, "dsp::Detection::checks", "invalid ndim=%d", ndim);
    output->set_npol(4 / ndim);
    output->set_ndim(ndim);
  } else {
    throw Error(InvalidState, "dsp::Detection::operate", "unsupported output state");
  }
  output->resize(input->get_ndat());
  if (state == Signal::Intensity || state == Signal::PPQQ)
    engine->square_law(input, output);
  else
    engine->polarimetry(ndim, input, output);
}

// ------------------------------------------------------------------- Fold

double Fold::get_phi(double seconds) const
{
  return std::fmod(seconds, folding_period) / folding_period - reference_phase;
}

void Fold::operate()
{
  if (input->get_ndat() == 0) return;
  if (folding_period <= 0) throw Error(InvalidState, "dsp::Fold::transformation", "no folding period");
  if (!folding_nbin) throw Error(InvalidState, "dsp::Fold::fold", "nbin not set");
  if (!engine) throw Error(InvalidState, "dsp::Fold::fold", "no engine");
  if (!limits_set) {  // Fold::set_limits (Fold.C:993-997)
    idat_start = 0;
    ndat_fold = input->get_ndat();
  }
  uint64_t idat_end = idat_start + ndat_fold;
  if (idat_end > input->get_ndat())
    throw Error(InvalidParam, "dsp::Fold:fold", "idat_start=%llu + ndat_fold=%llu > ndat=%llu",
                (unsigned long long) idat_start, (unsigned long long) ndat_fold,
                (unsigned long long) input->get_ndat());
  if (output->get_nbin() != folding_nbin || output->get_nchan() != input->get_nchan() ||
      output->get_npol() != input->get_npol() || output->get_ndim() != input->get_ndim()) {
    output->copy_configuration(input);
    output->resize(folding_nbin);
    output->zero();
    // zeroed data: hits per channel (PhaseSeries::mixable, PhaseSeries.C:402-404)
    output->set_hits_nchan(input->get_zeroed_data() ? input->get_nchan() : 1);
    output->resize_hits();
    engine->zero();
  }
  // midpoint of the first sample (Fold.C:686-693)
  double start = input->get_start_time() + (double(idat_start) + 0.5) / input->get_rate();
  double phi = get_phi(start);
  double pps = (1.0 / input->get_rate()) / folding_period;
  engine->set_input(input);
  engine->set_profiles(output);
  engine->set_nbin(folding_nbin);
  engine->set_ndat(ndat_fold, idat_start);
  uint64_t ndat_folded = engine->set_bins(phi, pps, ndat_fold, idat_start);
  unsigned* hits = output->get_hits();
  for (unsigned ibin = 0; ibin < folding_nbin; ibin++) hits[ibin] += unsigned(engine->get_bin_hits(int(ibin)));
  // with zeroed data ndat_folded is 0 here; the samples the engine counts
  // are added by get_result (Fold.C:862-887, 931-942)
  output->integration_length += double(ndat_folded) / input->get_rate();
  output->ndat_total += ndat_fold;
  output->folding_period = folding_period;
  engine->fold();
  if (input->get_zeroed_data()) {
    zeroed_rate = input->get_rate();
    zeroed_nchan = input->get_nchan();
  }
}

PhaseSeries* Fold::get_result()
{
  if (engine) {
    engine->synch(output);
    // zeroed data: every channel's non-zero samples were counted, so the time
    // folded is their number over nchan (Fold.C:931-942)
    if (zeroed_nchan) {
      output->integration_length += double(engine->get_ndat_folded()) / (zeroed_rate * zeroed_nchan);
      zeroed_nchan = 0;
    }
  }
  return output;
}

}  // namespace dsp
