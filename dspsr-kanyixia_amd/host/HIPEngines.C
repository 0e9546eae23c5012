// HIPEngines.C -- dsp::HIP engines over the libdspx C ABI.
#include "dsp/HIPEngines.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>

namespace dsp {
namespace HIP {

Context::Context(int device, void* stream)
{
  int rc = dspx_ctx_create(device, stream, &ctx);
  if (rc != DSPX_OK) throw Error(FailedCall, "dsp::HIP::Context", "dspx_ctx_create failed (%d)", rc);
}

Context::~Context() { dspx_ctx_destroy(ctx); }

void Context::synchronize() { check(dspx_ctx_synchronize(ctx), "dsp::HIP::Context::synchronize"); }

void Context::check(int status, const char* where) const
{
  if (status == DSPX_OK) return;
  ErrorCode code = status == DSPX_E_PARAM ? InvalidParam : status == DSPX_E_STATE ? InvalidState : FailedCall;
  throw Error(code, where, "%s", dspx_last_error(ctx));
}

static void require_device(const void* p, bool on_device, const char* where)
{
  if (!on_device) throw Error(InvalidState, where, "buffer is not in device memory");
  (void) p;
}

// ---------------------------------------------------------------- Unpacker

void Unpacker::unpack(const BitSeries* in, TimeSeries* out)
{
  require_device(in->get_rawptr(), in->get_memory()->on_device, "dsp::HIP::Unpacker::unpack");
  out->copy_configuration(in);
  out->set_state(Signal::Analytic);
  out->set_ndim(2);
  out->resize(in->get_ndat());
  require_device(out->get_datptr(), out->get_memory()->on_device, "dsp::HIP::Unpacker::unpack");
  ctx->check(dspx_unpack(ctx->get(), in->get_rawptr(), in->get_nbit(), in->offset_binary ? 1 : 0, in->get_nchan(),
                         in->get_npol(), in->get_ndat(), scale, out->get_datptr(), out->chan_span(),
                         out->get_pol_span()),
             "dsp::HIP::Unpacker::unpack");
}

// ------------------------------------------------------ ConvolutionEngine

ConvolutionEngine::~ConvolutionEngine() { dspx_conv_destroy(conv); }

void ConvolutionEngine::prepare(Convolution* convolution)
{
  parent = convolution;
  create(convolution->get_response());
}

void ConvolutionEngine::create(const Response* response)
{
  if (conv) {
    dspx_conv_destroy(conv);
    conv = nullptr;
  }
  const TimeSeries* in = parent->get_input();
  dspx_conv_desc d{};
  d.nchan = in->get_nchan();
  d.npol = in->get_npol();
  d.nfft = response->get_ndat();  // complex points; 2*nfft real samples of Nyquist input
  d.nyquist = in->get_state() == Signal::Nyquist ? 1u : 0u;
  d.nfilt_pos = parent->get_nfilt_pos();
  d.nfilt_neg = parent->get_nfilt_neg();
  d.form = form;
  auto* dd = dynamic_cast<const Dedispersion*>(response);
  if (dd) {
    // the chirp is evaluated on the device from the band and DM
    d.chirp_mode = DSPX_CHIRP_PARITY;
    d.dedisp = dspx_dedisp_desc{dd->get_centre_frequency(), dd->get_bandwidth(), dd->get_dispersion_measure(),
                                dd->get_nchan(), dd->get_dc_centred() ? 1u : 0u, d.nfft,
                                dd->get_interchan() ? 1u : 0u};
  } else {
    d.chirp_mode = DSPX_CHIRP_TABLE;
  }
  ctx->check(dspx_conv_create(ctx->get(), &d, &conv), "dsp::HIP::ConvolutionEngine::prepare");
}

void ConvolutionEngine::copy_response(const Response* response)
{
  if (dynamic_cast<const Dedispersion*>(response)) {
    // a DM change rebuilds the on-chip chirp parameters
    create(response);
    return;
  }
  const auto& data = response->get_data();
  if (data.size() != size_t(response->get_nchan()) * response->get_ndat() * 2)
    throw Error(InvalidState, "dsp::HIP::ConvolutionEngine::copy_response", "response has no host data");
  ctx->check(dspx_conv_set_response(conv, data.data()), "dsp::HIP::ConvolutionEngine::copy_response");
}

void ConvolutionEngine::perform(const TimeSeries* in, TimeSeries* out, unsigned npart)
{
  perform(in, out, nullptr, npart);
}

void ConvolutionEngine::perform(const TimeSeries* in, TimeSeries* out, TimeSeries* zdm, unsigned npart)
{
  ctx->check(dspx_conv_perform(conv, in->get_datptr(), in->chan_span(), in->get_pol_span(), in->get_ndat(),
                               out->get_datptr(), out->chan_span(), out->get_pol_span(),
                               zdm ? zdm->get_datptr() : nullptr, npart),
             "dsp::HIP::ConvolutionEngine::perform");
}

// ------------------------------------------------------- FilterbankEngine

FilterbankEngine::~FilterbankEngine() { dspx_fb_destroy(fb); }

void FilterbankEngine::setup(Filterbank* filterbank)
{
  parent = filterbank;
  const TimeSeries* in = filterbank->get_input();
  dspx_fb_destroy(fb);
  fb = nullptr;
  desc = dspx_fb_desc{};
  desc.nchan = in->get_nchan();
  desc.npol = in->get_npol();
  desc.nsub = filterbank->get_nchan_subband();
  desc.fres = filterbank->get_freq_res();
  desc.nfilt_pos = filterbank->get_nfilt_pos();
  desc.nfilt_neg = filterbank->get_nfilt_neg();
  ctx->check(dspx_fb_create(ctx->get(), &desc, &fb), "dsp::HIP::FilterbankEngine::setup");
  ctx->check(dspx_fb_set_response(fb, filterbank->get_response_table().data()), "dsp::HIP::FilterbankEngine::setup");
  filterbank->set_response_table_changed(false);
}

void FilterbankEngine::perform(const TimeSeries* in, TimeSeries* out, uint64_t npart, uint64_t in_step,
                               uint64_t out_step)
{
  perform(in, out, nullptr, npart, in_step, out_step);
}

void FilterbankEngine::perform(const TimeSeries* in, TimeSeries* out, TimeSeries* zdm, uint64_t npart,
                               uint64_t in_step, uint64_t out_step)
{
  const char* where = "dsp::HIP::FilterbankEngine::perform";
  require_device(in->get_datptr(), in->get_memory()->on_device, where);
  require_device(out->get_datptr(), out->get_memory()->on_device, where);
  if (zdm) {
    require_device(zdm->get_datptr(), zdm->get_memory()->on_device, where);
    // the library writes the zero-DM output with the output's strides
    if (zdm->chan_span() != out->chan_span() || zdm->get_pol_span() != out->get_pol_span())
      throw Error(InvalidParam, where, "zero-DM output layout differs from the output's");
  }
  const uint64_t nkeep = desc.fres - desc.nfilt_pos - desc.nfilt_neg;
  if (in_step != 2 * uint64_t(desc.nsub) * nkeep || out_step != 2 * nkeep)
    throw Error(InvalidParam, where, "in_step=%llu out_step=%llu do not match nsub=%u nkeep=%llu",
                (unsigned long long) in_step, (unsigned long long) out_step, desc.nsub, (unsigned long long) nkeep);
  if (parent && parent->response_table_changed()) {
    ctx->check(dspx_fb_set_response(fb, parent->get_response_table().data()), where);
    parent->set_response_table_changed(false);
  }
  // the library derives the part count from the samples it is given: hand it
  // exactly the span of npart parts
  const uint64_t nfft = uint64_t(desc.nsub) * desc.fres;
  const uint64_t span = (npart - 1) * (in_step / 2) + nfft;
  if (span > in->get_ndat()) throw Error(InvalidParam, where, "npart=%llu exceeds the input", (unsigned long long) npart);
  if (zdm)
    ctx->check(dspx_fb_perform_zdm(fb, in->get_datptr(), in->chan_span(), in->get_pol_span(), span,
                                   out->get_datptr(), out->chan_span(), out->get_pol_span(), zdm->get_datptr()),
               where);
  else
    ctx->check(dspx_fb_perform(fb, in->get_datptr(), in->chan_span(), in->get_pol_span(), span, out->get_datptr(),
                               out->chan_span(), out->get_pol_span()),
               where);
}

// ------------------------------------------------- SpectralKurtosisEngine

SpectralKurtosisEngine::~SpectralKurtosisEngine() { dspx_sk_destroy(sk); }

void SpectralKurtosisEngine::setup(SpectralKurtosis* s)
{
  const TimeSeries* in = s->get_input();
  dspx_sk_destroy(sk);
  sk = nullptr;
  nchan = in->get_nchan();
  dspx_sk_desc d{nchan, in->get_npol(), s->get_M(), 1u, s->get_omit_outliers() ? 1u : 0u, s->get_noverlap()};
  setup_M = s->get_M();
  tscr_valid = false;
  ctx->check(dspx_sk_create(ctx->get(), &d, &sk), "dsp::HIP::SpectralKurtosisEngine::setup");
  ctx->check(dspx_sk_set_limits(sk, s->get_lower().data(), s->get_upper().data(), s->get_symmetric().data(),
                                s->get_pc_std_devs()),
             "dsp::HIP::SpectralKurtosisEngine::setup");
  const std::vector<uint8_t> chans = s->get_channels(nchan);
  ctx->check(dspx_sk_set_channels(sk, chans.data()), "dsp::HIP::SpectralKurtosisEngine::setup");
}

void SpectralKurtosisEngine::compute(const TimeSeries* in, float* sums, uint32_t* counts, uint64_t npart)
{
  require_device(in->get_datptr(), in->get_memory()->on_device, "dsp::HIP::SpectralKurtosisEngine::compute");
  ctx->check(dspx_sk_compute(sk, in->get_datptr(), in->chan_span(), in->get_pol_span(), npart,
                             in->get_zeroed_data() ? 1 : 0, sums, counts),
             "dsp::HIP::SpectralKurtosisEngine::compute");
}

void SpectralKurtosisEngine::reset_mask(uint8_t* mask, uint64_t npart)
{
  hipStream_t s = static_cast<hipStream_t>(dspx_ctx_stream(ctx->get()));
  Memory::device()->zero(mask, size_t(npart) * nchan, s);
}

void SpectralKurtosisEngine::detect(float* sums, uint32_t* counts, uint64_t npart, unsigned flags, uint8_t* mask)
{
  ctx->check(dspx_sk_detect(sk, sums, counts, npart, flags, mask), "dsp::HIP::SpectralKurtosisEngine::detect");
}

std::vector<uint64_t> SpectralKurtosisEngine::tscr_totals(const uint32_t* counts, uint64_t npart)
{
  std::vector<uint64_t> t(nchan);
  ctx->check(dspx_sk_tscr_totals(sk, counts, npart, t.data()), "dsp::HIP::SpectralKurtosisEngine::tscr_totals");
  return t;
}

void SpectralKurtosisEngine::detect_tscr(const float* sums, const uint32_t* counts, uint64_t npart, uint8_t* mask,
                                         const std::map<uint64_t, std::pair<float, float>>& limits)
{
  const char* where = "dsp::HIP::SpectralKurtosisEngine::detect_tscr";
  if (!tscr_valid || limits != tscr_uploaded) {
    std::vector<uint64_t> tot;
    std::vector<float> lo, up;
    for (const auto& kv : limits) {
      if (kv.first <= setup_M) continue;
      tot.push_back(kv.first);
      lo.push_back(kv.second.first);
      up.push_back(kv.second.second);
    }
    ctx->check(dspx_sk_set_tscr_limits(sk, uint32_t(tot.size()), tot.data(), lo.data(), up.data()), where);
    tscr_uploaded = limits;
    tscr_valid = true;
  }
  ctx->check(dspx_sk_detect_tscr(sk, sums, counts, npart, mask), where);
}

void SpectralKurtosisEngine::mask(const uint8_t* m, const TimeSeries* in, TimeSeries* out, uint64_t npart)
{
  require_device(out->get_datptr(), out->get_memory()->on_device, "dsp::HIP::SpectralKurtosisEngine::mask");
  ctx->check(dspx_sk_mask(sk, m, in->get_datptr(), in->chan_span(), in->get_pol_span(), out->get_datptr(),
                          out->chan_span(), out->get_pol_span(), npart),
             "dsp::HIP::SpectralKurtosisEngine::mask");
}

// ------------------------------------------------------ SampleDelayEngine

SampleDelayEngine::~SampleDelayEngine() { dspx_delay_destroy(delay); }

void SampleDelayEngine::set_delays(unsigned npol, unsigned nchan, const std::vector<int64_t>& applied)
{
  dspx_delay_destroy(delay);
  delay = nullptr;
  if (applied.size() != size_t(nchan) * npol)
    throw Error(InvalidParam, "dsp::HIP::SampleDelayEngine::set_delays", "%zu delays for %u x %u", applied.size(),
                nchan, npol);
  ctx->check(dspx_delay_create(ctx->get(), nchan, npol, applied.data(), &delay),
             "dsp::HIP::SampleDelayEngine::set_delays");
}

void SampleDelayEngine::retard(const TimeSeries* in, TimeSeries* out, uint64_t output_ndat)
{
  const char* where = "dsp::HIP::SampleDelayEngine::retard";
  if (!delay) throw Error(InvalidState, where, "no delays");
  require_device(in->get_datptr(), in->get_memory()->on_device, where);
  require_device(out->get_datptr(), out->get_memory()->on_device, where);
  ctx->check(dspx_delay_retard(delay, in->get_datptr(), in->chan_span(), in->get_pol_span(), in->get_ndim(),
                               output_ndat, out->get_datptr(), out->chan_span(), out->get_pol_span()),
             where);
}

// -------------------------------------------------------- DetectionEngine

void DetectionEngine::square_law(const TimeSeries* in, TimeSeries* out)
{
  uint32_t st = out->get_npol() == 1 ? DSPX_INTENSITY : DSPX_PPQQ;
  ctx->check(dspx_detect(ctx->get(), st, 1, in->get_datptr(), in->chan_span(), in->get_pol_span(), in->get_nchan(),
                         in->get_npol(), in->get_ndat(), out->get_datptr(), out->chan_span(), out->get_pol_span()),
             "dsp::HIP::DetectionEngine::square_law");
}

void DetectionEngine::polarimetry(unsigned ndim, const TimeSeries* in, TimeSeries* out)
{
  uint32_t st = out->get_state() == Signal::Stokes ? DSPX_STOKES : DSPX_COHERENCE;
  ctx->check(dspx_detect(ctx->get(), st, ndim, in->get_datptr(), in->chan_span(), in->get_pol_span(),
                         in->get_nchan(), in->get_npol(), in->get_ndat(), out->get_datptr(), out->chan_span(),
                         out->get_pol_span()),
             "dsp::HIP::DetectionEngine::polarimetry");
}

// ------------------------------------------------------------- FoldEngine

FoldEngine::~FoldEngine()
{
  dspx_fold_destroy(plan);
  if (d_prof) (void) hipFree(d_prof);
  if (d_hits) (void) hipFree(d_hits);
  if (d_nfolded) (void) hipFree(d_nfolded);
}

void FoldEngine::set_nbin(unsigned n)
{
  if (n == nbin) return;
  nbin = n;
  hits_rows = 0;  // reallocated for the new bin count by the next set_bins
}

void FoldEngine::reserve_hits(unsigned rows)
{
  if (rows == hits_rows && d_hits) return;
  if (d_hits) (void) hipFree(d_hits);
  d_hits = nullptr;
  if (hipMalloc(&d_hits, size_t(rows) * nbin * sizeof(uint32_t)) != hipSuccess)
    throw Error(FailedCall, "dsp::HIP::FoldEngine::set_bins", "hipMalloc");
  hits_rows = rows;
  hipStream_t s = static_cast<hipStream_t>(dspx_ctx_stream(ctx->get()));
  if (hipMemsetAsync(d_hits, 0, size_t(rows) * nbin * sizeof(uint32_t), s) != hipSuccess)
    throw Error(FailedCall, "dsp::HIP::FoldEngine::set_bins", "hipMemsetAsync");
}

void FoldEngine::set_bin(uint64_t, double, double)
{
  throw Error(InvalidState, "dsp::HIP::FoldEngine::set_bin", "this engine builds its plan with set_bins");
}

void FoldEngine::set_weights(const uint16_t* w, uint64_t nweights, unsigned ndpw, unsigned widat)
{
  if (w && ndpw)
    weights.assign(w, w + nweights);
  else
    weights.clear();
  ndatperweight = w ? ndpw : 0;
  weight_idat = widat;
}

uint64_t FoldEngine::set_bins(double phi, double pps, uint64_t ndat, uint64_t idat_start)
{
  // the plan covers the input block; only [idat_start, idat_start + ndat) folds
  zeroed = in->get_zeroed_data();
  reserve_hits(zeroed ? in->get_nchan() : 1);  // PhaseSeries::mixable (PhaseSeries.C:402-404)
  uint64_t block = in->get_ndat();
  dspx_fold_chunk chunk{idat_start, ndat, phi, 0, 0};
  dspx_fold_opts opts{weights.empty() ? nullptr : weights.data(), weights.size(), ndatperweight, weight_idat,
                      zeroed ? 1u : 0u, 0};
  if (!plan)
    ctx->check(dspx_fold_create_opts(ctx->get(), block, pps, nbin, &chunk, 1, &opts, &plan),
               "dsp::HIP::FoldEngine::set_bins");
  else
    ctx->check(dspx_fold_replan(plan, block, pps, nbin, &chunk, 1, &opts), "dsp::HIP::FoldEngine::set_bins");
  if (!zeroed) ctx->check(dspx_fold_hits(plan, d_hits), "dsp::HIP::FoldEngine::set_bins");
  ctx->check(dspx_fold_ndat_folded(plan, &ndat_folded), "dsp::HIP::FoldEngine::set_bins");
  return ndat_folded;
}

uint64_t FoldEngine::get_bin_hits(int) { return 0; }

uint64_t FoldEngine::get_ndat_folded() const
{
  if (!zeroed || !d_nfolded) return ndat_folded;
  ctx->synchronize();
  unsigned long long total = 0;
  if (hipMemcpy(&total, d_nfolded, sizeof(total), hipMemcpyDeviceToHost) != hipSuccess)
    throw Error(FailedCall, "dsp::HIP::FoldEngine::get_ndat_folded", "hipMemcpy");
  uint64_t n = total - nfolded_taken;
  nfolded_taken = total;
  return n;
}

void FoldEngine::fold()
{
  size_t need = size_t(in->get_nchan()) * in->get_npol() * nbin * in->get_ndim();
  if (need != prof_floats) {
    if (d_prof) (void) hipFree(d_prof);
    d_prof = nullptr;
    if (hipMalloc(&d_prof, need * sizeof(float)) != hipSuccess)
      throw Error(FailedCall, "dsp::HIP::FoldEngine::fold", "hipMalloc");
    prof_floats = need;
    hipStream_t s = static_cast<hipStream_t>(dspx_ctx_stream(ctx->get()));
    if (hipMemsetAsync(d_prof, 0, prof_floats * sizeof(float), s) != hipSuccess)
      throw Error(FailedCall, "dsp::HIP::FoldEngine::fold", "hipMemsetAsync");
  }
  if (zeroed) {
    if (!d_nfolded) {
      if (hipMalloc(&d_nfolded, sizeof(unsigned long long)) != hipSuccess)
        throw Error(FailedCall, "dsp::HIP::FoldEngine::fold", "hipMalloc");
      hipStream_t s = static_cast<hipStream_t>(dspx_ctx_stream(ctx->get()));
      if (hipMemsetAsync(d_nfolded, 0, sizeof(unsigned long long), s) != hipSuccess)
        throw Error(FailedCall, "dsp::HIP::FoldEngine::fold", "hipMemsetAsync");
      nfolded_taken = 0;
    }
    ctx->check(dspx_fold_execute_zeroed(plan, in->get_datptr(), in->chan_span(), in->get_pol_span(),
                                        in->get_nchan(), in->get_npol(), in->get_ndim(), d_prof, d_hits, hits_rows,
                                        d_nfolded),
               "dsp::HIP::FoldEngine::fold");
  } else {
    ctx->check(dspx_fold_execute(plan, in->get_datptr(), in->chan_span(), in->get_pol_span(), in->get_nchan(),
                                 in->get_npol(), in->get_ndim(), d_prof),
               "dsp::HIP::FoldEngine::fold");
  }
}

void FoldEngine::zero()
{
  hipStream_t s = static_cast<hipStream_t>(dspx_ctx_stream(ctx->get()));
  if (d_prof && hipMemsetAsync(d_prof, 0, prof_floats * sizeof(float), s) != hipSuccess)
    throw Error(FailedCall, "dsp::HIP::FoldEngine::zero", "hipMemsetAsync");
  if (d_hits && hipMemsetAsync(d_hits, 0, size_t(hits_rows) * nbin * sizeof(uint32_t), s) != hipSuccess)
    throw Error(FailedCall, "dsp::HIP::FoldEngine::zero", "hipMemsetAsync");
}

void FoldEngine::synch(PhaseSeries* output)
{
  // TransferPhaseSeriesCUDA D2H (FoldCUDA.cu:121-146); host profiles are FPT
  ctx->synchronize();
  if (d_prof) {
    std::vector<float> h(prof_floats);
    if (hipMemcpy(h.data(), d_prof, prof_floats * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      throw Error(FailedCall, "dsp::HIP::FoldEngine::synch", "hipMemcpy");
    size_t per = size_t(nbin) * output->get_ndim();
    for (unsigned c = 0; c < output->get_nchan(); c++)
      for (unsigned p = 0; p < output->get_npol(); p++)
        std::copy(h.begin() + (size_t(c) * output->get_npol() + p) * per,
                  h.begin() + (size_t(c) * output->get_npol() + p + 1) * per, output->get_datptr(c, p));
  }
  if (d_hits) {
    if (output->get_hits_nchan() != hits_rows) {
      output->set_hits_nchan(hits_rows);
      output->resize_hits();
    }
    if (hipMemcpy(output->get_hits(), d_hits, size_t(hits_rows) * nbin * sizeof(uint32_t), hipMemcpyDeviceToHost) !=
        hipSuccess)
      throw Error(FailedCall, "dsp::HIP::FoldEngine::synch", "hipMemcpy");
  }
}

// ------------------------------------------------------- LoadToFoldEngine

LoadToFoldEngine::~LoadToFoldEngine()
{
  dspx_ltf_destroy(ltf);
  if (d_prof) (void) hipFree(d_prof);
  if (d_hits) (void) hipFree(d_hits);
}

static uint32_t state_code(Signal::State s)
{
  switch (s) {
    case Signal::Intensity: return DSPX_INTENSITY;
    case Signal::PPQQ: return DSPX_PPQQ;
    case Signal::Coherence: return DSPX_COHERENCE;
    case Signal::Stokes: return DSPX_STOKES;
    default: throw Error(InvalidState, "dsp::HIP::LoadToFoldEngine", "invalid output state");
  }
}

void LoadToFoldEngine::prepare(const BitSeries* raw)
{
  dspx_dedisp_desc dd{raw->get_centre_frequency(), raw->get_bandwidth(), raw->get_dispersion_measure(),
                      raw->get_nchan(), raw->get_dc_centred() ? 1u : 0u, cfg.nfft, 0};
  ctx->check(dspx_dedisp_geometry(&dd, &geom), "dsp::HIP::LoadToFoldEngine::prepare");
  dspx_ltf_desc d{};
  d.nbit = raw->get_nbit();
  d.offset_binary = raw->offset_binary ? 1 : 0;
  d.scale = cfg.scale;
  d.nchan = raw->get_nchan();
  d.npol = raw->get_npol();
  d.ndat = raw->get_ndat();
  d.dedisp = dd;
  d.nfft = geom.nfft;
  d.nfilt_pos = geom.impulse_pos;
  d.nfilt_neg = geom.impulse_neg;
  d.chirp_mode = cfg.chirp_mode;
  d.state = state_code(cfg.state);
  d.ndim = cfg.ndim;
  d.nbin = cfg.nbin;
  d.form = cfg.form;
  dspx_ltf_destroy(ltf);
  ltf = nullptr;
  ctx->check(dspx_ltf_create(ctx->get(), &d, &ltf), "dsp::HIP::LoadToFoldEngine::prepare");
  dspx_ltf_geometry(ltf, &npart, &ndat_out);
  // sub-integrations: equal spans of the block, each folded from the phase
  // of its first sample's midpoint (Fold.C:686-693, TimeDivide)
  double rate = raw->get_rate();
  std::vector<dspx_fold_chunk> chunks(cfg.nsub);
  for (unsigned k = 0; k < cfg.nsub; k++) {
    uint64_t i0 = uint64_t(std::rint(double(ndat_out) * k / cfg.nsub));
    uint64_t i1 = uint64_t(std::rint(double(ndat_out) * (k + 1) / cfg.nsub));
    double t = raw->get_start_time() + (double(geom.impulse_pos) + double(i0) + 0.5) / rate;
    double phi = std::fmod(t, cfg.folding_period) / cfg.folding_period - cfg.reference_phase;
    chunks[k] = dspx_fold_chunk{i0, i1 - i0, phi, k, 0};
  }
  ctx->check(dspx_ltf_set_fold(ltf, (1.0 / rate) / cfg.folding_period, chunks.data(), cfg.nsub),
             "dsp::HIP::LoadToFoldEngine::prepare");
  unsigned nval = cfg.state == Signal::Intensity ? 1 : cfg.state == Signal::PPQQ ? raw->get_npol() : 4;
  prof_floats = size_t(cfg.nsub) * raw->get_nchan() * nval * cfg.nbin;
  hits_count = size_t(cfg.nsub) * cfg.nbin;
  if (d_prof) (void) hipFree(d_prof);
  if (d_hits) (void) hipFree(d_hits);
  d_prof = nullptr;
  d_hits = nullptr;
  if (hipMalloc(&d_prof, prof_floats * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_hits, hits_count * sizeof(uint32_t)) != hipSuccess)
    throw Error(FailedCall, "dsp::HIP::LoadToFoldEngine::prepare", "hipMalloc");
  hipStream_t s = static_cast<hipStream_t>(dspx_ctx_stream(ctx->get()));
  if (hipMemsetAsync(d_prof, 0, prof_floats * sizeof(float), s) != hipSuccess ||
      hipMemsetAsync(d_hits, 0, hits_count * sizeof(uint32_t), s) != hipSuccess)
    throw Error(FailedCall, "dsp::HIP::LoadToFoldEngine::prepare", "hipMemsetAsync");
}

void LoadToFoldEngine::operate(const BitSeries* raw)
{
  if (!ltf) prepare(raw);
  ctx->check(dspx_ltf_execute(ltf, raw->get_rawptr(), d_prof, d_hits), "dsp::HIP::LoadToFoldEngine::operate");
}

void LoadToFoldEngine::get_result(std::vector<float>& prof, std::vector<uint32_t>& hits)
{
  ctx->synchronize();
  prof.resize(prof_floats);
  hits.resize(hits_count);
  if (hipMemcpy(prof.data(), d_prof, prof_floats * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hits.data(), d_hits, hits_count * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
    throw Error(FailedCall, "dsp::HIP::LoadToFoldEngine::get_result", "hipMemcpy");
}

}  // namespace HIP
}  // namespace dsp
