// Operations.h -- the Transformations of the hot path and their Engine
// interfaces, restated from the reference's call sites:
//   Convolution + Convolution::Engine  Signal/General/Convolution.C:316-813
//     (engine calls :446-451, :656-667, :690-697; Engine methods as used by
//      ConvolutionCUDA.cu:174-216, 219, 567-595)
//   Filterbank + Filterbank::Engine    Signal/General/Filterbank.C:75-566
//     (engine calls :249, :520-566; FilterbankEngineCPU.C:44-386)
//   SpectralKurtosis + its Engine      Signal/General/SpectralKurtosis.C:263-281,
//     674-909, 1014-1566 (engine calls :438, 701, 1133, 1228, 1338, 1523)
//   Detection + Detection::Engine      Signal/General/Detection.C:225-573
//     (engine calls :237-243, :483-490)
//   Fold + Fold::Engine                Signal/Pulsar/Fold.C:546-990,
//                                      Signal/Pulsar/dsp/Fold.h:262-345
// The CPU branches of these Transformations are not restated here: on this
// tree they are the oracle (oracle/), and the product requires an engine.
#pragma once
#include <map>
#include <utility>

#include "dsp/Types.h"

namespace dsp {

//! Frequency response applied by the Convolution, [nchan][ndat] complex
//! float in FFT-bin order (Response.C, after swap_as_needed)
class Response {
 public:
  virtual ~Response() = default;
  unsigned get_nchan() const { return nchan; }
  unsigned get_ndat() const { return ndat; }
  unsigned get_impulse_pos() const { return impulse_pos; }
  unsigned get_impulse_neg() const { return impulse_neg; }
  //! Response::prepare: fix nchan/ndat/impulse for this input
  virtual void prepare(const Observation* input, unsigned channels = 0) = 0;
  //! host copy of the response (empty when the engine builds it on chip)
  const std::vector<float>& get_data() const { return data; }
  bool has_changed() const { return changed; }
  void set_changed(bool c) { changed = c; }

 protected:
  unsigned nchan = 0, ndat = 0, impulse_pos = 0, impulse_neg = 0;
  std::vector<float> data;
  bool changed = true;
};

//! Coherent dedispersion kernel (Dedispersion.C / PlasmaResponse.C).  The
//! geometry follows the reference rule; the chirp itself is evaluated on the
//! device by the HIP engine (FP64 phase rounded to float, as the reference
//! stores it, Dedispersion.C:104-111).
class Dedispersion : public Response {
 public:
  void set_dispersion_measure(double dm)
  {
    if (dm != dispersion_measure) changed = true;
    dispersion_measure = dm;
  }
  double get_dispersion_measure() const { return dispersion_measure; }
  //! user-set frequency resolution (DedispersionPipeConfig -x)
  void set_frequency_resolution(unsigned n) { user_ndat = n; }
  void prepare(const Observation* input, unsigned channels = 0) override;
  double get_centre_frequency() const { return centre_frequency; }
  double get_bandwidth() const { return bandwidth; }
  bool get_dc_centred() const { return dc_centred; }
  //! Dedispersion::set_sample_delay (Dedispersion.C:61-67): with inter-channel
  //! delays removed by a SampleDelay (-K), each channel's fractional delay
  //! goes into the chirp (Dedispersion.C:142-178)
  void set_interchan(bool on)
  {
    if (on != interchan) changed = true;
    interchan = on;
  }
  bool get_interchan() const { return interchan; }

 protected:
  double dispersion_measure = 0, centre_frequency = 0, bandwidth = 0;
  bool dc_centred = false, interchan = false;
  unsigned user_ndat = 0;
};

//! A response whose table is built elsewhere (a ResponseProduct, a filter
//! designed by the caller): [nchan][ndat] complex float as Response holds it
class TableResponse : public Response {
 public:
  void set_table(unsigned nchan_, unsigned ndat_, unsigned impulse_pos_, unsigned impulse_neg_,
                 std::vector<float> table);
  void prepare(const Observation*, unsigned) override {}
};

class Convolution {
 public:
  class Engine {
   public:
    virtual ~Engine() = default;
    virtual void set_scratch(void*) {}
    virtual void prepare(Convolution* convolution) = 0;
    virtual void copy_response(const Response* response) = 0;
    virtual void perform(const TimeSeries* in, TimeSeries* out, unsigned npart) = 0;
    virtual void perform(const TimeSeries* in, TimeSeries* out, TimeSeries* zero_DM_out, unsigned npart) = 0;
  };

  void set_input(const TimeSeries* in) { input = in; }
  void set_output(TimeSeries* out) { output = out; }
  void set_response(Response* r) { response = r; }
  Response* get_response() const { return response; }
  void set_engine(Engine* e) { engine.reset(e); }
  void set_zero_DM(bool z) { zero_DM = z; }
  void set_zero_DM_output(TimeSeries* z) { zero_DM_output = z; }
  const TimeSeries* get_input() const { return input; }
  TimeSeries* get_output() const { return output; }

  //! Convolution.C:316-455 (Analytic or real-sampled Nyquist input)
  void prepare();
  //! Transformation::operate -> Convolution::transformation (Convolution.C:646-698)
  void operate();

  unsigned get_minimum_samples() const { return nsamp_fft; }
  unsigned get_minimum_samples_lost() const { return nsamp_overlap; }
  unsigned get_nfilt_pos() const { return nfilt_pos; }
  unsigned get_nfilt_neg() const { return nfilt_neg; }
  uint64_t get_npart() const { return npart; }

 protected:
  void prepare_output();
  const TimeSeries* input = nullptr;
  TimeSeries* output = nullptr;
  TimeSeries* zero_DM_output = nullptr;
  Response* response = nullptr;
  std::unique_ptr<Engine> engine;
  bool prepared = false, zero_DM = false;
  unsigned nsamp_fft = 0, nsamp_overlap = 0, nsamp_step = 0, nfilt_pos = 0, nfilt_neg = 0;
  uint64_t npart = 0;
};

//! Convolving filterbank, Analytic input (Signal/General/Filterbank.C:75-200,
//! 405-566).  Each input channel becomes nchan/input_nchan output channels of
//! freq_res bins; the response, when given, is applied on the way.
class Filterbank {
 public:
  //! The interface FilterbankEngineCPU implements (FilterbankEngineCPU.C:44,
  //! 150, 177, 184-191, 379-386) and Filterbank::filterbank calls
  //! (Filterbank.C:520-566)
  class Engine {
   public:
    virtual ~Engine() = default;
    virtual void setup(Filterbank* filterbank) = 0;
    virtual void set_scratch(float*) {}
    //! in_step / out_step are in floats: nsamp_step*ndim and nkeep*2
    virtual void perform(const TimeSeries* in, TimeSeries* out, uint64_t npart, uint64_t in_step,
                         uint64_t out_step) = 0;
    virtual void perform(const TimeSeries* in, TimeSeries* out, TimeSeries* zero_DM_out, uint64_t npart,
                         uint64_t in_step, uint64_t out_step) = 0;
    virtual void finish() {}
  };

  void set_input(const TimeSeries* in) { input = in; }
  void set_output(TimeSeries* out) { output = out; }
  //! total output channels (a multiple of the input's)
  void set_nchan(unsigned n) { nchan = n; }
  //! bins per output channel when no response sets it
  void set_frequency_resolution(unsigned n) { freq_res = n; }
  void set_response(Response* r) { response = r; }
  void set_engine(Engine* e) { engine.reset(e); }
  const TimeSeries* get_input() const { return input; }
  TimeSeries* get_output() const { return output; }
  //! the zero-DM output: the sub-bands without the response, shaped as the
  //! output (Filterbank.C:388-396, 549-555; set by LoadToFold1.C:291-303)
  void set_zero_DM(bool z) { zero_DM = z; }
  void set_zero_DM_output(TimeSeries* z) { zero_DM_output = z; }
  bool get_zero_DM() const { return zero_DM; }

  //! Filterbank::do_prepare (Filterbank.C:75-200)
  void prepare();
  //! Filterbank::transformation + filterbank (Filterbank.C:460-566)
  void operate();

  unsigned get_nchan_subband() const { return nchan_subband; }
  unsigned get_freq_res() const { return freq_res; }
  unsigned get_nfilt_pos() const { return nfilt_pos; }
  unsigned get_nfilt_neg() const { return nfilt_neg; }
  unsigned get_minimum_samples() const { return nsamp_fft; }
  uint64_t get_npart() const { return npart; }
  //! [input nchan][nchan_subband*freq_res] complex float: the response
  //! (unity without one) times the normaliser 1/sqrt(nsamp_fft*freq_res)
  //! that Filterbank.C:176-185 appends to it
  const std::vector<float>& get_response_table() const { return table; }
  bool response_table_changed() const { return table_changed; }
  void set_response_table_changed(bool c) { table_changed = c; }

 protected:
  void build_table();
  const TimeSeries* input = nullptr;
  TimeSeries* output = nullptr;
  Response* response = nullptr;
  std::unique_ptr<Engine> engine;
  std::vector<float> table;
  TimeSeries* zero_DM_output = nullptr;
  bool prepared = false, table_changed = true, zero_DM = false;
  unsigned nchan = 0, nchan_subband = 0, freq_res = 0, nfilt_pos = 0, nfilt_neg = 0;
  unsigned nsamp_fft = 0, nsamp_overlap = 0, nsamp_step = 0;
  uint64_t npart = 0;
};

//! Spectral kurtosis RFI excision at one resolution: block sums,
//! the detectors of SpectralKurtosis::detect and the mask
//! (Signal/General/SpectralKurtosis.C:674-909, 1014-1566).  The thresholds
//! come from SKLimits in the Transformation (SpectralKurtosis.C:800-822,
//! 989-1009) and are handed to the engine as tables by block count.
class SpectralKurtosis {
 public:
  //! The engine calls SpectralKurtosis.C makes (:438, 701, 1133, 1228, 1338,
  //! 1523), with the detectors taking the transformation's thresholds
  class Engine {
   public:
    virtual ~Engine() = default;
    virtual void setup(SpectralKurtosis* sk) = 0;
    //! sums [npart][nchan][npol][2] (S1, S2) and counts [npart][nchan]
    virtual void compute(const TimeSeries* input, float* sums, uint32_t* counts, uint64_t npart) = 0;
    virtual void reset_mask(uint8_t* mask, uint64_t npart) = 0;
    //! detect_ft / detect_fscr / detect_tscr as the flags select, after reset_mask
    virtual void detect(float* sums, uint32_t* counts, uint64_t npart, unsigned flags, uint8_t* mask) = 0;
    //! the per-channel totals that detect_tscr tests (SpectralKurtosis.C:1036-1040)
    virtual std::vector<uint64_t> tscr_totals(const uint32_t* counts, uint64_t npart) = 0;
    //! detect_tscr alone, OR-ed into the mask, with the thresholds of the
    //! totals above M given by total
    virtual void detect_tscr(const float* sums, const uint32_t* counts, uint64_t npart, uint8_t* mask,
                             const std::map<uint64_t, std::pair<float, float>>& limits) = 0;
    virtual void mask(const uint8_t* mask, const TimeSeries* input, TimeSeries* output, uint64_t npart) = 0;
  };

  void set_input(const TimeSeries* in) { input = in; }
  void set_output(TimeSeries* out) { output = out; }
  void set_engine(Engine* e) { engine.reset(e); }
  void set_M(unsigned m) { M = m; }
  unsigned get_M() const { return M; }
  //! blocks of M samples every M / noverlap samples (Resolution::set_noverlap)
  void set_noverlap(unsigned n)
  {
    noverlap = n ? n : 1;
    prepared = false;
  }
  unsigned get_noverlap() const { return noverlap; }
  //! the averaging detectors' channels [start, end) (set_channel_range,
  //! SpectralKurtosis.C:848-851), and ranges [first, last] excluded from them
  //! (Resolution::add_exclude)
  void set_channel_range(unsigned start, unsigned end) { include.emplace_back(start, end - 1); prepared = false; }
  void add_exclude(unsigned first, unsigned last) { exclude.emplace_back(first, last); prepared = false; }
  //! Resolution::get_channels (SpectralKurtosis.C:342-360): 1 where included
  std::vector<uint8_t> get_channels(unsigned nchan) const;
  void set_omit_outliers(bool o) { omit_outliers = o; }
  bool get_omit_outliers() const { return omit_outliers; }
  //! which detectors run (DSPX_SK_SKFB | DSPX_SK_FSCR | DSPX_SK_TSCR)
  void set_detectors(unsigned f) { detectors = f; }
  //! SKLimits thresholds by block count 0..M and the closed-form sigma of
  //! detect_tscr (SKLimits.C:152-160)
  void set_limits(std::vector<float> lower, std::vector<float> upper, std::vector<float> symmetric, double pcsd);
  //! detect_tscr's thresholds (lower, upper) of SKLimits(total) for channel
  //! totals above M (get_limits, SpectralKurtosis.C:989-1012, 1045); operate
  //! throws InvalidState, naming the total, when one in (M, 32768) is missing
  void set_tscr_limits(std::map<uint64_t, std::pair<float, float>> l) { tscr_limits = std::move(l); }
  const std::vector<float>& get_lower() const { return lower; }
  const std::vector<float>& get_upper() const { return upper; }
  const std::vector<float>& get_symmetric() const { return symmetric; }
  double get_pc_std_devs() const { return pc_std_devs; }
  const TimeSeries* get_input() const { return input; }
  //! compute, detect and mask the input blocks (SpectralKurtosis::transformation)
  void operate();
  uint64_t get_npart() const { return npart; }
  //! device arrays of the last operate
  const float* get_sums() const { return sums; }
  const uint32_t* get_counts() const { return counts; }
  const uint8_t* get_zapmask() const { return zapmask; }
  ~SpectralKurtosis();

 protected:
  const TimeSeries* input = nullptr;
  TimeSeries* output = nullptr;
  std::unique_ptr<Engine> engine;
  unsigned M = 128, detectors = 1, noverlap = 1;
  std::vector<std::pair<unsigned, unsigned>> include, exclude;
  bool omit_outliers = false, prepared = false;
  std::vector<float> lower, upper, symmetric;
  std::map<uint64_t, std::pair<float, float>> tscr_limits;
  double pc_std_devs = 0;
  uint64_t npart = 0, capacity = 0;
  float* sums = nullptr;
  uint32_t* counts = nullptr;
  uint8_t* zapmask = nullptr;
};

//! SampleDelay with the dispersion delay function (Signal/General/
//! SampleDelay.C:56-300, DedispersionSampleDelay.C:25-135; one span, every
//! pol alike): removes the integer inter-channel dispersion delays.
class SampleDelay {
 public:
  //! SampleDelay::Engine (SampleDelayCUDA.cu:58-160), given the applied
  //! offsets that set_delays derives from zero_delays and the function
  class Engine {
   public:
    virtual ~Engine() = default;
    virtual void set_delays(unsigned npol, unsigned nchan, const std::vector<int64_t>& applied) = 0;
    virtual void retard(const TimeSeries* in, TimeSeries* out, uint64_t output_ndat) = 0;
  };
  void set_input(const TimeSeries* in) { input = in; }
  void set_output(TimeSeries* out) { output = out; }
  void set_engine(Engine* e) { engine.reset(e); }
  //! SampleDelay::transformation (SampleDelay.C:150-300)
  void operate();
  int64_t get_zero_delay() const { return zero_delay; }
  uint64_t get_total_delay() const { return total_delay; }
  const std::vector<int64_t>& get_applied() const { return applied; }

 protected:
  void build();
  const TimeSeries* input = nullptr;
  TimeSeries* output = nullptr;
  std::unique_ptr<Engine> engine;
  std::vector<int64_t> applied;
  int64_t zero_delay = 0;
  uint64_t total_delay = 0;
  // the observation the delays were built for (Dedispersion::SampleDelay::match)
  double built_cfreq = 0, built_bw = 0, built_dm = -1, built_rate = 0;
  unsigned built_nchan = 0, built_npol = 0;
};

class Detection {
 public:
  class Engine {
   public:
    virtual ~Engine() = default;
    virtual void square_law(const TimeSeries* in, TimeSeries* out) = 0;
    virtual void polarimetry(unsigned ndim, const TimeSeries* in, TimeSeries* out) = 0;
  };
  void set_input(const TimeSeries* in) { input = in; }
  void set_output(TimeSeries* out) { output = out; }
  void set_output_state(Signal::State s) { state = s; }
  Signal::State get_output_state() const { return state; }
  void set_output_ndim(unsigned n) { ndim = n; }
  unsigned get_output_ndim() const { return ndim; }
  void set_engine(Engine* e) { engine.reset(e); }
  //! Detection::transformation: output shape, then square_law / polarimetry
  void operate();

 protected:
  const TimeSeries* input = nullptr;
  TimeSeries* output = nullptr;
  Signal::State state = Signal::Intensity;
  unsigned ndim = 4;
  std::unique_ptr<Engine> engine;
};

class Fold {
 public:
  //! Signal/Pulsar/dsp/Fold.h:262-345
  class Engine {
   public:
    virtual ~Engine() = default;
    virtual void set_input(const TimeSeries* input) { in = input; }
    virtual void set_profiles(PhaseSeries* o) { out = o; }
    virtual void set_nbin(unsigned nbin) = 0;
    virtual void set_bin(uint64_t idat, double ibin, double bins_per_samp) = 0;
    virtual uint64_t set_bins(double phi, double phase_per_sample, uint64_t ndat, uint64_t idat_start) = 0;
    bool use_set_bins = false;
    virtual uint64_t get_bin_hits(int ibin) = 0;
    virtual uint64_t get_ndat_folded() const = 0;
    virtual PhaseSeries* get_profiles() = 0;
    virtual void fold() = 0;
    virtual void synch(PhaseSeries*) = 0;
    virtual void zero() = 0;
    virtual void set_ndat(uint64_t, uint64_t) {}

   protected:
    const TimeSeries* in = nullptr;
    PhaseSeries* out = nullptr;
  };

  void set_input(const TimeSeries* in) { input = in; }
  void set_output(PhaseSeries* out) { output = out; }
  void set_engine(Engine* e) { engine.reset(e); }
  Engine* get_engine() const { return engine.get(); }
  void set_folding_period(double p) { folding_period = p; }
  double get_folding_period() const { return folding_period; }
  void set_nbin(unsigned n) { folding_nbin = n; }
  unsigned get_nbin() const { return folding_nbin; }
  void set_reference_phase(double r) { reference_phase = r; }
  //! fold samples [idat_start, idat_start + ndat_fold) of the input
  void set_limits(uint64_t idat_start_, uint64_t ndat_fold_)
  {
    idat_start = idat_start_;
    ndat_fold = ndat_fold_;
    limits_set = true;
  }
  //! Fold::transformation + Fold::fold (Fold.C:546-779, 862-887)
  void operate();
  //! Fold::get_result: engine->synch into the output (Fold.C:138-150)
  PhaseSeries* get_result();
  //! Fold::get_phi for a constant period (Fold.C:975-982)
  double get_phi(double seconds) const;

 protected:
  const TimeSeries* input = nullptr;
  PhaseSeries* output = nullptr;
  std::unique_ptr<Engine> engine;
  double folding_period = 0, reference_phase = 0;
  unsigned folding_nbin = 0;
  uint64_t idat_start = 0, ndat_fold = 0;
  bool limits_set = false;
  double zeroed_rate = 0;
  unsigned zeroed_nchan = 0;
};

}  // namespace dsp
