// HIPEngines.h -- MI355X engines behind the dsp Transformation interfaces.
//
// Each engine is a thin C++ shell over the C ABI of libdspx (include/dspx.h):
// it holds a dspx handle, converts dsp:: objects into pointers and strides,
// and turns a non-zero dspx status into
//   Error(FailedCall, "dsp::HIP::<Engine>::<method>", dspx_last_error(ctx))
// like the reference's CUDA engines throw on cuFFT / CUDA failures
// (ConvolutionCUDA.cu:105-108).  Attachment mirrors the reference builders:
//   convolution->set_engine (new dsp::HIP::ConvolutionEngine (ctx));
//     (DedispersionPipe.C:413-427)
//   detect->set_engine (new dsp::HIP::DetectionEngine (ctx));
//     (LoadToFold1.C:843-852)
//   fold->set_engine (new dsp::HIP::FoldEngine (ctx));
//     (LoadToFold1.C:969-978)
//   filterbank->set_engine (new dsp::HIP::FilterbankEngine (ctx));
//     (FilterbankConfig.C:115-133)
//   skestimator->set_engine (new dsp::HIP::SpectralKurtosisEngine (ctx));
//     (LoadToFold1.C:341-349)
//   sample_delay->set_engine (new dsp::HIP::SampleDelayEngine (ctx));
//     (DedispersionPipe.C:435-472)
#pragma once
#include "dsp/Operations.h"
#include "dspx.h"

namespace dsp {
namespace HIP {

//! One per pipeline thread: device, stream and shared tables
class Context {
 public:
  explicit Context(int device = 0, void* stream = nullptr);
  ~Context();
  dspx_ctx* get() const { return ctx; }
  void synchronize();
  void check(int status, const char* where) const;

 private:
  dspx_ctx* ctx = nullptr;
};

//! GenericEightBitUnpacker::unpack_on_gpu / SixteenBitUnpacker (device)
class Unpacker {
 public:
  explicit Unpacker(Context* c) : ctx(c) {}
  void set_scale(float s) { scale = s; }
  void unpack(const BitSeries* in, TimeSeries* out);

 private:
  Context* ctx;
  float scale = 1.0f;
};

class ConvolutionEngine : public Convolution::Engine {
 public:
  explicit ConvolutionEngine(Context* c, unsigned form = DSPX_FORM_FUSED) : ctx(c), form(form) {}
  ~ConvolutionEngine() override;
  void prepare(Convolution* convolution) override;
  void copy_response(const Response* response) override;
  void perform(const TimeSeries* in, TimeSeries* out, unsigned npart) override;
  void perform(const TimeSeries* in, TimeSeries* out, TimeSeries* zero_DM_out, unsigned npart) override;

 private:
  void create(const Response* response);
  Context* ctx;
  unsigned form;
  dspx_conv* conv = nullptr;
  Convolution* parent = nullptr;
};

//! Filterbank::Engine over dspx_fb_* (FFT-library form)
class FilterbankEngine : public Filterbank::Engine {
 public:
  explicit FilterbankEngine(Context* c) : ctx(c) {}
  ~FilterbankEngine() override;
  void setup(Filterbank* filterbank) override;
  void perform(const TimeSeries* in, TimeSeries* out, uint64_t npart, uint64_t in_step, uint64_t out_step) override;
  void perform(const TimeSeries* in, TimeSeries* out, TimeSeries* zero_DM_out, uint64_t npart, uint64_t in_step,
               uint64_t out_step) override;

 private:
  Context* ctx;
  Filterbank* parent = nullptr;
  dspx_fb* fb = nullptr;
  dspx_fb_desc desc{};
};

//! SpectralKurtosis::Engine over dspx_sk_* (one resolution, noverlap 1)
class SpectralKurtosisEngine : public SpectralKurtosis::Engine {
 public:
  explicit SpectralKurtosisEngine(Context* c) : ctx(c) {}
  ~SpectralKurtosisEngine() override;
  void setup(SpectralKurtosis* sk) override;
  void compute(const TimeSeries* input, float* sums, uint32_t* counts, uint64_t npart) override;
  void reset_mask(uint8_t* mask, uint64_t npart) override;
  void detect(float* sums, uint32_t* counts, uint64_t npart, unsigned flags, uint8_t* mask) override;
  std::vector<uint64_t> tscr_totals(const uint32_t* counts, uint64_t npart) override;
  void detect_tscr(const float* sums, const uint32_t* counts, uint64_t npart, uint8_t* mask,
                   const std::map<uint64_t, std::pair<float, float>>& limits) override;
  void mask(const uint8_t* mask, const TimeSeries* input, TimeSeries* output, uint64_t npart) override;

 private:
  Context* ctx;
  dspx_sk* sk = nullptr;
  std::map<uint64_t, std::pair<float, float>> tscr_uploaded;  // the table last handed to the library
  bool tscr_valid = false;
  unsigned setup_M = 0;
  unsigned nchan = 0;
};

//! SampleDelay::Engine over dspx_delay_*
class SampleDelayEngine : public SampleDelay::Engine {
 public:
  explicit SampleDelayEngine(Context* c) : ctx(c) {}
  ~SampleDelayEngine() override;
  void set_delays(unsigned npol, unsigned nchan, const std::vector<int64_t>& applied) override;
  void retard(const TimeSeries* in, TimeSeries* out, uint64_t output_ndat) override;

 private:
  Context* ctx;
  dspx_delay* delay = nullptr;
};

class DetectionEngine : public Detection::Engine {
 public:
  explicit DetectionEngine(Context* c) : ctx(c) {}
  void square_law(const TimeSeries* in, TimeSeries* out) override;
  void polarimetry(unsigned ndim, const TimeSeries* in, TimeSeries* out) override;
  //! the output state square_law forms (Intensity or PPQQ)
  void set_state(Signal::State s) { state = s; }

 private:
  Context* ctx;
  Signal::State state = Signal::Intensity;
  friend class dsp::Detection;
};

//! Fold::Engine with the bin plan evaluated on the device (use_set_bins)
class FoldEngine : public Fold::Engine {
 public:
  explicit FoldEngine(Context* c) : ctx(c) { use_set_bins = true; }
  ~FoldEngine() override;
  void set_nbin(unsigned nbin) override;
  void set_bin(uint64_t idat, double ibin, double bins_per_samp) override;
  //! Replans the engine's fold for this block (no allocation once the
  //! buffers fit, no wait for the stream) and adds the plan's hits into the
  //! device totals; returns Fold::fold's ndat_folded (0 with zeroed data).
  uint64_t set_bins(double phi, double phase_per_sample, uint64_t ndat, uint64_t idat_start) override;
  //! 0: the hits are accumulated on the device and handed over by synch.
  uint64_t get_bin_hits(int ibin) override;
  //! With zeroed data: the samples counted on the device since the last
  //! call (waits for the stream); otherwise the last set_bins count.
  uint64_t get_ndat_folded() const override;
  PhaseSeries* get_profiles() override { return out; }
  void fold() override;
  //! Copies the profiles and the accumulated hits ([hits_nchan][nbin]) into
  //! the output (FoldCUDA.cu:121-146).
  void synch(PhaseSeries* output) override;
  void zero() override;
  //! Weights of the next set_bins' block (WeightedTimeSeries), or none.
  void set_weights(const uint16_t* w, uint64_t nweights, unsigned ndatperweight, unsigned weight_idat);

 private:
  void reserve_hits(unsigned hits_nchan);
  Context* ctx;
  unsigned nbin = 0;
  dspx_fold* plan = nullptr;
  uint64_t ndat_folded = 0;
  bool zeroed = false;
  std::vector<uint16_t> weights;
  unsigned ndatperweight = 0, weight_idat = 0;
  float* d_prof = nullptr;  // [nchan][npol][nbin*ndim] accumulated on the device
  size_t prof_floats = 0;
  uint32_t* d_hits = nullptr;  // [hits_nchan][nbin] accumulated on the device
  unsigned hits_rows = 0;
  unsigned long long* d_nfolded = nullptr;  // zeroed data: samples counted by the device
  mutable unsigned long long nfolded_taken = 0;
};

//! The LoadToFold chain Unpack -> Convolution -> Detection -> Fold fused
//! into three kernels (or the plain form), for one block of raw samples.
class LoadToFoldEngine {
 public:
  struct Config {
    double folding_period = 0, reference_phase = 0;
    unsigned nbin = 1024, nsub = 1;
    Signal::State state = Signal::Intensity;
    unsigned ndim = 1;
    unsigned form = DSPX_FORM_FUSED;
    unsigned chirp_mode = DSPX_CHIRP_PARITY;
    float scale = 1.0f;
    unsigned nfft = 0;  // 0: reference rule
  };
  LoadToFoldEngine(Context* c, const Config& cfg) : ctx(c), cfg(cfg) {}
  ~LoadToFoldEngine();
  //! build for the shape of this raw block (dedispersion of its band at its DM)
  void prepare(const BitSeries* raw);
  //! fold one block; profiles accumulate in the engine
  void operate(const BitSeries* raw);
  //! copy the profiles [nsub][nchan][npol_out][nbin*ndim] and hits [nsub][nbin] to the host
  void get_result(std::vector<float>& prof, std::vector<uint32_t>& hits);
  uint64_t get_ndat_out() const { return ndat_out; }
  const dspx_geometry& get_geometry() const { return geom; }

 private:
  Context* ctx;
  Config cfg;
  dspx_ltf* ltf = nullptr;
  dspx_geometry geom{};
  uint64_t npart = 0, ndat_out = 0;
  float* d_prof = nullptr;
  uint32_t* d_hits = nullptr;
  size_t prof_floats = 0, hits_count = 0;
};

}  // namespace HIP
}  // namespace dsp
