// Types.h -- the parts of the dspsr data model that the HIP engines touch,
// restated so that the engines compile and run without the dspsr tree.
//
// Names, argument meanings and layouts follow the reference:
//   dsp::Observation    Kernel/Classes/Observation.C (get_centre_frequency :521-535)
//   dsp::TimeSeries     Kernel/Classes/TimeSeries.h (get_datptr(ichan, ipol), FPT order)
//   dsp::PhaseSeries    Signal/Pulsar/dsp/PhaseSeries.h (nbin = ndat, get_hits(ichan))
//   Error               PSRCHIVE Error(InvalidState|InvalidParam|FailedCall, "Class::method", ...)
// Only what the hot path needs is here; MJD start times are seconds relative
// to the folding reference epoch (PSRCHIVE's MJD class is not part of this tree).
#pragma once
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

enum ErrorCode { InvalidParam, InvalidState, FailedCall };

class Error : public std::exception {
 public:
  Error(ErrorCode c, const std::string& func, const char* fmt = "", ...) : code(c), where(func)
  {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    message = buf;
    text = where + ": " + message;
  }
  Error& operator+=(const std::string& ctx)
  {
    text += "\n\tcalled from " + ctx;
    return *this;
  }
  const char* what() const noexcept override { return text.c_str(); }
  ErrorCode get_code() const { return code; }
  ErrorCode code;
  std::string where, message, text;
};

namespace dsp {

namespace Signal {
enum State { Nyquist, Analytic, Intensity, PPQQ, Coherence, Stokes };
}

//! Where a buffer lives (dsp::Memory / CUDA::DeviceMemory in the reference)
struct Memory {
  bool on_device = false;
  void* allocate(size_t bytes);
  void free(void* p);
  void zero(void* p, size_t bytes, void* stream = nullptr);
  static Memory* host();
  static Memory* device();
};

class Observation {
 public:
  virtual ~Observation() = default;
  unsigned get_nchan() const { return nchan; }
  void set_nchan(unsigned n) { nchan = n; }
  unsigned get_npol() const { return npol; }
  void set_npol(unsigned n) { npol = n; }
  unsigned get_ndim() const { return ndim; }
  void set_ndim(unsigned n) { ndim = n; }
  Signal::State get_state() const { return state; }
  void set_state(Signal::State s) { state = s; }
  bool get_detected() const { return state != Signal::Nyquist && state != Signal::Analytic; }
  double get_centre_frequency() const { return centre_frequency; }
  void set_centre_frequency(double f) { centre_frequency = f; }
  double get_bandwidth() const { return bandwidth; }
  void set_bandwidth(double b) { bandwidth = b; }
  double get_rate() const { return rate; }
  void set_rate(double r) { rate = r; }
  double get_dispersion_measure() const { return dispersion_measure; }
  void set_dispersion_measure(double d) { dispersion_measure = d; }
  bool get_dc_centred() const { return dc_centred; }
  void set_dc_centred(bool d) { dc_centred = d; }
  //! seconds from the folding reference epoch to the first sample
  double get_start_time() const { return start_time; }
  void set_start_time(double t) { start_time = t; }
  //! Observation::change_start_time: shift by a number of samples
  void change_start_time(int64_t samples) { start_time += double(samples) / rate; }
  //! Observation.C:521-535
  double get_centre_frequency(unsigned ichan) const
  {
    double base = dc_centred ? centre_frequency - 0.5 * bandwidth
                             : centre_frequency - 0.5 * bandwidth + 0.5 * bandwidth / double(nchan);
    return base + double(ichan) * bandwidth / double(nchan);
  }
  void copy_configuration(const Observation* o) { *this = *o; }

 protected:
  unsigned nchan = 1, npol = 1, ndim = 2;
  Signal::State state = Signal::Analytic;
  double centre_frequency = 0, bandwidth = 0, rate = 0, dispersion_measure = 0;
  bool dc_centred = false;
  double start_time = 0;
};

//! FPT time series: get_datptr(ichan, ipol) -> ndat*ndim floats
class TimeSeries : public Observation {
 public:
  explicit TimeSeries(Memory* m = nullptr) : memory(m ? m : Memory::host()) {}
  ~TimeSeries() override { release(); }
  TimeSeries(const TimeSeries&) = delete;
  TimeSeries& operator=(const TimeSeries&) = delete;
  void set_memory(Memory* m)
  {
    release();
    memory = m;
  }
  Memory* get_memory() const { return memory; }
  uint64_t get_ndat() const { return ndat; }
  void set_ndat(uint64_t n) { ndat = n; }
  //! allocate room for ndat samples of every channel and polarisation
  void resize(uint64_t n);
  void zero();
  float* get_datptr(unsigned ichan = 0, unsigned ipol = 0) { return buffer + ichan * chan_span() + ipol * pol_span; }
  const float* get_datptr(unsigned ichan = 0, unsigned ipol = 0) const
  {
    return buffer + ichan * chan_span() + ipol * pol_span;
  }
  //! floats between consecutive polarisations / channels
  uint64_t get_pol_span() const { return pol_span; }
  uint64_t chan_span() const { return pol_span * npol; }
  bool get_zeroed_data() const { return zeroed_data; }
  void set_zeroed_data(bool z) { zeroed_data = z; }
  int64_t get_input_sample() const { return input_sample; }
  void set_input_sample(int64_t s) { input_sample = s; }

 protected:
  void release();
  Memory* memory;
  float* buffer = nullptr;
  uint64_t ndat = 0, capacity = 0, pol_span = 0;
  bool zeroed_data = false;
  int64_t input_sample = 0;
};

//! Folded profiles: get_datptr(ichan, ipol) -> nbin*ndim floats; hits[hits_nchan][nbin]
class PhaseSeries : public TimeSeries {
 public:
  using TimeSeries::TimeSeries;
  unsigned get_nbin() const { return unsigned(get_ndat()); }
  unsigned* get_hits(unsigned ichan = 0) { return hits.data() + ichan * get_nbin(); }
  unsigned get_hits_nchan() const { return hits_nchan; }
  void set_hits_nchan(unsigned n) { hits_nchan = n; }
  void resize_hits() { hits.assign(size_t(hits_nchan) * get_nbin(), 0); }
  double integration_length = 0;
  uint64_t ndat_total = 0;
  double folding_period = 0;

 protected:
  std::vector<unsigned> hits;
  unsigned hits_nchan = 1;
};

//! Raw digitised samples: TFP [ndat][nchan][npol][2] of nbit-bit integers
class BitSeries : public Observation {
 public:
  explicit BitSeries(Memory* m = nullptr) : memory(m ? m : Memory::host()) {}
  ~BitSeries() override;
  void resize(uint64_t ndat, unsigned nbit);
  uint64_t get_ndat() const { return ndat; }
  unsigned get_nbit() const { return nbit; }
  unsigned char* get_rawptr() { return data; }
  const unsigned char* get_rawptr() const { return data; }
  Memory* get_memory() const { return memory; }
  bool offset_binary = false;

 private:
  Memory* memory;
  unsigned char* data = nullptr;
  uint64_t ndat = 0;
  unsigned nbit = 8;
};

}  // namespace dsp
