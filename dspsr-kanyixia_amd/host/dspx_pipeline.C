// dspx_pipeline -- a LoadToFold-style command over the dsp::HIP engines.
//
// Reads one block of raw TFP samples from a file (--raw with the shape given,
// or --dada with the shape, band and rate from the DADA header), runs
//   filterbank : Unpacker -> Filterbank[HIP::FilterbankEngine] (writes the channels)
//   chain : Unpacker -> Convolution[HIP::ConvolutionEngine] ->
//           (--interchan 1: SampleDelay[HIP::SampleDelayEngine] ->)
//           Detection[HIP::DetectionEngine] -> Fold[HIP::FoldEngine] per sub-integration
//   fused : HIP::LoadToFoldEngine (three fused kernels)
//   plain : HIP::LoadToFoldEngine, plain FP32 form
// and writes the folded profiles [nsub][nchan][npol][nbin*ndim] (float) and
// hits [nsub][nbin] (uint32) to the output file (and, with --npz FILE, as a
// NumPy archive with the fold's parameters).  The construction mirrors
// LoadToFold::construct (LoadToFold1.C:109-258) with engines attached as
// DedispersionPipe.C:413-427 and LoadToFold1.C:843-852, 969-978 attach them.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>

#include "dsp/DADAFile.h"
#include "dsp/HIPEngines.h"
#include "dsp/ProfileWriter.h"

using namespace dsp;

static std::map<std::string, std::string> parse(int argc, char** argv)
{
  std::map<std::string, std::string> a;
  for (int i = 1; i + 1 < argc; i += 2) {
    if (std::strncmp(argv[i], "--", 2)) throw Error(InvalidParam, "dspx_pipeline", "bad option %s", argv[i]);
    a[argv[i] + 2] = argv[i + 1];
  }
  return a;
}

static Signal::State parse_state(const std::string& s)
{
  if (s == "Intensity") return Signal::Intensity;
  if (s == "PPQQ") return Signal::PPQQ;
  if (s == "Coherence") return Signal::Coherence;
  if (s == "Stokes") return Signal::Stokes;
  throw Error(InvalidParam, "dspx_pipeline", "unknown state %s", s.c_str());
}

int main(int argc, char** argv)
try {
  auto a = parse(argc, argv);
  auto need = [&](const char* k) -> std::string {
    if (!a.count(k)) throw Error(InvalidParam, "dspx_pipeline", "missing --%s", k);
    return a[k];
  };
  // --dada FILE: shape, band and sample rate from the file's header (DADAFile);
  // otherwise --raw FILE with the shape and band given
  std::unique_ptr<DADAFile> dada;
  if (a.count("dada")) dada.reset(new DADAFile(a["dada"]));
  const dspx_dada_info* hi = dada ? &dada->get_info() : nullptr;
  unsigned nbit = hi ? hi->nbit : std::stoul(need("nbit"));
  unsigned nchan = hi ? hi->nchan : std::stoul(need("nchan"));
  unsigned npol = hi ? hi->npol : std::stoul(need("npol"));
  uint64_t ndat = a.count("ndat") || !dada ? std::stoull(need("ndat")) : dada->get_ndat();
  double cfreq = hi ? hi->centre_frequency : std::stod(need("cfreq"));
  double bw = hi ? hi->bandwidth : std::stod(need("bw"));
  double dm = std::stod(need("dm"));
  std::string mode = need("mode");
  // folding options (not used by the filterbank mode)
  auto fold_opt = [&](const char* k, const char* dflt) { return mode == "filterbank" && !a.count(k) ? dflt : need(k); };
  double period = std::stod(fold_opt("period", "1"));
  unsigned nbin = std::stoul(fold_opt("nbin", "1")), nsub = std::stoul(fold_opt("nsub", "1"));
  Signal::State state = parse_state(fold_opt("state", "Intensity"));
  unsigned ndim = std::stoul(a.count("ndim") ? a["ndim"] : "1");

  HIP::Context ctx(0);
  BitSeries raw(Memory::device());
  if (dada) {
    dada->load(&raw, 0, ndat);
    raw.set_dispersion_measure(dm);
  } else {
    raw.set_nchan(nchan);
    raw.set_npol(npol);
    raw.set_centre_frequency(cfreq);
    raw.set_bandwidth(bw);
    raw.set_dispersion_measure(dm);
    raw.set_rate(std::fabs(bw) / nchan * 1e6);
    raw.set_start_time(0.0);
    raw.resize(ndat, nbit);
  }
  if (!dada) {
    size_t bytes = size_t(ndat) * nchan * npol * 2 * nbit / 8;
    std::vector<char> h(bytes);
    std::ifstream f(need("raw"), std::ios::binary);
    if (!f.read(h.data(), std::streamsize(bytes))) throw Error(FailedCall, "dspx_pipeline", "short read");
    if (hipMemcpy(raw.get_rawptr(), h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
      throw Error(FailedCall, "dspx_pipeline", "hipMemcpy");
  }

  std::vector<float> prof;
  std::vector<uint32_t> hits;
  if (mode == "filterbank") {
    // Unpacker -> Filterbank[HIP::FilterbankEngine] (FilterbankConfig.C:115-133);
    // the output is the FPT time series [nchan_out][npol][ndat_out][2]
    TimeSeries unpacked(Memory::device()), channels(Memory::device()), zdm(Memory::device());
    HIP::Unpacker unpacker(&ctx);
    unpacker.unpack(&raw, &unpacked);
    Filterbank filterbank;
    filterbank.set_input(&unpacked);
    filterbank.set_output(&channels);
    if (a.count("zdm")) {  // the zero-DM output too (LoadToFold1.C:291-303)
      filterbank.set_zero_DM(true);
      filterbank.set_zero_DM_output(&zdm);
    }
    unsigned nchan_out = std::stoul(need("fb_nchan"));
    filterbank.set_nchan(nchan_out);
    TableResponse table;
    if (a.count("response")) {
      // [nchan_out][fres] complex float in the Response's order
      unsigned fres = std::stoul(need("fres"));
      std::vector<float> t(size_t(nchan_out) * fres * 2);
      std::ifstream f(a["response"], std::ios::binary);
      if (!f.read(reinterpret_cast<char*>(t.data()), std::streamsize(t.size() * sizeof(float))))
        throw Error(FailedCall, "dspx_pipeline", "short read of the response");
      table.set_table(nchan_out, fres, std::stoul(need("nfilt_pos")), std::stoul(need("nfilt_neg")), std::move(t));
      filterbank.set_response(&table);
    } else {
      filterbank.set_frequency_resolution(std::stoul(need("fres")));
    }
    filterbank.set_engine(new HIP::FilterbankEngine(&ctx));
    filterbank.operate();
    ctx.synchronize();
    // [nchan_out][npol][ndat_out][2] floats of a device time series to a file
    auto save = [&](const TimeSeries& ts, const std::string& path) {
      const size_t per = size_t(ts.get_ndat()) * 2;
      std::vector<float> h(size_t(nchan_out) * npol * per);
      for (unsigned c = 0; c < nchan_out; c++)
        for (unsigned p = 0; p < npol; p++)
          if (per && hipMemcpy(&h[(size_t(c) * npol + p) * per], ts.get_datptr(c, p), per * sizeof(float),
                               hipMemcpyDeviceToHost) != hipSuccess)
            throw Error(FailedCall, "dspx_pipeline", "hipMemcpy");
      std::ofstream o(path, std::ios::binary);
      o.write(reinterpret_cast<const char*>(h.data()), std::streamsize(h.size() * sizeof(float)));
    };
    save(channels, need("out"));
    if (a.count("zdm")) save(zdm, a["zdm"]);
    std::cout << "dspx_pipeline filterbank: " << nchan_out << " channels x " << channels.get_ndat() << " samples"
              << std::endl;
    return 0;
  }
  if (mode == "chain") {
    TimeSeries unpacked(Memory::device()), filtered(Memory::device()), detected(Memory::device());
    HIP::Unpacker unpacker(&ctx);
    unpacker.unpack(&raw, &unpacked);

    // --interchan 1: inter-channel delays removed as dspsr -K does
    // (DedispersionPipe.C:435-472): fractional part in the chirp, integer
    // part by a SampleDelay after the convolution
    const bool interchan = a.count("interchan") && a["interchan"] != "0";
    Dedispersion kernel;
    kernel.set_dispersion_measure(dm);
    kernel.set_interchan(interchan);
    Convolution convolution;
    convolution.set_input(&unpacked);
    convolution.set_output(&filtered);
    convolution.set_response(&kernel);
    convolution.set_engine(new HIP::ConvolutionEngine(&ctx));
    convolution.operate();

    TimeSeries delayed(Memory::device());
    TimeSeries* to_detect = &filtered;
    if (interchan) {
      SampleDelay sample_delay;
      sample_delay.set_input(&filtered);
      sample_delay.set_output(&delayed);
      sample_delay.set_engine(new HIP::SampleDelayEngine(&ctx));
      sample_delay.operate();
      to_detect = &delayed;
    }

    Detection detect;
    detect.set_input(to_detect);
    detect.set_output(&detected);
    detect.set_output_state(state);
    detect.set_output_ndim(ndim);
    detect.set_engine(new HIP::DetectionEngine(&ctx));
    detect.operate();

    unsigned npol_out = detected.get_npol(), ndim_out = detected.get_ndim();
    uint64_t nout = detected.get_ndat();
    size_t per_sub = size_t(nchan) * npol_out * nbin * ndim_out;
    prof.assign(per_sub * nsub, 0.f);
    hits.assign(size_t(nsub) * nbin, 0);
    for (unsigned k = 0; k < nsub; k++) {
      uint64_t i0 = uint64_t(std::rint(double(nout) * k / nsub));
      uint64_t i1 = uint64_t(std::rint(double(nout) * (k + 1) / nsub));
      PhaseSeries sub(Memory::host());
      Fold fold;
      fold.set_input(&detected);
      fold.set_output(&sub);
      fold.set_folding_period(period);
      fold.set_nbin(nbin);
      fold.set_limits(i0, i1 - i0);
      fold.set_engine(new HIP::FoldEngine(&ctx));
      fold.operate();
      fold.get_result();
      for (unsigned c = 0; c < nchan; c++)
        for (unsigned p = 0; p < npol_out; p++)
          std::memcpy(&prof[k * per_sub + (size_t(c) * npol_out + p) * nbin * ndim_out], sub.get_datptr(c, p),
                      size_t(nbin) * ndim_out * sizeof(float));
      std::memcpy(&hits[size_t(k) * nbin], sub.get_hits(), nbin * sizeof(uint32_t));
    }
  } else if (mode == "fused" || mode == "plain") {
    HIP::LoadToFoldEngine::Config cfg;
    cfg.folding_period = period;
    cfg.nbin = nbin;
    cfg.nsub = nsub;
    cfg.state = state;
    cfg.ndim = ndim;
    cfg.form = mode == "fused" ? DSPX_FORM_FUSED : DSPX_FORM_PLAIN;
    HIP::LoadToFoldEngine ltf(&ctx, cfg);
    ltf.prepare(&raw);
    ltf.operate(&raw);
    ltf.get_result(prof, hits);
  } else {
    throw Error(InvalidParam, "dspx_pipeline", "unknown mode %s", mode.c_str());
  }
  if (a.count("npz")) {
    // Archiver stand-in: the profiles, hits and fold parameters as named arrays
    const unsigned ndim_out = (state == Signal::Intensity || state == Signal::PPQQ) ? 1 : ndim;
    const unsigned npol_out = state == Signal::Intensity ? 1 : state == Signal::PPQQ ? npol : 4 / ndim;
    ProfileWriter w;
    w.add("profiles", prof.data(), {nsub, nchan, npol_out, uint64_t(nbin) * ndim_out});
    w.add("hits", hits.data(), {nsub, nbin});
    w.add("folding_period", period);
    w.add("dispersion_measure", dm);
    w.add("centre_frequency", cfreq);
    w.add("bandwidth", bw);
    w.add("nbin", double(nbin));
    w.add("state", double(state));
    w.write(a["npz"]);
  }
  std::ofstream o(need("out"), std::ios::binary);
  o.write(reinterpret_cast<const char*>(prof.data()), std::streamsize(prof.size() * sizeof(float)));
  o.write(reinterpret_cast<const char*>(hits.data()), std::streamsize(hits.size() * sizeof(uint32_t)));
  std::cout << "dspx_pipeline " << mode << ": " << prof.size() << " profile values, " << hits.size() << " hits"
            << std::endl;
  return 0;
} catch (const Error& e) {
  std::cerr << "Error: " << e.what() << std::endl;
  return 1;
}
