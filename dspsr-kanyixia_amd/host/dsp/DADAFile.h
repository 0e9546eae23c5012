// DADAFile.h -- the DADA raw-file source for the HIP engines.
//
// Reference: dsp::DADAFile (Kernel/Classes/DADAFile.C:26-68),
// dsp::DADAHeader::load_header (Kernel/Classes/DADAHeader.C:50-140) and
// dsp::ASCIIObservation::load (Kernel/Classes/ASCIIObservation.C:24-460).
// The header is parsed by dspx_dada_parse; the raw samples (TFP, complex,
// 8 or 16 bits) are read block by block into a BitSeries in device memory,
// configured with the header's band and sample rate.  Start times are
// seconds from UTC_START (MJD arithmetic is not part of this tree).
#pragma once
#include <cstdio>
#include <string>

#include "dsp/Types.h"
#include "dspx.h"

namespace dsp {

class DADAFile {
 public:
  explicit DADAFile(const std::string& filename);
  ~DADAFile();
  DADAFile(const DADAFile&) = delete;
  DADAFile& operator=(const DADAFile&) = delete;
  const dspx_dada_info& get_info() const { return info; }
  const std::string& get_header() const { return header; }
  //! samples in the file after the header (limited by NDAT when given)
  uint64_t get_ndat() const { return ndat; }
  uint64_t get_bytes_per_sample() const { return bytes_per_sample; }
  //! configure out from the header and load samples [start, start + n)
  void load(BitSeries* out, uint64_t start, uint64_t n);

 private:
  std::string filename, header;
  dspx_dada_info info{};
  FILE* fp = nullptr;
  uint64_t bytes_per_sample = 0, ndat = 0;
};

}  // namespace dsp
