// DADAFile.C -- see dsp/DADAFile.h.
#include "dsp/DADAFile.h"

#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <cstring>
#include <vector>

namespace dsp {

DADAFile::DADAFile(const std::string& fname) : filename(fname)
{
  const char* where = "dsp::DADAFile::open_file";
  fp = std::fopen(filename.c_str(), "rb");
  if (!fp) throw Error(FailedCall, where, "fopen (%s)", filename.c_str());
  // DADAHeader::load_header: the default 4096 bytes, read again while
  // HDR_SIZE asks for more (DADAHeader.C:61-83)
  uint64_t size = 4096, hdr = 0;
  std::vector<char> buf;
  for (;;) {
    buf.assign(size + 1, '\0');
    std::rewind(fp);
    if (std::fread(buf.data(), 1, size, fp) != size)
      throw Error(FailedCall, where, "fread (nbyte=%llu) of %s", (unsigned long long) size, filename.c_str());
    buf[size - 1] = '\0';
    char tok[64];
    hdr = 0;
    if (dspx_ascii_header_get(buf.data(), "HDR_SIZE", tok, sizeof tok) == 1) hdr = std::strtoull(tok, nullptr, 10);
    if (hdr <= size) break;
    size = hdr;
  }
  if (hdr == 0) throw Error(InvalidState, where, "HDR_SIZE undefined in %s", filename.c_str());
  header.assign(buf.data());
  char missing[32] = {};
  int rc = dspx_dada_parse(header.c_str(), &info, missing);
  if (rc == DSPX_E_STATE) throw Error(InvalidState, "ASCIIObservation", "invalid NDIM=%u", info.ndim);
  if (rc != DSPX_OK) throw Error(InvalidState, "ASCIIObservation", "failed load %s", missing);
  if (info.ndim != 2) throw Error(InvalidState, where, "only complex (NDIM 2) samples are on this path");
  if (info.nbit != 8 && info.nbit != 16) throw Error(InvalidState, where, "NBIT=%u (8 or 16)", info.nbit);
  bytes_per_sample = uint64_t(info.nchan) * info.npol * info.ndim * info.nbit / 8;
  if (std::fseek(fp, 0, SEEK_END) != 0) throw Error(FailedCall, where, "fseek");
  const long end = std::ftell(fp);
  if (end < 0 || uint64_t(end) < hdr) throw Error(FailedCall, where, "ftell");
  ndat = (uint64_t(end) - hdr) / bytes_per_sample;
  if (info.ndat && info.ndat < ndat) ndat = info.ndat;
}

DADAFile::~DADAFile()
{
  if (fp) std::fclose(fp);
}

void DADAFile::load(BitSeries* out, uint64_t start, uint64_t n)
{
  const char* where = "dsp::DADAFile::load";
  if (start + n > ndat)
    throw Error(InvalidParam, where, "samples [%llu, %llu) beyond ndat=%llu", (unsigned long long) start,
                (unsigned long long) (start + n), (unsigned long long) ndat);
  out->set_nchan(info.nchan);
  out->set_npol(info.npol);
  out->set_ndim(2);
  out->set_state(Signal::Analytic);
  out->set_centre_frequency(info.centre_frequency);
  out->set_bandwidth(info.bandwidth);
  out->set_rate(1e6 / info.tsamp);  // TSAMP in microseconds (ASCIIObservation.C:306-317)
  const double tsamp_s = info.tsamp * 1e-6;
  out->set_start_time(double(info.obs_offset / bytes_per_sample) * tsamp_s + double(info.picoseconds) / 1e12 +
                      double(start) * tsamp_s);
  out->resize(n, info.nbit);
  const uint64_t bytes = n * bytes_per_sample;
  std::vector<unsigned char> h(bytes);
  if (std::fseek(fp, long(info.hdr_size + start * bytes_per_sample), SEEK_SET) != 0)
    throw Error(FailedCall, where, "fseek");
  if (std::fread(h.data(), 1, bytes, fp) != bytes) throw Error(FailedCall, where, "short read of %s", filename.c_str());
  if (out->get_memory()->on_device) {
    if (hipMemcpy(out->get_rawptr(), h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
      throw Error(FailedCall, where, "hipMemcpy");
  } else {
    std::memcpy(out->get_rawptr(), h.data(), bytes);
  }
}

}  // namespace dsp
