// ProfileWriter.C -- see dsp/ProfileWriter.h.  The .npy layout is NumPy's
// format 1.0 (magic, version, little-endian header length, a Python dict
// literal padded to 64 bytes); the .npz container is a zip archive with
// stored entries (local headers, data, central directory, end record).
#include "dsp/ProfileWriter.h"

#include <cstdio>
#include <cstring>

#include "dsp/Types.h"

namespace dsp {

namespace {

uint32_t crc32(const unsigned char* p, size_t n, uint32_t crc = 0)
{
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  crc = ~crc;
  for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xff] ^ (crc >> 8);
  return ~crc;
}

void put16(std::vector<unsigned char>& b, uint16_t v)
{
  b.push_back(v & 0xff);
  b.push_back(v >> 8);
}

void put32(std::vector<unsigned char>& b, uint32_t v)
{
  for (int i = 0; i < 4; i++) b.push_back((v >> (8 * i)) & 0xff);
}

std::vector<unsigned char> npy(const std::string& descr, const std::vector<uint64_t>& shape,
                               const std::vector<unsigned char>& data)
{
  std::string dict = "{'descr': '" + descr + "', 'fortran_order': False, 'shape': (";
  for (size_t i = 0; i < shape.size(); i++) dict += std::to_string(shape[i]) + (shape.size() == 1 ? "," : i + 1 < shape.size() ? ", " : "");
  dict += "), }";
  size_t total = 10 + dict.size() + 1;
  dict.append((64 - total % 64) % 64, ' ');
  dict += '\n';
  std::vector<unsigned char> out = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  put16(out, uint16_t(dict.size()));
  out.insert(out.end(), dict.begin(), dict.end());
  out.insert(out.end(), data.begin(), data.end());
  return out;
}

}  // namespace

void ProfileWriter::add_bytes(const std::string& name, const char* descr, const void* data, size_t elem,
                              const std::vector<uint64_t>& shape)
{
  uint64_t n = 1;
  for (auto d : shape) n *= d;
  Entry e{name, descr, shape, {}};
  e.bytes.resize(n * elem);
  if (n) std::memcpy(e.bytes.data(), data, n * elem);
  entries.push_back(std::move(e));
}

void ProfileWriter::add(const std::string& name, const float* data, const std::vector<uint64_t>& shape)
{
  add_bytes(name, "<f4", data, sizeof(float), shape);
}

void ProfileWriter::add(const std::string& name, const uint32_t* data, const std::vector<uint64_t>& shape)
{
  add_bytes(name, "<u4", data, sizeof(uint32_t), shape);
}

void ProfileWriter::add(const std::string& name, double value) { add_bytes(name, "<f8", &value, sizeof value, {}); }

void ProfileWriter::write(const std::string& filename) const
{
  std::vector<unsigned char> zip, central;
  for (const auto& e : entries) {
    const std::vector<unsigned char> body = npy(e.descr, e.shape, e.bytes);
    if (body.size() >= 0xffffffffull || zip.size() >= 0xffffffffull)
      throw Error(InvalidState, "dsp::ProfileWriter::write", "array %s too large", e.name.c_str());
    const std::string fname = e.name + ".npy";
    const uint32_t crc = crc32(body.data(), body.size());
    const uint32_t offset = uint32_t(zip.size());
    // local file header
    put32(zip, 0x04034b50u);
    put16(zip, 20);  // version needed
    put16(zip, 0);   // flags
    put16(zip, 0);   // stored
    put16(zip, 0);   // time
    put16(zip, 0x21);  // date: 1980-01-01
    put32(zip, crc);
    put32(zip, uint32_t(body.size()));
    put32(zip, uint32_t(body.size()));
    put16(zip, uint16_t(fname.size()));
    put16(zip, 0);
    zip.insert(zip.end(), fname.begin(), fname.end());
    zip.insert(zip.end(), body.begin(), body.end());
    // central directory entry
    put32(central, 0x02014b50u);
    put16(central, 20);  // version made by
    put16(central, 20);  // version needed
    put16(central, 0);
    put16(central, 0);
    put16(central, 0);
    put16(central, 0x21);
    put32(central, crc);
    put32(central, uint32_t(body.size()));
    put32(central, uint32_t(body.size()));
    put16(central, uint16_t(fname.size()));
    put16(central, 0);  // extra
    put16(central, 0);  // comment
    put16(central, 0);  // disk
    put16(central, 0);  // internal attributes
    put32(central, 0);  // external attributes
    put32(central, offset);
    central.insert(central.end(), fname.begin(), fname.end());
  }
  const uint32_t cd_offset = uint32_t(zip.size());
  zip.insert(zip.end(), central.begin(), central.end());
  put32(zip, 0x06054b50u);
  put16(zip, 0);
  put16(zip, 0);
  put16(zip, uint16_t(entries.size()));
  put16(zip, uint16_t(entries.size()));
  put32(zip, uint32_t(central.size()));
  put32(zip, cd_offset);
  put16(zip, 0);
  FILE* f = std::fopen(filename.c_str(), "wb");
  if (!f) throw Error(FailedCall, "dsp::ProfileWriter::write", "fopen (%s)", filename.c_str());
  const size_t w = std::fwrite(zip.data(), 1, zip.size(), f);
  const int c = std::fclose(f);
  if (w != zip.size() || c != 0) throw Error(FailedCall, "dsp::ProfileWriter::write", "write (%s)", filename.c_str());
}

}  // namespace dsp
