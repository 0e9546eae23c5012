// ProfileWriter.h -- folded profiles to a NumPy .npz archive.
//
// Stands in for dsp::Archiver (Signal/Pulsar/Archiver.C), which writes
// PSRCHIVE archives and needs PSRCHIVE: the profiles, hits and parameters
// of a fold are written as named arrays in a .npz file (a zip archive of
// .npy arrays, stored without compression), readable by numpy.load.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace dsp {

class ProfileWriter {
 public:
  //! float32 array of the given shape (C order)
  void add(const std::string& name, const float* data, const std::vector<uint64_t>& shape);
  //! uint32 array of the given shape (C order)
  void add(const std::string& name, const uint32_t* data, const std::vector<uint64_t>& shape);
  //! float64 scalar (0-d array)
  void add(const std::string& name, double value);
  //! write every array added so far; throws Error on I/O failure
  void write(const std::string& filename) const;

 private:
  struct Entry {
    std::string name;
    std::string descr;  // numpy dtype string
    std::vector<uint64_t> shape;
    std::vector<unsigned char> bytes;
  };
  void add_bytes(const std::string& name, const char* descr, const void* data, size_t elem,
                 const std::vector<uint64_t>& shape);
  std::vector<Entry> entries;
};

}  // namespace dsp
