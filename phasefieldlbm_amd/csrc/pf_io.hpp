// Field output: raw .bin dumps and VTK XML files (ImageData .vti and
// RectilinearGrid .vtr) with raw appended data.  Host code only.
//
// The reference writes one <name><step6>.bin per field
// (PostProcess::bin, src/postProcess/PostProcess.cuh:43-97) and then a
// step_<step6>.vti whose appended block repeats those bytes
// (PostProcess::vti, :99-261).  The writers here stream each field once,
// from one host buffer, into every requested file.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace pfio {

enum Format { kBin = 1, kVti = 2, kVtr = 4 };

// "%06llu", the reference's stepSuffix6.
std::string step_suffix(uint64_t step);
std::string join(const std::string& dir, const std::string& name);

// One VTK XML file with float32 point arrays of nx*ny*nz values each,
// written array by array: open() writes the XML header (offsets follow from
// the sizes alone), put() appends the next array, close() the trailer.
class VtkWriter {
public:
    // kind: kVti or kVtr.
    bool open(const std::string& path, int kind, int nx, int ny, int nz, const std::vector<std::string>& names);
    bool put(const float* data);
    bool close();
    ~VtkWriter();

private:
    bool put_bytes(const void* data, uint64_t nbytes);
    std::FILE* f_ = nullptr;
    int kind_ = 0, nx_ = 0, ny_ = 0, nz_ = 0;
    bool wide_ = false;  // UInt64 block headers (an array of 4 GiB or more)
    size_t expected_ = 0, written_ = 0;
};

bool write_bin(const std::string& path, const float* data, size_t n);

}  // namespace pfio
