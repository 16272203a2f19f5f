// Small .npy reader / writer (format version 1.0 / 2.0, C order, little endian),
// standing in for fast-wave's NpyCpp dependency (fast-wave/CMakeLists.txt:18-81).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rs {

struct NpyArray {
  std::string descr;             // e.g. "<c16", "<c8", "|i1", "<f8"
  std::vector<int64_t> shape;
  std::vector<char> data;        // raw little-endian bytes
  int64_t count() const {
    int64_t n = 1;
    for (auto s : shape) n *= s;
    return n;
  }
};

// header of a .npy file: element type, shape, where the data starts and the
// file's size in bytes
struct NpyHeader {
  std::string descr;
  std::vector<int64_t> shape;
  int64_t data_offset = 0;
  int64_t file_bytes = 0;
};

NpyHeader npy_header(const std::string &path);
// the bytes of a version 1.0 .npy header for a C-order array (the data
// follows them directly)
std::string npy_header_bytes(const std::string &descr, const std::vector<int64_t> &shape);
NpyArray npy_load(const std::string &path);
void npy_save(const std::string &path, const std::string &descr, const std::vector<int64_t> &shape,
              const void *data, size_t bytes);

}  // namespace rs
