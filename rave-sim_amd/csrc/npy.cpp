#include "npy.hpp"

#include <cstring>
#include <fstream>
#include <sstream>

#include "engine.hpp"

namespace rs {

namespace {
std::string dict_value(const std::string &h, const std::string &key) {
  const size_t k = h.find("'" + key + "'");
  if (k == std::string::npos) throw Error(3, "npy header lacks " + key);
  size_t c = h.find(':', k);
  if (c == std::string::npos) throw Error(3, "bad npy header");
  ++c;
  while (c < h.size() && h[c] == ' ') ++c;
  if (h[c] == '\'') {
    const size_t e = h.find('\'', c + 1);
    return h.substr(c + 1, e - c - 1);
  }
  if (h[c] == '(') {
    const size_t e = h.find(')', c);
    return h.substr(c, e - c + 1);
  }
  size_t e = c;
  while (e < h.size() && h[e] != ',' && h[e] != '}') ++e;
  return h.substr(c, e - c);
}
}  // namespace

namespace {
// reads the header of the open file f; leaves f at the first data byte
NpyHeader read_header(std::ifstream &f, const std::string &path) {
  char magic[8];
  f.read(magic, 8);
  if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) throw Error(3, path + " is not a .npy file");
  const int major = (unsigned char)magic[6];
  uint32_t hlen = 0;
  if (major == 1) {
    unsigned char b[2];
    f.read((char *)b, 2);
    hlen = b[0] | (b[1] << 8);
  } else {
    unsigned char b[4];
    f.read((char *)b, 4);
    hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
  }
  std::string h(hlen, ' ');
  f.read(&h[0], hlen);
  if (!f) throw Error(3, path + ": truncated header");
  NpyHeader a;
  a.descr = dict_value(h, "descr");
  if (dict_value(h, "fortran_order").find("True") != std::string::npos)
    throw Error(3, path + ": Fortran-ordered arrays are not supported");
  const std::string shp = dict_value(h, "shape");
  std::string num;
  for (char c : shp) {
    if (c >= '0' && c <= '9') num += c;
    else if (!num.empty()) {
      a.shape.push_back(std::stoll(num));
      num.clear();
    }
  }
  a.data_offset = (int64_t)f.tellg();
  return a;
}

int64_t item_size(const std::string &d, const std::string &path) {
  if (d == "<c16") return 16;
  if (d == "<c8" || d == "<f8" || d == "<i8") return 8;
  if (d == "<f4") return 4;
  if (d == "|i1" || d == "|u1" || d == "<i1") return 1;
  throw Error(3, path + ": unsupported dtype " + d);
}
}  // namespace

NpyHeader npy_header(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(3, "cannot open " + path);
  NpyHeader h = read_header(f, path);
  f.seekg(0, std::ios::end);
  h.file_bytes = (int64_t)f.tellg();
  return h;
}

NpyArray npy_load(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(3, "cannot open " + path);
  const NpyHeader h = read_header(f, path);
  NpyArray a;
  a.descr = h.descr;
  a.shape = h.shape;
  a.data.resize((size_t)(a.count() * item_size(a.descr, path)));
  f.read(a.data.data(), (std::streamsize)a.data.size());
  if (!f) throw Error(3, path + ": truncated data");
  return a;
}

std::string npy_header_bytes(const std::string &descr, const std::vector<int64_t> &shape) {
  std::ostringstream h;
  h << "{'descr': '" << descr << "', 'fortran_order': False, 'shape': (";
  for (size_t i = 0; i < shape.size(); ++i) {
    h << shape[i];
    if (shape.size() == 1 || i + 1 < shape.size()) h << ",";
    if (i + 1 < shape.size()) h << " ";
  }
  h << "), }";
  std::string hs = h.str();
  const size_t total = 10 + hs.size() + 1;
  const size_t pad = (64 - total % 64) % 64;
  hs += std::string(pad, ' ');
  hs += '\n';
  const char pre[10] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0, (char)(hs.size() & 0xff), (char)(hs.size() >> 8)};
  return std::string(pre, 10) + hs;
}

void npy_save(const std::string &path, const std::string &descr, const std::vector<int64_t> &shape,
              const void *data, size_t bytes) {
  const std::string hdr = npy_header_bytes(descr, shape);
  std::ofstream f(path, std::ios::binary);
  if (!f) throw Error(3, "cannot write " + path);
  f.write(hdr.data(), (std::streamsize)hdr.size());
  f.write((const char *)data, (std::streamsize)bytes);
  if (!f) throw Error(3, "write failed: " + path);
}

}  // namespace rs
