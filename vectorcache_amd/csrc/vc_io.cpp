// NPY float32 matrices and the VCEX exact top-k cache (see vc_io.hpp).
#include "vc_io.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cctype>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string_view>

namespace vcio {
namespace fs = std::filesystem;

// ---- NPY (src/datasets/npy.cpp:28-110 for the accepted headers) ----------
void parse_npy_f32(const char* data, std::size_t size, std::size_t& count, std::size_t& dim,
                   std::size_t& offset) {
  if (size < 10) throw IoError("NPY file too small");
  if (std::memcmp(data, "\x93NUMPY", 6) != 0) throw IoError("invalid NPY magic bytes");
  const auto byte = [&](std::size_t i) { return static_cast<std::size_t>(static_cast<unsigned char>(data[i])); };
  const unsigned major = unsigned(byte(6)), minor = unsigned(byte(7));
  std::size_t hlen = 0, hstart = 0;
  if (major == 1 && minor == 0) {  // u16 header length
    hlen = byte(8) | byte(9) << 8;
    hstart = 10;
  } else if (major == 2 && minor == 0) {  // u32 header length
    if (size < 12) throw IoError("NPY v2 header truncated");
    hlen = byte(8) | byte(9) << 8 | byte(10) << 16 | byte(11) << 24;
    hstart = 12;
  } else {
    throw IoError("unsupported NPY version " + std::to_string(major) + "." + std::to_string(minor));
  }
  const std::size_t hend = hstart + hlen;
  if (hend > size) throw IoError("NPY header extends past end of file");
  const std::string_view h(data + hstart, hlen);
  if (h.find("'descr': '<f4'") == std::string_view::npos && h.find("\"descr\": \"<f4\"") == std::string_view::npos)
    throw IoError("expected float32 ('<f4') NPY array");
  const std::size_t s = h.find("shape");
  if (s == std::string_view::npos) throw IoError("missing 'shape' in NPY header");
  const std::size_t lp = h.find('(', s);
  const std::size_t rp = lp == std::string_view::npos ? lp : h.find(')', lp);
  if (lp == std::string_view::npos || rp == std::string_view::npos) throw IoError("failed to parse NPY shape");
  // comma-separated dimensions; blanks around them and empty fields (the
  // trailing comma of a 1-tuple) are skipped
  std::vector<std::size_t> shape;
  std::size_t i = lp + 1;
  while (i <= rp) {
    std::size_t j = h.find(',', i);
    if (j == std::string_view::npos || j > rp) j = rp;
    std::string_view tok = h.substr(i, j - i);
    while (!tok.empty() && (tok.front() == ' ' || tok.front() == '\t')) tok.remove_prefix(1);
    while (!tok.empty() && (tok.back() == ' ' || tok.back() == '\t')) tok.remove_suffix(1);
    if (!tok.empty()) shape.push_back(std::stoull(std::string(tok)));
    i = j + 1;
  }
  if (shape.size() != 2) throw IoError("expected 2-D NPY matrix");
  count = shape[0];
  dim = shape[1];
  offset = hend;
  if (size < hend + count * dim * 4) throw IoError("NPY data truncated");
}

NpyMatrix::NpyMatrix(const std::string& path) {
  const int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) throw IoError("failed to open " + path);
  struct stat st {};
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    throw IoError("failed to stat " + path);
  }
  size_ = static_cast<std::size_t>(st.st_size);
  void* m = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (m == MAP_FAILED) throw IoError("failed to mmap " + path);
  map_ = m;
  try {
    parse_npy_f32(static_cast<const char*>(map_), size_, count_, dim_, offset_);
  } catch (...) {
    ::munmap(map_, size_);
    throw;
  }
}

NpyMatrix::~NpyMatrix() {
  if (map_ != nullptr) ::munmap(map_, size_);
}

const float* NpyMatrix::data() const {
  return reinterpret_cast<const float*>(static_cast<const char*>(map_) + offset_);
}

std::size_t NpyMatrix::copy_rows(std::size_t first, std::size_t n, float* out) const {
  if (first >= count_ || n == 0) return 0;
  const std::size_t take = std::min(n, count_ - first);
  std::memcpy(out, static_cast<const char*>(map_) + offset_ + first * dim_ * 4, take * dim_ * 4);
  return take;
}

// The reference writer's layout: version 1.0, u16 header length, the dict
// padded with blanks so that 10 + dict is a multiple of 16, then a newline.
void write_npy_f32(const std::string& path, std::size_t rows, std::size_t cols, const float* data) {
  const fs::path p(path);
  if (p.has_parent_path()) fs::create_directories(p.parent_path());
  std::string hdr = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(rows) + ", " +
                    std::to_string(cols) + "), }";
  hdr.append((16 - (10 + hdr.size()) % 16) % 16, ' ');
  hdr.push_back('\n');
  std::ofstream f(path, std::ios::binary);
  if (!f) throw IoError("failed to create " + path);
  const char pre[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
  f.write(pre, 8);
  const unsigned char len[2] = {static_cast<unsigned char>(hdr.size() & 0xff),
                                static_cast<unsigned char>(hdr.size() >> 8)};
  f.write(reinterpret_cast<const char*>(len), 2);
  f.write(hdr.data(), static_cast<std::streamsize>(hdr.size()));
  f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(rows * cols * 4));
  if (!f) throw IoError("failed to write " + path);
}

// ---- VCEX (tools/query-bench.cpp:146-340) ----------------------------------
std::uint64_t fnv1a64(const std::string& s) {
  std::uint64_t h = 14695981039346656037ULL;
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

std::string cache_source_label(const std::string& npy_path, const std::string& dataset) {
  if (npy_path.empty()) return dataset;
  const fs::path p(npy_path);
  std::ostringstream o;
  o << "npy-" << p.stem().string() << '-' << std::hex << fnv1a64(p.lexically_normal().string());
  return o.str();
}

static std::string token(std::string s) {
  for (char& c : s)
    if (!std::isalnum(static_cast<unsigned char>(c)) && c != '-' && c != '_') c = '_';
  return s;
}

std::string exact_topk_cache_path(const ExactTopkKey& key) {
  const std::string name = token(key.source_label) + '_' + token(key.split) + "_n" + std::to_string(key.index_n) +
                           "_qs-" + token(key.query_split) + "_ql" + std::to_string(key.query_n) + "_k" +
                           std::to_string(key.k) + "_seed" + std::to_string(key.seed) + ".bin";
  return (fs::path(".cache") / "exact_topk" / name).string();
}

template <typename T>
static bool get(std::istream& in, T& v) {
  in.read(reinterpret_cast<char*>(&v), sizeof(T));
  return static_cast<bool>(in);
}
static bool get_str(std::istream& in, std::string& s) {
  std::uint32_t n = 0;
  if (!get(in, n)) return false;
  s.assign(n, '\0');
  if (n != 0) in.read(s.data(), n);
  return static_cast<bool>(in);
}
template <typename T>
static bool put(std::ostream& out, T v) {
  out.write(reinterpret_cast<const char*>(&v), sizeof(T));
  return static_cast<bool>(out);
}
static bool put_str(std::ostream& out, const std::string& s) {
  if (s.size() > 0xffffffffu || !put<std::uint32_t>(out, static_cast<std::uint32_t>(s.size()))) return false;
  out.write(s.data(), static_cast<std::streamsize>(s.size()));
  return static_cast<bool>(out);
}

bool load_exact_topk(const std::string& path, const ExactTopkKey& key, std::vector<std::uint64_t>& ids) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  char magic[4] = {};
  in.read(magic, 4);
  std::uint32_t version = 0;
  if (!in || std::memcmp(magic, "VCEX", 4) != 0 || !get(in, version) || version != 1) return false;
  std::uint64_t nq = 0, k = 0, n = 0, seed = 0;
  if (!get(in, nq) || !get(in, k) || !get(in, n) || !get(in, seed)) return false;
  if (nq != key.query_n || k != key.k || n != key.index_n || seed != key.seed) return false;
  std::string label, split, qsplit;
  if (!get_str(in, label) || !get_str(in, split) || !get_str(in, qsplit)) return false;
  if (label != key.source_label || split != key.split || qsplit != key.query_split) return false;
  ids.assign(static_cast<std::size_t>(nq * k), 0);
  for (auto& id : ids)
    if (!get(in, id)) return false;
  return true;
}

bool save_exact_topk(const std::string& path, const ExactTopkKey& key, const std::uint64_t* ids) {
  std::error_code ec;
  const fs::path p(path);
  if (p.has_parent_path()) {
    fs::create_directories(p.parent_path(), ec);
    if (ec) return false;
  }
  const std::string tmp = path + ".tmp";
  {
    std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
    if (!out) return false;
    out.write("VCEX", 4);
    if (!put<std::uint32_t>(out, 1) || !put(out, key.query_n) || !put(out, key.k) || !put(out, key.index_n) ||
        !put(out, key.seed) || !put_str(out, key.source_label) || !put_str(out, key.split) ||
        !put_str(out, key.query_split))
      return false;
    for (std::uint64_t i = 0; i < key.query_n * key.k; ++i)
      if (!put(out, ids[i])) return false;
    if (!out.flush()) return false;
  }
  fs::rename(tmp, p, ec);
  if (ec) {  // (a target that cannot be replaced in place is removed first)
    fs::remove(p, ec);
    ec.clear();
    fs::rename(tmp, p, ec);
  }
  if (ec) {
    fs::remove(tmp, ec);
    return false;
  }
  return true;
}

}  // namespace vcio
