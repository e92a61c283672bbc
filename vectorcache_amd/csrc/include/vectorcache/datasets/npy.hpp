// vectorcache::datasets::NpyReader and write_npy_f32_matrix over the C ABI of
// libvectorcache_hip.so: the reference's interface (include/vectorcache/
// datasets/npy.hpp:33-65, reader.hpp:16-27) for float32 (N, dim) NPY files
// read through mmap. File and format errors throw std::runtime_error with the
// reference's messages.
#pragma once

#include <cstddef>
#include <cstring>
#include <filesystem>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>

#include "vectorcache_hip.h"

namespace vectorcache::datasets {

struct DatasetMeta {
  std::size_t dim = 0;
  std::size_t count = 0;
  const char* label = "";
};

class NpyReader {
 public:
  static NpyReader open(const std::filesystem::path& path, const char* label) {
    vc_npy* h = nullptr;
    std::size_t count = 0, dim = 0;
    if (vc_npy_open(path.c_str(), &h, &count, &dim) != VC_OK) throw std::runtime_error(vc_last_error());
    return NpyReader(h, label, count, dim);
  }

  NpyReader(NpyReader&& o) noexcept
      : h_(std::exchange(o.h_, nullptr)), label_(o.label_), count_(o.count_), dim_(o.dim_), index_(o.index_) {}
  NpyReader& operator=(NpyReader&& o) noexcept {
    if (this != &o) {
      if (h_) vc_npy_close(h_);
      h_ = std::exchange(o.h_, nullptr);
      label_ = o.label_;
      count_ = o.count_;
      dim_ = o.dim_;
      index_ = o.index_;
    }
    return *this;
  }
  NpyReader(const NpyReader&) = delete;
  NpyReader& operator=(const NpyReader&) = delete;
  ~NpyReader() {
    if (h_) vc_npy_close(h_);
  }

  DatasetMeta meta() const { return DatasetMeta{dim_, count_, label_}; }

  /// Zero-copy view of the next vector, or nullopt after the last one.
  std::optional<std::span<const float>> next_vector_span() {
    if (index_ >= count_) return std::nullopt;
    const float* row = vc_npy_data(h_) + index_ * dim_;
    ++index_;
    return std::span<const float>(row, dim_);
  }

  bool next_vector_into(std::span<float> out) {
    if (out.size() != dim_)
      throw std::runtime_error("buffer dimension mismatch: expected " + std::to_string(dim_) + ", got " +
                               std::to_string(out.size()));
    auto v = next_vector_span();
    if (!v) return false;
    std::memcpy(out.data(), v->data(), dim_ * sizeof(float));
    return true;
  }

  /// Copies the next (up to) n vectors into out (n * dim floats); returns the
  /// rows copied and advances by them.
  std::size_t copy_rows_into(std::span<float> out, std::size_t n) {
    if (n == 0) return 0;
    const std::size_t take = std::min(n, count_ - index_);
    if (out.size() < take * dim_) throw std::runtime_error("copy_rows_into buffer too small");
    std::size_t got = 0;
    if (vc_npy_read_rows(h_, index_, take, out.data(), &got) != VC_OK) throw std::runtime_error(vc_last_error());
    index_ += got;
    return got;
  }

 private:
  NpyReader(vc_npy* h, const char* label, std::size_t count, std::size_t dim)
      : h_(h), label_(label), count_(count), dim_(dim) {}
  vc_npy* h_ = nullptr;
  const char* label_ = "";
  std::size_t count_ = 0, dim_ = 0, index_ = 0;
};

/// Writes a float32 NPY matrix (rows x cols) to path (the reference writer's bytes).
inline void write_npy_f32_matrix(const std::filesystem::path& path, std::size_t rows, std::size_t cols,
                                 const float* data) {
  if (vc_npy_write_f32(path.c_str(), rows, cols, data) != VC_OK) throw std::runtime_error(vc_last_error());
}

}  // namespace vectorcache::datasets
