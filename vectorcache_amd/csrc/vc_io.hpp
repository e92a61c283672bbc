// On-disk formats of the reference's recall runs (host code):
//  * NPY float32 matrices, read through mmap (include/vectorcache/datasets/
//    npy.hpp:33-65, src/datasets/npy.cpp) and written with the reference's
//    header layout (write_npy_f32_matrix, src/datasets/npy.cpp:336-366);
//  * the VCEX exact top-k cache (tools/query-bench.cpp:146-340).
// Errors are thrown as IoError (a std::runtime_error, like the reference's
// vectorcache::Error) with the reference's messages.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace vcio {

struct IoError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// A read-only mapping of a float32 (rows, dim) NPY file.
class NpyMatrix {
 public:
  explicit NpyMatrix(const std::string& path);
  ~NpyMatrix();
  NpyMatrix(const NpyMatrix&) = delete;
  NpyMatrix& operator=(const NpyMatrix&) = delete;

  std::size_t count() const { return count_; }
  std::size_t dim() const { return dim_; }
  // row-major count x dim floats inside the mapping
  const float* data() const;
  // copies rows [first, first + n) clipped to count(); returns the rows copied
  std::size_t copy_rows(std::size_t first, std::size_t n, float* out) const;

 private:
  void* map_ = nullptr;
  std::size_t size_ = 0;
  std::size_t count_ = 0, dim_ = 0, offset_ = 0;
};

// Parses the header of an in-memory NPY float32 matrix (exposed for tests):
// count, dim and the byte offset of the data.
void parse_npy_f32(const char* data, std::size_t size, std::size_t& count, std::size_t& dim,
                   std::size_t& offset);

void write_npy_f32(const std::string& path, std::size_t rows, std::size_t cols, const float* data);

struct ExactTopkKey {
  std::string source_label, split, query_split;
  std::uint64_t index_n = 0, query_n = 0, k = 0, seed = 0;
};

// "npy-<stem>-<FNV-1a 64 of the normalized path, hex>", or `dataset` when the
// path is empty (cache_source_label, tools/query-bench.cpp:166-178)
std::string cache_source_label(const std::string& npy_path, const std::string& dataset);
// .cache/exact_topk/<label>_<split>_n<N>_qs-<qsplit>_ql<nq>_k<k>_seed<s>.bin
std::string exact_topk_cache_path(const ExactTopkKey& key);
// false when the file is missing, not VCEX version 1, or made for another key
bool load_exact_topk(const std::string& path, const ExactTopkKey& key, std::vector<std::uint64_t>& ids);
// query_n x k ids, written to path.tmp and renamed; false on any failure
bool save_exact_topk(const std::string& path, const ExactTopkKey& key, const std::uint64_t* ids);

std::uint64_t fnv1a64(const std::string& s);

}  // namespace vcio
