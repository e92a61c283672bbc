// The reference's search free functions over the C ABI of
// libvectorcache_hip.so (include/vectorcache_hip.h), with the reference's
// names, argument meaning and exceptions
// (reference include/vectorcache/search/search.hpp:18-192):
//  * UninitBuffer / SearchResults: row-major nq x k scores / ids, rows sorted
//    by (score desc, id asc), `counts` empty when every row has k hits;
//  * PreparedQueries + prepare_queries / prepared_query_at: the query state
//    built once on the GPU (rotation, TQ+, LUTs) and kept in device memory;
//  * score_prepared / score_prepared_into / heaps_to_search_results /
//    search_flat / merge_search_results over FastScan PERM0 blocks.
// Rotation tables and codebooks are fixed by (dim, bits), as in the
// reference; the Rotation and centroids arguments are checked against them.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vectorcache/quantize/codebook.hpp"
#include "vectorcache/transform/rotation.hpp"
#include "vectorcache_hip.h"

namespace vectorcache {

namespace detail {
inline void throw_on(vc_status st) {
  if (st == VC_OK) return;
  const std::string msg = vc_last_error();
  switch (st) {
    case VC_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case VC_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    default: throw std::runtime_error(msg);
  }
}

// The centroids a caller passes must be the library's codebook for (bits, dim).
inline void check_centroids(std::span<const float> centroids, std::size_t bits, std::size_t dim) {
  if (centroids.empty()) return;
  const auto cb = codebook(bits, dim);
  if (centroids.size() != cb.second.size() || !std::equal(centroids.begin(), centroids.end(), cb.second.begin()))
    throw std::invalid_argument("centroids must be codebook(bits, dim)");
}

inline std::size_t bits_of_centroids(std::size_t levels) {
  if (levels == 4) return 2;
  if (levels == 8) return 3;
  if (levels == 16) return 4;
  throw std::invalid_argument("bits must be 2, 3, or 4");
}
}  // namespace detail

/// Contiguous buffer that grows without initializing its elements (trivial T):
/// the reference's result storage, so callers that fill results themselves
/// (resize_uninitialized, then data()) compile unchanged.
template <typename T>
struct UninitBuffer {
  std::unique_ptr<T[]> data_;
  std::size_t size_ = 0;

  UninitBuffer() = default;
  UninitBuffer(UninitBuffer&&) noexcept = default;
  UninitBuffer& operator=(UninitBuffer&&) noexcept = default;
  UninitBuffer(const UninitBuffer& o) { copy_from(o); }
  UninitBuffer& operator=(const UninitBuffer& o) {
    if (this != &o) copy_from(o);
    return *this;
  }

  void clear() noexcept {
    data_.reset();
    size_ = 0;
  }
  void resize_uninitialized(std::size_t n) {
    if (n == 0) {
      clear();
    } else {
      data_ = std::make_unique_for_overwrite<T[]>(n);
      size_ = n;
    }
  }
  void assign(std::size_t n, const T& value) {
    resize_uninitialized(n);
    std::fill(begin(), end(), value);
  }

  T* data() noexcept { return data_.get(); }
  const T* data() const noexcept { return data_.get(); }
  T* begin() noexcept { return data_.get(); }
  const T* begin() const noexcept { return data_.get(); }
  T* end() noexcept { return data_.get() + size_; }
  const T* end() const noexcept { return data_.get() + size_; }
  std::size_t size() const noexcept { return size_; }
  bool empty() const noexcept { return size_ == 0; }
  T& operator[](std::size_t i) noexcept { return data_[i]; }
  const T& operator[](std::size_t i) const noexcept { return data_[i]; }

 private:
  void copy_from(const UninitBuffer& o) {
    resize_uninitialized(o.size_);
    if (o.size_ != 0) std::copy(o.begin(), o.end(), begin());
  }
};

struct SearchResults {
  UninitBuffer<float> scores;       // nq * k
  UninitBuffer<std::uint64_t> ids;  // nq * k
  std::size_t k = 0;
  std::size_t nq = 0;
  std::vector<std::size_t> counts;  // empty: every row has k hits
};

namespace detail {
// rows of `k_have` entries narrowed (or kept) to `k_want` entries each
inline void narrow_rows(SearchResults& r, std::size_t k_have, std::size_t k_want) {
  if (k_have == k_want) return;
  UninitBuffer<float> s;
  UninitBuffer<std::uint64_t> i;
  s.resize_uninitialized(r.nq * k_want);
  i.resize_uninitialized(r.nq * k_want);
  for (std::size_t q = 0; q < r.nq; ++q) {
    std::copy_n(r.scores.data() + q * k_have, k_want, s.data() + q * k_want);
    std::copy_n(r.ids.data() + q * k_have, k_want, i.data() + q * k_want);
  }
  r.scores = std::move(s);
  r.ids = std::move(i);
}
}  // namespace detail

/// Query-side state shared across one or more blocked code ranges (buckets),
/// held on the GPU.
class PreparedQueries {
 public:
  PreparedQueries() = default;
  explicit PreparedQueries(vc_prepared* h) : h_(h) { shape(); }
  ~PreparedQueries() {
    if (h_ != nullptr) vc_prepared_destroy(h_);
  }
  PreparedQueries(const PreparedQueries&) = delete;
  PreparedQueries& operator=(const PreparedQueries&) = delete;
  PreparedQueries(PreparedQueries&& o) noexcept
      : nq(o.nq), dim(o.dim), bits(o.bits), n_byte_groups(o.n_byte_groups), h_(std::exchange(o.h_, nullptr)) {}
  PreparedQueries& operator=(PreparedQueries&& o) noexcept {
    if (this != &o) {
      if (h_ != nullptr) vc_prepared_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
      nq = o.nq;
      dim = o.dim;
      bits = o.bits;
      n_byte_groups = o.n_byte_groups;
    }
    return *this;
  }

  std::size_t nq = 0;
  std::size_t dim = 0;
  std::size_t bits = 0;
  std::size_t n_byte_groups = 0;

  /// The per-query LUT state as prepare_queries builds it: nq x G*32 LUT bytes,
  /// and per query the LUT scale, LUT bias and bias_corr.
  void export_state(std::vector<std::uint8_t>& luts, std::vector<float>& lut_scale, std::vector<float>& lut_bias,
                    std::vector<float>& bias_corr) const {
    luts.resize(nq * n_byte_groups * 32);
    lut_scale.resize(nq);
    lut_bias.resize(nq);
    bias_corr.resize(nq);
    detail::throw_on(vc_prepared_export(h_, luts.data(), lut_scale.data(), lut_bias.data(), bias_corr.data()));
  }

  vc_prepared* handle() const { return h_; }

 private:
  void shape() { detail::throw_on(vc_prepared_shape(h_, &nq, &dim, &bits, &n_byte_groups)); }
  vc_prepared* h_ = nullptr;
};

inline PreparedQueries prepare_queries(std::span<const float> queries, std::size_t nq, std::size_t dim,
                                       const Rotation& rotation, std::span<const float> centroids, std::size_t bits,
                                       std::span<const float> tqplus_shift, std::span<const float> tqplus_scale) {
  if (rotation.dim() != dim) throw std::invalid_argument("prepare_queries: rotation dim mismatch");
  if (queries.size() < nq * dim) throw std::invalid_argument("prepare_queries: queries too short");
  if (tqplus_shift.empty() != tqplus_scale.empty() ||
      (!tqplus_shift.empty() && (tqplus_shift.size() != dim || tqplus_scale.size() != dim)))
    throw std::invalid_argument("prepare_queries: TQ+ shift / scale must both hold dim values");
  detail::check_centroids(centroids, bits, dim);
  vc_prepared* h = nullptr;
  detail::throw_on(vc_prepare_queries(queries.data(), nq, dim, bits,
                                      tqplus_shift.empty() ? nullptr : tqplus_shift.data(),
                                      tqplus_scale.empty() ? nullptr : tqplus_scale.data(), &h));
  return PreparedQueries(h);
}

/// Copy prepared state for a single query index (for per-query IVF probing).
inline PreparedQueries prepared_query_at(const PreparedQueries& prep, std::size_t qi) {
  vc_prepared* h = nullptr;
  detail::throw_on(vc_prepared_query_at(prep.handle(), qi, &h));
  return PreparedQueries(h);
}

/// Score a range of PERM0 blocks (n_blocks x G x 192 bytes; scales.size()
/// vectors) with prepared queries; the arguments of the reference's AVX2 path.
inline SearchResults score_prepared(const PreparedQueries& prep, std::size_t k,
                                    std::span<const std::uint8_t> blocked_codes, std::size_t n_blocks,
                                    std::span<const float> scales, std::span<const std::uint64_t> id_map = {},
                                    std::span<const std::size_t> query_indices = {},
                                    std::span<const float> residual_norms = {},
                                    std::span<const float> residual_scales = {}) {
  const std::size_t n = scales.size();
  if (blocked_codes.size() < n_blocks * prep.n_byte_groups * 192 || n_blocks * 192 < n)
    throw std::invalid_argument("score_prepared: blocked codes too short");
  if (!id_map.empty() && id_map.size() != n) throw std::invalid_argument("id_map size must equal scales size");
  SearchResults r;
  r.nq = query_indices.empty() ? prep.nq : query_indices.size();
  const std::size_t kk = std::min(k, n);
  r.scores.resize_uninitialized(r.nq * kk);
  r.ids.resize_uninitialized(r.nq * kk);
  std::size_t k_out = 0;
  const bool resid = !residual_norms.empty() && !residual_scales.empty();
  detail::throw_on(vc_score_prepared(prep.handle(), k, blocked_codes.data(), scales.data(), n,
                                     id_map.empty() ? nullptr : id_map.data(),
                                     query_indices.empty() ? nullptr : query_indices.data(), query_indices.size(),
                                     resid ? residual_norms.data() : nullptr, resid ? residual_scales.data() : nullptr,
                                     r.scores.data(), r.ids.data(), &k_out));
  r.k = k_out;
  detail::narrow_rows(r, kk, k_out);
  return r;
}

/// Continue scoring into existing per-query top-k heaps (indexed by global
/// qi); only query_indices are updated; score_offsets[qi] is added after
/// alpha x LUT.
inline void score_prepared_into(const PreparedQueries& prep, std::size_t k,
                                std::span<const std::uint8_t> blocked_codes, std::size_t n_blocks,
                                std::span<const float> scales, std::span<const std::uint64_t> id_map,
                                std::span<const std::size_t> query_indices, float* heap_s, std::uint64_t* heap_i,
                                std::size_t* heap_sz, float* heap_min, std::size_t* heap_mi,
                                const float* score_offsets = nullptr) {
  const std::size_t n = scales.size();
  if (blocked_codes.size() < n_blocks * prep.n_byte_groups * 192 || n_blocks * 192 < n)
    throw std::invalid_argument("score_prepared_into: blocked codes too short");
  if (!id_map.empty() && id_map.size() != n) throw std::invalid_argument("id_map size must equal scales size");
  detail::throw_on(vc_score_prepared_into(prep.handle(), k, blocked_codes.data(), scales.data(), n,
                                          id_map.empty() ? nullptr : id_map.data(), query_indices.data(),
                                          query_indices.size(), heap_s, heap_i, heap_sz, heap_min, heap_mi,
                                          score_offsets));
}

/// Sort unordered top-k heaps into rows (score desc, id asc); the rest of each
/// row is (0, 0).
inline void heaps_to_search_results(SearchResults& out, std::size_t nq, std::size_t k, const float* heap_s,
                                    const std::uint64_t* heap_i, const std::size_t* heap_sz) {
  out.nq = nq;
  out.k = k;
  out.scores.assign(nq * k, 0.f);
  out.ids.assign(nq * k, 0);
  out.counts.clear();
  if (nq == 0 || k == 0) return;
  detail::throw_on(vc_heaps_to_results(nq, k, heap_s, heap_i, heap_sz, out.scores.data(), out.ids.data()));
}

/// prepare_queries + score_prepared over one range of PERM0 blocks.
inline SearchResults search_flat(std::span<const float> queries, std::size_t nq, std::size_t dim, std::size_t k,
                                 const Rotation& rotation, std::span<const float> centroids, std::size_t bits,
                                 std::span<const std::uint8_t> blocked_codes, std::size_t n_blocks,
                                 std::span<const float> scales, std::span<const float> tqplus_shift,
                                 std::span<const float> tqplus_scale) {
  const PreparedQueries prep =
      prepare_queries(queries, nq, dim, rotation, centroids, bits, tqplus_shift, tqplus_scale);
  return score_prepared(prep, k, blocked_codes, n_blocks, scales);
}

/// Keep the top dst.k of dst U src per query (higher score wins, tie -> lower
/// id, a duplicated id keeps its higher score).
inline void merge_search_results(SearchResults& dst, const SearchResults& src) {
  if (src.nq == 0 || src.k == 0) return;
  if (dst.nq == 0 || dst.k == 0) {
    dst = src;
    return;
  }
  if (dst.nq != src.nq) throw std::invalid_argument("merge_search_results: nq mismatch");
  detail::throw_on(vc_merge_results(dst.nq, dst.k, dst.scores.data(), dst.ids.data(), src.k, src.scores.data(),
                                    src.ids.data()));
}

}  // namespace vectorcache
