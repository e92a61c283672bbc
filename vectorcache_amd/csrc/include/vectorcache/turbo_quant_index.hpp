// C++20 host API of the MI355X TurboQuant path, over the C ABI of
// libvectorcache_hip.so (include/vectorcache_hip.h).
//
// Mirrors the reference's shapes so callers switch by changing the include:
//  * TurboQuantIndex: ctor(dim, bits), add(span), prepare(), search(span, k),
//    dim(), size(), bit_width(), codes_bytes(), index_bytes()
//    (reference include/vectorcache/index.hpp:18-78) plus calibrate(span)
//    (README.md:105-111; fit_calibration semantics, encode.hpp:33-35);
//  * SearchResults: row-major nq x k scores/ids, rows sorted by
//    (score desc, id asc) (reference include/vectorcache/search/search.hpp:77-84);
//  * score_perm0_into, the score_prepared_into of one bucket given as PERM0
//    blocks with the query rows (search.hpp:163-177); SearchResults, the
//    prepared-query free functions, heaps_to_search_results and
//    merge_search_results come from vectorcache/search/search.hpp.
// Exceptions follow the reference: std::invalid_argument for bad dim/bits or
// a span that is not a multiple of dim, std::runtime_error for search before
// prepare and for device failures, std::out_of_range for index errors.
#pragma once

#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vectorcache/search/search.hpp"
#include "vectorcache_hip.h"

namespace vectorcache {

class TurboQuantIndex {
 public:
  TurboQuantIndex(std::size_t dim, std::size_t bits) { detail::throw_on(vc_index_create(dim, bits, &h_)); }
  ~TurboQuantIndex() {
    if (h_ != nullptr) vc_index_destroy(h_);
  }
  TurboQuantIndex(const TurboQuantIndex&) = delete;
  TurboQuantIndex& operator=(const TurboQuantIndex&) = delete;
  TurboQuantIndex(TurboQuantIndex&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
  TurboQuantIndex& operator=(TurboQuantIndex&& o) noexcept {
    if (this != &o) {
      if (h_ != nullptr) vc_index_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
    }
    return *this;
  }

  std::size_t dim() const { return vc_index_dim(h_); }
  std::size_t bit_width() const { return vc_index_bits(h_); }
  std::size_t size() const { return vc_index_size(h_); }
  std::size_t codes_bytes() const { return vc_index_codes_bytes(h_); }
  std::size_t index_bytes() const { return vc_index_index_bytes(h_); }
  bool prepared() const { return vc_index_is_prepared(h_) != 0; }

  void calibrate(std::span<const float> sample) {
    check_rows(sample);
    detail::throw_on(vc_index_calibrate(h_, sample.data(), sample.size() / dim()));
  }
  void add(std::span<const float> vectors) {
    check_rows(vectors);
    detail::throw_on(vc_index_add(h_, vectors.data(), vectors.size() / dim()));
  }
  void prepare() { detail::throw_on(vc_index_prepare(h_)); }

  SearchResults search(std::span<const float> queries, std::size_t k) const {
    check_rows(queries);
    SearchResults r;
    r.nq = queries.size() / dim();
    const std::size_t kk = std::min(k, size());
    r.scores.resize_uninitialized(r.nq * kk);
    r.ids.resize_uninitialized(r.nq * kk);
    std::size_t k_out = 0;
    detail::throw_on(vc_index_search(h_, queries.data(), r.nq, k, r.scores.data(), r.ids.data(), &k_out));
    r.k = k_out;
    detail::narrow_rows(r, kk, k_out);
    return r;
  }

  vc_index* handle() const { return h_; }

 private:
  void check_rows(std::span<const float> v) const {
    if (v.size() % dim() != 0) throw std::invalid_argument("vectors size must be a multiple of dim");
  }
  vc_index* h_ = nullptr;
};

// score_prepared_into (reference search.hpp:163-177) for one bucket given as
// PERM0 blocks + per-vector scales: `queries` (nq x dim rows; TQ+ shift/scale
// or both empty) are the batch the reference's PreparedQueries would be built
// from. Only query_indices are scored, each into its heap (heap_s / heap_i
// [nq][k], heap_sz / heap_min / heap_mi [nq]), updated in place as the
// reference updates them; score_offsets (length nq) is added after alpha x LUT.
// Throws as the reference: std::invalid_argument for empty query_indices or
// an id_map of the wrong size, std::out_of_range for an index >= nq.
inline void score_perm0_into(std::span<const float> queries, std::size_t dim, std::size_t bits, std::size_t k,
                             std::span<const std::uint8_t> perm0, std::span<const float> scales,
                             std::span<const std::uint64_t> id_map, std::span<const std::size_t> query_indices,
                             float* heap_s, std::uint64_t* heap_i, std::size_t* heap_sz, float* heap_min,
                             std::size_t* heap_mi, const float* score_offsets = nullptr,
                             std::span<const float> shift = {}, std::span<const float> scale = {}) {
  if (dim == 0 || queries.size() % dim != 0) throw std::invalid_argument("queries size must be a multiple of dim");
  if (!id_map.empty() && id_map.size() != scales.size())
    throw std::invalid_argument("id_map size must equal number of vectors in range");
  detail::throw_on(vc_score_perm0_into(queries.data(), queries.size() / dim, dim, bits, k, perm0.data(),
                                       scales.data(), scales.size(), id_map.empty() ? nullptr : id_map.data(),
                                       shift.empty() ? nullptr : shift.data(), scale.empty() ? nullptr : scale.data(),
                                       query_indices.data(), query_indices.size(), heap_s, heap_i, heap_sz, heap_min,
                                       heap_mi, score_offsets));
}

}  // namespace vectorcache
