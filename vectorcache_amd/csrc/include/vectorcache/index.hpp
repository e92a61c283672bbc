// The reference's LandmarkIndex (include/vectorcache/index.hpp:16-68,
// src/index.cpp:120-289) over the C ABI of libvectorcache_hip.so: the same
// constructor arguments, accessors, exceptions and search() signature. Every
// step runs on the GPU; results equal the reference's AVX2 path bit for bit.
//
// The reference aliases TurboQuantIndex to this class; here the plain flat
// index keeps the name TurboQuantIndex (vectorcache/turbo_quant_index.hpp),
// so a caller that used the alias spells LandmarkIndex.
#pragma once

#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <utility>

#include "vectorcache/turbo_quant_index.hpp"
#include "vectorcache_hip.h"

namespace vectorcache {

inline constexpr std::size_t kDefaultLandmarks = 1024;  // reference include/vectorcache/constants.hpp:28

class LandmarkIndex {
 public:
  LandmarkIndex(std::size_t dim, std::size_t scan_bits, std::size_t n_landmarks = kDefaultLandmarks,
                std::uint64_t landmark_seed = 0, std::size_t rerank_bits = 0) {
    detail::throw_on(vc_lm_index_create(dim, scan_bits, n_landmarks, landmark_seed, rerank_bits, &h_));
  }
  ~LandmarkIndex() {
    if (h_ != nullptr) vc_lm_index_destroy(h_);
  }
  LandmarkIndex(const LandmarkIndex&) = delete;
  LandmarkIndex& operator=(const LandmarkIndex&) = delete;
  LandmarkIndex(LandmarkIndex&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
  LandmarkIndex& operator=(LandmarkIndex&& o) noexcept {
    if (this != &o) {
      if (h_ != nullptr) vc_lm_index_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
    }
    return *this;
  }

  std::size_t dim() const { return vc_lm_index_dim(h_); }
  std::size_t bit_width() const { return vc_lm_index_scan_bits(h_); }
  std::size_t scan_bits() const { return vc_lm_index_scan_bits(h_); }
  std::size_t rerank_bits() const { return vc_lm_index_rerank_bits(h_); }
  std::size_t n_landmarks() const { return vc_lm_index_n_landmarks(h_); }
  std::size_t scan_dim() const { return vc_lm_index_scan_dim(h_); }
  std::size_t size() const { return vc_lm_index_size(h_); }
  std::size_t codes_bytes() const { return vc_lm_index_codes_bytes(h_); }
  std::size_t index_bytes() const { return vc_lm_index_index_bytes(h_); }

  void add(std::span<const float> vectors) {
    if (vectors.size() % dim() != 0) throw std::invalid_argument("vectors length must be multiple of dim");
    detail::throw_on(vc_lm_index_add(h_, vectors.data(), vectors.size() / dim()));
  }
  void prepare() { detail::throw_on(vc_lm_index_prepare(h_)); }
  /// The reference's behaviour on an AVX-512 VNNI host: sign planes from
  /// 32768 prepared vectors and the staged sign-plane search.
  void set_planes(bool on) { detail::throw_on(vc_lm_index_set_planes(h_, on ? 1 : 0)); }
  bool planes_active() const { return vc_lm_index_planes_active(h_) != 0; }

  SearchResults search(std::span<const float> queries, std::size_t k,
                       std::span<const float> float_corpus = {}) const {
    if (queries.size() % dim() != 0) throw std::invalid_argument("queries length must be multiple of dim");
    if (!float_corpus.empty() && float_corpus.size() < size() * dim())
      throw std::invalid_argument("float_corpus must have at least size() * dim rows");
    SearchResults r;
    r.nq = queries.size() / dim();
    const std::size_t kk = std::min(k, size());
    r.scores.resize_uninitialized(r.nq * kk);
    r.ids.resize_uninitialized(r.nq * kk);
    std::size_t k_out = 0;
    detail::throw_on(vc_lm_index_search(h_, queries.data(), r.nq, k,
                                        float_corpus.empty() ? nullptr : float_corpus.data(), r.scores.data(),
                                        r.ids.data(), &k_out));
    r.k = k_out;
    detail::narrow_rows(r, kk, k_out);
    return r;
  }

  vc_lm_index* handle() const { return h_; }

 private:
  vc_lm_index* h_ = nullptr;
};

}  // namespace vectorcache
