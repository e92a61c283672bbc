// The README's BucketedTurboQuantIndex and BucketParams (reference README.md:
// 56,103-121, which includes this header path; the reference has no code for
// it) over the C ABI of libvectorcache_hip.so (vc_bucketed_*): L2-normalized
// rows grouped into buckets by a spherical k-means coarse quantizer with hard
// fission at max_bucket_size, each bucket's rows encoded with the flat
// index's encode (residuals x_hat - c against frozen encode centroids), a
// query opening its n_probe best buckets by routing score and ranking their
// vectors by <q_hat, c> + alpha <q_hat, r_hat>. Each opened bucket is scored
// as the reference's score_prepared_into scores one bucket; build, routing
// and the scan run on the GPU. Exceptions as TurboQuantIndex's.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <utility>

#include "vectorcache/search/search.hpp"
#include "vectorcache_hip.h"

namespace vectorcache {

struct BucketParams {
  std::size_t n_probe = 64;           // buckets a query opens
  std::size_t max_bucket_size = 800;  // hard cap (fission by spherical 2-means)
  std::size_t energy_soft = 400;      // mean bucket size the coarse quantizer aims at
  std::size_t expected_n = 0;         // corpus size hint (capacity)
  std::size_t kmeans_iters = 8;       // Lloyd iterations of the coarse quantizer
  std::size_t kmeans_sample = 64;     // training rows per list
  std::uint64_t seed = 4001;
  bool residual = true;               // residual codes with score offsets (false: codes of the rows)
};

class BucketedTurboQuantIndex {
 public:
  BucketedTurboQuantIndex(std::size_t dim, std::size_t bits, const BucketParams& p = {}) : params_(p) {
    vc_bucket_params c{};
    vc_bucket_params_default(&c);
    c.n_probe = p.n_probe;
    c.max_bucket_size = p.max_bucket_size;
    c.energy_soft = p.energy_soft;
    c.expected_n = p.expected_n;
    c.kmeans_iters = p.kmeans_iters;
    c.kmeans_sample = p.kmeans_sample;
    c.seed = p.seed;
    c.residual = p.residual ? 1 : 0;
    detail::throw_on(vc_bucketed_create(dim, bits, &c, &h_));
  }
  ~BucketedTurboQuantIndex() {
    if (h_ != nullptr) vc_bucketed_destroy(h_);
  }
  BucketedTurboQuantIndex(const BucketedTurboQuantIndex&) = delete;
  BucketedTurboQuantIndex& operator=(const BucketedTurboQuantIndex&) = delete;
  BucketedTurboQuantIndex(BucketedTurboQuantIndex&& o) noexcept
      : params_(o.params_), h_(std::exchange(o.h_, nullptr)) {}
  BucketedTurboQuantIndex& operator=(BucketedTurboQuantIndex&& o) noexcept {
    if (this != &o) {
      if (h_ != nullptr) vc_bucketed_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
      params_ = o.params_;
    }
    return *this;
  }

  std::size_t dim() const { return vc_bucketed_dim(h_); }
  std::size_t bit_width() const { return vc_bucketed_bits(h_); }
  std::size_t size() const { return vc_bucketed_size(h_); }
  std::size_t n_buckets() const { return vc_bucketed_n_buckets(h_); }
  bool prepared() const { return vc_bucketed_is_prepared(h_) != 0; }
  const BucketParams& params() const { return params_; }

  void calibrate(std::span<const float> sample) {
    check_rows(sample);
    detail::throw_on(vc_bucketed_calibrate(h_, sample.data(), sample.size() / dim()));
  }
  void add(std::span<const float> vectors) {
    check_rows(vectors);
    detail::throw_on(vc_bucketed_add(h_, vectors.data(), vectors.size() / dim()));
  }
  void prepare() { detail::throw_on(vc_bucketed_prepare(h_)); }

  /// Top k of the opened buckets' vectors per query, by (score desc, id asc);
  /// a row with fewer than k opened vectors ends with (0, 0) entries.
  SearchResults search(std::span<const float> queries, std::size_t k) const {
    check_rows(queries);
    SearchResults r;
    r.nq = queries.size() / dim();
    const std::size_t kk = std::min(k, size());
    r.scores.resize_uninitialized(r.nq * kk);
    r.ids.resize_uninitialized(r.nq * kk);
    std::size_t k_out = 0;
    detail::throw_on(vc_bucketed_search(h_, queries.data(), r.nq, k, r.scores.data(), r.ids.data(), &k_out));
    r.k = k_out;
    detail::narrow_rows(r, kk, k_out);
    return r;
  }

  vc_bucketed* handle() const { return h_; }

 private:
  void check_rows(std::span<const float> v) const {
    if (v.size() % dim() != 0) throw std::invalid_argument("vectors size must be a multiple of dim");
  }
  BucketParams params_;
  vc_bucketed* h_ = nullptr;
};

}  // namespace vectorcache
