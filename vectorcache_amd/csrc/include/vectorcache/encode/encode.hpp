// The reference's encode free functions (include/vectorcache/encode/
// encode.hpp:14-35) over the C ABI of libvectorcache_hip.so: the encode runs on
// the GPU (norm, rotation, Lloyd-Max quantization, alpha, bit-plane codes) and
// equals the reference's CPU encode bit for bit; outputs are appended to the
// caller's vectors, as the reference appends them.
#pragma once

#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <vector>

#include "vectorcache/search/search.hpp"
#include "vectorcache/transform/rotation.hpp"
#include "vectorcache_hip.h"

namespace vectorcache {

struct Calibration {
  std::vector<float> shift;
  std::vector<float> scale_tq;
};

/// vectors (n x dim) -> packed_out += n x bit_width*dim/8 bit-plane bytes,
/// scales_out += n alphas. boundaries / centroids must be codebook(bit_width,
/// dim); rotated_scratch is not needed on the GPU path and is left as given.
inline void encode(std::span<const float> vectors, std::size_t n, std::size_t dim, const Rotation& rotation,
                   std::span<const float> boundaries, std::span<const float> centroids, std::size_t bit_width,
                   const Calibration* calibration, std::vector<float>& rotated_scratch,
                   std::vector<std::uint8_t>& packed_out, std::vector<float>& scales_out) {
  (void)rotated_scratch;
  if (dim == 0 || dim % 8 != 0) throw std::invalid_argument("encode requires dim multiple of 8");
  if (bit_width < 2 || bit_width > 4) throw std::invalid_argument("bit_width must be 2, 3, or 4");
  if (rotation.dim() != dim) throw std::invalid_argument("encode: rotation dim mismatch");
  if (vectors.size() < n * dim) throw std::invalid_argument("encode: vectors too short");
  detail::check_centroids(centroids, bit_width, dim);
  if (!boundaries.empty()) {
    const auto cb = codebook(bit_width, dim);
    if (boundaries.size() != cb.first.size() || !std::equal(boundaries.begin(), boundaries.end(), cb.first.begin()))
      throw std::invalid_argument("boundaries must be codebook(bits, dim)");
  }
  const float* sh = nullptr;
  const float* sc = nullptr;
  if (calibration != nullptr) {
    if (calibration->shift.size() != dim || calibration->scale_tq.size() != dim)
      throw std::invalid_argument("encode: calibration must hold dim shifts and scales");
    sh = calibration->shift.data();
    sc = calibration->scale_tq.data();
  }
  const std::size_t bpr = bit_width * dim / 8, p0 = packed_out.size(), s0 = scales_out.size();
  packed_out.resize(p0 + n * bpr);
  scales_out.resize(s0 + n);
  detail::throw_on(
      vc_encode(vectors.data(), n, dim, bit_width, sh, sc, packed_out.data() + p0, scales_out.data() + s0));
}

/// TQ+ calibration (per-dimension shift / scale from order statistics of the
/// rotated sample) on the GPU; bits follow from centroids.size().
inline Calibration fit_calibration(std::span<const float> vectors, std::size_t n, std::size_t dim,
                                   const Rotation& rotation, std::span<const float> centroids,
                                   std::vector<float>& rotated_scratch) {
  (void)rotated_scratch;
  if (n < 2) throw std::invalid_argument("calibration needs at least 2 rows");
  if (rotation.dim() != dim) throw std::invalid_argument("fit_calibration: rotation dim mismatch");
  if (vectors.size() < n * dim) throw std::invalid_argument("fit_calibration: vectors too short");
  const std::size_t bits = detail::bits_of_centroids(centroids.size());
  detail::check_centroids(centroids, bits, dim);
  Calibration c;
  c.shift.resize(dim);
  c.scale_tq.resize(dim);
  detail::throw_on(vc_fit_calibration(vectors.data(), n, dim, bits, c.shift.data(), c.scale_tq.data()));
  return c;
}

}  // namespace vectorcache
