// The reference's codebook() (include/vectorcache/quantize/codebook.hpp:8-12):
// Lloyd-Max levels of Beta((dim-1)/2, (dim-1)/2) on [-1, 1] as f32 centroids,
// boundaries the f32 midpoints; computed by libvectorcache_hip.so's host code,
// bit-identical to src/quantize/codebook.cpp:135-226.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <utility>
#include <vector>

#include "vectorcache_hip.h"

namespace vectorcache {

/// Returns (boundaries, centroids).
inline std::pair<std::vector<float>, std::vector<float>> codebook(std::size_t bits, std::size_t dim) {
  if (bits < 2 || bits > 4) throw std::invalid_argument("bits must be 2, 3, or 4");
  std::pair<std::vector<float>, std::vector<float>> cb;
  cb.first.resize((std::size_t(1) << bits) - 1);
  cb.second.resize(std::size_t(1) << bits);
  if (vc_codebook(bits, dim, cb.first.data(), cb.second.data()) != VC_OK) throw std::invalid_argument(vc_last_error());
  return cb;
}

}  // namespace vectorcache
