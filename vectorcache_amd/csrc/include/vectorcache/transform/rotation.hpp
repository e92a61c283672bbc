// The reference's Rotation (include/vectorcache/transform/rotation.hpp:13-38)
// as the tables the GPU path uses: ChaCha8-seeded signs and Fisher-Yates
// permutation of one round, block size B = dim & -dim and 1/sqrt(B), read
// from libvectorcache_hip.so (bit-identical to src/transform/rotation.cpp:
// 139-160). The transform itself runs inside the GPU kernels; callers pass a
// Rotation to the free functions to name the dimension, as in the reference.
#pragma once

#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "vectorcache_hip.h"

namespace vectorcache {

class Rotation {
 public:
  explicit Rotation(std::size_t dim) : dim_(dim), perm_(dim), signs_(dim) {
    if (vc_rotation_tables(dim, perm_.data(), signs_.data(), &inv_sqrt_block_, &block_) != VC_OK)
      throw std::invalid_argument(vc_last_error());
  }

  std::size_t dim() const { return dim_; }
  std::size_t block() const { return block_; }
  float inv_sqrt_block() const { return inv_sqrt_block_; }
  /// Round r of kRotationRounds = 1 (reference constants.hpp:20).
  std::span<const float> signs(std::size_t r) const {
    if (r != 0) throw std::out_of_range("rotation round out of range");
    return signs_;
  }
  std::span<const std::uint32_t> perm(std::size_t r) const {
    if (r != 0) throw std::out_of_range("rotation round out of range");
    return perm_;
  }

 private:
  std::size_t dim_ = 0, block_ = 0;
  float inv_sqrt_block_ = 0.f;
  std::vector<std::uint32_t> perm_;
  std::vector<float> signs_;
};

}  // namespace vectorcache
