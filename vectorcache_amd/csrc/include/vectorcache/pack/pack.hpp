// The reference's pack free functions (include/vectorcache/pack/pack.hpp:
// 14-56) over the C ABI of libvectorcache_hip.so: bit-plane codes -> flat rows,
// PERM0 or sequential 192 x 1 blocks, repacked on the GPU; blocked geometry
// and single-code reads on the host. repack() is the reference's AVX2 host
// choice, PERM0 (the sequential layout is its VNNI-host choice).
#pragma once

#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <utility>
#include <vector>

#include "vectorcache/search/search.hpp"
#include "vectorcache_hip.h"

namespace vectorcache {

inline constexpr std::size_t kBlock = 192;       // vectors per FastScan block (reference constants.hpp:7)
inline constexpr std::size_t kFlushChunk = 32;   // PERM0 sub-tile (constants.hpp:9)
// PERM0 order of a 32-lane sub-tile's bytes (reference pack.hpp:14-15)
inline constexpr std::size_t kPerm0[16] = {0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15};
inline constexpr std::size_t kInvPerm0[16] = {0, 2, 4, 6, 8, 10, 12, 14, 1, 3, 5, 7, 9, 11, 13, 15};

struct BlockedGeometry {
  std::size_t n_blocks;
  std::size_t n_byte_groups;
  std::size_t blocked_len;
};

inline BlockedGeometry blocked_geometry(std::size_t n_vectors, std::size_t bits, std::size_t dim) {
  if (bits < 2 || bits > 4) throw std::invalid_argument("bits must be 2, 3, or 4");
  const std::size_t g = dim / (8 / bits), nb = (n_vectors + kBlock - 1) / kBlock;
  return {nb, g, nb * g * kBlock};
}

namespace detail {
inline void check_packed(std::span<const std::uint8_t> packed, std::size_t n, std::size_t bits, std::size_t dim) {
  if (packed.size() < n * bits * dim / 8) throw std::invalid_argument("packed codes too short");
}
}  // namespace detail

/// Bit-plane codes -> row-major flat codes (n_vectors x n_byte_groups).
inline std::vector<std::uint8_t> codes_flat_from_packed(std::span<const std::uint8_t> packed_codes,
                                                        std::size_t n_vectors, std::size_t bits, std::size_t dim) {
  detail::check_packed(packed_codes, n_vectors, bits, dim);
  const auto geo = blocked_geometry(n_vectors, bits, dim);
  std::vector<std::uint8_t> out(n_vectors * geo.n_byte_groups);
  detail::throw_on(vc_codes_flat_from_packed(packed_codes.data(), n_vectors, dim, bits, out.data()));
  return out;
}

/// Always the AVX2 PERM0 blocked layout; returns (bytes, n_blocks).
inline std::pair<std::vector<std::uint8_t>, std::size_t> repack_perm0(std::span<const std::uint8_t> packed_codes,
                                                                      std::size_t n_vectors, std::size_t bits,
                                                                      std::size_t dim) {
  detail::check_packed(packed_codes, n_vectors, bits, dim);
  const auto geo = blocked_geometry(n_vectors, bits, dim);
  std::vector<std::uint8_t> out(geo.blocked_len);
  detail::throw_on(vc_repack_perm0(packed_codes.data(), n_vectors, dim, bits, out.data()));
  return {std::move(out), geo.n_blocks};
}

/// Native search layout of an AVX2 host: PERM0; returns (bytes, n_blocks).
inline std::pair<std::vector<std::uint8_t>, std::size_t> repack(std::span<const std::uint8_t> packed_codes,
                                                                std::size_t n_vectors, std::size_t bits,
                                                                std::size_t dim) {
  return repack_perm0(packed_codes, n_vectors, bits, dim);
}

/// Sequential 192 x 1 blocked layout.
inline std::vector<std::uint8_t> repack_seq(std::span<const std::uint8_t> packed_codes, std::size_t n_vectors,
                                            std::size_t bits, std::size_t dim) {
  detail::check_packed(packed_codes, n_vectors, bits, dim);
  const auto geo = blocked_geometry(n_vectors, bits, dim);
  std::vector<std::uint8_t> out(geo.blocked_len);
  detail::throw_on(vc_repack_seq(packed_codes.data(), n_vectors, dim, bits, out.data()));
  return out;
}

/// Byte group g of vector (block_idx, lane) from PERM0 blocks: in each
/// 32-lane sub-tile, byte j holds the high nibbles of lanes kPerm0[j] (bits
/// 0-3) and kPerm0[j] + 16 (bits 4-7), byte 16 + j their low nibbles
/// (reference pack.cpp:157-203,403-420 on an AVX2 host).
inline std::uint8_t read_code(std::span<const std::uint8_t> blocked, std::size_t bits, std::size_t n_byte_groups,
                              std::size_t block_idx, std::size_t g, std::size_t lane) {
  (void)bits;
  const std::size_t base = block_idx * n_byte_groups * kBlock + g * kBlock + (lane / kFlushChunk) * kFlushChunk;
  const std::size_t l = lane % kFlushChunk, j = kInvPerm0[l & 15];
  const std::uint8_t hi_plane = blocked[base + j], lo_plane = blocked[base + 16 + j];
  const unsigned shift = l < 16 ? 0 : 4;
  return static_cast<std::uint8_t>((((hi_plane >> shift) & 0x0F) << 4) | ((lo_plane >> shift) & 0x0F));
}

}  // namespace vectorcache
