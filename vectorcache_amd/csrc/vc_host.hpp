// Host-side precompute for the TurboQuant path: rotation tables, Lloyd-Max
// codebook, TQ+ anchor probabilities. Everything here is bit-identical to the
// reference's host code (see vc_host.cpp for the per-function citations).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace vchip {

inline constexpr std::size_t kFastScanBlock = 192;   // reference kBlock
inline constexpr std::size_t kWindowGroups = 256;    // reference kFlushEvery
inline constexpr std::size_t kMaxDim = 16384;
inline constexpr float kMinInputNorm = 1e-10f;
inline constexpr double kDegenerateInnerEps = 1e-12;

struct RotationTables {
  std::size_t dim = 0;
  std::size_t block = 0;            // largest power of two dividing dim
  float inv_sqrt_block = 0.f;
  std::vector<std::uint32_t> perm;  // y[i] = x[perm[i]] ...
  std::vector<std::uint32_t> inv_perm;
  std::vector<float> signs;
};

struct Codebook {
  std::vector<float> boundaries;  // 2^b - 1
  std::vector<float> centroids;   // 2^b
};

// Throws std::invalid_argument on bad dim.
const RotationTables& rotation_tables(std::size_t dim);
// Throws std::invalid_argument on bad bits/dim. Memoized, thread-safe.
const Codebook& codebook(std::size_t bits, std::size_t dim);
// I_t(a, a), the regularized incomplete beta function.
double beta_cdf_symmetric(double a, double t);

// Quantizer cell table: a monotone cell function c(v) = clamp(trunc((v - lo) *
// inv_w), 0, 255) evaluated in float exactly as the device does; every cell
// holds at most one boundary, so code(v) = base[c] + (v > bound[c]) equals the
// reference's count of boundaries below v for every float v (NaN -> 0).
struct CellEntry {
  std::uint32_t base;
  float bound;   // +inf when the cell holds no boundary
  float c_lo;    // centroid[base]
  float c_hi;    // centroid[base + 1] (c_lo when base is the top code)
};
inline constexpr int kCells = 256;
struct CellTable {
  float lo = 0.f, inv_w = 0.f, off = 0.f;  // cell of v: floor(fmaf(v, inv_w, off)) clamped to [0, kCells)
  bool ok = false;  // false: the device counts boundaries one by one
  CellEntry cells[kCells];
};
const CellTable& cell_table(std::size_t bits, std::size_t dim);

struct TqAnchors {
  float qc_lo, qc_span;     // -c_outer, 2 c_outer
  std::size_t lo_idx, hi_idx;
};
// Order-statistic ranks and target span for TQ+ from n sample rows.
TqAnchors tq_anchors(std::size_t bits, std::size_t dim, std::size_t n);

inline std::size_t byte_groups(std::size_t bits, std::size_t dim) { return dim / (8 / bits); }

// ---- landmark basis (vc_landmark_host.cpp) ----------------------------------
inline constexpr std::size_t kDefaultLandmarks = 1024;

// Orthonormal landmark matrix L (dim x m floats, column j at L[j * dim]):
// seeded Gaussian matrix, modified Gram-Schmidt in double. Memoized per
// (dim, m, seed), thread-safe. Throws std::invalid_argument unless
// 0 < m <= dim, std::runtime_error on a degenerate column.
const std::vector<float>& ortho_landmarks(std::size_t dim, std::size_t m, std::uint64_t seed);

// Per-column absmax 4-bit quantization of L. codes: m columns of dim_pad / 2
// bytes (dim_pad = dim rounded up to 16), byte i of a column holds dims 2i
// (low nibble) and 2i + 1 (high nibble) as level + 7, levels in [-7, 7].
struct QuantizedLandmarks {
  std::vector<std::uint8_t> codes;
  std::vector<float> scales;  // m
  std::size_t dim = 0, dim_pad = 0, m = 0;
};
QuantizedLandmarks quantize_landmarks(const float* L, std::size_t dim, std::size_t m);

}  // namespace vchip
