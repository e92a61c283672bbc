// Host precompute, compiled with -ffp-contract=off: every fused multiply-add
// the reference's compiled code performs (it builds with -mfma
// -ffp-contract=fast) is an explicit std::fma here, so results are
// bit-identical to it on any x86-64 host.
#include "vc_host.hpp"

#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <utility>

namespace vchip {
namespace {

// ---- ChaCha8 keystream (reference src/chacha8.cpp:12-85) -----------------
// IETF state layout, 64-bit block counter in words 12-13, zero nonce,
// 4 double rounds, output = permuted state + input state.
class ChaChaStream {
 public:
  explicit ChaChaStream(const std::array<std::uint8_t, 32>& key) {
    in_ = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int w = 0; w < 8; ++w) {
      in_[4 + w] = std::uint32_t(key[4 * w]) | std::uint32_t(key[4 * w + 1]) << 8 |
                   std::uint32_t(key[4 * w + 2]) << 16 | std::uint32_t(key[4 * w + 3]) << 24;
    }
  }

  std::uint32_t u32() {
    if (used_ == 16) next_block();
    return out_[used_++];
  }
  std::uint64_t u64() {
    const std::uint64_t lo = u32();
    return lo | (std::uint64_t(u32()) << 32);
  }

 private:
  static std::uint32_t rot(std::uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }
  static void quarter(std::array<std::uint32_t, 16>& s, int a, int b, int c, int d) {
    s[a] += s[b]; s[d] = rot(s[d] ^ s[a], 16);
    s[c] += s[d]; s[b] = rot(s[b] ^ s[c], 12);
    s[a] += s[b]; s[d] = rot(s[d] ^ s[a], 8);
    s[c] += s[d]; s[b] = rot(s[b] ^ s[c], 7);
  }
  void next_block() {
    std::array<std::uint32_t, 16> s = in_;
    for (int dr = 0; dr < 4; ++dr) {
      quarter(s, 0, 4, 8, 12); quarter(s, 1, 5, 9, 13);
      quarter(s, 2, 6, 10, 14); quarter(s, 3, 7, 11, 15);
      quarter(s, 0, 5, 10, 15); quarter(s, 1, 6, 11, 12);
      quarter(s, 2, 7, 8, 13); quarter(s, 3, 4, 9, 14);
    }
    for (int i = 0; i < 16; ++i) out_[i] = s[i] + in_[i];
    if (++in_[12] == 0) ++in_[13];
    used_ = 0;
  }
  std::array<std::uint32_t, 16> in_{};
  std::array<std::uint32_t, 16> out_{};
  int used_ = 16;
};

// The rotation key is a format constant: codes are only interchangeable with
// the reference's when the same key seeds the rotation (rotation.cpp:18-20).
constexpr std::array<std::uint8_t, 32> kRotationKey = {
    164, 143, 161, 123, 88, 50, 61, 10, 234, 184, 161, 204, 105, 1, 20, 184,
    43, 140, 200, 117, 24, 180, 247, 84, 141, 68, 110, 161, 228, 223, 32, 242};

std::unique_ptr<RotationTables> make_rotation(std::size_t dim) {
  auto t = std::make_unique<RotationTables>();
  t->dim = dim;
  ChaChaStream rng(kRotationKey);
  t->signs.resize(dim);
  for (auto& s : t->signs) s = (rng.u32() & 1u) ? -1.f : 1.f;  // signs first
  t->perm.resize(dim);
  for (std::size_t i = 0; i < dim; ++i) t->perm[i] = std::uint32_t(i);
  for (std::size_t i = dim - 1; i > 0; --i) {  // Fisher-Yates, top down
    const std::size_t j = std::size_t(rng.u64() % (std::uint64_t(i) + 1));
    std::swap(t->perm[i], t->perm[j]);
  }
  t->inv_perm.resize(dim);
  for (std::size_t i = 0; i < dim; ++i) t->inv_perm[t->perm[i]] = std::uint32_t(i);
  t->block = dim & (~dim + 1);
  t->inv_sqrt_block = 1.f / std::sqrt(float(t->block));
  return t;
}

// ---- Beta((d-1)/2, (d-1)/2) numerics (reference src/quantize/codebook.cpp) --
// Lanczos log-gamma, g = 7, n = 9 (published coefficients; identical values
// are required for identical codebooks).
constexpr double kLanczosG7[9] = {0.99999999999980993,  676.5203681218851,   -1259.1392167224028,
                                  771.32342877765313,   -176.61502916214059, 12.507343278686905,
                                  -0.13857109526572012, 9.984369654078861e-6, 1.5056327351493116e-7};
constexpr double kPi = 3.14159265358979323846;

double log_gamma(double z) {
  if (z < 0.5) return std::log(kPi) - std::log(std::sin(kPi * z)) - log_gamma(1.0 - z);
  z -= 1.0;
  double series = kLanczosG7[0];
  for (int i = 1; i < 9; ++i) series += kLanczosG7[i] / (z + double(i));
  const double t = z + 7.5;
  return std::fma(z + 0.5, std::log(t), 0.5 * std::log(2.0 * kPi)) - t + std::log(series);
}

double guard(double v) { return std::fabs(v) < 1e-30 ? 1e-30 : v; }

// Continued fraction for the incomplete beta (modified Lentz).
double beta_cf(double a, double b, double x) {
  const double apb = a + b, ap1 = a + 1.0, am1 = a - 1.0;
  double c = 1.0;
  double d = 1.0 / guard(1.0 - apb * x / ap1);
  double h = d;
  for (int m = 1; m <= 200; ++m) {
    const double two_m = 2.0 * m;
    double num = double(m) * (b - double(m)) * x / ((am1 + two_m) * (a + two_m));
    d = 1.0 / guard(std::fma(num, d, 1.0));
    c = guard(1.0 + num / c);
    h *= d * c;
    num = -(a + double(m)) * (apb + double(m)) * x / ((a + two_m) * (ap1 + two_m));
    d = 1.0 / guard(std::fma(num, d, 1.0));
    c = guard(1.0 + num / c);
    const double step = d * c;
    h *= step;
    if (std::fabs(step - 1.0) < 3e-14) break;
  }
  return h;
}

double incomplete_beta(double a, double b, double x) {
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  const double log_b = log_gamma(a) + log_gamma(b) - log_gamma(a + b);
  const double pre = std::exp(std::fma(std::log(x), a, std::log(1.0 - x) * b) - log_b);
  if (x < (a + 1.0) / (a + b + 2.0)) return pre / a * beta_cf(a, b, x);
  return std::fma(-(pre / b), beta_cf(b, a, 1.0 - x), 1.0);
}

struct BetaDensity {
  double a, log_norm;
  double operator()(double t) const {
    if (t <= 0.0 || t >= 1.0) return 0.0;
    return std::exp(std::fma(a - 1.0, std::log(t), (a - 1.0) * std::log(1.0 - t)) - log_norm);
  }
};

// first moment integrand on [-1, 1]: x * f((x + 1) / 2) / 2
struct Moment {
  const BetaDensity* f;
  double operator()(double x) const { return x * (*f)((x + 1.0) / 2.0) / 2.0; }
};

double simpson_step(const Moment& g, double lo, double hi, double glo, double ghi, double gmid,
                    double whole, double tol, int depth) {
  const double mid = (lo + hi) / 2.0;
  const double q1 = (lo + mid) / 2.0, q3 = (mid + hi) / 2.0;
  const double g1 = g(q1), g3 = g(q3);
  const double left = (mid - lo) / 6.0 * (std::fma(4.0, g1, glo) + gmid);
  const double right = (hi - mid) / 6.0 * (std::fma(4.0, g3, gmid) + ghi);
  const double both = left + right;
  if (depth == 0 || std::fabs(both - whole) < 15.0 * tol) return both + (both - whole) / 15.0;
  return simpson_step(g, lo, mid, glo, gmid, g1, left, tol / 2.0, depth - 1) +
         simpson_step(g, mid, hi, gmid, ghi, g3, right, tol / 2.0, depth - 1);
}

double simpson(const Moment& g, double lo, double hi) {
  const double mid = (lo + hi) / 2.0;
  const double glo = g(lo), ghi = g(hi), gmid = g(mid);
  const double whole = (hi - lo) / 6.0 * (std::fma(4.0, gmid, glo) + ghi);
  return simpson_step(g, lo, hi, glo, ghi, gmid, whole, 1e-14, 50);
}

// Lloyd-Max on Beta(a, a) mapped to [-1, 1] (codebook.cpp:135-195).
std::unique_ptr<Codebook> lloyd_max(std::size_t bits, std::size_t dim) {
  const double a = (double(dim) - 1.0) / 2.0;
  const BetaDensity f{a, 2.0 * log_gamma(a) - log_gamma(2.0 * a)};
  const Moment g{&f};
  const std::size_t levels = std::size_t(1) << bits;
  const double spread = 3.0 * std::sqrt(2.0 * a / ((2.0 * a + 1.0) * 4.0 * a));
  std::vector<double> c(levels), next(levels), edge(levels + 1);
  for (std::size_t i = 0; i < levels; ++i)
    c[i] = -spread + 2.0 * spread * double(i) / double(levels - 1);
  for (int it = 0; it < 200; ++it) {
    edge[0] = -1.0;
    edge[levels] = 1.0;
    for (std::size_t i = 1; i < levels; ++i) edge[i] = (c[i - 1] + c[i]) / 2.0;
    double moved = 0.0;
    for (std::size_t i = 0; i < levels; ++i) {
      const double mass = incomplete_beta(a, a, (edge[i + 1] + 1.0) / 2.0) -
                          incomplete_beta(a, a, (edge[i] + 1.0) / 2.0);
      next[i] = mass < 1e-15 ? c[i] : simpson(g, edge[i], edge[i + 1]) / mass;
    }
    for (std::size_t i = 0; i < levels; ++i) moved = std::max(moved, std::fabs(c[i] - next[i]));
    c.swap(next);
    if (moved < 1e-12) break;
  }
  auto cb = std::make_unique<Codebook>();
  cb->centroids.resize(levels);
  for (std::size_t i = 0; i < levels; ++i) cb->centroids[i] = float(c[i]);
  cb->boundaries.resize(levels - 1);
  for (std::size_t i = 0; i + 1 < levels; ++i)
    cb->boundaries[i] = (cb->centroids[i] + cb->centroids[i + 1]) * 0.5f;
  return cb;
}

// Device cell index of a value (the same float operation as the kernels: one
// fused multiply-add, monotone in v since inv_w > 0).
int cell_of(float v, float inv_w, float off) {
  const float f = std::fmaf(v, inv_w, off);
  if (f >= float(kCells - 1)) return kCells - 1;
  if (f >= 0.f) return int(f);
  return 0;
}

std::unique_ptr<CellTable> make_cell_table(const Codebook& cb) {
  auto t = std::make_unique<CellTable>();
  const auto& b = cb.boundaries;
  const std::size_t nb = b.size();
  const float w = (b[nb - 1] - b[0]) / float(kCells - 2);
  t->lo = b[0] - w;
  t->inv_w = 1.f / w;
  t->off = -t->lo * t->inv_w;
  std::vector<int> cell(nb);
  bool ok = w > 0.f && std::isfinite(t->inv_w);
  for (std::size_t i = 0; ok && i < nb; ++i) {
    cell[i] = cell_of(b[i], t->inv_w, t->off);
    if (i > 0 && cell[i] <= cell[i - 1]) ok = false;  // two boundaries share a cell
  }
  t->ok = ok;
  if (!ok) return t;
  const std::size_t top = cb.centroids.size() - 1;
  for (int c = 0; c < kCells; ++c) {
    std::uint32_t base = 0;
    float bound = std::numeric_limits<float>::infinity();
    for (std::size_t i = 0; i < nb; ++i) {
      if (cell[i] < c) ++base;
      if (cell[i] == c) bound = b[i];
    }
    t->cells[c] = CellEntry{base, bound, cb.centroids[base], cb.centroids[std::min<std::size_t>(base + 1, top)]};
  }
  return t;
}

std::mutex g_mu;
std::map<std::pair<std::size_t, std::size_t>, std::unique_ptr<CellTable>> g_cells;
std::map<std::size_t, std::unique_ptr<RotationTables>> g_rot;
std::map<std::pair<std::size_t, std::size_t>, std::unique_ptr<Codebook>> g_cb;

}  // namespace

const RotationTables& rotation_tables(std::size_t dim) {
  if (dim == 0 || dim % 8 != 0) throw std::invalid_argument("dim must be a positive multiple of 8");
  if (dim > kMaxDim) throw std::invalid_argument("dim exceeds the maximum of 16384");
  std::lock_guard lock(g_mu);
  auto& slot = g_rot[dim];
  if (!slot) slot = make_rotation(dim);
  return *slot;
}

const Codebook& codebook(std::size_t bits, std::size_t dim) {
  if (bits < 2 || bits > 4) throw std::invalid_argument("bits must be 2, 3, or 4");
  if (dim < 2) throw std::invalid_argument("dim must be >= 2");
  std::lock_guard lock(g_mu);
  auto& slot = g_cb[{bits, dim}];
  if (!slot) slot = lloyd_max(bits, dim);
  return *slot;
}

const CellTable& cell_table(std::size_t bits, std::size_t dim) {
  const Codebook& cb = codebook(bits, dim);
  std::lock_guard lock(g_mu);
  auto& slot = g_cells[{bits, dim}];
  if (!slot) slot = make_cell_table(cb);
  return *slot;
}

double beta_cdf_symmetric(double a, double t) { return incomplete_beta(a, a, t); }

// TQ+ anchors (reference src/encode/encode.cpp:233-240,282-303).
TqAnchors tq_anchors(std::size_t bits, std::size_t dim, std::size_t n) {
  const Codebook& cb = codebook(bits, dim);
  float c_outer = 0.f;
  for (float c : cb.centroids) c_outer = std::max(c_outer, std::fabs(c));
  const double a = (double(dim) - 1.0) / 2.0;
  const double p_hi = incomplete_beta(a, a, (double(c_outer) + 1.0) / 2.0);
  const double p_lo = 1.0 - p_hi;
  TqAnchors t{};
  t.qc_lo = -c_outer;
  t.qc_span = c_outer - (-c_outer);
  t.lo_idx = std::size_t(double(n) * p_lo);
  t.hi_idx = std::min(std::size_t(double(n) * p_hi), n - 1);
  t.hi_idx = std::max(t.hi_idx, t.lo_idx + 1);
  return t;
}

}  // namespace vchip
