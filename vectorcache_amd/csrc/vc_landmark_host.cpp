// Host-side landmark basis of the landmark index: the seeded orthonormal
// matrix (reference src/transform/landmarks.cpp:134-198) and its 4-bit
// column quantization (src/transform/landmarks.cpp:236-269).
//
// Bit-exactness notes (the reference is built with -O3 -mavx2 -mfma
// -ffp-contract=fast; its object code fixes the float semantics restated
// here with contraction off and explicit std::fma):
//  * std::normal_distribution<double> over std::mt19937_64 is libstdc++'s
//    Marsaglia polar method: u = double(x) * 2^-64 (clamped below 1),
//    a = 2u - 1, b = 2u' - 1, r2 = fma(a, a, b * b), rejected unless
//    0 < r2 <= 1, f = sqrt(-2 log(r2) / r2); it returns b * f and keeps
//    a * f for the next draw. The mean 0 is added (+0.0).
//  * Gram-Schmidt dot products and norms are sums of ROUNDED products in
//    element order (the compiler vectorizes the products and folds the sum
//    in order); the update a_i -= dot * a_j is one fused multiply-add.
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <tuple>

#include "vc_host.hpp"

namespace vchip {
namespace {

class PolarGaussian {
 public:
  explicit PolarGaussian(std::uint64_t seed) : rng_(seed) {}
  double next() {
    if (have_saved_) {
      have_saved_ = false;
      return saved_ + 0.0;
    }
    double a, b, r2;
    do {
      a = std::fma(2.0, uniform(), -1.0);
      b = std::fma(2.0, uniform(), -1.0);
      r2 = std::fma(a, a, b * b);
    } while (r2 > 1.0 || r2 == 0.0);
    const double f = std::sqrt(-2.0 * std::log(r2) / r2);
    saved_ = a * f;
    have_saved_ = true;
    return b * f + 0.0;
  }

 private:
  double uniform() {
    double u = static_cast<double>(rng_()) * 0x1p-64;
    if (u >= 1.0) u = std::nextafter(1.0, 0.0);
    return u;
  }
  std::mt19937_64 rng_;
  double saved_ = 0.0;
  bool have_saved_ = false;
};

std::vector<float> make_landmarks(std::size_t dim, std::size_t m, std::uint64_t seed) {
  PolarGaussian g(seed);
  std::vector<double> A(dim * m);
  for (double& v : A) v = g.next();
  for (std::size_t i = 0; i < m; ++i) {
    double* ai = A.data() + i * dim;
    for (std::size_t j = 0; j < i; ++j) {
      const double* aj = A.data() + j * dim;
      double dot = 0.0;
      for (std::size_t r = 0; r < dim; ++r) dot = dot + ai[r] * aj[r];
      for (std::size_t r = 0; r < dim; ++r) ai[r] = std::fma(-dot, aj[r], ai[r]);
    }
    double nrm = 0.0;
    for (std::size_t r = 0; r < dim; ++r) nrm = nrm + ai[r] * ai[r];
    nrm = std::sqrt(nrm);
    if (nrm < 1e-15) throw std::runtime_error("degenerate landmark column");
    const double inv = 1.0 / nrm;
    for (std::size_t r = 0; r < dim; ++r) ai[r] = ai[r] * inv;
  }
  std::vector<float> L(A.size());
  for (std::size_t i = 0; i < A.size(); ++i) L[i] = static_cast<float>(A[i]);
  return L;
}

std::mutex g_lm_mu;
std::map<std::tuple<std::size_t, std::size_t, std::uint64_t>, std::unique_ptr<std::vector<float>>> g_lm;

}  // namespace

const std::vector<float>& ortho_landmarks(std::size_t dim, std::size_t m, std::uint64_t seed) {
  if (dim == 0 || m == 0 || m > dim) throw std::invalid_argument("landmarks require 0 < m <= dim");
  if (dim % 8 != 0) throw std::invalid_argument("landmarks require dim to be a multiple of 8");
  const auto key = std::make_tuple(dim, m, seed);
  {
    std::lock_guard lock(g_lm_mu);
    auto it = g_lm.find(key);
    if (it != g_lm.end()) return *it->second;
  }
  auto L = std::make_unique<std::vector<float>>(make_landmarks(dim, m, seed));
  std::lock_guard lock(g_lm_mu);
  auto [it, inserted] = g_lm.emplace(key, std::move(L));
  (void)inserted;
  return *it->second;
}

QuantizedLandmarks quantize_landmarks(const float* L, std::size_t dim, std::size_t m) {
  if (L == nullptr || dim == 0 || m == 0) throw std::invalid_argument("quantize_landmarks: bad dim/m/L");
  QuantizedLandmarks q;
  q.dim = dim;
  q.dim_pad = (dim + 15) / 16 * 16;
  q.m = m;
  q.scales.assign(m, 0.f);
  const std::size_t bpc = q.dim_pad / 2;
  q.codes.assign(bpc * m, 0);
  auto level = [](float x, float inv) {
    const float r = std::round(x * inv);
    return static_cast<int>(std::fmax(-7.f, std::fmin(7.f, r)));
  };
  for (std::size_t j = 0; j < m; ++j) {
    const float* col = L + j * dim;
    float amax = 0.f;
    for (std::size_t i = 0; i < dim; ++i) amax = std::fmax(amax, std::fabs(col[i]));
    if (amax < 1e-30f) continue;  // scale 0, codes stay 0
    q.scales[j] = amax / 7.f;
    const float inv = 7.f / amax;
    std::uint8_t* dst = q.codes.data() + j * bpc;
    for (std::size_t i = 0; i < dim; i += 2) {
      const int e = level(col[i], inv);
      const int o = i + 1 < dim ? level(col[i + 1], inv) : 0;
      dst[i / 2] = static_cast<std::uint8_t>(((o + 7) << 4) | (e + 7));
    }
  }
  return q;
}

}  // namespace vchip
