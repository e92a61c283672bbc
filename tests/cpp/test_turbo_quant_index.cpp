// C++ API checks of vectorcache::TurboQuantIndex (the host side above the C ABI).
// Restates the reference's own index tests for the plain flat path
// (reference tests/test_index.cpp:11-113): ids in range, rows sorted, a
// self-query returns its own id first, batched search equals per-query
// search, plus the reference's error behaviour (std::invalid_argument,
// std::runtime_error) and merge_search_results semantics; and the same
// tests for vectorcache::LandmarkIndex, which is what they exercise in the
// reference (its TurboQuantIndex aliases LandmarkIndex).
// Prints "OK <checks>" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <limits>
#include <random>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "vectorcache/datasets/npy.hpp"
#include "vectorcache/index.hpp"
#include "vectorcache/turbo_quant_index.hpp"

namespace {
int g_checks = 0;
void expect(bool ok, const std::string& what) {
  ++g_checks;
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what.c_str());
    std::exit(1);
  }
}

template <typename E, typename F>
void expect_throw(F&& f, const std::string& what) {
  try {
    f();
  } catch (const E&) {
    ++g_checks;
    return;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "FAIL: %s threw the wrong type: %s\n", what.c_str(), e.what());
    std::exit(1);
  }
  std::fprintf(stderr, "FAIL: %s did not throw\n", what.c_str());
  std::exit(1);
}

std::vector<float> gaussian(std::size_t n, std::size_t dim, unsigned seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  std::vector<float> v(n * dim);
  for (auto& x : v) x = nd(rng);
  return v;
}
// The reference's LandmarkIndex tests (tests/test_index.cpp:11-113).
void landmark_checks() {
  using vectorcache::LandmarkIndex;
  {  // AddAndSearchReturnsIds
    const std::size_t dim = 64, n = 256, nq = 4, k = 5;
    const auto db = gaussian(n, dim, 42);
    const auto queries = gaussian(nq, dim, 43);
    LandmarkIndex index(dim, 4, 48);
    expect(index.scan_dim() == 48 && index.n_landmarks() == 48 && index.rerank_bits() == 4, "landmark accessors");
    index.add(db);
    index.prepare();
    const auto res = index.search(queries, k);
    expect(res.nq == nq && res.k == k, "landmark result shape");
    for (std::size_t qi = 0; qi < nq; ++qi)
      for (std::size_t j = 0; j < k; ++j) {
        expect(res.ids[qi * k + j] < n, "landmark ids in range");
        if (j > 0) expect(res.scores[qi * k + j - 1] >= res.scores[qi * k + j], "landmark rows sorted");
      }
  }
  {  // SelfQueryHighScore
    const std::size_t dim = 128, n = 64;
    const auto db = gaussian(n, dim, 7);
    LandmarkIndex index(dim, 4, 96);
    index.add(db);
    index.prepare();
    const auto res = index.search(std::span<const float>(db.data(), dim), 1);
    expect(res.k == 1 && res.ids[0] == 0 && res.scores[0] > 0.7f, "landmark self-query");
  }
  {  // FloatRescoreSelfQuery
    const std::size_t dim = 128, n = 256, k = 16;
    auto db = gaussian(n, dim, 11);
    for (std::size_t i = 0; i < n; ++i) {
      double s = 0.0;
      float* row = db.data() + i * dim;
      for (std::size_t d = 0; d < dim; ++d) s += static_cast<double>(row[d]) * row[d];
      const float inv = s > 0.0 ? static_cast<float>(1.0 / std::sqrt(s)) : 0.f;
      for (std::size_t d = 0; d < dim; ++d) row[d] *= inv;
    }
    LandmarkIndex index(dim, 4, 64);
    index.add(db);
    index.prepare();
    const auto res = index.search(std::span<const float>(db.data(), dim), k, db);
    expect(res.k == k && res.ids[0] == 0 && std::fabs(res.scores[0] - 1.0f) < 1e-4f, "landmark float re-rank");
    for (std::size_t j = 1; j < k; ++j) expect(res.scores[j - 1] >= res.scores[j], "landmark re-ranked sorted");
  }
  {  // MultiQueryMatchesSingle
    const std::size_t dim = 64, n = 512, nq = 8, k = 8;
    const auto db = gaussian(n, dim, 123);
    const auto queries = gaussian(nq, dim, 124);
    LandmarkIndex index(dim, 4, 48);
    index.add(db);
    index.prepare();
    const auto batch = index.search(queries, k);
    for (std::size_t qi = 0; qi < nq; ++qi) {
      const auto one = index.search(std::span<const float>(queries.data() + qi * dim, dim), k);
      for (std::size_t j = 0; j < k; ++j)
        expect(batch.ids[qi * k + j] == one.ids[j] && batch.scores[qi * k + j] == one.scores[j],
               "landmark batch equals single");
    }
  }
  expect_throw<std::invalid_argument>([] { LandmarkIndex bad(64, 4, 65); }, "landmarks > dim");
  expect_throw<std::invalid_argument>([] { LandmarkIndex bad(64, 4, 16, 0, 5); }, "rerank bits");
  LandmarkIndex idx(64, 4, 16);
  idx.add(gaussian(10, 64, 5));
  expect_throw<std::runtime_error>([&] { (void)idx.search(gaussian(1, 64, 6), 3); }, "landmark search before prepare");
}
}  // namespace

int main() {
  using vectorcache::SearchResults;
  using vectorcache::TurboQuantIndex;
  const std::size_t dim = 128, n = 2000, nq = 16, k = 10;
  auto data = gaussian(n, dim, 42);

  expect_throw<std::invalid_argument>([] { TurboQuantIndex bad(100, 4); }, "dim not a multiple of 8");
  expect_throw<std::invalid_argument>([] { TurboQuantIndex bad(128, 5); }, "bits out of range");

  TurboQuantIndex index(dim, 4);
  expect(index.dim() == dim && index.bit_width() == 4, "accessors");
  expect_throw<std::invalid_argument>([&] { index.add(std::span<const float>(data.data(), 130)); },
                                      "span not a multiple of dim");
  index.add(data);
  expect(index.size() == n, "size after add");
  expect_throw<std::runtime_error>([&] { (void)index.search(std::span<const float>(data.data(), dim), k); },
                                   "search before prepare");
  index.prepare();
  expect(index.codes_bytes() == n * dim / 2, "codes_bytes");

  std::span<const float> queries(data.data(), nq * dim);  // the first nq vectors query themselves
  SearchResults res = index.search(queries, k);
  expect(res.nq == nq && res.k == k && res.ids.size() == nq * k, "result shape");
  for (std::size_t q = 0; q < nq; ++q) {
    expect(res.ids[q * k] == q, "self-query returns its own id first");
    for (std::size_t j = 0; j < k; ++j) {
      expect(res.ids[q * k + j] < n, "ids in range");
      if (j > 0) {
        const float a = res.scores[q * k + j - 1], b = res.scores[q * k + j];
        expect(a > b || (a == b && res.ids[q * k + j - 1] < res.ids[q * k + j]), "rows sorted (score desc, id asc)");
      }
    }
  }
  for (std::size_t q = 0; q < nq; ++q) {
    SearchResults one = index.search(std::span<const float>(data.data() + q * dim, dim), k);
    for (std::size_t j = 0; j < k; ++j) {
      expect(one.ids[j] == res.ids[q * k + j], "batched ids == per-query ids");
      expect(one.scores[j] == res.scores[q * k + j], "batched scores == per-query scores");
    }
  }
  SearchResults big = index.search(queries, n + 50);
  expect(big.k == n, "k clipped to size()");

  // merge: dst and src share ids; the union keeps the k best, once per id
  SearchResults dst = index.search(std::span<const float>(data.data(), dim), 5);
  SearchResults src = index.search(std::span<const float>(data.data(), dim), 8);
  vectorcache::merge_search_results(dst, src);
  expect(dst.k == 5, "merge keeps dst.k");
  for (std::size_t j = 0; j < 5; ++j) expect(dst.ids[j] == src.ids[j], "merge of overlapping rows");

  // score_prepared_into over the index's own codes as one bucket, into empty
  // heaps (threshold -inf): ids increase with the scan order, so the heaps
  // hold the exact top-k -- the same rows as search()
  {
    const std::size_t G = dim / 2, nb = (n + 191) / 192;
    std::vector<std::uint8_t> perm0(nb * G * 192);
    std::vector<float> alpha(n);
    expect(vc_index_export_perm0(index.handle(), perm0.data()) == VC_OK, "export perm0");
    expect(vc_index_export_alpha(index.handle(), alpha.data()) == VC_OK, "export alpha");
    for (std::size_t kk : {std::size_t(7), std::size_t(200)}) {
      std::vector<float> hs(nq * kk), hmin(nq, -std::numeric_limits<float>::infinity());
      std::vector<std::uint64_t> hi(nq * kk);
      std::vector<std::size_t> hsz(nq, 0), hmi(nq, 0), qidx(nq);
      for (std::size_t q = 0; q < nq; ++q) qidx[q] = nq - 1 - q;
      vectorcache::score_perm0_into(queries, dim, 4, kk, perm0, alpha, {}, qidx, hs.data(), hi.data(), hsz.data(),
                                    hmin.data(), hmi.data());
      SearchResults hr;
      vectorcache::heaps_to_search_results(hr, nq, kk, hs.data(), hi.data(), hsz.data());
      SearchResults sr = index.search(queries, kk);
      expect(hr.ids == sr.ids && hr.scores == sr.scores, "score_prepared_into into empty heaps == search");
      std::vector<std::size_t> none;
      expect_throw<std::invalid_argument>([&] {
        vectorcache::score_perm0_into(queries, dim, 4, kk, perm0, alpha, {}, none, hs.data(), hi.data(), hsz.data(),
                                      hmin.data(), hmi.data());
      }, "score_prepared_into: empty query_indices");
      std::vector<std::size_t> bad{nq};
      expect_throw<std::out_of_range>([&] {
        vectorcache::score_perm0_into(queries, dim, 4, kk, perm0, alpha, {}, bad, hs.data(), hi.data(), hsz.data(),
                                      hmin.data(), hmi.data());
      }, "score_prepared_into: query index out of range");
    }
  }

  // TQ+ calibrated index
  TurboQuantIndex tq(dim, 2);
  auto sample = gaussian(500, dim, 7);
  tq.calibrate(sample);
  tq.add(data);
  tq.prepare();
  SearchResults rt = tq.search(queries, k);
  for (std::size_t q = 0; q < nq; ++q) expect(rt.ids[q * k] == q, "TQ+ self-query");
  expect_throw<std::runtime_error>([&] { tq.calibrate(sample); }, "calibrate after prepare");

  // NpyReader / write_npy_f32_matrix (reference include/vectorcache/datasets/npy.hpp:33-65)
  {
    namespace ds = vectorcache::datasets;
    const std::filesystem::path p = std::filesystem::temp_directory_path() / "vc_cpp_api_check" / "m.npy";
    std::vector<float> m(5 * 3);
    for (std::size_t i = 0; i < m.size(); ++i) m[i] = float(i) * 0.5f;
    ds::write_npy_f32_matrix(p, 5, 3, m.data());
    auto r = ds::NpyReader::open(p, "m");
    expect(r.meta().count == 5 && r.meta().dim == 3 && std::string(r.meta().label) == "m", "npy meta");
    auto v = r.next_vector_span();
    expect(v && v->size() == 3 && (*v)[2] == 1.0f, "npy first row");
    std::vector<float> row(3);
    expect(r.next_vector_into(row) && row[0] == 1.5f, "npy second row");
    std::vector<float> wrong(4);
    expect_throw<std::runtime_error>([&] { (void)r.next_vector_into(wrong); }, "npy row size mismatch");
    std::vector<float> rest(9);
    expect(r.copy_rows_into(rest, 10) == 3 && rest[8] == 7.0f, "npy bulk copy");
    expect(!r.next_vector_span(), "npy end");
    expect_throw<std::runtime_error>([&] { (void)ds::NpyReader::open(p.parent_path() / "none.npy", "x"); },
                                     "npy missing file");
    std::filesystem::remove_all(p.parent_path());
  }
  landmark_checks();
  std::printf("OK %d\n", g_checks);
  return 0;
}
