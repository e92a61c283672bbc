// C++ API checks of vectorcache::TurboQuantIndex (the host side above the C ABI).
// Restates the reference's own index tests for the plain flat path
// (reference tests/test_index.cpp:11-113): ids in range, rows sorted, a
// self-query returns its own id first, batched search equals per-query
// search, plus the reference's error behaviour (std::invalid_argument,
// std::runtime_error) and merge_search_results semantics; and the same
// tests for vectorcache::LandmarkIndex, which is what they exercise in the
// reference (its TurboQuantIndex aliases LandmarkIndex).
// Prints "OK <checks>" and exits 0 on success.
#include <algorithm>
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

#include "vectorcache/cluster/kmeans_buckets.hpp"
#include "vectorcache/datasets/npy.hpp"
#include "vectorcache/encode/encode.hpp"
#include "vectorcache/index.hpp"
#include "vectorcache/pack/pack.hpp"
#include "vectorcache/quantize/codebook.hpp"
#include "vectorcache/search/search.hpp"
#include "vectorcache/transform/rotation.hpp"
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

bool same_rows(const vectorcache::SearchResults& a, const vectorcache::SearchResults& b) {
  return a.nq == b.nq && a.k == b.k && a.ids.size() == b.ids.size() && a.scores.size() == b.scores.size() &&
         std::equal(a.ids.begin(), a.ids.end(), b.ids.begin()) &&
         std::equal(a.scores.begin(), a.scores.end(), b.scores.begin());
}

std::vector<float> gaussian(std::size_t n, std::size_t dim, unsigned seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  std::vector<float> v(n * dim);
  for (auto& x : v) x = nd(rng);
  return v;
}
// The reference's free-function path (encode -> repack -> prepare_queries ->
// score_prepared / search_flat, score_prepared_into chains over buckets) against
// the index built from the same rows, and the README's bucketed index.
void free_function_checks() {
  using namespace vectorcache;
  const std::size_t dim = 128, n = 2000, nq = 6, k = 10, bits = 4;
  auto data = gaussian(n, dim, 21);
  std::vector<float> queries(data.begin(), data.begin() + nq * dim);
  TurboQuantIndex index(dim, bits);
  index.add(data);
  index.prepare();
  const SearchResults want = index.search(queries, k);

  const Rotation rot(dim);
  expect(rot.dim() == dim && rot.block() == 128 && rot.perm(0).size() == dim, "rotation tables");
  const auto cb = codebook(bits, dim);
  expect(cb.first.size() == 15 && cb.second.size() == 16, "codebook sizes");
  for (std::size_t i = 0; i + 1 < cb.second.size(); ++i)
    expect(cb.first[i] == (cb.second[i] + cb.second[i + 1]) * 0.5f, "boundaries are f32 midpoints");

  std::vector<float> scratch, alpha;
  std::vector<std::uint8_t> packed;
  encode(data, n, dim, rot, cb.first, cb.second, bits, nullptr, scratch, packed, alpha);
  expect(packed.size() == n * bits * dim / 8 && alpha.size() == n, "encode output sizes");
  std::vector<float> alpha_ix(n);
  expect(vc_index_export_alpha(index.handle(), alpha_ix.data()) == VC_OK, "export alpha");
  expect(std::equal(alpha.begin(), alpha.end(), alpha_ix.begin()), "encode alpha == index alpha");
  encode(std::span<const float>(data.data(), dim), 1, dim, rot, cb.first, cb.second, bits, nullptr, scratch, packed,
         alpha);
  expect(packed.size() == (n + 1) * bits * dim / 8 && alpha.size() == n + 1 && alpha[n] == alpha[0],
         "encode appends");
  packed.resize(n * bits * dim / 8);
  alpha.resize(n);

  const auto geo = blocked_geometry(n, bits, dim);
  expect(geo.n_blocks == 11 && geo.n_byte_groups == 64 && geo.blocked_len == 11 * 64 * 192, "blocked geometry");
  const auto flat = codes_flat_from_packed(packed, n, bits, dim);
  auto [perm0, nb] = repack_perm0(packed, n, bits, dim);
  expect(nb == geo.n_blocks && perm0.size() == geo.blocked_len, "repack_perm0 shape");
  std::vector<std::uint8_t> perm0_ix(perm0.size());
  expect(vc_index_export_perm0(index.handle(), perm0_ix.data()) == VC_OK, "export perm0");
  expect(perm0 == perm0_ix, "repack_perm0 == index PERM0 export");
  bool codes_ok = true;
  for (std::size_t v = 0; v < n; v += 97)
    for (std::size_t g = 0; g < geo.n_byte_groups; g += 7)
      codes_ok = codes_ok && read_code(perm0, bits, geo.n_byte_groups, v / kBlock, g, v % kBlock) ==
                                 flat[v * geo.n_byte_groups + g];
  expect(codes_ok, "read_code(PERM0) == codes_flat_from_packed");
  expect(repack_seq(packed, n, bits, dim).size() == geo.blocked_len, "repack_seq shape");

  SearchResults sf = search_flat(queries, nq, dim, k, rot, cb.second, bits, perm0, nb, alpha, {}, {});
  expect(same_rows(sf, want), "search_flat == TurboQuantIndex::search");

  PreparedQueries prep = prepare_queries(queries, nq, dim, rot, cb.second, bits, {}, {});
  expect(prep.nq == nq && prep.n_byte_groups == 64, "prepared shape");
  // two buckets (halves of the rows), each scored into persistent heaps with
  // its global ids: the union equals the flat search
  const std::size_t h = 1000;
  auto [p0, nb0] = repack_perm0(std::span<const std::uint8_t>(packed.data(), h * bits * dim / 8), h, bits, dim);
  auto [p1, nb1] = repack_perm0(std::span<const std::uint8_t>(packed.data() + h * bits * dim / 8,
                                                              (n - h) * bits * dim / 8), n - h, bits, dim);
  std::vector<std::uint64_t> ids0(h), ids1(n - h);
  for (std::size_t i = 0; i < h; ++i) ids0[i] = i;
  for (std::size_t i = 0; i < n - h; ++i) ids1[i] = h + i;
  std::vector<float> hs(nq * k), hmin(nq, -std::numeric_limits<float>::infinity());
  std::vector<std::uint64_t> hi(nq * k);
  std::vector<std::size_t> hsz(nq, 0), hmi(nq, 0), all(nq);
  for (std::size_t q = 0; q < nq; ++q) all[q] = q;
  score_prepared_into(prep, k, p0, nb0, std::span<const float>(alpha.data(), h), ids0, all, hs.data(), hi.data(),
                      hsz.data(), hmin.data(), hmi.data());
  score_prepared_into(prep, k, p1, nb1, std::span<const float>(alpha.data() + h, n - h), ids1, all, hs.data(),
                      hi.data(), hsz.data(), hmin.data(), hmi.data());
  SearchResults hr;
  heaps_to_search_results(hr, nq, k, hs.data(), hi.data(), hsz.data());
  expect(same_rows(hr, want), "two-bucket score_prepared_into chain == flat search");
  SearchResults one = score_prepared(prepared_query_at(prep, 2), k, perm0, nb, alpha);
  expect(one.nq == 1 && std::equal(one.ids.begin(), one.ids.end(), want.ids.begin() + 2 * k),
         "prepared_query_at(2) == row 2");
  std::vector<std::size_t> sub{5, 0, 5};
  SearchResults rs = score_prepared(prep, k, perm0, nb, alpha, {}, sub);
  expect(rs.nq == 3 && std::equal(rs.ids.begin(), rs.ids.begin() + k, want.ids.begin() + 5 * k),
         "score_prepared query subset");
  expect_throw<std::out_of_range>([&] { (void)prepared_query_at(prep, nq); }, "prepared_query_at out of range");
  std::vector<float> wrong(cb.second);
  wrong[0] += 1.f;
  expect_throw<std::invalid_argument>([&] { (void)prepare_queries(queries, nq, dim, rot, wrong, bits, {}, {}); },
                                      "prepare_queries: foreign centroids");
  Calibration cal = fit_calibration(data, 500, dim, rot, cb.second, scratch);
  expect(cal.shift.size() == dim && cal.scale_tq.size() == dim, "fit_calibration sizes");

  // the README's bucketed index: with every bucket open the residual scores
  // rank the rows themselves first (a stored vector queried by itself)
  BucketParams bp;
  bp.max_bucket_size = 200;
  bp.energy_soft = 100;
  bp.n_probe = 512;
  BucketedTurboQuantIndex bix(dim, bits, bp);
  bix.add(data);
  expect_throw<std::runtime_error>([&] { (void)bix.search(queries, k); }, "bucketed search before prepare");
  bix.prepare();
  expect(bix.size() == n && bix.n_buckets() >= n / 200, "bucketed build");
  SearchResults br = bix.search(queries, k);
  expect(br.nq == nq && br.k == k, "bucketed result shape");
  for (std::size_t q = 0; q < nq; ++q) {
    expect(br.ids[q * k] == q, "bucketed self-query");
    for (std::size_t j = 1; j < k; ++j) expect(br.scores[q * k + j - 1] >= br.scores[q * k + j], "bucketed rows sorted");
  }
  expect_throw<std::invalid_argument>([&] { BucketedTurboQuantIndex bad(12, bits, bp); }, "bucketed bad dim");
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
      expect(same_rows(hr, sr), "score_prepared_into into empty heaps == search");
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
  free_function_checks();
  landmark_checks();
  std::printf("OK %d\n", g_checks);
  return 0;
}
