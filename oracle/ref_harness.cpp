// Test infrastructure only: a thin extern "C" harness over the reference
// library's own plain-flat-path functions, compiled together with the
// reference sources (read in place, never copied) into oracle/_ref/.
// Only tests/, tests/golden/ and bench.py's CPU-baseline leg load it, and
// only as a checker / timed baseline — never as part of the product path.
//
// Every entry forces the reference's AVX2 PERM0 uint8-LUT scan
// (TURBOVEC_NO_VNNI=1, reference src/pack/pack.cpp:214-233), which is the
// deterministic, ISA-portable path this repository is bit-exact against.
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <span>
#include <string>
#include <vector>

#include "vectorcache/constants.hpp"
#include "vectorcache/encode/encode.hpp"
#include "vectorcache/index.hpp"
#include "vectorcache/pack/pack.hpp"
#include "vectorcache/quantize/codebook.hpp"
#include "vectorcache/search/float_rescore.hpp"
#include "vectorcache/search/search.hpp"
#include "vectorcache/transform/landmarks.hpp"
#include "vectorcache/transform/rotation.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace vc = vectorcache;

namespace {
thread_local std::string g_err;

void force_avx2_path() { setenv("TURBOVEC_NO_VNNI", "1", 1); }

template <typename F>
int guarded(F&& f, bool avx2 = true) {
  try {
    if (avx2) force_avx2_path();
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
}  // namespace

extern "C" {

const char* vcref_last_error() { return g_err.c_str(); }

void vcref_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int vcref_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Rotation tables (reference src/transform/rotation.cpp:139-160).
int vcref_rotation_tables(size_t dim, uint32_t* perm, float* signs, float* inv_sqrt_block,
                          size_t* block) {
  return guarded([&] {
    vc::Rotation rot(dim);
    auto p = rot.perm(0);
    auto s = rot.signs(0);
    std::memcpy(perm, p.data(), dim * sizeof(uint32_t));
    std::memcpy(signs, s.data(), dim * sizeof(float));
    *inv_sqrt_block = rot.inv_sqrt_block();
    *block = rot.block();
  });
}

// Codebook (reference src/quantize/codebook.cpp:203-226).
int vcref_codebook(size_t bits, size_t dim, float* boundaries, float* centroids) {
  return guarded([&] {
    auto [b, c] = vc::codebook(bits, dim);
    std::memcpy(boundaries, b.data(), b.size() * sizeof(float));
    std::memcpy(centroids, c.data(), c.size() * sizeof(float));
  });
}

float vcref_row_norm(const float* row, size_t dim) {
  return vc::row_norm(std::span<const float>(row, dim));
}

double vcref_beta_cdf_aa(double a, double t) { return vc::beta_cdf_aa(a, t); }

// Rotation::apply_scaled_into (reference src/transform/rotation.cpp:185-205).
int vcref_rotate_scaled(size_t dim, const float* src, float inv, float* dst) {
  return guarded([&] {
    vc::Rotation rot(dim);
    std::vector<float> scratch(dim);
    rot.apply_scaled_into(std::span<const float>(src, dim), inv, std::span<float>(dst, dim),
                          scratch);
  });
}

// encode() (reference src/encode/encode.cpp:294-307). shift/scale may be null
// (no calibration). rotated_out may be null.
int vcref_encode(const float* x, size_t n, size_t dim, size_t bits, const float* shift,
                 const float* scale, uint8_t* packed_out, float* alpha_out, float* rotated_out) {
  return guarded([&] {
    vc::Rotation rot(dim);
    auto [b, c] = vc::codebook(bits, dim);
    vc::Calibration cal;
    const vc::Calibration* calp = nullptr;
    if (shift != nullptr && scale != nullptr) {
      cal.shift.assign(shift, shift + dim);
      cal.scale_tq.assign(scale, scale + dim);
      calp = &cal;
    }
    std::vector<float> scratch;
    std::vector<uint8_t> packed;
    std::vector<float> alphas;
    vc::encode(std::span<const float>(x, n * dim), n, dim, rot, b, c, bits, calp, scratch, packed,
               alphas);
    std::memcpy(packed_out, packed.data(), packed.size());
    std::memcpy(alpha_out, alphas.data(), alphas.size() * sizeof(float));
    if (rotated_out != nullptr) {
      std::memcpy(rotated_out, scratch.data(), n * dim * sizeof(float));
    }
  });
}

// fit_calibration() (reference src/encode/encode.cpp:327-335).
int vcref_fit_calibration(const float* x, size_t n, size_t dim, size_t bits, float* shift_out,
                          float* scale_out) {
  return guarded([&] {
    vc::Rotation rot(dim);
    auto [b, c] = vc::codebook(bits, dim);
    std::vector<float> scratch;
    auto cal = vc::fit_calibration(std::span<const float>(x, n * dim), n, dim, rot, c, scratch);
    std::memcpy(shift_out, cal.shift.data(), dim * sizeof(float));
    std::memcpy(scale_out, cal.scale_tq.data(), dim * sizeof(float));
  });
}

// codes_flat_from_packed() (reference src/pack/pack.cpp:352-356).
int vcref_codes_flat(const uint8_t* packed, size_t n, size_t bits, size_t dim, uint8_t* out) {
  return guarded([&] {
    const size_t bpr = bits * (dim / 8);
    auto flat = vc::codes_flat_from_packed(std::span<const uint8_t>(packed, n * bpr), n, bits, dim);
    std::memcpy(out, flat.data(), flat.size());
  });
}

// repack_perm0() (reference src/pack/pack.cpp:388-401). Returns n_blocks via out param.
int vcref_repack_perm0(const uint8_t* packed, size_t n, size_t bits, size_t dim, uint8_t* out,
                       size_t* n_blocks) {
  return guarded([&] {
    const size_t bpr = bits * (dim / 8);
    auto [blocked, nb] =
        vc::repack_perm0(std::span<const uint8_t>(packed, n * bpr), n, bits, dim);
    std::memcpy(out, blocked.data(), blocked.size());
    *n_blocks = nb;
  });
}

// repack_seq() (reference src/pack/pack.cpp:358-364).
int vcref_repack_seq(const uint8_t* packed, size_t n, size_t bits, size_t dim, uint8_t* out) {
  return guarded([&] {
    const size_t bpr = bits * (dim / 8);
    auto blocked = vc::repack_seq(std::span<const uint8_t>(packed, n * bpr), n, bits, dim);
    std::memcpy(out, blocked.data(), blocked.size());
  });
}

// prepare_queries() (reference src/search/search.cpp:762-860) on the AVX2 backend.
// luts_out: nq * G * 32 bytes.
int vcref_prepare_queries(const float* q, size_t nq, size_t dim, size_t bits, const float* shift,
                          const float* scale, float* q_rot_out, float* bias_corr_out,
                          uint8_t* luts_out, float* lut_scale_out, float* lut_bias_out) {
  return guarded([&] {
    vc::Rotation rot(dim);
    auto [b, c] = vc::codebook(bits, dim);
    std::span<const float> sh, sc;
    if (shift != nullptr && scale != nullptr) {
      sh = std::span<const float>(shift, dim);
      sc = std::span<const float>(scale, dim);
    }
    auto prep = vc::prepare_queries(std::span<const float>(q, nq * dim), nq, dim, rot, c, bits, sh,
                                    sc);
    if (prep.backend != vc::SearchBackendKind::Avx2) {
      throw std::runtime_error("reference did not select the AVX2 PERM0 backend");
    }
    std::memcpy(q_rot_out, prep.q_rot.data(), nq * dim * sizeof(float));
    std::memcpy(bias_corr_out, prep.bias_corrs.data(), nq * sizeof(float));
    const size_t lut_bytes = prep.n_byte_groups * 32;
    for (size_t i = 0; i < nq; ++i) {
      std::memcpy(luts_out + i * lut_bytes, prep.luts[i].uint8_luts.data(), lut_bytes);
      lut_scale_out[i] = prep.luts[i].scale;
      lut_bias_out[i] = prep.luts[i].bias;
    }
  });
}

// search_flat() (reference src/search/search.cpp:1293-1302) over PERM0 blocked
// codes. Writes nq * k_eff scores/ids, returns k_eff through *k_out.
int vcref_search_flat(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                      const uint8_t* blocked, size_t n_blocks, const float* alphas, size_t n,
                      const float* shift, const float* scale, float* scores_out,
                      uint64_t* ids_out, size_t* k_out) {
  return guarded([&] {
    vc::Rotation rot(dim);
    auto [b, c] = vc::codebook(bits, dim);
    std::span<const float> sh, sc;
    if (shift != nullptr && scale != nullptr) {
      sh = std::span<const float>(shift, dim);
      sc = std::span<const float>(scale, dim);
    }
    const auto geo = vc::blocked_geometry(n, bits, dim);
    auto res = vc::search_flat(std::span<const float>(q, nq * dim), nq, dim, k, rot, c, bits,
                               std::span<const uint8_t>(blocked, geo.blocked_len), n_blocks,
                               std::span<const float>(alphas, n), sh, sc);
    *k_out = res.k;
    if (res.k > 0) {
      std::memcpy(scores_out, res.scores.data(), nq * res.k * sizeof(float));
      std::memcpy(ids_out, res.ids.data(), nq * res.k * sizeof(uint64_t));
    }
  });
}

// score_prepared() with an id_map (reference src/search/search.cpp:1072-1154):
// the multi-bucket plumbing used per bucket.
int vcref_search_flat_idmap(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                            const uint8_t* blocked, size_t n_blocks, const float* alphas,
                            size_t n, const uint64_t* id_map, float* scores_out,
                            uint64_t* ids_out, size_t* k_out) {
  return guarded([&] {
    vc::Rotation rot(dim);
    auto [b, c] = vc::codebook(bits, dim);
    auto prep = vc::prepare_queries(std::span<const float>(q, nq * dim), nq, dim, rot, c, bits, {},
                                    {});
    const auto geo = vc::blocked_geometry(n, bits, dim);
    auto res = vc::score_prepared(prep, k, std::span<const uint8_t>(blocked, geo.blocked_len),
                                  n_blocks, std::span<const float>(alphas, n),
                                  std::span<const uint64_t>(id_map, n));
    *k_out = res.k;
    if (res.k > 0) {
      std::memcpy(scores_out, res.scores.data(), nq * res.k * sizeof(float));
      std::memcpy(ids_out, res.ids.data(), nq * res.k * sizeof(uint64_t));
    }
  });
}

// score_prepared() with every argument of the AVX2 path (reference
// src/search/search.cpp:1072-1154): id_map, query_indices and the residual
// post-correction (residual_norms per vector, residual_scales per query).
// Writes nq_out * k_eff scores/ids, nq_out = n_indices or nq.
int vcref_score_prepared(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                         const uint8_t* blocked, size_t n_blocks, const float* alphas, size_t n,
                         const float* shift, const float* scale, const uint64_t* id_map,
                         const size_t* query_indices, size_t n_indices, const float* residual_norms,
                         const float* residual_scales, float* scores_out, uint64_t* ids_out,
                         size_t* k_out) {
  return guarded([&] {
    vc::Rotation rot(dim);
    auto [b, c] = vc::codebook(bits, dim);
    std::span<const float> sh, sc;
    if (shift != nullptr && scale != nullptr) {
      sh = std::span<const float>(shift, dim);
      sc = std::span<const float>(scale, dim);
    }
    auto prep = vc::prepare_queries(std::span<const float>(q, nq * dim), nq, dim, rot, c, bits, sh, sc);
    const auto geo = vc::blocked_geometry(n, bits, dim);
    std::span<const uint64_t> im;
    if (id_map != nullptr) im = std::span<const uint64_t>(id_map, n);
    std::span<const size_t> qi;
    if (query_indices != nullptr) qi = std::span<const size_t>(query_indices, n_indices);
    std::span<const float> rn, rs;
    if (residual_norms != nullptr) rn = std::span<const float>(residual_norms, n);
    if (residual_scales != nullptr) rs = std::span<const float>(residual_scales, nq);
    auto res = vc::score_prepared(prep, k, std::span<const uint8_t>(blocked, geo.blocked_len), n_blocks,
                                  std::span<const float>(alphas, n), im, qi, rn, rs);
    *k_out = res.k;
    if (res.k > 0) {
      std::memcpy(scores_out, res.scores.data(), res.nq * res.k * sizeof(float));
      std::memcpy(ids_out, res.ids.data(), res.nq * res.k * sizeof(uint64_t));
    }
  });
}

// merge_search_results() (reference src/search/search.cpp:1551-1600).
int vcref_merge_results(size_t nq, size_t k_dst, float* dst_scores, uint64_t* dst_ids,
                        size_t k_src, const float* src_scores, const uint64_t* src_ids) {
  return guarded([&] {
    vc::SearchResults dst, src;
    dst.nq = nq;
    dst.k = k_dst;
    dst.scores.resize_uninitialized(nq * k_dst);
    dst.ids.resize_uninitialized(nq * k_dst);
    std::memcpy(dst.scores.data(), dst_scores, nq * k_dst * sizeof(float));
    std::memcpy(dst.ids.data(), dst_ids, nq * k_dst * sizeof(uint64_t));
    src.nq = nq;
    src.k = k_src;
    src.scores.resize_uninitialized(nq * k_src);
    src.ids.resize_uninitialized(nq * k_src);
    std::memcpy(src.scores.data(), src_scores, nq * k_src * sizeof(float));
    std::memcpy(src.ids.data(), src_ids, nq * k_src * sizeof(uint64_t));
    vc::merge_search_results(dst, src);
    std::memcpy(dst_scores, dst.scores.data(), nq * k_dst * sizeof(float));
    std::memcpy(dst_ids, dst.ids.data(), nq * k_dst * sizeof(uint64_t));
  });
}

// score_prepared_into() (reference src/search/search.cpp:1156-1267): the nq
// queries are prepared as one batch; the heaps [nq][k] (heap_s, heap_i) and
// their state [nq] (heap_sz, heap_min, heap_mi) are updated in place for the
// queries listed in qidx only; offsets may be null.
int vcref_score_into(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                     const uint8_t* blocked, size_t n_blocks, const float* alphas, size_t n,
                     const uint64_t* id_map, const float* shift, const float* scale,
                     const size_t* qidx, size_t n_qidx, float* heap_s, uint64_t* heap_i,
                     size_t* heap_sz, float* heap_min, size_t* heap_mi, const float* offsets) {
  return guarded([&] {
    vc::Rotation rot(dim);
    auto [b, c] = vc::codebook(bits, dim);
    std::span<const float> sh, sc;
    if (shift != nullptr && scale != nullptr) {
      sh = std::span<const float>(shift, dim);
      sc = std::span<const float>(scale, dim);
    }
    auto prep = vc::prepare_queries(std::span<const float>(q, nq * dim), nq, dim, rot, c, bits, sh,
                                    sc);
    if (prep.backend != vc::SearchBackendKind::Avx2) {
      throw std::runtime_error("reference did not select the AVX2 PERM0 backend");
    }
    const auto geo = vc::blocked_geometry(n, bits, dim);
    std::span<const uint64_t> im;
    if (id_map != nullptr) im = std::span<const uint64_t>(id_map, n);
    vc::score_prepared_into(prep, k, std::span<const uint8_t>(blocked, geo.blocked_len), n_blocks,
                            std::span<const float>(alphas, n), im,
                            std::span<const size_t>(qidx, n_qidx), heap_s, heap_i, heap_sz,
                            heap_min, heap_mi, offsets);
  });
}

// heaps_to_search_results() (reference src/search/search.cpp:1269-1291).
int vcref_heaps_to_results(size_t nq, size_t k, const float* heap_s, const uint64_t* heap_i,
                           const size_t* heap_sz, float* scores_out, uint64_t* ids_out) {
  return guarded([&] {
    vc::SearchResults out;
    vc::heaps_to_search_results(out, nq, k, heap_s, heap_i, heap_sz);
    std::memcpy(scores_out, out.scores.data(), nq * k * sizeof(float));
    std::memcpy(ids_out, out.ids.data(), nq * k * sizeof(uint64_t));
  });
}

// float_rescore_topk() (reference src/search/float_rescore.cpp:80-164): the
// nq x k rows re-scored in place; counts may be null (every row has k hits).
int vcref_float_rescore(size_t nq, size_t k, float* scores, uint64_t* ids, const size_t* counts,
                        const float* queries, const float* corpus, size_t n_corpus, size_t dim) {
  return guarded([&] {
    vc::SearchResults out;
    out.nq = nq;
    out.k = k;
    out.scores.resize_uninitialized(nq * k);
    out.ids.resize_uninitialized(nq * k);
    std::memcpy(out.scores.data(), scores, nq * k * sizeof(float));
    std::memcpy(out.ids.data(), ids, nq * k * sizeof(uint64_t));
    if (counts != nullptr) out.counts.assign(counts, counts + nq);
    vc::float_rescore_topk(out, std::span<const float>(queries, nq * dim),
                           std::span<const float>(corpus, n_corpus * dim), dim);
    std::memcpy(scores, out.scores.data(), nq * k * sizeof(float));
    std::memcpy(ids, out.ids.data(), nq * k * sizeof(uint64_t));
  });
}

// make_ortho_landmarks() (reference src/transform/landmarks.cpp:180-198):
// L is dim x m, column j at L[j * dim].
int vcref_ortho_landmarks(size_t dim, size_t m, uint64_t seed, float* L) {
  return guarded([&] {
    auto v = vc::make_ortho_landmarks(dim, m, seed);
    std::memcpy(L, v.data(), dim * m * sizeof(float));
  });
}

// quantize_landmarks() (reference src/transform/landmarks.cpp:236-269):
// codes m x (pad16(dim) / 2) bytes, scales m.
int vcref_quantize_landmarks(const float* L, size_t dim, size_t m, uint8_t* codes, float* scales) {
  return guarded([&] {
    auto q = vc::quantize_landmarks(std::span<const float>(L, dim * m), dim, m);
    std::memcpy(codes, q.codes.data(), q.codes.size());
    std::memcpy(scales, q.scales.data(), m * sizeof(float));
  });
}

// project_rows_onto_landmarks() (reference src/transform/landmarks.cpp:200-234).
int vcref_project_rows(const float* Y, size_t n, size_t dim, const float* L, size_t m, float* out,
                       size_t out_stride) {
  return guarded([&] { vc::project_rows_onto_landmarks(Y, n, dim, L, m, out, out_stride); });
}

// project_onto_landmarks_q4() (reference src/transform/landmarks.cpp:271-348)
// with the 4-bit basis quantize_landmarks() makes of L.
int vcref_project_q4(const float* q_rot, size_t nq, size_t dim, const float* L, size_t m, float* q_land,
                     size_t scan_dim) {
  return guarded([&] {
    auto q = vc::quantize_landmarks(std::span<const float>(L, dim * m), dim, m);
    vc::project_onto_landmarks_q4(q_rot, nq, dim, q, q_land, scan_dim);
  });
}

// LandmarkIndex end to end (reference src/index.cpp:120-289): the rows of x
// added in batches of batch_rows[b], prepare() after every batch when
// prepare_between, else once at the end, then search() with the optional
// float corpus.
// info[0..2] = scan_dim, codes_bytes(), index_bytes(); search_seconds (may
// be null) = wall time of the search() call.
static int lm_search_run(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                         const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                         size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                         uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds, bool avx2) {
  return guarded([&] {
    vc::LandmarkIndex ix(dim, scan_bits, m, seed, rerank_bits);
    size_t off = 0;
    for (size_t b = 0; b < n_batches; ++b) {
      ix.add(std::span<const float>(x + off * dim, batch_rows[b] * dim));
      off += batch_rows[b];
      if (prepare_between) ix.prepare();
    }
    if (!prepare_between) ix.prepare();  // a second prepare() with no add reads released rows
    std::span<const float> fc;
    if (float_corpus != nullptr) fc = std::span<const float>(float_corpus, off * dim);
    const auto t0 = std::chrono::steady_clock::now();
    auto res = ix.search(std::span<const float>(q, nq * dim), k, fc);
    if (search_seconds != nullptr)
      *search_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *k_out = res.k;
    if (res.k > 0 && nq > 0) {
      std::memcpy(scores_out, res.scores.data(), res.nq * res.k * sizeof(float));
      std::memcpy(ids_out, res.ids.data(), res.nq * res.k * sizeof(uint64_t));
    }
    if (info != nullptr) {
      info[0] = ix.scan_dim();
      info[1] = ix.codes_bytes();
      info[2] = ix.index_bytes();
    }
  }, avx2);
}

int vcref_lm_search(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                    const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                    size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                    uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds) {
  return lm_search_run(x, batch_rows, n_batches, prepare_between, q, nq, dim, scan_bits, m, seed, rerank_bits, k,
                       float_corpus, scores_out, ids_out, k_out, info, search_seconds, true);
}

// The sign-plane scan alone on an AVX-512 VNNI host: prepare_sign_from_rows
// of the rows (nq x S) and score_prepared over planes_repack_sign of the
// bit-plane codes (n x 4 S / 8), top k (src/search/search.cpp:916-971,
// src/pack/pack.cpp:490-541, the VNNI sign kernels).
int vcref_sign_topk(const float* rows, size_t nq, size_t S, const uint8_t* packed, size_t n, size_t k,
                    float* scores_out, uint64_t* ids_out, size_t* k_out) {
  unsetenv("TURBOVEC_NO_VNNI");
  const int rc = guarded([&] {
    auto sign = vc::planes_repack_sign(std::span<const uint8_t>(packed, n * 4 * S / 8), n, 4, S);
    auto prep = vc::prepare_sign_from_rows(std::span<const float>(rows, nq * S), nq, S);
    std::vector<float> ones(n, 1.f);
    auto res = vc::score_prepared(prep, k, sign.first, sign.second, std::span<const float>(ones));
    *k_out = res.k;
    if (res.k > 0 && nq > 0) {
      std::memcpy(scores_out, res.scores.data(), res.nq * res.k * sizeof(float));
      std::memcpy(ids_out, res.ids.data(), res.nq * res.k * sizeof(uint64_t));
    }
  }, false);
  force_avx2_path();
  return rc;
}

// vcref_lm_search as the reference runs on an AVX-512 VNNI host: the call
// sees the environment without TURBOVEC_NO_VNNI, so prepare() keeps the sign
// planes (n >= 32768) and search() takes the staged branch
// (src/search/search.cpp:1379-1510); every other entry point keeps forcing
// the AVX2 path.
int vcref_lm_search_planes(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                           const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                           size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                           uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds) {
  unsetenv("TURBOVEC_NO_VNNI");
  const int rc = lm_search_run(x, batch_rows, n_batches, prepare_between, q, nq, dim, scan_bits, m, seed,
                               rerank_bits, k, float_corpus, scores_out, ids_out, k_out, info, search_seconds, false);
  force_avx2_path();
  return rc;
}

}  // extern "C"
