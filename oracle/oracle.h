/* Test infrastructure only — the CPU oracle for the plain flat TurboQuant
 * encode + search path. Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it, and only as the checker. The product path
 * (libvectorcache_hip.so) never links or calls it.
 *
 * All entry points return 0 on success, a negative value on bad arguments. */
#ifndef VC_ORACLE_H
#define VC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int oc_rotation_tables(size_t dim, uint32_t* perm, float* signs, float* inv_sqrt_block,
                       size_t* block);
int oc_codebook(size_t bits, size_t dim, float* boundaries, float* centroids);
double oc_beta_cdf_aa(double a, double t);
float oc_row_norm(const float* row, size_t dim);
int oc_rotate_scaled(size_t dim, const float* src, float inv, float* dst);
int oc_encode(const float* x, size_t n, size_t dim, size_t bits, const float* shift,
              const float* scale, uint8_t* packed_out, float* alpha_out, float* rotated_out);
int oc_fit_calibration(const float* x, size_t n, size_t dim, size_t bits, float* shift_out,
                       float* scale_out);
int oc_codes_flat(const uint8_t* packed, size_t n, size_t bits, size_t dim, uint8_t* out);
int oc_pack_perm0(const uint8_t* flat, size_t n, size_t n_groups, uint8_t* out);
int oc_pack_seq(const uint8_t* flat, size_t n, size_t n_groups, uint8_t* out);
int oc_unpack_perm0(const uint8_t* blocked, size_t n, size_t n_groups, uint8_t* flat_out);
int oc_flat_to_gpu32(const uint8_t* flat, size_t n, size_t n_groups, uint32_t* words);
int oc_gpu32_to_flat(const uint32_t* words, size_t n, size_t n_groups, uint8_t* flat_out);
int oc_prepare_queries(const float* q, size_t nq, size_t dim, size_t bits, const float* shift,
                       const float* scale, float* q_rot_out, float* bias_corr_out,
                       uint8_t* luts_out, float* lut_scale_out, float* lut_bias_out);
int oc_window_sums(const uint8_t* lut, const uint8_t* flat, size_t n, size_t n_groups,
                   uint32_t* sums_out);
int oc_scan_topk(const uint8_t* luts, const float* lut_scale, const float* lut_bias_total,
                 size_t nq, const uint8_t* flat, size_t n, size_t n_groups, const float* alphas,
                 const uint64_t* id_map, size_t k, float* scores_out, uint64_t* ids_out,
                 size_t* k_out);
int oc_search_flat(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                   const uint8_t* blocked_perm0, const float* alphas, size_t n,
                   const float* shift, const float* scale, float* scores_out, uint64_t* ids_out,
                   size_t* k_out);
int oc_score_into(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                  const uint8_t* blocked_perm0, const float* alphas, size_t n,
                  const uint64_t* id_map, const float* shift, const float* scale,
                  const size_t* qidx, size_t n_qidx, float* heap_s, uint64_t* heap_i,
                  size_t* heap_sz, float* heap_min, size_t* heap_mi, const float* offsets);
int oc_heaps_to_results(size_t nq, size_t k, const float* heap_s, const uint64_t* heap_i,
                        const size_t* heap_sz, float* scores_out, uint64_t* ids_out);
int oc_float_rescore(size_t nq, size_t k, float* scores, uint64_t* ids, const size_t* counts,
                     const float* queries, const float* corpus, size_t n_corpus, size_t dim);
int oc_score_prepared(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                      const uint8_t* blocked_perm0, const float* alphas, size_t n, const float* shift,
                      const float* scale, const uint64_t* id_map, const size_t* query_indices,
                      size_t n_indices, const float* residual_norms, const float* residual_scales,
                      float* scores_out, uint64_t* ids_out, size_t* k_out);
int oc_merge_results(size_t nq, size_t k_dst, float* dst_scores, uint64_t* dst_ids, size_t k_src,
                     const float* src_scores, const uint64_t* src_ids);
void oc_set_threads(int n);
/* landmark index (reference src/index.cpp, src/transform/landmarks.cpp) */
int oc_ortho_landmarks(size_t dim, size_t m, uint64_t seed, float* L);
int oc_quantize_landmarks(const float* L, size_t dim, size_t m, uint8_t* codes, float* scales);
int oc_project_rows(const float* Y, size_t n, size_t dim, const float* L, size_t m, float* out,
                    size_t stride);
int oc_project_q4(const float* q_rot, size_t nq, size_t dim, const float* L, size_t m, float* q_land,
                  size_t stride);
int oc_lm_search(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                 const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                 size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                 uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds);
/* as oc_lm_search on an AVX-512 VNNI host: the staged sign-plane branch
 * (n >= 32768 prepared vectors, scan_bits 4); -2 when that branch is not taken */
int oc_lm_search_planes(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                        const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                        size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                        uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds);
size_t oc_planes_exact_len(size_t k);
int oc_sign_topk(const float* rows, size_t nq, size_t S, const uint8_t* packed, size_t n, size_t k,
                 float* scores_out, uint64_t* ids_out, size_t* k_out);
void oc_set_staged_variant(int v);
/* bucketed index (README.md:56: rows L2-normalized, residuals x_hat - c,
 * scores <q_hat, c> + alpha <q_hat, r_hat>) */
int oc_bucket_residuals(const float* x, size_t dim, const uint64_t* pos_rows, const uint32_t* pos_bucket,
                        size_t n, const float* cent, float* out);
int oc_probe_offsets(const float* q, size_t nq, size_t dim, const float* cent, const int32_t* probe_b,
                     size_t P, const int32_t* probe_n, float* off);
/* the ChaCha block function (8 rounds for the rotation, 20 for RFC 7539 tests) */
void oc_chacha_block(const uint32_t in[16], uint32_t out[16], int rounds);

#ifdef __cplusplus
}
#endif

#endif
