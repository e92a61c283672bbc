/* vectorcache_hip.h — C ABI of libvectorcache_hip.so, the MI355X (gfx950)
 * implementation of VectorCache's plain flat TurboQuant encode + search path.
 *
 * Plain C: pointers and sizes only. Every call returns a vc_status; on failure
 * vc_last_error() holds a thread-local message. The C++20 TurboQuantIndex
 * wrapper (vectorcache_amd/csrc/include/vectorcache/turbo_quant_index.hpp)
 * maps statuses back onto the reference's exception types:
 *   VC_ERR_INVALID_ARGUMENT -> std::invalid_argument
 *   VC_ERR_STATE            -> std::runtime_error   (e.g. search before prepare)
 *   VC_ERR_OUT_OF_RANGE     -> std::out_of_range
 *   VC_ERR_DEVICE / OOM     -> std::runtime_error
 *
 * Numerics: codes, PERM0 blocks, alpha, per-query LUTs, LUT-scan integer sums,
 * scores and top-k ids are bit-identical to the reference CPU path
 * (AVX2 PERM0 uint8-LUT scan) on the same inputs. Result rows are ordered by
 * (score desc, id asc); a -0.0f score is reported as +0.0f.
 *
 * "host" pointers are ordinary CPU memory; "device" pointers are HIP device
 * memory on the index's device; `stream` is a hipStream_t (NULL = the null
 * stream, as in every HIP call; host-buffer calls use the index's own stream).
 */
#ifndef VECTORCACHE_HIP_H
#define VECTORCACHE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  VC_OK = 0,
  VC_ERR_INVALID_ARGUMENT = 1,
  VC_ERR_STATE = 2,
  VC_ERR_OUT_OF_RANGE = 3,
  VC_ERR_DEVICE = 4,
  VC_ERR_OOM = 5
} vc_status;

typedef struct vc_index vc_index;
typedef struct vc_lm_index vc_lm_index;

/* Thread-local message of the last failing call on this thread. */
const char* vc_last_error(void);

/* Library version string. */
const char* vc_version(void);

/* ---- TurboQuantIndex ------------------------------------------------------
 * Replaces the reference's index class shape (include/vectorcache/index.hpp:18-78:
 * ctor(dim, bits), add, prepare, search, dim/size/bit_width/codes_bytes/
 * index_bytes) for the plain flat TurboQuant path, plus calibrate()
 * (README.md:108; semantics of fit_calibration, include/vectorcache/encode/
 * encode.hpp:33-35).
 * dim: positive multiple of 8, <= 16384 (src/index.cpp:129-131);
 * bits in {2, 3, 4} (src/index.cpp:132-137). */
vc_status vc_index_create(size_t dim, size_t bits, vc_index** out);
vc_status vc_index_destroy(vc_index* index);

/* TQ+ calibration from n sample rows (host, row-major n x dim). Must precede
 * the first prepare() that encodes vectors. n >= 2 (src/encode/encode.cpp:327-335). */
vc_status vc_index_calibrate(vc_index* index, const float* sample, size_t n);
/* Same, sample already in device memory. */
vc_status vc_index_calibrate_device(vc_index* index, const float* sample_dev, size_t n);

/* Append n row-major vectors (host) — replaces add(std::span<const float>)
 * (src/index.cpp:151-213). Vectors are staged on the device; prepare() encodes
 * them. add() after prepare() is allowed and marks the index unprepared. */
vc_status vc_index_add(vc_index* index, const float* vectors, size_t n);
/* Same, rows in device memory: they are read after the work enqueued so far
 * on `stream` (NULL: the null stream), and work enqueued on `stream` after
 * the call runs after the copy, so the caller may reuse the rows' memory at
 * once. No device-wide synchronization. */
vc_status vc_index_add_device(vc_index* index, const float* vectors_dev, size_t n, void* stream);

/* Encode every staged vector on the device (replaces prepare(),
 * src/index.cpp:215-265). */
vc_status vc_index_prepare(vc_index* index);

/* Top-k search — replaces search(queries, k) (src/index.cpp:267-289) and
 * search_flat (src/search/search.cpp:1293-1302). queries: nq x dim host rows.
 * k is clipped to size(); *k_out receives the clipped k. scores / ids are
 * nq x (*k_out) row-major, caller-allocated with room for nq * min(k, size()).
 * Search before prepare -> VC_ERR_STATE. */
vc_status vc_index_search(vc_index* index, const float* queries, size_t nq, size_t k,
                          float* scores, uint64_t* ids, size_t* k_out);
/* Same with device-resident queries and outputs, ordered on `stream`
 * (NULL: the null stream). The call returns once the work is enqueued. */
vc_status vc_index_search_device(vc_index* index, const float* queries_dev, size_t nq, size_t k,
                                 float* scores_dev, uint64_t* ids_dev, size_t* k_out,
                                 void* stream);

/* Search of the first n_prefix vectors only (n_prefix <= size()): the result
 * a search issued while the index held n_prefix vectors returned, since a
 * vector's code depends only on its own row and the calibration. For
 * checking a request served during concurrent adds against the same request
 * alone. No counterpart in the reference. */
vc_status vc_index_search_prefix_device(vc_index* index, const float* queries_dev, size_t nq, size_t k,
                                        size_t n_prefix, float* scores_dev, uint64_t* ids_dev, size_t* k_out,
                                        void* stream);

/* vc_index_search_device without allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results as vc_index_search_device; the caller passes inputs
 * and outputs, and the search's scratch (LUTs, pair-table sources, per-range
 * top-k lists) lives in one fixed array per GPU of vc_noalloc_scratch_bytes()
 * bytes (16 MiB), held until the process ends. A search that needs more
 * (many queries, a large k or dim) returns VC_ERR_INVALID_ARGUMENT and changes
 * nothing. Nothing in the scratch needs to start at zero. Every call on one
 * GPU shares that array: calls that overlap on one GPU (two streams, two host
 * threads on different indexes) race and corrupt each other's results. The
 * index's device must be the current device. */
vc_status vc_index_search_device_noalloc(vc_index* index, const float* queries_dev, size_t nq, size_t k,
                                         float* scores_dev, uint64_t* ids_dev, size_t* k_out, void* stream);
size_t vc_noalloc_scratch_bytes(void);

/* Streams: the vc_index_* calls of one index share its scratch and staging
 * buffers, so one index takes work from one stream at a time (calls are
 * serialized by a mutex, but the enqueued work of two streams may overlap on
 * the device). Concurrent requests go through a vc_server. */

/* Accessors (index.hpp:23-40). */
size_t vc_index_dim(const vc_index* index);
size_t vc_index_size(const vc_index* index);
size_t vc_index_bits(const vc_index* index);
int vc_index_is_prepared(const vc_index* index);
int vc_index_is_calibrated(const vc_index* index);
size_t vc_index_codes_bytes(const vc_index* index);
size_t vc_index_index_bytes(const vc_index* index);

/* ---- exports in the reference's formats (parity and interop) --------------
 * packed: bit-plane codes exactly as encode() writes them (n x bits*dim/8,
 *         src/encode/encode.cpp:112-157);
 * flat:   row-major byte groups (n x G, codes_flat_from_packed, src/pack/pack.cpp:98-131);
 * perm0:  192x1 PERM0 blocked bytes (ceil(n/192) x G x 192, repack_perm0,
 *         src/pack/pack.cpp:157-186,388-401);
 * alpha:  per-vector scale (n floats);
 * calibration: TQ+ shift/scale (dim floats each; identity when uncalibrated). */
vc_status vc_index_export_packed(vc_index* index, uint8_t* packed_host);
vc_status vc_index_export_flat(vc_index* index, uint8_t* flat_host);
vc_status vc_index_export_perm0(vc_index* index, uint8_t* perm0_host);
/* seq: 192x1 sequential blocked bytes (ceil(n/192) x G x 192, repack_seq,
 *      src/pack/pack.cpp:133-154,358-364). */
vc_status vc_index_export_seq(vc_index* index, uint8_t* seq_host);
vc_status vc_index_export_alpha(vc_index* index, float* alpha_host);
/* gpu32: the reference CUDA index's code layout, u32 words
 *        [ceil(n/32)][ceil(G/4)][32] (src/gpu/kernels.cuh:18-21, written by
 *        encode_coop, src/gpu/kernels.cu:111-133): word (tile, g4, lane) holds
 *        byte groups 4 g4 .. 4 g4 + 3 of vector 32 tile + lane, little-endian;
 *        groups past G and vectors past n are zero. vc_gpu32_words(n, dim,
 *        bits) words (0 for a bad dim / bits). */
vc_status vc_index_export_gpu32(vc_index* index, uint32_t* codes_host);
size_t vc_gpu32_words(size_t n, size_t dim, size_t bits);
vc_status vc_index_get_calibration(const vc_index* index, float* shift_host, float* scale_host);

/* Per-query state as prepare_queries builds it (src/search/search.cpp:762-860):
 * luts nq x G*32 u8, lut_scale, lut_bias (without bias_corr), bias_corr. */
vc_status vc_index_prepare_queries(vc_index* index, const float* queries, size_t nq,
                                   uint8_t* luts, float* lut_scale, float* lut_bias,
                                   float* bias_corr);

/* ---- landmark index --------------------------------------------------------
 * The reference's LandmarkIndex (include/vectorcache/index.hpp:16-68,
 * src/index.cpp:120-289), which its TurboQuantIndex aliases, on the AVX2
 * path: add() encodes alpha * y-hat, projects it onto an orthonormal landmark
 * basis and keeps the residual norm; prepare() encodes the landmark rows at
 * scan_dim = n_landmarks rounded up to 8; search() rotates the queries,
 * projects them onto the 4-bit basis, scans and applies the residual
 * post-correction (src/search/search.cpp:1304-1546). Results equal the
 * reference's bit for bit. Rows added after prepare() find the earlier rows
 * zeroed at the next prepare(), as in the reference.
 * create: dim a multiple of 8 (<= 16384), scan_bits and rerank_bits in 2..4
 * (rerank_bits 0 = scan_bits), 1 <= n_landmarks <= dim (the reference's
 * default is 1024). */
vc_status vc_lm_index_create(size_t dim, size_t scan_bits, size_t n_landmarks, uint64_t landmark_seed,
                             size_t rerank_bits, vc_lm_index** out);
vc_status vc_lm_index_destroy(vc_lm_index* index);
vc_status vc_lm_index_add(vc_lm_index* index, const float* vectors, size_t n);
/* Rows in device memory, read after the work enqueued so far on `stream`
 * (NULL: the null stream); the call returns once they are encoded. */
vc_status vc_lm_index_add_device(vc_lm_index* index, const float* vectors_dev, size_t n, void* stream);
vc_status vc_lm_index_prepare(vc_lm_index* index);
/* search(queries, k, float_corpus): k_eff = min(k, size()) <= 2048; with a
 * float corpus (size() x dim rows, may be NULL) the top k is re-scored by exact
 * inner products (float_rescore_topk). The host call reads the corpus on the
 * host and uploads only the nq x k_eff candidate rows (one synchronization
 * between the search and the re-score). VC_ERR_STATE before prepare(). */
vc_status vc_lm_index_search(vc_lm_index* index, const float* queries, size_t nq, size_t k,
                             const float* float_corpus, float* scores, uint64_t* ids, size_t* k_out);
vc_status vc_lm_index_search_device(vc_lm_index* index, const float* queries_dev, size_t nq, size_t k,
                                    const float* float_corpus_dev, float* scores_dev, uint64_t* ids_dev,
                                    size_t* k_out, void* stream);
/* The query side alone: landmark rows q_land (nq x scan_dim, zero past
 * n_landmarks) and residual scales q_perp (nq), as search() computes them. */
vc_status vc_lm_index_query_rows(vc_lm_index* index, const float* queries, size_t nq, float* q_land,
                                 float* q_perp);
/* Per-vector state: residual norms (size()), and after prepare() the scan
 * scales (alpha) and flat codes (n x scan_dim / (8 / scan_bits) bytes); each
 * output may be NULL. */
vc_status vc_lm_index_export(vc_lm_index* index, float* residual_norms, float* scales, uint8_t* codes_flat);
size_t vc_lm_index_dim(const vc_lm_index* index);
size_t vc_lm_index_size(const vc_lm_index* index);
size_t vc_lm_index_scan_bits(const vc_lm_index* index);
size_t vc_lm_index_rerank_bits(const vc_lm_index* index);
size_t vc_lm_index_n_landmarks(const vc_lm_index* index);
size_t vc_lm_index_scan_dim(const vc_lm_index* index);
int vc_lm_index_is_prepared(const vc_lm_index* index);
/* codes_bytes() / index_bytes() as the reference reports them (its PERM0
 * blocks of 192 vectors; index_bytes adds residual norms, scales and the
 * scale/residual pairs). */
/* The reference's behaviour on an AVX-512 VNNI host (on = 1; default 0, its
 * AVX2 behaviour): prepare() keeps the sign plane of the 4-bit scan codes (the
 * top bit of each code) once at least 32768 vectors are prepared
 * (src/pack/pack.cpp:428-449,490-541), and search() then takes the staged
 * branch of search_landmark_rerank (src/search/search.cpp:1379-1510) whenever
 * its shortlist is shorter than the index: the E = max(256, 6k) best sign
 * scores by (score desc, id asc) (build_sign_lut LUTs, score = fma(A,
 * lut_scale, lut_bias), no alpha; the reference's buffered collector returns
 * exactly that set, ties included), then those E rescored with
 * build_permute_dot's int8 weights and levels plus the residual term, into a
 * k-heap. Results equal the reference's VNNI host's staged search; k up to 341
 * (E <= 2048, else VC_ERR_INVALID_ARGUMENT) and at most 2048 landmarks.
 * Planes are kept only when n_landmarks rounds to a multiple of 32 (the
 * reference has no sign kernel for other widths and reads past its sign
 * codes there); below 32768 vectors a VNNI host scans with its permute-dot
 * kernels, which are not rebuilt. In both cases the index keeps the AVX2
 * scan. Takes effect at the next prepare() (or at once on a prepared index). */
vc_status vc_lm_index_set_planes(vc_lm_index* index, int on);
int vc_lm_index_planes_active(const vc_lm_index* index);
size_t vc_lm_index_codes_bytes(const vc_lm_index* index);
size_t vc_lm_index_index_bytes(const vc_lm_index* index);

/* ---- bucketed index ----------------------------------------------------------
 * The README's BucketedTurboQuantIndex (README.md:56,114-120; the reference has
 * prose and an API sketch, no code): rows are L2-normalized and grouped into
 * buckets; each bucket's rows are encoded with the flat index's encode, as
 * residuals x_hat - c_b against the bucket's encode centroid (residual = 1) or
 * as themselves (residual = 0); a query opens its n_probe best buckets by
 * routing score and ranks their vectors with <q_hat, c_b> + alpha <q_hat, r_hat>.
 * Each opened bucket is scored exactly as score_prepared_into scores one
 * bucket (src/search/search.cpp:1156-1267, score_offsets = <q_hat, c_b>); the
 * result rows are the top k of the opened vectors by (score desc, id asc).
 * Build (prepare): a coarse quantizer of ceil(n / energy_soft) lists trained
 * by spherical k-means (kmeans_iters Lloyd iterations on kmeans_sample rows
 * per list, strided over the rows; the seed offsets the initial centroids),
 * every row assigned to its nearest list, and every bucket above
 * max_bucket_size split by one iteration of spherical 2-means until none is.
 * Routing and assignment run on bf16 matrix cores: they choose buckets only.
 * The offset is (float) of a sequential double sum over the dimensions of
 * q_hat[d] c_b[d]; q_hat = q * (1 / ||q||) with the encode's row norm.
 * prepare() rebuilds from every row added so far (ids = insertion order). */
typedef struct vc_bucketed vc_bucketed;
typedef struct {
  size_t n_probe;          /* buckets a query opens without a learned or per-query count (64) */
  size_t max_bucket_size;  /* hard cap at the build (800) */
  size_t energy_soft;      /* mean bucket size the coarse quantizer aims at (400) */
  size_t expected_n;       /* capacity hint, rows (0) */
  size_t kmeans_iters;     /* Lloyd iterations of the coarse quantizer (8) */
  size_t kmeans_sample;    /* training rows per list (64) */
  uint64_t seed;           /* (4001) */
  int residual;            /* 1: residual codes with score offsets; 0: codes of the rows (1) */
} vc_bucket_params;
void vc_bucket_params_default(vc_bucket_params* params);
/* dim a multiple of 8, <= 4096; bits in {2, 3, 4}; params NULL = defaults. */
vc_status vc_bucketed_create(size_t dim, size_t bits, const vc_bucket_params* params, vc_bucketed** out);
vc_status vc_bucketed_destroy(vc_bucketed* index);
/* TQ+ calibration: fitted at prepare() on the sample's residuals (or rows). */
vc_status vc_bucketed_calibrate(vc_bucketed* index, const float* sample, size_t n);
vc_status vc_bucketed_calibrate_device(vc_bucketed* index, const float* sample_dev, size_t n, void* stream);
vc_status vc_bucketed_add(vc_bucketed* index, const float* vectors, size_t n);
/* Device rows, copied in order after `stream` (NULL: the null stream). */
vc_status vc_bucketed_add_device(vc_bucketed* index, const float* vectors_dev, size_t n, void* stream);
vc_status vc_bucketed_prepare(vc_bucketed* index);
/* Host queries; probe counts from the learned model when one is enabled, else
 * params.n_probe. k_eff = min(k, size()) <= 2048. Rows with fewer than k_eff
 * opened vectors end with (0, 0). */
vc_status vc_bucketed_search(vc_bucketed* index, const float* queries, size_t nq, size_t k, float* scores,
                             uint64_t* ids, size_t* k_out);
/* Device queries and outputs, ordered on `stream` (NULL: the null stream).
 * n_probe > 0: every query opens min(n_probe, 512, n_buckets) buckets, or with
 * probe_counts_dev (int32 [nq]) query q opens clamp(counts[q], 1, that bound);
 * n_probe = 0: the learned model when enabled, else params.n_probe.
 * counts_dev (u64 [nq], may be NULL) receives the entries found per row.
 * Batches that put several queries on a bucket (nq * probe width >= 8 *
 * n_buckets, k <= 64) are scanned bucket-major (each bucket's codes read once
 * per tile of up to 8 of its queries), smaller ones query-major; the results
 * are the same. */
vc_status vc_bucketed_search_device(vc_bucketed* index, const float* queries_dev, size_t nq, size_t k, size_t n_probe,
                                    const int32_t* probe_counts_dev, float* scores_dev, uint64_t* ids_dev,
                                    uint64_t* counts_dev, size_t* k_out, void* stream);
/* residual = 0 only: the plain flat scan (the flat index's byte-pair LUT
 * scan and exact top-k) over every code of the index, ids by insertion order:
 * the flat index's search of the normalized rows, bit for bit, and what a
 * bucketed search with every bucket open returns. The judge of the probe
 * budget's recall. */
vc_status vc_bucketed_search_exhaustive_device(vc_bucketed* index, const float* queries_dev, size_t nq, size_t k,
                                               float* scores_dev, uint64_t* ids_dev, size_t* k_out, void* stream);
/* Each query's P best buckets by routing score (bucket ids int32, scores f32,
 * nq x P, sorted by (score desc, bucket asc)); 1 <= P <= min(512, n_buckets). */
vc_status vc_bucketed_route_device(vc_bucketed* index, const float* queries_dev, size_t nq, size_t P,
                                   int32_t* probe_b_dev, float* probe_s_dev, void* stream);
size_t vc_bucketed_dim(const vc_bucketed* index);
size_t vc_bucketed_size(const vc_bucketed* index);
size_t vc_bucketed_bits(const vc_bucketed* index);
size_t vc_bucketed_n_buckets(const vc_bucketed* index);
int vc_bucketed_is_prepared(const vc_bucketed* index);
int vc_bucketed_is_calibrated(const vc_bucketed* index);
/* The built state (each output may be NULL): bucket sizes (n_buckets), encode
 * centroids (n_buckets x dim), the id of every position (size(); positions
 * run bucket by bucket), flat codes by position (size() x G, the reference's
 * codes_flat_from_packed bytes), alpha by position, TQ+ shift / scale (dim). */
vc_status vc_bucketed_export(vc_bucketed* index, uint32_t* bucket_sizes, float* centroids, uint64_t* pos_ids,
                             uint8_t* flat_codes, float* alpha, float* shift, float* scale);

/* ---- learned probe budget (no reference counterpart) -----------------------
 * An MLP 96 -> 256 -> 256 -> 1 (ReLU) per bucketed index, trained and run on
 * the GPU with hand-written bf16 matrix-core kernels (fp32 accumulation, fp32
 * master weights, Adam). Inputs of a query (vc_bucketed_probe_features): its
 * best routing score, 16 x the gaps to its next 63, and log2(1 + size) / 10 of
 * its 32 best buckets. Output: log2 of the buckets it needs; once enabled, a
 * search with n_probe = 0 opens clamp(ceil(2^(y + shift)), 1, 256) buckets,
 * the model running in one kernel on the routing lists. Parameters: fp32
 * vc_probe_param_count() values, W1 [256][96], b1, W2 [256][256], b2, w3, b3. */
size_t vc_probe_param_count(void);
vc_status vc_bucketed_probe_features(vc_bucketed* index, const float* queries_dev, size_t nq, float* feats_dev,
                                     void* stream);
vc_status vc_bucketed_probe_init(vc_bucketed* index, uint64_t seed);
vc_status vc_bucketed_probe_set_weights(vc_bucketed* index, const float* weights_host);
vc_status vc_bucketed_probe_get_weights(vc_bucketed* index, float* weights_host);
/* n_steps Adam steps (lr, beta1 0.9, beta2 0.999, eps 1e-8) on minibatches of
 * 64 rows: rows order_dev[64 s .. 64 s + 63] of feats_dev (n x 96) and
 * targets_dev (n); losses_dev (n_steps, may be NULL) the minibatch MSEs. */
vc_status vc_bucketed_probe_train(vc_bucketed* index, const float* feats_dev, const float* targets_dev, size_t n,
                                  const int32_t* order_dev, size_t n_steps, float lr, float* losses_dev, void* stream);
/* The gradients of one minibatch (order_dev[0..63]) without an update. */
vc_status vc_bucketed_probe_grads(vc_bucketed* index, const float* feats_dev, const float* targets_dev, size_t n,
                                  const int32_t* order_dev, float* grads_host);
vc_status vc_bucketed_probe_predict(vc_bucketed* index, const float* feats_dev, size_t n, float* y_dev, void* stream);
/* Probe table: a prediction y opens counts[b] buckets, b = clamp(floor((y -
 * y0) / y_step), 0, n_bins - 1) (n_bins <= 64, counts >= 1; n_bins = 0 goes
 * back to ceil(2^(y + shift))). The counts are chosen by the caller, e.g. to
 * meet a mean recall at the least scanned vectors on validation queries. */
vc_status vc_bucketed_probe_set_table(vc_bucketed* index, const int32_t* counts, size_t n_bins, float y0,
                                      float y_step);
/* on = 1: searches with n_probe = 0 use the model with this log2 shift. */
vc_status vc_bucketed_probe_enable(vc_bucketed* index, int on, float shift);

/* ---- stateless entry points (host buffers) ---------------------------------
 * Host precompute, bit-identical to the reference:
 * Rotation tables (src/transform/rotation.cpp:139-160), codebook
 * (src/quantize/codebook.cpp:203-226). */
vc_status vc_rotation_tables(size_t dim, uint32_t* perm, float* signs, float* inv_sqrt_block,
                             size_t* block);
vc_status vc_codebook(size_t bits, size_t dim, float* boundaries, float* centroids);
/* make_ortho_landmarks (src/transform/landmarks.cpp:180-198): L dim x m floats,
 * column j at L[j * dim]; 0 < m <= dim, dim a multiple of 8. */
vc_status vc_landmarks_make(size_t dim, size_t m, uint64_t seed, float* L);
/* quantize_landmarks (src/transform/landmarks.cpp:236-269): codes
 * m x (pad16(dim) / 2) bytes (level + 7 per nibble, low nibble = even dim),
 * scales m. */
vc_status vc_landmarks_quantize(const float* L, size_t dim, size_t m, uint8_t* codes, float* scales);

/* encode() on the GPU (src/encode/encode.cpp:294-307): n x dim host rows ->
 * bit-plane codes and alpha. shift/scale: TQ+ calibration or both NULL. */
vc_status vc_encode(const float* vectors, size_t n, size_t dim, size_t bits, const float* shift,
                    const float* scale, uint8_t* packed_out, float* alpha_out);

/* rotate_batch_into (src/encode/encode.cpp:159-183) on the GPU: each row
 * scaled by 1/||row|| (0 below 1e-10) and rotated; rotated_out n x dim. */
vc_status vc_rotate(const float* vectors, size_t n, size_t dim, float* rotated_out);

/* fit_calibration() on the GPU (src/encode/encode.cpp:327-335). */
vc_status vc_fit_calibration(const float* sample, size_t n, size_t dim, size_t bits,
                             float* shift_out, float* scale_out);

/* score_prepared with an id map over PERM0 blocked codes (the multi-bucket
 * plumbing, src/search/search.cpp:1072-1154): queries (host) scored against n
 * vectors given as PERM0 blocks + alpha; ids reported through id_map (may be
 * NULL = identity). Output nq x min(k, n). */
vc_status vc_search_perm0(const float* queries, size_t nq, size_t dim, size_t bits, size_t k,
                          const uint8_t* perm0, const float* alpha, size_t n,
                          const uint64_t* id_map, const float* shift, const float* scale,
                          float* scores, uint64_t* ids, size_t* k_out);

/* Import of codes in the reference CUDA index's layout (see gpu32 above) as
 * PERM0 blocks (ceil(n/192) x G x 192 bytes), the input of vc_search_perm0 /
 * vc_score_perm0_into: the same bytes the reference's repack_perm0 makes from
 * those codes' flat rows (src/pack/pack.cpp:388-401). Bytes of groups past G
 * in the last word of a row are ignored. */
vc_status vc_gpu32_to_perm0(const uint32_t* codes, size_t n, size_t dim, size_t bits, uint8_t* perm0_out);

/* score_prepared with every argument the reference's AVX2 path honours
 * (src/search/search.cpp:1072-1154, include/vectorcache/search/search.hpp:
 * 156-165): as vc_search_perm0, over the queries listed in query_indices
 * (n_indices of them, repeats allowed; NULL = all nq; an index >= nq is
 * VC_ERR_OUT_OF_RANGE), and, when both residual_norms (n) and residual_scales
 * (nq) are given, each kept score s of vector v becomes
 * fmaf(residual_scales[query], residual_norms[v], s) and the row is ranked
 * again by (score desc, v asc) before ids go through id_map (the landmark
 * path's correction: the top-k membership is decided on the LUT scores).
 * Output nq_out x min(k, n), nq_out = n_indices or nq; *k_out = 0 when n or
 * nq_out is 0. scale_resid, hit bitmaps, buffered collectors and seeds are
 * arguments of the reference's VNNI backends only and have no counterpart. */
vc_status vc_score_perm0(const float* queries, size_t nq, size_t dim, size_t bits, size_t k,
                         const uint8_t* perm0, const float* alpha, size_t n, const uint64_t* id_map,
                         const float* shift, const float* scale, const size_t* query_indices,
                         size_t n_indices, const float* residual_norms, const float* residual_scales,
                         float* scores, uint64_t* ids, size_t* k_out);

/* score_prepared_into (src/search/search.cpp:1156-1267, include/vectorcache/
 * search/search.hpp:163-177) over one bucket given as PERM0 blocks + alpha:
 * the nq queries (host, nq x dim) are the prepared batch; only the queries in
 * query_indices (non-empty, each < nq, no repeats) are scored, each into its
 * persistent top-k heap: heap_s / heap_i [nq][k] and heap_sz / heap_min /
 * heap_mi [nq], the reference's arrays, updated in place exactly as the
 * reference's AVX2 PERM0 path updates them (strict admission above heap_min,
 * the k < 128 rescan heap and the k >= 128 binary heap; ids = id_map[v], or v
 * when id_map is NULL; score_offsets[qi], when not NULL, added after
 * alpha x LUT). n == 0 or k == 0 changes nothing. Not supported: the
 * buffered collector (heap_mi == SIZE_MAX) and k > 4096 (VC_ERR_INVALID_ARGUMENT).
 * Empty query_indices -> VC_ERR_INVALID_ARGUMENT, an index >= nq ->
 * VC_ERR_OUT_OF_RANGE, as the reference throws. */
vc_status vc_score_perm0_into(const float* queries, size_t nq, size_t dim, size_t bits, size_t k,
                              const uint8_t* perm0, const float* alpha, size_t n, const uint64_t* id_map,
                              const float* shift, const float* scale, const size_t* query_indices,
                              size_t n_indices, float* heap_s, uint64_t* heap_i, size_t* heap_sz,
                              float* heap_min, size_t* heap_mi, const float* score_offsets);

/* ---- prepared queries (PreparedQueries) ------------------------------------
 * prepare_queries (src/search/search.cpp:762-860) run once on the GPU and kept
 * in device memory (LUT rows, LUT scales and biases, bias_corr), so that the
 * score calls of a multi-bucket search reuse them as the reference reuses its
 * PreparedQueries (include/vectorcache/search/search.hpp:95-112, 143-177).
 * vc_prepare_queries takes host query rows (nq x dim) and TQ+ shift / scale
 * (both NULL: none); vc_prepare_queries_device takes device rows and device
 * shift / scale and is ordered on `stream` (NULL: the null stream); later
 * calls on the handle wait for it. nq may be 0. */
typedef struct vc_prepared vc_prepared;
vc_status vc_prepare_queries(const float* queries, size_t nq, size_t dim, size_t bits, const float* shift,
                             const float* scale, vc_prepared** out);
vc_status vc_prepare_queries_device(const float* queries_dev, size_t nq, size_t dim, size_t bits,
                                    const float* shift_dev, const float* scale_dev, void* stream, vc_prepared** out);
vc_status vc_prepared_destroy(vc_prepared* prepared);
vc_status vc_prepared_shape(const vc_prepared* prepared, size_t* nq, size_t* dim, size_t* bits,
                            size_t* n_byte_groups);
/* The state as host arrays (any may be NULL): luts nq x G*32 u8, lut_scale,
 * lut_bias (without bias_corr), bias_corr; G = n_byte_groups. */
vc_status vc_prepared_export(const vc_prepared* prepared, uint8_t* luts, float* lut_scale, float* lut_bias,
                             float* bias_corr);
/* prepared_query_at (src/search/search.cpp:1045-1068): a new handle holding
 * query qi alone; qi >= nq -> VC_ERR_OUT_OF_RANGE. */
vc_status vc_prepared_query_at(const vc_prepared* prepared, size_t qi, vc_prepared** out);
/* score_prepared / score_prepared_into over one range of PERM0 blocks + alpha
 * with the handle's queries: the arguments, results and errors of
 * vc_score_perm0 / vc_score_perm0_into, without rebuilding the query LUTs. */
vc_status vc_score_prepared(const vc_prepared* prepared, size_t k, const uint8_t* perm0, const float* alpha, size_t n,
                            const uint64_t* id_map, const size_t* query_indices, size_t n_indices,
                            const float* residual_norms, const float* residual_scales, float* scores, uint64_t* ids,
                            size_t* k_out);
vc_status vc_score_prepared_into(const vc_prepared* prepared, size_t k, const uint8_t* perm0, const float* alpha,
                                 size_t n, const uint64_t* id_map, const size_t* query_indices, size_t n_indices,
                                 float* heap_s, uint64_t* heap_i, size_t* heap_sz, float* heap_min, size_t* heap_mi,
                                 const float* score_offsets);

/* Repack free functions over host buffers, on the GPU (src/pack/pack.cpp:
 * 98-131, 157-186, 352-401): bit-plane codes (n x bits*dim/8) -> flat rows
 * (n x G, codes_flat_from_packed), PERM0 blocks (ceil(n/192) x G x 192,
 * repack_perm0) or sequential blocks (same size, repack_seq). */
vc_status vc_codes_flat_from_packed(const uint8_t* packed, size_t n, size_t dim, size_t bits, uint8_t* flat_out);
vc_status vc_repack_perm0(const uint8_t* packed, size_t n, size_t dim, size_t bits, uint8_t* perm0_out);
vc_status vc_repack_seq(const uint8_t* packed, size_t n, size_t dim, size_t bits, uint8_t* seq_out);

/* heaps_to_search_results (src/search/search.cpp:1269-1291): each of the nq
 * heaps' heap_sz entries by (score desc, id asc) into scores / ids [nq][k],
 * the rest of each row (0, 0). */
vc_status vc_heaps_to_results(size_t nq, size_t k, const float* heap_s, const uint64_t* heap_i,
                              const size_t* heap_sz, float* scores, uint64_t* ids);

/* float_rescore_topk (src/search/float_rescore.cpp:80-164): each of the nq
 * result rows (scores / ids, nq x k) re-scored in place with the float inner
 * product of its query (nq x dim) and the corpus row of each id (n_corpus x
 * dim; -inf for an id >= n_corpus), summed in the reference's AVX2 order, then
 * its first min(k, counts[q]) entries (k when counts is NULL) ordered by
 * (score desc, id asc); entries past that are left as they are. k <= 4096. */
vc_status vc_float_rescore(size_t nq, size_t k, float* scores, uint64_t* ids, const size_t* counts,
                           const float* queries, const float* corpus, size_t n_corpus, size_t dim);
/* Same on device buffers (counts_dev: u64 or NULL), ordered on `stream`
 * (NULL: the default stream); the call returns once the work is enqueued. */
vc_status vc_float_rescore_device(size_t nq, size_t k, float* scores_dev, uint64_t* ids_dev,
                                  const uint64_t* counts_dev, const float* queries_dev, const float* corpus_dev,
                                  size_t n_corpus, size_t dim, void* stream);

/* merge_search_results (src/search/search.cpp:1551-1600) on the GPU:
 * dst (nq x k_dst) <- top k_dst of dst U src by (score desc, id asc), a
 * duplicated id keeping its higher score, short rows padded with (0, 0). */
vc_status vc_merge_results(size_t nq, size_t k_dst, float* dst_scores, uint64_t* dst_ids,
                           size_t k_src, const float* src_scores, const uint64_t* src_ids);

/* ---- tuning / introspection ------------------------------------------------
 * Query-tile size T the scan kernel uses for (dim, bits, nq) — 0 = automatic.
 * vc_set_scan_tile caps T at 1, 2, 4 or 8 for later searches (0 restores the
 * automatic choice); a k too large for that tile's LDS budget takes a smaller one. */
vc_status vc_set_scan_tile(size_t tile);
size_t vc_scan_tile_for(size_t dim, size_t bits, size_t nq);

/* Instrumentation (no reference counterpart): when profiling is on, every
 * search brackets its scan kernel (the bucketed index: its bucket scan, tile
 * 1, ranges 0) with HIP events on the launch stream.
 * vc_last_scan_stats waits for the last such kernel and reports its duration
 * in ms, its query tile and the number of vector ranges (work split) it used. */
vc_status vc_set_profiling(int on);
/* Test hook (no reference counterpart): flags bit 0 makes every later encode
 * take the exact sequential alpha path instead of the certified fast one;
 * bit 2 makes indexes created later encode with the 8-rows-per-wavefront tile
 * kernel instead of the streaming kernels; bit 6 makes them encode with the
 * two-pass streaming kernels (a norm pass, then a row pass) instead of the
 * one-pass kernel of the common shapes;
 * bit 1 makes searches use the nibble-table scan kernel (two 16-entry lookups
 * per code byte) instead of the byte-pair table kernel; bit 3 makes the
 * byte-pair kernel take the other of its two vectors-per-thread variants
 * (1 <-> 2) from the one the planner picks; bit 4 makes query preparation
 * of indexes created later always take the reference's sequential bias_corr
 * chain instead of the certified parallel sum; bit 5 makes landmark searches
 * rotate their queries in a separate kernel instead of inside the query
 * projection kernel; bit 7 / bit 8 make bucketed searches take the
 * bucket-major tile scan (tiles of up to 8 queries of one bucket; k <= 64) /
 * the query-major scan whatever the batch size; bits 16-31, when not
 * zero, set the number of vector ranges each query tile is split into (at
 * most 512) in place of the automatic split. */
vc_status vc_set_debug(unsigned flags);
vc_status vc_last_scan_stats(double* ms, size_t* tile, size_t* n_ranges);

/* ---- on-disk formats of the recall runs (host code; no device) -------------
 * NPY float32 (rows, dim) matrices through mmap (NpyReader, include/
 * vectorcache/datasets/npy.hpp:33-65, src/datasets/npy.cpp): vc_npy_open maps
 * the file and checks it as the reference does (version 1.0 or 2.0, descr
 * '<f4', two dimensions, data complete); vc_npy_data is the zero-copy
 * row-major view, valid until vc_npy_close; vc_npy_read_rows copies rows
 * [first, first + n) clipped to the row count (rows copied in *n_read).
 * File and format errors return VC_ERR_STATE with the reference's message
 * (its vectorcache::Error, a std::runtime_error). */
typedef struct vc_npy vc_npy;
vc_status vc_npy_open(const char* path, vc_npy** out, size_t* count, size_t* dim);
const float* vc_npy_data(const vc_npy* npy);
vc_status vc_npy_read_rows(const vc_npy* npy, size_t first, size_t n, float* out, size_t* n_read);
vc_status vc_npy_close(vc_npy* npy);
/* write_npy_f32_matrix (src/datasets/npy.cpp:336-366): the same bytes
 * (version 1.0; 10 + the header dict padded with blanks to a multiple of 16,
 * then a newline); parent directories are created. */
vc_status vc_npy_write_f32(const char* path, size_t rows, size_t cols, const float* data);

/* VCEX exact top-k cache of query-bench (tools/query-bench.cpp:146-340):
 * "VCEX", u32 version 1, u64 query_n, k, index_n, seed, three u32-length
 * strings (source label, split, query split), then query_n x k u64 ids.
 * vc_vcex_save writes path.tmp and renames it (VC_ERR_STATE when any step
 * fails); vc_vcex_load sets *loaded = 1 and fills ids (query_n x k) only when
 * the file exists, is version 1 and was made for exactly this key, else 0.
 * vc_vcex_source_label (cache_source_label: "npy-<stem>-<FNV-1a 64 of the
 * normalized path, hex>", or `dataset` for an empty path) and
 * vc_vcex_cache_path (exact_topk_cache_path: ".cache/exact_topk/<label>_
 * <split>_n<N>_qs-<query split>_ql<nq>_k<k>_seed<seed>.bin", characters
 * other than [A-Za-z0-9_-] in the strings replaced by '_') report the length
 * in *len and write the NUL-terminated string when cap > *len. */
typedef struct {
  const char* source_label;
  const char* split;
  const char* query_split;
  uint64_t index_n;
  uint64_t query_n;
  uint64_t k;
  uint64_t seed;
} vc_vcex_key;
vc_status vc_vcex_save(const char* path, const vc_vcex_key* key, const uint64_t* ids);
vc_status vc_vcex_load(const char* path, const vc_vcex_key* key, uint64_t* ids, int* loaded);
vc_status vc_vcex_source_label(const char* npy_path, const char* dataset, char* out, size_t cap, size_t* len);
vc_status vc_vcex_cache_path(const vc_vcex_key* key, char* out, size_t cap, size_t* len);

/* ---- many collections in one launch -----------------------------------------
 * No counterpart in the reference (one index per search call,
 * src/index.cpp:267-289). A set holds prepared flat indexes of any dim, bits
 * and size (each at most once), with a query batch size nq[j] and k[j] per
 * collection, fixed at create. Its chain: per collection the query
 * preparation and LUTs (prepare_queries), then ONE persistent scan launch over
 * work items (collection, query tile, vector range) of every collection, then
 * per collection the exact merge into its top-k rows; every collection's rows
 * are bit-identical to its own vc_index_search_device of the same queries.
 * Queries are read from, and results written to, device rows the set owns
 * (vc_multi_rows): nq[j] x dim queries; nq[j] x k_eff scores / ids with
 * k_eff = min(k[j], size()). While a set exists its indexes refuse to add,
 * prepare, calibrate, be served or be destroyed. The set's work on a stream
 * must finish before vc_multi_destroy (which synchronizes the device). */
typedef struct vc_multi vc_multi;
typedef struct vc_multi_plan_info {
  size_t tile;              /* largest query tile: a collection takes the smallest tile (1, 2, 4, 8) that
                               holds its batch, at most this, halved while its top-k state does not fit */
  size_t vec_per_thread;    /* vectors per thread per scan step (tiles >= 2) */
  float item_scale;         /* work-item size relative to the default (about 4 items per workgroup slot) */
  size_t workgroups_per_cu; /* 2: every collection's LDS fits half a CU; 1: one workgroup per CU */
  size_t items, grid;    /* work items; persistent workgroups */
  size_t lds_bytes;      /* LDS per workgroup */
  double measured_ms;    /* the chain's time when the plan was tuned (0: not measured) */
} vc_multi_plan_info;

vc_status vc_multi_create(vc_index* const* indexes, size_t n_sets, const size_t* nq, const size_t* k,
                          vc_multi** out);
vc_status vc_multi_destroy(vc_multi* set);
vc_status vc_multi_rows(vc_multi* set, size_t j, float** queries_dev, float** scores_dev, uint64_t** ids_dev,
                        size_t* k_eff);
/* Stream-ordered copies into collection j's query rows (nq[j] x dim from
 * device memory) and out of its result rows (nq[j] x k_eff each). */
vc_status vc_multi_set_queries(vc_multi* set, size_t j, const float* queries_dev, void* stream);
vc_status vc_multi_get_results(vc_multi* set, size_t j, float* scores_dev, uint64_t* ids_dev, void* stream);
/* Enqueue the chain on `stream` (NULL: the null stream). */
vc_status vc_multi_search(vc_multi* set, void* stream);
/* Capture the chain on `stream` (not NULL) into a HIP graph, replacing a
 * previous one; vc_multi_replay launches it on any stream. The graph holds no
 * memset nodes. A new plan (set_plan, tune) drops the graph. */
vc_status vc_multi_capture(vc_multi* set, void* stream);
vc_status vc_multi_replay(vc_multi* set, void* stream);
/* Plan: tile 1/2/4/8, vec_per_thread 1/2, item_scale > 0, workgroups per CU
 * 1/2. The set starts with the measured plan of its signature when one is
 * loaded (vc_tune_load), else tile 8, 2 vectors per thread, item_scale 1, 2
 * workgroups per CU. */
vc_status vc_multi_set_plan(vc_multi* set, size_t tile, size_t vec_per_thread, float item_scale,
                            size_t workgroups_per_cu);
vc_status vc_multi_get_plan(vc_multi* set, vc_multi_plan_info* out);
/* Collection j's query tile and vector ranges (lists per query) under the plan. */
vc_status vc_multi_collection_plan(vc_multi* set, size_t j, size_t* tile, size_t* lists);
/* Times the chain for every plan (2 or 1 workgroups per CU, largest tile 8
 * or 4, 2 or 1 vectors per thread, item scales 0.5/1/2; median of reps runs
 * each) on the set's current queries, keeps the fastest and records it under
 * the set's signature. */
vc_status vc_multi_tune(vc_multi* set, size_t reps, vc_multi_plan_info* best);

/* ---- scan-plan autotuner ------------------------------------------------------
 * No counterpart in the reference. The single-index scan's plan (query tile,
 * vectors per thread, vector ranges) comes from a fixed rule unless a measured
 * entry exists for its (groups, log2 n, log2 nq, log2 k) bucket. vc_tune_scan
 * measures every candidate (tiles 1/2/4/8, 1 or 2 vectors per thread, ranges
 * x0.5/1/2) against the rule on pseudo-random codes of the shape and records
 * the fastest; vc_tune_save / vc_tune_load write / read all measured entries
 * (scan plans and multi-collection plans) as a text file, one per line. */
typedef struct vc_scan_choice {
  size_t tile, vec_per_thread;
  float range_scale;
  double ms;        /* the chosen plan's median search time */
  double rule_ms;   /* the fixed rule's */
  size_t rule_tile;
} vc_scan_choice;
vc_status vc_tune_scan(size_t dim, size_t bits, size_t n, size_t nq, size_t k, size_t reps, vc_scan_choice* out);
vc_status vc_tune_save(const char* path);
/* Adds the file's entries to the table; VC_ERR_STATE for a missing file or a
 * malformed line (then nothing is added). */
vc_status vc_tune_load(const char* path, size_t* n_entries);
vc_status vc_tune_clear(void);
size_t vc_tune_count(void);

/* ---- request server: continuous batching over one flat index --------------
 * No counterpart in the reference (its GPU search mutates shared device
 * buffers and is not safe to call concurrently, src/gpu/kernels.cu:602-627).
 * For serving loops: many callers submit search requests (host query rows,
 * any k <= max_k) and add batches (host rows) from any thread; each submit
 * copies its rows and returns a ticket at once.
 *
 * Batching. `slots` workers, each with its own HIP stream, take waiting
 * search requests into one batch whenever they are free (at most
 * max_batch_queries queries), so requests that arrive while others are on
 * the device join the next batch a slot starts. A batch holds the requests
 * of the oldest waiting request's k class (k in (2^(c-1), 2^c]), in arrival
 * order, and runs at the largest k it holds; a request keeps the first
 * min(k, n) entries of its rows, which are its own top-k (rows are sorted by
 * score, ties by lower id, exactly as search() orders them,
 * src/index.cpp:267-289).
 * Admission. A batch holds device memory for its query rows, LUTs, per-range
 * lists and results, carved from a pool of hbm_budget_bytes allocated at
 * create; it starts only when that memory is free, taking fewer requests if
 * need be, else it waits for a running batch to finish. A request that
 * cannot fit even alone fails with VC_ERR_OOM.
 * Adds. Add batches are encoded on a stream of their own, add_chunk rows per
 * launch, into the index's capacity, which is reserved at create for
 * reserve_vectors vectors (growing it later waits for the running batches to
 * finish). An add beyond max_vectors (when not 0) fails with VC_ERR_OOM.
 * A search batch scans the vectors whose encode had finished when it started
 * (vc_request_info.n_snapshot); rows being encoded are never read.
 *
 * The index must be prepared; while the server exists the index's own calls
 * that add, prepare, calibrate, search, export or destroy fail with
 * VC_ERR_STATE. vc_server_destroy finishes every submitted request first and
 * must not run concurrently with other calls on the server. */
typedef struct vc_server vc_server;
typedef struct vc_server_params {
  size_t slots;             /* batches in flight at once, one HIP stream each (default 4, <= 64) */
  size_t max_batch_queries; /* queries per batch and per request (default 512) */
  size_t max_k;             /* largest k of a request (default 100, <= 2048) */
  size_t hbm_budget_bytes;  /* device pool for the batches in flight (default 1 GiB) */
  size_t reserve_vectors;   /* index capacity reserved at create (0: size() + 1,000,000) */
  size_t max_vectors;       /* adds past this size fail (0: no limit but device memory) */
  size_t add_chunk;         /* rows encoded per add launch (default 65,536) */
} vc_server_params;

typedef struct vc_request_info {
  double t_submit, t_start, t_done; /* seconds on CLOCK_MONOTONIC */
  size_t n_snapshot;    /* search: vectors the batch scanned; add: size() after it */
  size_t k_eff;         /* search: min(k, n_snapshot); add: rows added */
  size_t batch_queries; /* queries in the request's batch */
  uint64_t batch;       /* batch sequence number */
  int slot;             /* worker slot that ran it (-1: add) */
} vc_request_info;

typedef struct vc_server_stats {
  uint64_t requests, queries, batches, adds, vectors_added, admission_waits;
  size_t in_flight_peak;  /* most batches on the device at once */
  size_t pool_peak_bytes; /* most pool bytes held at once */
  size_t pool_bytes;      /* the pool's size */
  size_t size;            /* vectors visible to the next batch */
} vc_server_stats;

vc_status vc_server_params_default(vc_server_params* params);
/* Pool bytes one batch of nq queries at top-k over n vectors of `index` holds
 * (search scratch, query rows, results); 0 for nq = 0 or an invalid k. For
 * sizing hbm_budget_bytes. */
size_t vc_server_batch_bytes(const vc_index* index, size_t n, size_t nq, size_t k);
/* params may be NULL (defaults). The index's device must be current. */
vc_status vc_server_create(vc_index* index, const vc_server_params* params, vc_server** out);
vc_status vc_server_destroy(vc_server* server);
/* queries: nq x dim host rows, 1 <= nq <= max_batch_queries, 1 <= k <= max_k. */
vc_status vc_server_submit_search(vc_server* server, const float* queries, size_t nq, size_t k, uint64_t* ticket);
vc_status vc_server_submit_add(vc_server* server, const float* vectors, size_t n, uint64_t* ticket);
/* Blocks until the request is done and releases its ticket. Search results go
 * to scores / ids (nq x info->k_eff, room for nq * k); add requests take NULL
 * outputs. Returns the request's own status (its message in vc_last_error). */
vc_status vc_server_wait(vc_server* server, uint64_t ticket, float* scores, uint64_t* ids, vc_request_info* info);
/* Blocks until every submitted request is done. */
vc_status vc_server_flush(vc_server* server);
vc_status vc_server_get_stats(vc_server* server, vc_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* VECTORCACHE_HIP_H */
