// Launch interface of the bucketed (IVF) index kernels (vc_ivf.hip) and of
// the bucketed scan (vc_kernels.hip). Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace vchip {

// Scores C = A B^T (bf16 rows of K, K a multiple of 32; rows 16-byte aligned).
hipError_t launch_gemm_scores(const __bf16* A, std::size_t M, int lda, const __bf16* B, std::size_t N, int ldb, int K,
                              float* C, std::size_t ldc, hipStream_t s);
// best_col[m] = argmax_n (A B^T)[m][n], ties to the lower column.
hipError_t launch_gemm_argmax(const __bf16* A, std::size_t M, int lda, const __bf16* B, std::size_t N, int ldb, int K,
                              int* best_col, hipStream_t s);
// Top P (<= 512) columns of each score row by (score desc, column asc), sorted;
// entries past min(P, N) are (-1, -inf).
hipError_t launch_topk_rows(const float* scores, std::size_t M, int N, std::size_t ld, int P, int* out_c,
                            float* out_s, hipStream_t s);
hipError_t launch_fixed_probes(std::size_t nq, int P, int n, const int* probe_b, int* probe_n, hipStream_t s);
// pre [nq][P + 1] running vector counts of the probe lists; item_len [nq],
// n_items [nq] and item_off [nq + 1] cut each query's vectors into items.
hipError_t launch_plan_probes(std::size_t nq, int P, const int* probe_b, const int* probe_n, const int* bsize,
                              int min_item, int max_items, int* pre, int* item_len, int* n_items, int* item_off,
                              hipStream_t s);
// cent null: offsets 0 (codes of the rows themselves).
hipError_t launch_probe_offsets(const float* qhat, std::size_t nq, int D, const float* cent, int P,
                                const int* probe_b, const int* probe_n, float* off, hipStream_t s);
hipError_t launch_rows_to_bf16(const float* x, std::size_t n, int D, const float* inv, int Kp, __bf16* out,
                               hipStream_t s);
// cent null: the normalized rows themselves.
hipError_t launch_residual_rows(const float* x, int D, const float* inv, const std::uint32_t* perm,
                                const int* pos_bucket, const float* cent, std::size_t p0, std::size_t n, float* out,
                                hipStream_t s);
hipError_t launch_count_buckets(const int* assign, std::size_t n, int* counts, hipStream_t s);
hipError_t launch_fill_pos_bucket(const long* boff, const int* bsize, int L, int* pos_bucket, hipStream_t s);
hipError_t launch_centroid_sums(const float* x, int D, const float* inv, const std::uint32_t* rows, const long* boff,
                                const int* bsize, int L, float* cent, hipStream_t s);
hipError_t launch_split_buckets(const float* x, int D, const float* inv, const std::uint32_t* rows, const long* boff,
                                const int* bsize, const int* list, int n_list, std::uint8_t* side, float* cent_a,
                                float* cent_b, hipStream_t s);
hipError_t launch_apply_split(const std::uint32_t* rows, const long* boff, const int* bsize, const int* list,
                              int n_list, const int* new_id, const std::uint8_t* side, int* assign, hipStream_t s);

hipError_t launch_gather_rows_bf16(const __bf16* src, int Kp, const std::uint32_t* rows, std::size_t n, __bf16* dst,
                                   hipStream_t s);
hipError_t launch_normalized_rows(const float* x, int D, const float* inv, const std::uint32_t* rows, std::size_t n,
                                  float* out, hipStream_t s);
hipError_t launch_scatter_children(float* cent, int D, const int* list, int n_list, int base, const float* cent_a,
                                   const float* cent_b, hipStream_t s);
hipError_t launch_widen_ids(const std::uint32_t* rows, std::size_t n, std::uint64_t base, std::uint64_t* ids,
                            hipStream_t s);

hipError_t launch_gather_rows_f32(const float* src, int D, const int* rows, std::size_t n, float* dst, hipStream_t s);
hipError_t launch_remap(int* v, std::size_t n, const int* map, hipStream_t s);

// ---- vc_kernels.hip ------------------------------------------------------------
// Row norms as the encode computes them (8 fused chains, tree) -> inv[r] =
// norm > 1e-10 ? 1 / norm : 0; when qhat is set, qhat = x * inv (n x D).
hipError_t launch_row_inv_norms(const float* x, std::size_t n, int D, float* inv, float* qhat, hipStream_t s);

struct BucketScanParams {
  const std::uint8_t* tiled;
  const float* alpha;
  int G16;
  const std::uint8_t* src;  // filled by launch_bucket_scan
  const float* lut_scale;
  const float* bias_total;
  long nq;
  int P;
  const int* probe_n;
  const int* probe_b;
  const float* probe_off;
  const int* probe_pre;
  const long* boff;
  const int* item_len;
  const int* item_off;
  int k, cap;
  float* part_s;
  std::uint32_t* part_i;
  int* part_cnt;
};
// Bytes of the per-query pair-table sources the bucketed scan stages.
std::size_t bucket_src_bytes(int groups16, std::size_t nq);
// Vectors of one work item (before lengthening) and the item cap per query.
constexpr int kBucketItem = 4096;
constexpr int kBucketMaxItems = 64;
// The bucketed scan over luts (nq x groups * 32); p.src is set to src.
hipError_t launch_bucket_scan(BucketScanParams p, const std::uint8_t* luts, int groups, std::uint8_t* src,
                              hipStream_t s);
hipError_t launch_merge_items(const float* part_s, const std::uint32_t* part_i, const int* part_cnt,
                              const int* item_off, std::size_t nq, int k, int k_out, const std::uint64_t* id_map,
                              float* out_s, std::uint64_t* out_i, std::uint64_t* counts, hipStream_t s);

// Bucket-major scan of a large query batch: the (query, probe) pairs grouped
// by bucket, cut into tiles of up to kTileQueries queries of one bucket; each
// tile reads its bucket's codes once for all its queries. Every pair (q, j)
// gets its own exact top-k list in slot q * P + j.
constexpr int kTileQueries = 8;
constexpr int kTileMaxK = 64;  // list length the tile scan keeps
struct BucketTileParams {
  const std::uint8_t* tiled;
  const float* alpha;
  int G, G16;
  const std::uint8_t* luts;  // nq x G x 32
  const float* lut_scale;
  const float* bias_total;
  long nq;
  int P;
  const int* probe_n;
  const int* probe_b;
  const float* probe_off;
  const long* boff;
  const int* bsize;
  int L;
  int k, cap;
  // planning buffers, filled by launch_bucket_tiles
  int* bcnt;    // [L] pairs per bucket
  int* bstart;  // [L + 1] first pair of each bucket
  int* tstart;  // [L + 1] first tile of each bucket
  int* n_tiles; // [1]
  int* prank;   // [nq * P] rank of pair (q, j) within its bucket
  int* pair_q;  // [nq * P] pairs grouped by bucket
  int* pair_j;
  int* tile_b;  // [nq * P] tiles: bucket, first pair, pair count
  int* tile_p0;
  int* tile_n;
  int* gbound;  // [nq] per-query score bound (float_order image)
  float* part_s;  // [nq * P * k]
  std::uint32_t* part_i;
  int* part_cnt;  // [nq * P]
};
// Plan (pairs by bucket, tiles), scan, and merge each query's probe_n[q]
// slots into out_s / out_i (ids through id_map; counts as launch_merge_items).
hipError_t launch_bucket_tiles(BucketTileParams p, int k_out, const std::uint64_t* id_map, float* out_s,
                               std::uint64_t* out_i, std::uint64_t* counts, hipStream_t s, hipEvent_t scan_start,
                               hipEvent_t scan_stop);

}  // namespace vchip
