// Internal launch interface between the C ABI (vc_capi.cpp) and the gfx950
// kernels (vc_kernels.hip). Not installed; not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <cstdint>

namespace vchip {

// Device-side constant tables of one (dim, bits) configuration.
struct CodeTables {
  float boundaries[15];
  float centroids[16];
  int dim, bits, block, groups, groups16;  // groups16 = ceil(groups / 16)
  float inv_sqrt_block;
  const std::uint32_t* inv_perm;  // device, dim
  const float* signs;             // device, dim (indexed by destination)
  const float* shift;             // device, dim; null = identity calibration
  const float* scale;             // device, dim
  const float* inv_scale;         // device, dim
  const std::uint32_t* perm_sign;  // device, dim: inv_perm[j] | (signs[inv_perm[j]] < 0) << 31
  const std::uint32_t* perm_gather;  // device, dim: perm[i] | (signs[i] < 0) << 31
  const void* cells;              // device, 256 x {u32 base, f32 bound, f32 c_lo, f32 c_hi}
  float cell_inv_w, cell_off;      // cell index of v: fmaf(v, cell_inv_w, cell_off), clamped to [0, 255]
  int use_cells;
  int debug_flags;                // bit 0: take the sequential alpha path for every vector
  // Landmark index modes (zero for the plain flat path):
  int prerotated;  // rows are already in the rotated basis: encode scales them by
                   // 1 / norm without rotating; prepare_queries builds LUTs from the
                   // rows as given (no norm, no rotation)
  int norm_seq;    // encode: the 8 norm chains are summed in order and rows with
                   // norm < 1e-10 become zero (the landmark l2-normalize)
};

// Internal scan layout ("tiled"): for vector v and byte group g the flat byte
// (hi nibble = first code of the group) lives at
//   ((v / 64) * groups16 + g / 16) * 1024 + (v % 64) * 16 + g % 16,
// i.e. 64-vector tiles, each lane's 16 consecutive groups contiguous, so one
// 16-byte load per lane covers 16 groups and a wavefront reads 1 KiB
// contiguously. Groups >= `groups` in the last 16-chunk are zero.
inline std::size_t tiled_bytes(std::size_t n_vectors_cap, int groups16) {
  return ((n_vectors_cap + 63) / 64) * std::size_t(groups16) * 1024;
}

// encode: norm -> rotation -> quantize -> alpha -> codes.
// Outputs may be null: tiled (written at vector offset `out_offset`),
// packed (bit-planes, row-major n x bits*dim/8), rotated (n x dim floats:
// the calibration path), alpha (written at out_offset).
hipError_t launch_encode(const CodeTables& t, const float* x, std::size_t n, std::size_t out_offset,
                         std::uint8_t* tiled, std::uint8_t* packed, float* rotated, float* alpha,
                         hipStream_t s);

// TQ+ per-dimension order statistics of rotated sample rows (n x dim):
// keys -> radix select -> shift/scale.
hipError_t launch_calibrate(const float* rotated, std::size_t n, int dim, std::size_t lo_idx,
                            std::size_t hi_idx, float qc_lo, float qc_span, std::uint32_t* keys_scratch,
                            float* shift, float* scale, float* inv_scale, hipStream_t s);

// Query preparation: per query rotation (+ TQ+), LUT build and quantization.
// luts: nq x groups*32 u8. bias_total = lut_bias + bias_corr.
hipError_t launch_prepare_queries(const CodeTables& t, const float* q, std::size_t nq,
                                  std::uint8_t* luts, float* lut_scale, float* lut_bias,
                                  float* bias_corr, float* bias_total, hipStream_t s);

struct ScanPlan {
  int tile;          // queries per workgroup (1, 2, 4 or 8)
  int pair;          // 1: byte-pair table kernel, 0: nibble-table kernel
  int n_ranges;      // vector ranges (workgroups per query tile)
  int range_len;     // vectors per range (multiple of 64)
  int lut_chunk16;   // 16-group chunks of LUT resident in LDS (all groups if it fits)
  int cand_cap;      // candidate buffer per query
  int vec_per_thread;  // vectors per thread per step (2 when the LUT is reloaded in chunks)
  std::size_t lds_bytes;
};

// Plan for nq queries over n vectors with top-k (k <= kMaxTopK).
inline constexpr int kMaxTopK = 2048;
// nibble = true selects the nibble-table kernel (two lookups per code byte).
ScanPlan plan_scan(int groups16, std::size_t n, std::size_t nq, int k, int forced_tile, bool nibble);
// The fixed rule alone (no lookup of measured plans).
ScanPlan plan_scan_rule(int groups16, std::size_t n, std::size_t nq, int k, int forced_tile, bool nibble);

// ---- measured plans ----------------------------------------------------------
// plan_scan's byte-pair branch first looks up a table of measured choices,
// keyed by (groups16, floor(log2 n), ceil(log2 nq), ceil(log2 k)); a hit
// replaces the fixed rule's tile, vectors per thread and range count (scaled
// from plan_ranges' choice). The table starts empty; the autotuner fills it
// and it round-trips through a text file, one entry per line:
//   scan g16=<G16> nlog=<a> nqlog=<b> kcls=<c> tile=<T> vpt=<V> rscale=<x> ms=<t>
struct TunedScan {
  int tile = 0;
  int vec_per_thread = 0;
  float range_scale = 1.f;
  double ms = 0.0;
};
struct TuneKey {
  int g16, nlog, nqlog, kcls;
  bool operator<(const TuneKey& o) const {
    return g16 != o.g16 ? g16 < o.g16 : nlog != o.nlog ? nlog < o.nlog : nqlog != o.nqlog ? nqlog < o.nqlog : kcls < o.kcls;
  }
};
TuneKey tune_key(int groups16, std::size_t n, std::size_t nq, int k);
void tune_put(const TuneKey& key, const TunedScan& v);
bool tune_get(const TuneKey& key, TunedScan* v);
void tune_clear();
std::size_t tune_count();
std::string tune_dump();                      // the table as text lines
int tune_parse(const std::string& text);      // "scan" entries read (other lines skipped), -1 if malformed
// The plan of one explicit choice (tile, vectors per thread, range scale),
// LDS permitting; tile 0 when it does not fit.
ScanPlan plan_scan_with(int groups16, std::size_t n, std::size_t nq, int k, int tile, int vec_per_thread,
                        float range_scale);

// ---- many code sets in one launch ------------------------------------------
// One code set of a multi-set scan, as the device reads it.
struct MultiColl {
  const std::uint8_t* tiled;
  const float* alpha;
  const std::uint8_t* src;  // pair-source rows of its query tiles (pair_source_bytes)
  const float* lut_scale;
  const float* bias_total;
  float* part_s;            // [nq][n_lists][k]
  std::uint32_t* part_i;
  int* part_cnt;            // [nq][n_lists]
  long n, nq;
  int G16, k, cap, n_lists;
  int tile;                 // queries per work item: 1, 2, 4 or 8
};
// One work item: query tile qt of code set `coll` over vectors [v_begin, v_end), writing list `list`.
struct MultiItem {
  int coll;
  int list;
  long qt;
  long v_begin, v_end;
};
// Bytes of the pair-source rows of nq queries for a tile-T scan.
std::size_t pair_source_bytes(int tile, int groups16, std::size_t nq);
hipError_t launch_pair_source(int tile, const std::uint8_t* luts, std::size_t nq, int groups, int groups16,
                              std::uint8_t* out, hipStream_t s);
// LDS the scan holds besides its top-k state (table + stage), and the top-k
// state of a T-query tile at top-k with a cap-candidate buffer.
std::size_t pair_fixed_lds();
std::size_t pair_topk_lds(int tile, int k, int cap);
inline constexpr int kLdsPerCu = 160 * 1024;
inline constexpr int kMaxCandCap = 512;
// Sets the kernel's dynamic-LDS limit (outside graph capture); the launch.
// vec_per_thread applies to the code sets of tile >= 2 (tile 1 runs 1).
hipError_t prepare_multi_scan(int vec_per_thread, std::size_t dyn_lds);
hipError_t launch_multi_scan(int vec_per_thread, const MultiColl* colls, const MultiItem* items, long n_items,
                             std::size_t dyn_lds, int grid, hipStream_t s);
// Deterministic pseudo-random rows (autotuning inputs): n x dim floats.
hipError_t launch_fill_rows(float* x, std::size_t n, int dim, std::uint64_t seed, hipStream_t s);
// x[0 .. n) = v
hipError_t launch_fill_value(float* x, std::size_t n, float v, hipStream_t s);

// Bytes of the per-query-tile LUT rows (interleaved / pair source) the scan needs.
std::size_t scan_lut_bytes(const ScanPlan& p, int groups16, std::size_t nq);

// Scan + per-range exact top-k. luts: nq x groups*32 (prepare_queries);
// lut_t: scratch of scan_lut_bytes. Partial lists: part_s/part_i
// [nq][n_ranges][k], part_cnt [nq][n_ranges]. phase 0 launches both steps;
// 1 only the table sources (pair sources or interleaved LUTs) into lut_t, 2
// only the scan kernel (for timing the scan alone).
hipError_t launch_scan(const ScanPlan& p, const std::uint8_t* tiled, const float* alpha,
                       std::size_t n, int groups, int groups16, const std::uint8_t* luts,
                       std::uint8_t* lut_t, const float* lut_scale, const float* bias_total,
                       std::size_t nq, int k, float* part_s, std::uint32_t* part_i, int* part_cnt,
                       hipStream_t s, int phase = 0);

// Merge per-range lists into final rows of k_out entries (padded with 0/0).
// id_map (device, may be null) maps vector positions to reported ids.
hipError_t launch_merge_ranges(const float* part_s, const std::uint32_t* part_i, const int* part_cnt,
                               std::size_t nq, int n_ranges, int k, int k_out,
                               const std::uint64_t* id_map, float* out_s, std::uint64_t* out_i,
                               hipStream_t s);

// merge_search_results on device buffers.
// score_prepared_into for one bucket: p [nq][n] = fa * alpha, then the heap
// replay into heaps [nq][k] (state arrays [nq]); k <= kMaxHeapK.
constexpr int kMaxHeapK = 4096;
hipError_t launch_bucket_scores(const std::uint8_t* tiled, const float* alpha, std::size_t n, int groups,
                                int groups16, const std::uint8_t* luts, const float* lut_scale,
                                const float* bias_total, std::size_t nq, float* p_out, hipStream_t s);
hipError_t launch_heap_replay(const float* p, std::size_t n, int k, const std::uint64_t* id_map,
                              const float* offsets, std::size_t nq, float* hs, std::uint64_t* hi,
                              std::uint64_t* hsz, float* hmin, std::uint64_t* hmi, hipStream_t s);
hipError_t launch_heaps_to_results(std::size_t nq, int k, const float* hs, const std::uint64_t* hi,
                                   const std::uint64_t* hsz, float* out_s, std::uint64_t* out_i, hipStream_t s);
// float_rescore_topk on device rows: scores / ids [nq][k], counts [nq] or null.
hipError_t launch_float_rescore(std::size_t nq, int k, float* scores, std::uint64_t* ids,
                                const std::uint64_t* counts, const float* queries, const float* corpus,
                                std::size_t n_corpus, int dim, hipStream_t s);
hipError_t launch_merge_results(std::size_t nq, int k_dst, float* dst_s, std::uint64_t* dst_i,
                                int k_src, const float* src_s, const std::uint64_t* src_i,
                                hipStream_t s);

// Format conversions.
hipError_t launch_tiled_to_flat(const std::uint8_t* tiled, std::size_t n, int groups, int groups16,
                                std::uint8_t* flat, hipStream_t s);
hipError_t launch_flat_to_tiled(const std::uint8_t* flat, std::size_t n, int groups, int groups16,
                                std::uint8_t* tiled, hipStream_t s);
// The reference CUDA index's layout: ceil(n/32) x ceil(G/4) x 32 u32 words.
inline std::size_t gpu32_words(std::size_t n, int groups) { return (n + 31) / 32 * ((groups + 3) / 4) * 32; }
hipError_t launch_flat_to_gpu32(const std::uint8_t* flat, std::size_t n, int groups, std::uint32_t* words,
                                hipStream_t s);
hipError_t launch_gpu32_to_flat(const std::uint32_t* words, std::size_t n, int groups, std::uint8_t* flat,
                                hipStream_t s);
// Residual post-correction + re-rank + id remap of rows x k results in place.
hipError_t launch_residual_post(std::size_t rows, int k, float* scores, std::uint64_t* ids, const float* rn,
                                const float* rs, std::size_t n, const std::uint64_t* id_map, hipStream_t s);
// ---- landmark index -------------------------------------------------------
// rows (n x dim, rotated unit rows in place) -> alpha * centroid rows
// (x >= boundary levels of t's codebook), vnorm2[n] = their squared norms.
hipError_t launch_lm_yhat(const CodeTables& t, float* rows, std::size_t n, float* vnorm2, hipStream_t s);
// out[r][j] = <Y_r, L_j> in the 8-lane order, j < m; zero for m <= j < stride.
hipError_t launch_lm_project(const float* Y, std::size_t n, int dim, const float* L, int m, float* out,
                             int stride, hipStream_t s);
// rn[r] = sqrt(max(0, vnorm2[r] - |rows_r[0:m)|^2)), rows with the given stride.
// Staged sign-plane search of the landmark index: sign bytes (top bit of 8
// codes) from the 4-bit tiled codes into a tiled layout of groups_s bytes;
// build_sign_lut LUTs of the query rows (S dims, S / 8 groups); the shortlist
// rescore (E positions per query in shortlist order) into sorted k rows.
hipError_t launch_lm_sign_codes(const std::uint8_t* tiled4, std::size_t n, int groups16_4, int groups_s,
                                int groups16_s, std::uint8_t* tiled_s, hipStream_t s);
hipError_t launch_sign_luts(const float* q_land, std::size_t nq, int S, std::uint8_t* luts, float* lut_scale,
                            float* bias_total, hipStream_t s);
std::size_t staged_lds_bytes(int S, int E, int k);
hipError_t launch_lm_staged(const float* q_land, std::size_t nq, int S, const float* q_perp, const float* centroids16,
                            const std::uint8_t* tiled4, int groups16_4, const float* alpha, const float* rn,
                            const std::uint64_t* cand, int E, int k, float* out_s, std::uint64_t* out_i,
                            hipStream_t s);
hipError_t launch_lm_resid(const float* rows, std::size_t n, int m, int stride, const float* vnorm2, float* rn,
                           hipStream_t s);
// Query projection onto the 4-bit basis: q_land nq x stride (zero past m) and
// q_perp[nq] from rotated query rows q_rot (nq x dim); with rot (the query
// tables) non-null, q_rot holds the raw queries and the kernel applies the
// row norm and rotation itself (only where lm_query_rotates(*rot, ...)).
// codes_t: the 4-bit basis as u32 words [dim_pad / 8][m] (word w of column j
// at w * m + j).
bool lm_query_rotates(const CodeTables& t, int dim, int dim_pad, int stride);
hipError_t launch_lm_query(const CodeTables* rot, const float* q_rot, std::size_t nq, int dim, int dim_pad,
                           const std::uint32_t* codes_t, const float* lscales, int m, float* q_land, int stride,
                           float* q_perp, hipStream_t s);

hipError_t launch_flat_to_perm0(const std::uint8_t* flat, std::size_t n, int groups,
                                std::uint8_t* perm0, hipStream_t s);
hipError_t launch_flat_to_seq(const std::uint8_t* flat, std::size_t n, int groups, std::uint8_t* seq,
                              hipStream_t s);
hipError_t launch_perm0_to_flat(const std::uint8_t* perm0, std::size_t n, int groups,
                                std::uint8_t* flat, hipStream_t s);
hipError_t launch_flat_to_packed(const std::uint8_t* flat, std::size_t n, int dim, int bits,
                                 std::uint8_t* packed, hipStream_t s);
// bit-planes -> flat bytes (codes_flat_from_packed)
hipError_t launch_packed_to_flat(const std::uint8_t* packed, std::size_t n, int dim, int bits, std::uint8_t* flat,
                                 hipStream_t s);
// dst row i = src row idx[i]; row_bytes a multiple of 4
hipError_t launch_gather_rows(const void* src, const std::uint64_t* idx, std::size_t m, std::size_t row_bytes,
                              void* dst, hipStream_t s);

}  // namespace vchip
