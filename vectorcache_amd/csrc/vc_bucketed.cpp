// C ABI of the bucketed TurboQuant index (vc_bucketed_*, include/vectorcache_hip.h):
// the README's BucketedTurboQuantIndex (README.md:56,114-120) on one GPU.
//
// Build (prepare): rows are L2-normalized; a coarse quantizer of
// ceil(n / energy_soft) lists is trained by spherical k-means on a strided
// sample and every row is assigned to its nearest list; a list above
// max_bucket_size is split by one iteration of spherical 2-means until none
// is; lists left empty are dropped. Each bucket's encode centroid is frozen at that point and its rows are
// encoded as residuals x_hat - c (residual mode) or as themselves, with the
// flat index's encode, into one bucket-ordered code set.
// Search: the query's routing scores against every bucket (matrix cores)
// pick its probe list; each probed bucket is scanned with the query's LUTs
// exactly as score_prepared_into scans a bucket (src/search/search.cpp:
// 1156-1267), with the offset <q_hat, c> added after alpha (residual mode).
// The probe count per query comes from the fixed n_probe, a per-query
// array, or the learned probe model.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "vc_common.hpp"
#include "vc_ivf.hpp"
#include "vc_probe.hpp"

using namespace vcapi;

namespace {

constexpr int kRouteMaxP = 512;
constexpr size_t kEncodeChunk = size_t(1) << 20;  // residual rows per encode pass
constexpr size_t kRouteScoreBytes = size_t(2) << 30;  // routing score matrix per query chunk

int pad32(size_t d) { return int((d + 31) / 32 * 32); }

void sort_pairs(const uint32_t* k_in, uint32_t* k_out, const uint32_t* v_in, uint32_t* v_out, size_t n, int end_bit,
                DevBuf& temp, hipStream_t s) {
  size_t tb = 0;
  hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k_in, k_out, v_in, v_out, int(n), 0, end_bit, s),
            "radix sort (size)");
  temp.reserve(tb + 256, s);
  hip_check(hipcub::DeviceRadixSort::SortPairs(temp.p, tb, k_in, k_out, v_in, v_out, int(n), 0, end_bit, s),
            "radix sort");
}

int bits_for(size_t v) {
  int b = 1;
  while ((size_t(1) << b) <= v) ++b;
  return b;
}

template <typename T>
void upload(DevBuf& b, const std::vector<T>& h, hipStream_t s) {
  b.reserve(std::max<size_t>(h.size() * sizeof(T), 16), s);
  if (!h.empty()) hip_check(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "copy");
}

template <typename T>
std::vector<T> download(const DevBuf& b, size_t n, hipStream_t s) {
  std::vector<T> h(n);
  if (n > 0) hip_check(hipMemcpyAsync(h.data(), b.p, n * sizeof(T), hipMemcpyDeviceToHost, s), "copy");
  hip_check(hipStreamSynchronize(s), "sync");
  return h;
}

}  // namespace

struct vc_bucketed {
  size_t dim = 0, bits = 0;
  int Kp = 0;
  vc_bucket_params prm{};
  vchip::CodeTables t{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  // staged rows (every row added; the build reads them all)
  DevBuf rows;
  size_t n_rows = 0;
  // calibration sample (TQ+ fitted on its residuals at the build)
  DevBuf cal_sample;
  size_t n_cal = 0;
  DevBuf calib;  // shift | scale | inv_scale
  bool calibrated = false;
  // built state
  bool prepared = false;
  size_t n = 0;  // encoded rows
  int L = 0;     // buckets
  DevBuf cent;       // [L][dim] encode centroids (f32, unit)
  DevBuf cent_bf16;  // [L][Kp] routing centroids
  DevBuf bsize;      // int [L]
  DevBuf boff;       // long [L]
  std::vector<int> h_bsize;
  DevBuf tiled, alpha, ids;  // codes by position, alpha, original ids [n]
  // probe model
  vchip::ProbeModel probe;
  bool use_probe = false;
  // scratch
  DevBuf scratch, sort_tmp;
  DevBuf io;
};

extern "C" {

void vc_bucket_params_default(vc_bucket_params* p) {
  if (p == nullptr) return;
  p->n_probe = 64;
  p->max_bucket_size = 800;
  p->energy_soft = 400;
  p->expected_n = 0;
  p->kmeans_iters = 8;
  p->kmeans_sample = 64;
  p->seed = 4001;
  p->residual = 1;
}

vc_status vc_bucketed_create(size_t dim, size_t bits, const vc_bucket_params* params, vc_bucketed** out) {
  return guarded([&] {
    if (out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "out is null");
    check_dim_bits(dim, bits);
    if (dim > 4096) fail(VC_ERR_INVALID_ARGUMENT, "the bucketed index supports dim <= 4096");
    vc_bucket_params p;
    vc_bucket_params_default(&p);
    if (params != nullptr) p = *params;
    if (p.n_probe == 0) fail(VC_ERR_INVALID_ARGUMENT, "n_probe must be positive");
    if (p.max_bucket_size < 2) fail(VC_ERR_INVALID_ARGUMENT, "max_bucket_size must be at least 2");
    if (p.energy_soft == 0 || p.energy_soft > p.max_bucket_size)
      fail(VC_ERR_INVALID_ARGUMENT, "energy_soft must be in [1, max_bucket_size]");
    if (p.kmeans_sample == 0) fail(VC_ERR_INVALID_ARGUMENT, "kmeans_sample must be positive");
    auto* ix = new vc_bucketed();
    try {
      ix->dim = dim;
      ix->bits = bits;
      ix->Kp = pad32(dim);
      ix->prm = p;
      hip_check(hipGetDevice(&ix->device), "hipGetDevice");
      hip_check(hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking), "hipStreamCreate");
      ix->t = make_tables(dim, bits);
    } catch (...) {
      if (ix->stream) (void)hipStreamDestroy(ix->stream);
      delete ix;
      throw;
    }
    *out = ix;
  });
}

vc_status vc_bucketed_destroy(vc_bucketed* ix) {
  return guarded([&] {
    if (ix == nullptr) return;
    (void)hipStreamSynchronize(ix->stream);
    for (DevBuf* b : {&ix->rows, &ix->cal_sample, &ix->calib, &ix->cent, &ix->cent_bf16, &ix->bsize, &ix->boff,
                      &ix->tiled, &ix->alpha, &ix->ids, &ix->scratch, &ix->sort_tmp, &ix->io})
      b->release();
    ix->probe.release();
    (void)hipStreamDestroy(ix->stream);
    delete ix;
  });
}

// Appends n rows; device rows are copied in order after the caller's
// producer stream (NULL: the null stream).
static void bucketed_stage(vc_bucketed* ix, DevBuf& dst, size_t& count, const float* v, size_t n, hipMemcpyKind kind,
                           hipStream_t producer) {
  if (n == 0) return;
  const size_t row = ix->dim * 4;
  dst.reserve((count + n) * row, ix->stream, /*keep=*/true);
  if (kind == hipMemcpyDeviceToDevice) {
    hipEvent_t ev;
    hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventRecord(ev, producer), "hipEventRecord");
    hip_check(hipStreamWaitEvent(ix->stream, ev, 0), "hipStreamWaitEvent");
    hip_check(hipEventDestroy(ev), "hipEventDestroy");
  }
  hip_check(hipMemcpyAsync(dst.as<char>(count * row), v, n * row, kind, ix->stream), "copy");
  if (kind == hipMemcpyHostToDevice) hip_check(hipStreamSynchronize(ix->stream), "sync");
  count += n;
}

vc_status vc_bucketed_add(vc_bucketed* ix, const float* vectors, size_t n) {
  return guarded([&] {
    if (ix == nullptr || (vectors == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    bucketed_stage(ix, ix->rows, ix->n_rows, vectors, n, hipMemcpyHostToDevice, nullptr);
    if (n > 0) ix->prepared = false;
  });
}

vc_status vc_bucketed_add_device(vc_bucketed* ix, const float* vectors_dev, size_t n, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (vectors_dev == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    bucketed_stage(ix, ix->rows, ix->n_rows, vectors_dev, n, hipMemcpyDeviceToDevice,
                   static_cast<hipStream_t>(stream));
    if (n > 0) ix->prepared = false;
  });
}

vc_status vc_bucketed_calibrate(vc_bucketed* ix, const float* sample, size_t n) {
  return guarded([&] {
    if (ix == nullptr || (sample == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 2) fail(VC_ERR_INVALID_ARGUMENT, "calibration needs at least 2 rows");
    std::lock_guard lock(ix->mu);
    ix->n_cal = 0;
    bucketed_stage(ix, ix->cal_sample, ix->n_cal, sample, n, hipMemcpyHostToDevice, nullptr);
    ix->prepared = false;
  });
}

vc_status vc_bucketed_calibrate_device(vc_bucketed* ix, const float* sample_dev, size_t n, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (sample_dev == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 2) fail(VC_ERR_INVALID_ARGUMENT, "calibration needs at least 2 rows");
    std::lock_guard lock(ix->mu);
    ix->n_cal = 0;
    bucketed_stage(ix, ix->cal_sample, ix->n_cal, sample_dev, n, hipMemcpyDeviceToDevice,
                   static_cast<hipStream_t>(stream));
    ix->prepared = false;
  });
}

}  // extern "C"

// ---- build ---------------------------------------------------------------------
namespace {

// Sorts row indices by their bucket (stable: rows ascending inside a bucket),
// and returns bucket sizes (host) with device sizes / offsets filled.
std::vector<int> group_by_bucket(vc_bucketed* ix, const int* assign, size_t n, int L, uint32_t* rows_sorted,
                                 DevBuf& bsize, DevBuf& boff, DevBuf& tmp) {
  hipStream_t s = ix->stream;
  Carve c;
  const size_t o_keys = c.take(n * 4), o_iota = c.take(n * 4), o_ko = c.take(n * 4);
  tmp.reserve(c.off, s);
  std::vector<uint32_t> iota(n);
  std::iota(iota.begin(), iota.end(), 0u);
  hip_check(hipMemcpyAsync(tmp.as<uint32_t>(o_iota), iota.data(), n * 4, hipMemcpyHostToDevice, s), "copy");
  hip_check(hipMemcpyAsync(tmp.as<uint32_t>(o_keys), assign, n * 4, hipMemcpyDeviceToDevice, s), "copy");
  sort_pairs(tmp.as<uint32_t>(o_keys), tmp.as<uint32_t>(o_ko), tmp.as<uint32_t>(o_iota), rows_sorted, n,
             bits_for(size_t(L)), ix->sort_tmp, s);
  bsize.reserve(size_t(L) * 4, s);
  hip_check(hipMemsetAsync(bsize.p, 0, size_t(L) * 4, s), "memset");
  hip_check(vchip::launch_count_buckets(assign, n, bsize.as<int>(), s), "count buckets");
  std::vector<int> h = download<int>(bsize, L, s);
  std::vector<long> off(L);
  long run = 0;
  for (int b = 0; b < L; ++b) {
    off[b] = run;
    run += h[b];
  }
  upload(boff, off, s);
  return h;
}

void build_buckets(vc_bucketed* ix) {
  hipStream_t s = ix->stream;
  const size_t N = ix->n_rows, D = ix->dim;
  const int Kp = ix->Kp;
  const float* x = ix->rows.as<float>();
  // normalized rows: inverse norms and bf16 copies
  DevBuf inv, xb, assign, perm, tmp;
  inv.reserve(N * 4, s);
  hip_check(vchip::launch_row_inv_norms(x, N, int(D), inv.as<float>(), nullptr, s), "row norms");
  xb.reserve(N * size_t(Kp) * 2, s);
  hip_check(vchip::launch_rows_to_bf16(x, N, int(D), inv.as<float>(), Kp, xb.as<__bf16>(), s), "bf16 rows");
  // coarse quantizer: spherical k-means on a strided sample
  int L = int(std::min<size_t>(N, (N + ix->prm.energy_soft - 1) / ix->prm.energy_soft));
  L = std::max(L, 1);
  const size_t S = std::min<size_t>(N, size_t(L) * ix->prm.kmeans_sample);
  std::vector<uint32_t> srows(S);
  for (size_t i = 0; i < S; ++i) srows[i] = uint32_t((unsigned __int128)i * N / S);
  // initial centroids: sample rows spread over the sample, offset by the seed
  std::vector<uint32_t> crows(L);
  const size_t shift = size_t(ix->prm.seed % std::max<size_t>(1, S / size_t(L)));
  for (int j = 0; j < L; ++j) crows[j] = srows[std::min(S - 1, (size_t)j * S / size_t(L) + shift)];
  DevBuf d_srows, d_crows, xs;
  upload(d_srows, srows, s);
  upload(d_crows, crows, s);
  ix->cent.reserve(size_t(L) * D * 4, s);
  hip_check(vchip::launch_normalized_rows(x, int(D), inv.as<float>(), d_crows.as<uint32_t>(), L, ix->cent.as<float>(),
                                          s),
            "initial centroids");
  xs.reserve(S * size_t(Kp) * 2, s);
  hip_check(vchip::launch_gather_rows_bf16(xb.as<__bf16>(), Kp, d_srows.as<uint32_t>(), S, xs.as<__bf16>(), s),
            "sample rows");
  assign.reserve(std::max(N, S) * 4, s);
  perm.reserve(std::max(N, S) * 4, s);
  DevBuf bsize, boff, srows_sorted;
  ix->cent_bf16.reserve(size_t(L) * Kp * 2, s);
  for (size_t it = 0; it < ix->prm.kmeans_iters; ++it) {
    hip_check(vchip::launch_rows_to_bf16(ix->cent.as<float>(), L, int(D), nullptr, Kp, ix->cent_bf16.as<__bf16>(), s),
              "bf16 centroids");
    hip_check(vchip::launch_gemm_argmax(xs.as<__bf16>(), S, Kp, ix->cent_bf16.as<__bf16>(), L, Kp, Kp,
                                        assign.as<int>(), s),
              "assign sample");
    group_by_bucket(ix, assign.as<int>(), S, L, perm.as<uint32_t>(), bsize, boff, tmp);
    // members as rows of x: the sorted sample positions mapped to rows
    srows_sorted.reserve(S * 4, s);
    std::vector<uint32_t> sp = download<uint32_t>(perm, S, s);
    for (auto& v : sp) v = srows[v];
    upload(srows_sorted, sp, s);
    hip_check(vchip::launch_centroid_sums(x, int(D), inv.as<float>(), srows_sorted.as<uint32_t>(), boff.as<long>(),
                                          bsize.as<int>(), L, ix->cent.as<float>(), s),
              "centroid update");
  }
  xs.release();
  // every row to its nearest list
  hip_check(vchip::launch_rows_to_bf16(ix->cent.as<float>(), L, int(D), nullptr, Kp, ix->cent_bf16.as<__bf16>(), s),
            "bf16 centroids");
  hip_check(vchip::launch_gemm_argmax(xb.as<__bf16>(), N, Kp, ix->cent_bf16.as<__bf16>(), L, Kp, Kp, assign.as<int>(),
                                      s),
            "assign rows");
  xb.release();
  // hard fission: split every bucket above max_bucket_size
  std::vector<int> h;
  DevBuf side, ca, cb, d_list, d_new;
  for (int round = 0;; ++round) {
    h = group_by_bucket(ix, assign.as<int>(), N, L, perm.as<uint32_t>(), bsize, boff, tmp);
    std::vector<int> list;
    for (int b = 0; b < L; ++b)
      if (size_t(h[b]) > ix->prm.max_bucket_size) list.push_back(b);
    if (list.empty()) break;
    if (round >= 64) fail(VC_ERR_DEVICE, "bucket fission did not converge");
    const int nl = int(list.size());
    std::vector<int> nid(nl);
    for (int i = 0; i < nl; ++i) nid[i] = L + i;
    upload(d_list, list, s);
    upload(d_new, nid, s);
    side.reserve(N, s);
    ca.reserve(size_t(nl) * D * 4, s);
    cb.reserve(size_t(nl) * D * 4, s);
    hip_check(vchip::launch_split_buckets(x, int(D), inv.as<float>(), perm.as<uint32_t>(), boff.as<long>(),
                                          bsize.as<int>(), d_list.as<int>(), nl, side.as<uint8_t>(), ca.as<float>(),
                                          cb.as<float>(), s),
              "split buckets");
    ix->cent.reserve(size_t(L + nl) * D * 4, s, /*keep=*/true);
    hip_check(vchip::launch_scatter_children(ix->cent.as<float>(), int(D), d_list.as<int>(), nl, L, ca.as<float>(),
                                             cb.as<float>(), s),
              "children");
    hip_check(vchip::launch_apply_split(perm.as<uint32_t>(), boff.as<long>(), bsize.as<int>(), d_list.as<int>(), nl,
                                        d_new.as<int>(), side.as<uint8_t>(), assign.as<int>(), s),
              "apply split");
    L += nl;
  }
  // lists the coarse quantizer left empty are dropped (buckets renumbered in order)
  if (std::find(h.begin(), h.end(), 0) != h.end()) {
    std::vector<int> map(L, 0), keep;
    for (int b = 0; b < L; ++b) {
      map[b] = int(keep.size());
      if (h[b] > 0) keep.push_back(b);
    }
    DevBuf d_map, d_keep, c2;
    upload(d_map, map, s);
    upload(d_keep, keep, s);
    hip_check(vchip::launch_remap(assign.as<int>(), N, d_map.as<int>(), s), "remap buckets");
    c2.reserve(keep.size() * D * 4, s);
    hip_check(vchip::launch_gather_rows_f32(ix->cent.as<float>(), int(D), d_keep.as<int>(), keep.size(), c2.as<float>(),
                                            s),
              "compact centroids");
    std::swap(ix->cent, c2);
    hip_check(hipStreamSynchronize(s), "sync");
    c2.release();
    L = int(keep.size());
    h = group_by_bucket(ix, assign.as<int>(), N, L, perm.as<uint32_t>(), bsize, boff, tmp);
  }
  ix->L = L;
  ix->h_bsize = h;
  std::swap(ix->bsize, bsize);
  std::swap(ix->boff, boff);
  bsize.release();
  boff.release();
  ix->cent_bf16.reserve(size_t(L) * Kp * 2, s);
  hip_check(vchip::launch_rows_to_bf16(ix->cent.as<float>(), L, int(D), nullptr, Kp, ix->cent_bf16.as<__bf16>(), s),
            "routing centroids");
  // TQ+ on the residuals of the calibration sample
  if (ix->n_cal >= 2) {
    const size_t nc = ix->n_cal;
    DevBuf cinv, cbf, cas, res, crows2;
    cinv.reserve(nc * 4, s);
    hip_check(vchip::launch_row_inv_norms(ix->cal_sample.as<float>(), nc, int(D), cinv.as<float>(), nullptr, s),
              "sample norms");
    cbf.reserve(nc * size_t(Kp) * 2, s);
    hip_check(vchip::launch_rows_to_bf16(ix->cal_sample.as<float>(), nc, int(D), cinv.as<float>(), Kp,
                                         cbf.as<__bf16>(), s),
              "sample bf16");
    cas.reserve(nc * 4, s);
    hip_check(vchip::launch_gemm_argmax(cbf.as<__bf16>(), nc, Kp, ix->cent_bf16.as<__bf16>(), L, Kp, Kp,
                                        cas.as<int>(), s),
              "assign sample");
    std::vector<uint32_t> iota(nc);
    std::iota(iota.begin(), iota.end(), 0u);
    upload(crows2, iota, s);
    res.reserve(nc * D * 4, s);
    hip_check(vchip::launch_residual_rows(ix->cal_sample.as<float>(), int(D), cinv.as<float>(), crows2.as<uint32_t>(),
                                          cas.as<int>(), ix->prm.residual ? ix->cent.as<float>() : nullptr, 0, nc,
                                          res.as<float>(), s),
              "sample residuals");
    ix->calib.reserve(3 * D * 4, s);
    float* sh = ix->calib.as<float>(0);
    calibrate_device(ix->t, res.as<float>(), nc, sh, sh + D, sh + 2 * D, ix->scratch, s);
    hip_check(hipStreamSynchronize(s), "sync");
    ix->t.shift = sh;
    ix->t.scale = sh + D;
    ix->t.inv_scale = sh + 2 * D;
    ix->calibrated = true;
  } else {
    ix->t.shift = ix->t.scale = ix->t.inv_scale = nullptr;
    ix->calibrated = false;
  }
  // encode every row by position (bucket order)
  DevBuf pos_bucket, res;
  pos_bucket.reserve(N * 4, s);
  hip_check(vchip::launch_fill_pos_bucket(ix->boff.as<long>(), ix->bsize.as<int>(), L, pos_bucket.as<int>(), s),
            "position buckets");
  const size_t cap = (N + 63) / 64 * 64;
  ix->tiled.reserve(vchip::tiled_bytes(cap, ix->t.groups16), s);
  ix->alpha.reserve(cap * 4, s);
  ix->ids.reserve(N * 8, s);
  ix->t.debug_flags = g_debug_flags.load();
  const size_t chunk = std::min(N, kEncodeChunk);
  res.reserve(chunk * D * 4, s);
  for (size_t p0 = 0; p0 < N; p0 += chunk) {
    const size_t m = std::min(chunk, N - p0);
    hip_check(vchip::launch_residual_rows(x, int(D), inv.as<float>(), perm.as<uint32_t>(), pos_bucket.as<int>(),
                                          ix->prm.residual ? ix->cent.as<float>() : nullptr, p0, m, res.as<float>(), s),
              "residual rows");
    hip_check(vchip::launch_encode(ix->t, res.as<float>(), m, p0, ix->tiled.as<uint8_t>(), nullptr, nullptr,
                                   ix->alpha.as<float>(), s),
              "encode");
  }
  hip_check(vchip::launch_widen_ids(perm.as<uint32_t>(), N, 0, ix->ids.as<uint64_t>(), s), "ids");
  hip_check(hipStreamSynchronize(s), "sync");
  ix->n = N;
}

}  // namespace

extern "C" {

vc_status vc_bucketed_prepare(vc_bucketed* ix) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    if (ix->prepared) return;
    if (ix->n_rows > 0xFFFFFFFFull) fail(VC_ERR_INVALID_ARGUMENT, "index exceeds 2^32 vectors");
    if (ix->n_rows == 0) {
      ix->n = 0;
      ix->L = 0;
      ix->prepared = true;
      return;
    }
    build_buckets(ix);
    ix->prepared = true;
  });
}

}  // extern "C"

// ---- search ---------------------------------------------------------------------
namespace {

// Per-query state of one bucketed search, carved from one scratch region.
struct SearchBufs {
  float *qinv, *qhat, *scores, *probe_s, *probe_off, *ls, *lb, *bc, *bt, *part_s;
  __bf16* qb;
  int *probe_b, *probe_n, *pre, *item_len, *n_items, *item_off, *part_cnt;
  uint32_t* part_i;
  uint8_t *luts, *src;
  size_t route_chunk;
  // the bucket-major tile scan (tiles set): its planning buffers
  bool tiles;
  int *bcnt, *bstart, *tstart, *n_tiles, *prank, *pair_q, *pair_j, *tile_b, *tile_p0, *tile_n, *gbound;
};

constexpr size_t kTilePartBytes = size_t(1) << 30;  // slot lists of the tile scan, at most

// The bucket-major tile scan serves batches that put several queries on a
// bucket (each tile then reads its bucket's codes once for up to eight
// queries); small batches keep the query-major scan. Debug flags bit 7 /
// bit 8 force one or the other (test hooks).
bool use_tile_scan(const vc_bucketed* ix, size_t nq, int P, size_t k) {
  if (k > size_t(vchip::kTileMaxK) || nq * size_t(P) * k * 8 > kTilePartBytes || nq > size_t(INT32_MAX) / size_t(P))
    return false;
  const int f = g_debug_flags.load();
  if ((f & 128) != 0) return true;
  if ((f & 256) != 0) return false;
  return nq * size_t(P) >= 8 * size_t(ix->L);
}

SearchBufs carve_search(vc_bucketed* ix, size_t nq, int P, int k, uint8_t* (*take)(void*, size_t), void* ctx,
                        size_t* need, bool tiles = false) {
  const size_t D = ix->dim, G = size_t(ix->t.groups);
  const int L = ix->L, Kp = ix->Kp;
  const size_t items = tiles ? nq * size_t(P) : nq * vchip::kBucketMaxItems;
  const size_t chunk = std::max<size_t>(1, std::min(nq, kRouteScoreBytes / (size_t(L) * 4)));
  Carve c;
  const size_t o_qinv = c.take(nq * 4), o_qhat = c.take(nq * D * 4), o_qb = c.take(nq * size_t(Kp) * 2),
               o_sc = c.take(chunk * size_t(L) * 4), o_pb = c.take(nq * P * 4), o_ps = c.take(nq * P * 4),
               o_pn = c.take(nq * 4), o_po = c.take(nq * P * 4), o_pre = c.take(nq * (P + 1) * 4),
               o_il = c.take(nq * 4), o_ni = c.take(nq * 4), o_io = c.take((nq + 1) * 4), o_l = c.take(nq * G * 32),
               o_ls = c.take(nq * 4), o_lb = c.take(nq * 4), o_bc = c.take(nq * 4), o_bt = c.take(nq * 4),
               o_src = c.take(vchip::bucket_src_bytes(ix->t.groups16, nq)), o_ps2 = c.take(items * k * 4),
               o_pi = c.take(items * k * 4), o_pc = c.take(items * 4);
  size_t o_tile[11] = {};
  if (tiles) {
    const size_t sizes[11] = {size_t(L), size_t(L) + 1, size_t(L) + 1, 1, nq * P, nq * P, nq * P,
                              nq * P,    nq * P,        nq * P,        nq};
    for (int i = 0; i < 11; ++i) o_tile[i] = c.take(sizes[i] * 4);
  }
  if (need != nullptr) *need = c.off;
  uint8_t* b = take(ctx, c.off);
  SearchBufs r;
  r.qinv = reinterpret_cast<float*>(b + o_qinv);
  r.qhat = reinterpret_cast<float*>(b + o_qhat);
  r.qb = reinterpret_cast<__bf16*>(b + o_qb);
  r.scores = reinterpret_cast<float*>(b + o_sc);
  r.probe_b = reinterpret_cast<int*>(b + o_pb);
  r.probe_s = reinterpret_cast<float*>(b + o_ps);
  r.probe_n = reinterpret_cast<int*>(b + o_pn);
  r.probe_off = reinterpret_cast<float*>(b + o_po);
  r.pre = reinterpret_cast<int*>(b + o_pre);
  r.item_len = reinterpret_cast<int*>(b + o_il);
  r.n_items = reinterpret_cast<int*>(b + o_ni);
  r.item_off = reinterpret_cast<int*>(b + o_io);
  r.luts = b + o_l;
  r.ls = reinterpret_cast<float*>(b + o_ls);
  r.lb = reinterpret_cast<float*>(b + o_lb);
  r.bc = reinterpret_cast<float*>(b + o_bc);
  r.bt = reinterpret_cast<float*>(b + o_bt);
  r.src = b + o_src;
  r.part_s = reinterpret_cast<float*>(b + o_ps2);
  r.part_i = reinterpret_cast<uint32_t*>(b + o_pi);
  r.part_cnt = reinterpret_cast<int*>(b + o_pc);
  r.route_chunk = chunk;
  r.tiles = tiles;
  int** tp[11] = {&r.bcnt, &r.bstart, &r.tstart, &r.n_tiles, &r.prank, &r.pair_q,
                  &r.pair_j, &r.tile_b, &r.tile_p0, &r.tile_n, &r.gbound};
  for (int i = 0; i < 11; ++i) *tp[i] = tiles ? reinterpret_cast<int*>(b + o_tile[i]) : nullptr;
  return r;
}

uint8_t* take_growable(void* ctx, size_t bytes) {
  auto* ix = static_cast<vc_bucketed*>(ctx);
  ix->scratch.reserve(bytes, ix->stream);
  return ix->scratch.as<uint8_t>();
}

// Routing: normalized queries, scores against every routing centroid, and
// each query's top-P buckets.
void route(vc_bucketed* ix, const float* q_dev, size_t nq, int P, const SearchBufs& b, hipStream_t s) {
  const int L = ix->L, Kp = ix->Kp;
  hip_check(vchip::launch_row_inv_norms(q_dev, nq, int(ix->dim), b.qinv, b.qhat, s), "query norms");
  hip_check(vchip::launch_rows_to_bf16(b.qhat, nq, int(ix->dim), nullptr, Kp, b.qb, s), "query bf16");
  for (size_t q0 = 0; q0 < nq; q0 += b.route_chunk) {
    const size_t m = std::min(b.route_chunk, nq - q0);
    hip_check(vchip::launch_gemm_scores(b.qb + q0 * Kp, m, Kp, ix->cent_bf16.as<__bf16>(), L, Kp, Kp, b.scores, L, s),
              "routing scores");
    hip_check(vchip::launch_topk_rows(b.scores, m, L, L, P, b.probe_b + q0 * P, b.probe_s + q0 * P, s),
              "routing top-P");
  }
}

// The search after routing: probe counts -> plan -> offsets -> LUTs -> scan -> merge.
size_t search_probed(vc_bucketed* ix, const float* q_dev, size_t nq, size_t k, int P, const SearchBufs& b,
                     float* out_s, uint64_t* out_i, uint64_t* counts, hipStream_t s) {
  const int ke = int(std::min(k, ix->n));
  if (b.tiles) {
    hip_check(vchip::launch_probe_offsets(b.qhat, nq, int(ix->dim), ix->prm.residual ? ix->cent.as<float>() : nullptr,
                                          P, b.probe_b, b.probe_n, b.probe_off, s),
              "probe offsets");
    hip_check(vchip::launch_prepare_queries(ix->t, q_dev, nq, b.luts, b.ls, b.lb, b.bc, b.bt, s), "prepare_queries");
    vchip::BucketTileParams p{};
    p.tiled = ix->tiled.as<uint8_t>();
    p.alpha = ix->alpha.as<float>();
    p.G = ix->t.groups;
    p.G16 = ix->t.groups16;
    p.luts = b.luts;
    p.lut_scale = b.ls;
    p.bias_total = b.bt;
    p.nq = long(nq);
    p.P = P;
    p.probe_n = b.probe_n;
    p.probe_b = b.probe_b;
    p.probe_off = b.probe_off;
    p.boff = ix->boff.as<long>();
    p.bsize = ix->bsize.as<int>();
    p.L = ix->L;
    p.k = ke;
    p.bcnt = b.bcnt;
    p.bstart = b.bstart;
    p.tstart = b.tstart;
    p.n_tiles = b.n_tiles;
    p.prank = b.prank;
    p.pair_q = b.pair_q;
    p.pair_j = b.pair_j;
    p.tile_b = b.tile_b;
    p.tile_p0 = b.tile_p0;
    p.tile_n = b.tile_n;
    p.gbound = b.gbound;
    p.part_s = b.part_s;
    p.part_i = b.part_i;
    p.part_cnt = b.part_cnt;
    std::unique_lock probe_lock(g_probe.mu);
    const bool probe = g_probe.on;
    if (!probe) probe_lock.unlock();
    hip_check(vchip::launch_bucket_tiles(p, ke, ix->ids.as<uint64_t>(), out_s, out_i, counts, s,
                                         probe ? g_probe.start : nullptr, probe ? g_probe.stop : nullptr),
              "bucket tile scan");
    if (probe) {
      g_probe.recorded = true;
      g_probe.tile = vchip::kTileQueries;
      g_probe.ranges = 0;
    }
    return size_t(ke);
  }
  hip_check(vchip::launch_plan_probes(nq, P, b.probe_b, b.probe_n, ix->bsize.as<int>(), vchip::kBucketItem,
                                      vchip::kBucketMaxItems, b.pre, b.item_len, b.n_items, b.item_off, s),
            "plan probes");
  hip_check(vchip::launch_probe_offsets(b.qhat, nq, int(ix->dim), ix->prm.residual ? ix->cent.as<float>() : nullptr,
                                        P, b.probe_b, b.probe_n, b.probe_off, s),
            "probe offsets");
  hip_check(vchip::launch_prepare_queries(ix->t, q_dev, nq, b.luts, b.ls, b.lb, b.bc, b.bt, s), "prepare_queries");
  vchip::BucketScanParams p{};
  p.tiled = ix->tiled.as<uint8_t>();
  p.alpha = ix->alpha.as<float>();
  p.G16 = ix->t.groups16;
  p.lut_scale = b.ls;
  p.bias_total = b.bt;
  p.nq = long(nq);
  p.P = P;
  p.probe_n = b.probe_n;
  p.probe_b = b.probe_b;
  p.probe_off = b.probe_off;
  p.probe_pre = b.pre;
  p.boff = ix->boff.as<long>();
  p.item_len = b.item_len;
  p.item_off = b.item_off;
  p.k = ke;
  p.part_s = b.part_s;
  p.part_i = b.part_i;
  p.part_cnt = b.part_cnt;
  std::unique_lock probe_lock(g_probe.mu);
  const bool probe = g_probe.on;
  if (!probe) probe_lock.unlock();
  if (probe) hip_check(hipEventRecord(g_probe.start, s), "hipEventRecord");
  hip_check(vchip::launch_bucket_scan(p, b.luts, ix->t.groups, b.src, s), "bucket scan");
  if (probe) {
    hip_check(hipEventRecord(g_probe.stop, s), "hipEventRecord");
    g_probe.recorded = true;
    g_probe.tile = 1;
    g_probe.ranges = 0;
    probe_lock.unlock();
  }
  hip_check(vchip::launch_merge_items(b.part_s, b.part_i, b.part_cnt, b.item_off, nq, ke, ke, ix->ids.as<uint64_t>(),
                                      out_s, out_i, counts, s),
            "merge items");
  return size_t(ke);
}

int probe_width(vc_bucketed* ix, size_t n_probe, bool learned) {
  size_t P = learned ? std::max<size_t>(n_probe, vchip::kProbeMaxP) : n_probe;
  P = std::min<size_t>(P, kRouteMaxP);
  return int(std::min<size_t>(P, size_t(ix->L)));
}

// mode: 0 = fixed n_probe, 1 = per-query counts given, 2 = learned
size_t bucketed_search_dev(vc_bucketed* ix, const float* q_dev, size_t nq, size_t k, size_t n_probe,
                           const int32_t* probe_counts, float* out_s, uint64_t* out_i, uint64_t* counts,
                           hipStream_t s) {
  if (!ix->prepared) fail(VC_ERR_STATE, "search() before prepare()");
  const size_t ke = std::min(k, ix->n);
  if (nq == 0 || ke == 0) return ke;
  if (ke > size_t(vchip::kMaxTopK)) fail(VC_ERR_INVALID_ARGUMENT, "k exceeds the GPU top-k capacity of 2048");
  const bool learned = probe_counts == nullptr && n_probe == 0 && ix->use_probe && ix->probe.trained;
  const size_t np = n_probe != 0 ? n_probe : ix->prm.n_probe;
  if (probe_counts != nullptr && n_probe == 0)
    fail(VC_ERR_INVALID_ARGUMENT, "per-query probe counts need n_probe as their upper bound");
  const int P = probe_width(ix, learned ? vchip::kProbeMaxP : np, learned);
  const SearchBufs b = carve_search(ix, nq, P, int(ke), take_growable, ix, nullptr, use_tile_scan(ix, nq, P, ke));
  route(ix, q_dev, nq, P, b, s);
  if (learned) {
    hip_check(vchip::probe_counts_from_model(ix->probe, nq, P, b.probe_s, b.probe_b, ix->bsize.as<int>(), b.probe_n,
                                             s),
              "probe model");
  } else if (probe_counts != nullptr) {
    hip_check(vchip::launch_clamp_counts(nq, P, probe_counts, b.probe_b, b.probe_n, s), "probe counts");
  } else {
    hip_check(vchip::launch_fixed_probes(nq, P, int(std::min<size_t>(np, size_t(P))), b.probe_b, b.probe_n, s),
              "fixed probes");
  }
  return search_probed(ix, q_dev, nq, ke, P, b, out_s, out_i, counts, s);
}

}  // namespace

extern "C" {

vc_status vc_bucketed_search(vc_bucketed* ix, const float* queries, size_t nq, size_t k, float* scores, uint64_t* ids,
                             size_t* k_out) {
  return guarded([&] {
    if (ix == nullptr || (queries == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (!ix->prepared) fail(VC_ERR_STATE, "search() before prepare()");
    const size_t ke = std::min(k, ix->n);
    if (k_out) *k_out = ke;
    if (nq == 0 || ke == 0) return;
    if (scores == nullptr || ids == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null output");
    Carve c;
    const size_t oq = c.take(nq * ix->dim * 4), os = c.take(nq * ke * 4), oi = c.take(nq * ke * 8);
    ix->io.reserve(c.off, ix->stream);
    hip_check(hipMemcpyAsync(ix->io.as<float>(oq), queries, nq * ix->dim * 4, hipMemcpyHostToDevice, ix->stream),
              "copy");
    bucketed_search_dev(ix, ix->io.as<float>(oq), nq, k, 0, nullptr, ix->io.as<float>(os), ix->io.as<uint64_t>(oi),
                        nullptr, ix->stream);
    hip_check(hipMemcpyAsync(scores, ix->io.as<float>(os), nq * ke * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipMemcpyAsync(ids, ix->io.as<uint64_t>(oi), nq * ke * 8, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_bucketed_search_device(vc_bucketed* ix, const float* queries_dev, size_t nq, size_t k, size_t n_probe,
                                    const int32_t* probe_counts_dev, float* scores_dev, uint64_t* ids_dev,
                                    uint64_t* counts_dev, size_t* k_out, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (queries_dev == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    const size_t ke = bucketed_search_dev(ix, queries_dev, nq, k, n_probe, probe_counts_dev, scores_dev, ids_dev,
                                          counts_dev, static_cast<hipStream_t>(stream));
    if (k_out) *k_out = ke;
  });
}

vc_status vc_bucketed_search_exhaustive_device(vc_bucketed* ix, const float* queries_dev, size_t nq, size_t k,
                                               float* scores_dev, uint64_t* ids_dev, size_t* k_out, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (queries_dev == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (!ix->prepared) fail(VC_ERR_STATE, "search() before prepare()");
    if (ix->prm.residual) fail(VC_ERR_STATE, "the exhaustive flat scan needs codes of the rows (residual = 0)");
    const auto r = search_tiled(ix->t, ix->tiled.as<uint8_t>(), ix->alpha.as<float>(), ix->n, queries_dev, nq, k,
                                ix->ids.as<uint64_t>(), scores_dev, ids_dev, ix->scratch,
                                static_cast<hipStream_t>(stream));
    if (k_out) *k_out = r.k_eff;
  });
}

vc_status vc_bucketed_route_device(vc_bucketed* ix, const float* queries_dev, size_t nq, size_t P, int32_t* probe_b_dev,
                                   float* probe_s_dev, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (queries_dev == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (!ix->prepared) fail(VC_ERR_STATE, "route() before prepare()");
    if (P == 0 || P > size_t(kRouteMaxP) || P > size_t(ix->L))
      fail(VC_ERR_INVALID_ARGUMENT, "P must be in [1, min(512, n_buckets)]");
    if (nq == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const SearchBufs b = carve_search(ix, nq, int(P), 1, take_growable, ix, nullptr);
    route(ix, queries_dev, nq, int(P), b, s);
    hip_check(hipMemcpyAsync(probe_b_dev, b.probe_b, nq * P * 4, hipMemcpyDeviceToDevice, s), "copy");
    hip_check(hipMemcpyAsync(probe_s_dev, b.probe_s, nq * P * 4, hipMemcpyDeviceToDevice, s), "copy");
  });
}

size_t vc_bucketed_dim(const vc_bucketed* ix) { return ix ? ix->dim : 0; }
size_t vc_bucketed_size(const vc_bucketed* ix) { return ix ? ix->n_rows : 0; }
size_t vc_bucketed_bits(const vc_bucketed* ix) { return ix ? ix->bits : 0; }
size_t vc_bucketed_n_buckets(const vc_bucketed* ix) { return ix && ix->prepared ? size_t(ix->L) : 0; }
int vc_bucketed_is_prepared(const vc_bucketed* ix) { return ix && ix->prepared ? 1 : 0; }
int vc_bucketed_is_calibrated(const vc_bucketed* ix) { return ix && ix->calibrated ? 1 : 0; }

vc_status vc_bucketed_export(vc_bucketed* ix, uint32_t* bucket_sizes, float* centroids, uint64_t* pos_ids,
                             uint8_t* flat_codes, float* alpha, float* shift, float* scale) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    if (!ix->prepared) fail(VC_ERR_STATE, "export before prepare()");
    hipStream_t s = ix->stream;
    const size_t n = ix->n, D = ix->dim, G = size_t(ix->t.groups), L = size_t(ix->L);
    if (bucket_sizes) for (size_t b = 0; b < L; ++b) bucket_sizes[b] = uint32_t(ix->h_bsize[b]);
    if (centroids && L) hip_check(hipMemcpyAsync(centroids, ix->cent.p, L * D * 4, hipMemcpyDeviceToHost, s), "copy");
    if (pos_ids && n) hip_check(hipMemcpyAsync(pos_ids, ix->ids.p, n * 8, hipMemcpyDeviceToHost, s), "copy");
    if (alpha && n) hip_check(hipMemcpyAsync(alpha, ix->alpha.p, n * 4, hipMemcpyDeviceToHost, s), "copy");
    if (flat_codes && n) {
      ix->io.reserve(n * G, s);
      hip_check(vchip::launch_tiled_to_flat(ix->tiled.as<uint8_t>(), n, int(G), ix->t.groups16, ix->io.as<uint8_t>(), s),
                "tiled_to_flat");
      hip_check(hipMemcpyAsync(flat_codes, ix->io.p, n * G, hipMemcpyDeviceToHost, s), "copy");
    }
    if (shift || scale) {
      if (ix->calibrated) {
        if (shift) hip_check(hipMemcpyAsync(shift, ix->calib.as<float>(0), D * 4, hipMemcpyDeviceToHost, s), "copy");
        if (scale) hip_check(hipMemcpyAsync(scale, ix->calib.as<float>(D * 4), D * 4, hipMemcpyDeviceToHost, s), "copy");
      } else {
        if (shift) std::fill(shift, shift + D, 0.f);
        if (scale) std::fill(scale, scale + D, 1.f);
      }
    }
    hip_check(hipStreamSynchronize(s), "sync");
  });
}

}  // extern "C"

// ---- learned probe budget ------------------------------------------------------
extern "C" {

size_t vc_probe_param_count(void) { return size_t(vchip::kProbeParams); }

vc_status vc_bucketed_probe_features(vc_bucketed* ix, const float* queries_dev, size_t nq, float* feats_dev,
                                     void* stream) {
  return guarded([&] {
    if (ix == nullptr || (queries_dev == nullptr && nq > 0) || (feats_dev == nullptr && nq > 0))
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (!ix->prepared) fail(VC_ERR_STATE, "probe_features() before prepare()");
    if (nq == 0 || ix->L == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int P = probe_width(ix, vchip::kProbeMaxP, true);
    const SearchBufs b = carve_search(ix, nq, P, 1, take_growable, ix, nullptr);
    route(ix, queries_dev, nq, P, b, s);
    hip_check(vchip::launch_probe_features(b.probe_s, b.probe_b, ix->bsize.as<int>(), nq, P, feats_dev, s),
              "probe features");
  });
}

static void require_model(vc_bucketed* ix) {
  if (ix->probe.mem == nullptr) fail(VC_ERR_STATE, "no probe model: call vc_bucketed_probe_init first");
}

vc_status vc_bucketed_probe_init(vc_bucketed* ix, uint64_t seed) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    hip_check(vchip::probe_model_init(ix->probe, seed, ix->stream), "probe model init");
    ix->use_probe = false;
  });
}

vc_status vc_bucketed_probe_set_weights(vc_bucketed* ix, const float* w) {
  return guarded([&] {
    if (ix == nullptr || w == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (ix->probe.mem == nullptr) hip_check(vchip::probe_model_init(ix->probe, 0, ix->stream), "probe model init");
    hip_check(vchip::probe_model_set(ix->probe, w, ix->stream), "probe weights");
    ix->probe.trained = true;
  });
}

vc_status vc_bucketed_probe_get_weights(vc_bucketed* ix, float* w) {
  return guarded([&] {
    if (ix == nullptr || w == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    require_model(ix);
    hip_check(vchip::probe_model_get(ix->probe, w, ix->stream), "probe weights");
  });
}

vc_status vc_bucketed_probe_train(vc_bucketed* ix, const float* feats_dev, const float* targets_dev, size_t n,
                                  const int32_t* order_dev, size_t n_steps, float lr, float* losses_dev,
                                  void* stream) {
  return guarded([&] {
    if (ix == nullptr || feats_dev == nullptr || targets_dev == nullptr || order_dev == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) fail(VC_ERR_INVALID_ARGUMENT, "no training rows");
    if (n_steps > 0x7fffffffull) fail(VC_ERR_INVALID_ARGUMENT, "too many steps");
    if (!(lr > 0.f)) fail(VC_ERR_INVALID_ARGUMENT, "lr must be positive");
    std::lock_guard lock(ix->mu);
    require_model(ix);
    hip_check(vchip::launch_probe_train(ix->probe, feats_dev, targets_dev, n, reinterpret_cast<const int*>(order_dev),
                                        int(n_steps), lr, 0.9f, 0.999f, 1e-8f, false, losses_dev,
                                        static_cast<hipStream_t>(stream)),
              "probe train");
    ix->probe.trained = true;
  });
}

vc_status vc_bucketed_probe_grads(vc_bucketed* ix, const float* feats_dev, const float* targets_dev, size_t n,
                                  const int32_t* order_dev, float* grads_host) {
  return guarded([&] {
    if (ix == nullptr || feats_dev == nullptr || targets_dev == nullptr || order_dev == nullptr ||
        grads_host == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) fail(VC_ERR_INVALID_ARGUMENT, "no training rows");
    std::lock_guard lock(ix->mu);
    require_model(ix);
    hip_check(vchip::launch_probe_train(ix->probe, feats_dev, targets_dev, n, reinterpret_cast<const int*>(order_dev), 1,
                                        1e-3f, 0.9f, 0.999f, 1e-8f, true, nullptr, ix->stream),
              "probe grads");
    hip_check(hipMemcpyAsync(grads_host, ix->probe.d.grad, size_t(vchip::kProbeParams) * 4, hipMemcpyDeviceToHost,
                             ix->stream),
              "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_bucketed_probe_predict(vc_bucketed* ix, const float* feats_dev, size_t n, float* y_dev, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (n > 0 && (feats_dev == nullptr || y_dev == nullptr)))
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    require_model(ix);
    hip_check(vchip::launch_probe_predict(ix->probe, feats_dev, n, y_dev, static_cast<hipStream_t>(stream)),
              "probe predict");
  });
}

vc_status vc_bucketed_probe_set_table(vc_bucketed* ix, const int32_t* counts, size_t n_bins, float y0, float y_step) {
  return guarded([&] {
    if (ix == nullptr || (counts == nullptr && n_bins > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (n_bins > size_t(vchip::kProbeMaxBins)) fail(VC_ERR_INVALID_ARGUMENT, "at most 64 bins");
    if (!(y_step > 0.f)) fail(VC_ERR_INVALID_ARGUMENT, "y_step must be positive");
    for (size_t b = 0; b < n_bins; ++b)
      if (counts[b] < 1) fail(VC_ERR_INVALID_ARGUMENT, "probe counts must be at least 1");
    std::lock_guard lock(ix->mu);
    require_model(ix);
    hip_check(vchip::probe_model_set_table(ix->probe, counts, int(n_bins), y0, y_step, ix->stream), "probe table");
  });
}

vc_status vc_bucketed_probe_enable(vc_bucketed* ix, int on, float shift) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    if (on) {
      require_model(ix);
      if (!ix->probe.trained) fail(VC_ERR_STATE, "the probe model is not trained");
    }
    ix->use_probe = on != 0;
    ix->probe.shift = shift;
  });
}

}  // extern "C"
