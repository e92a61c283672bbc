// Prepared queries as a device handle (vc_prepared_*, vc_score_prepared*) and
// the reference's repack free functions over host buffers
// (vc_codes_flat_from_packed, vc_repack_perm0, vc_repack_seq).
//
// The reference's PreparedQueries (include/vectorcache/search/search.hpp:
// 95-112) is built once by prepare_queries (src/search/search.cpp:762-860) and
// reused by every score_prepared / score_prepared_into call over the blocked
// code ranges of a multi-bucket search (src/search/search.cpp:1045-1267). The
// handle keeps that state in HBM (LUT rows, LUT scales and biases, bias_corr),
// so a chain of score calls builds the query LUTs once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "vc_common.hpp"
#include "vc_internal.hpp"
#include "vectorcache_hip.h"

using namespace vcapi;

struct vc_prepared {
  int device = 0;
  size_t nq = 0, dim = 0, bits = 0, G = 0;
  DevBuf buf;  // [luts nq x G*32][lut_scale nq][lut_bias nq][bias_corr nq][bias_total nq]
  size_t o_luts = 0, o_ls = 0, o_lb = 0, o_bc = 0, o_bt = 0;
  hipEvent_t ready = nullptr;  // recorded after the state is written

  const uint8_t* luts() const { return buf.as<uint8_t>(o_luts); }
  const float* ls() const { return buf.as<float>(o_ls); }
  const float* bt() const { return buf.as<float>(o_bt); }
};

namespace {

void layout(vc_prepared* p) {
  Carve c;
  p->o_luts = c.take(p->nq * p->G * 32);
  p->o_ls = c.take(p->nq * 4);
  p->o_lb = c.take(p->nq * 4);
  p->o_bc = c.take(p->nq * 4);
  p->o_bt = c.take(p->nq * 4);
  hipStream_t s0 = nullptr;
  p->buf.reserve(std::max<size_t>(c.off, 256), s0);
}

vc_prepared* new_prepared(size_t nq, size_t dim, size_t bits) {
  auto* p = new vc_prepared();
  try {
    hip_check(hipGetDevice(&p->device), "hipGetDevice");
    p->nq = nq;
    p->dim = dim;
    p->bits = bits;
    p->G = size_t(vchip::byte_groups(bits, dim));
    layout(p);
    hip_check(hipEventCreateWithFlags(&p->ready, hipEventDisableTiming), "hipEventCreate");
  } catch (...) {
    p->buf.release();
    delete p;
    throw;
  }
  return p;
}

// Runs prepare_queries for device query rows into the handle on stream s.
void prepare_into(vc_prepared* p, const float* q_dev, const float* shift_dev, const float* scale_dev,
                  hipStream_t s) {
  vchip::CodeTables t = make_tables(p->dim, p->bits);
  if (shift_dev != nullptr) {
    t.shift = shift_dev;
    t.scale = scale_dev;
    t.inv_scale = nullptr;  // not used by query preparation
  }
  uint8_t* b = p->buf.as<uint8_t>();
  if (p->nq > 0)
    hip_check(vchip::launch_prepare_queries(t, q_dev, p->nq, b + p->o_luts, p->buf.as<float>(p->o_ls),
                                            p->buf.as<float>(p->o_lb), p->buf.as<float>(p->o_bc),
                                            p->buf.as<float>(p->o_bt), s),
              "prepare_queries");
  hip_check(hipEventRecord(p->ready, s), "hipEventRecord");
}

void check_prepared(const vc_prepared* p) {
  if (p == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null prepared queries");
  int dev = -1;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != p->device) fail(VC_ERR_STATE, "prepared queries belong to another device");
}

// The listed queries' LUT rows, scales and total biases, gathered in list
// order into `a` (the handle's own rows when idx is NULL).
PreLuts gather(const vc_prepared* p, const size_t* idx, size_t m, DevBuf& a, size_t o_idx, size_t o_rows,
               size_t o_ls, size_t o_bt, hipStream_t s) {
  PreLuts pl{p->luts(), p->ls(), p->bt()};
  if (idx == nullptr) return pl;
  std::vector<uint64_t> hidx(idx, idx + m);
  hip_check(hipMemcpyAsync(a.as<uint64_t>(o_idx), hidx.data(), m * 8, hipMemcpyHostToDevice, s), "copy");
  hip_check(hipStreamSynchronize(s), "sync");  // hidx leaves scope
  hip_check(vchip::launch_gather_rows(p->luts(), a.as<uint64_t>(o_idx), m, p->G * 32, a.as<uint8_t>(o_rows), s),
            "gather");
  hip_check(vchip::launch_gather_rows(p->ls(), a.as<uint64_t>(o_idx), m, 4, a.as<float>(o_ls), s), "gather");
  hip_check(vchip::launch_gather_rows(p->bt(), a.as<uint64_t>(o_idx), m, 4, a.as<float>(o_bt), s), "gather");
  return PreLuts{a.as<uint8_t>(o_rows), a.as<float>(o_ls), a.as<float>(o_bt)};
}

// PERM0 blocks of n vectors (host) -> the scan's tiled layout in `a` at o_t.
void upload_perm0_tiled(const vchip::CodeTables& t, const uint8_t* perm0, size_t n, DevBuf& a, size_t o_p,
                        size_t o_f, size_t o_t, hipStream_t s) {
  const size_t G = size_t(t.groups);
  const size_t pbytes = (n + 191) / 192 * G * 192;
  hip_check(hipMemcpyAsync(a.as<uint8_t>(o_p), perm0, pbytes, hipMemcpyHostToDevice, s), "copy");
  hip_check(vchip::launch_perm0_to_flat(a.as<uint8_t>(o_p), n, int(G), a.as<uint8_t>(o_f), s), "perm0_to_flat");
  hip_check(vchip::launch_flat_to_tiled(a.as<uint8_t>(o_f), n, int(G), t.groups16, a.as<uint8_t>(o_t), s),
            "flat_to_tiled");
}

}  // namespace

extern "C" {

vc_status vc_prepare_queries(const float* queries, size_t nq, size_t dim, size_t bits, const float* shift,
                             const float* scale, vc_prepared** out) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    if (out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (nq > 0 && queries == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if ((shift == nullptr) != (scale == nullptr)) fail(VC_ERR_INVALID_ARGUMENT, "shift and scale go together");
    Scoped sc;
    vc_prepared* p = new_prepared(nq, dim, bits);
    try {
      Carve c;
      const size_t oq = c.take(nq * dim * 4), ocal = c.take(2 * dim * 4);
      sc.a.reserve(c.off, sc.s);
      if (nq > 0)
        hip_check(hipMemcpyAsync(sc.a.as<float>(oq), queries, nq * dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
      if (shift != nullptr) {
        hip_check(hipMemcpyAsync(sc.a.as<float>(ocal), shift, dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
        hip_check(hipMemcpyAsync(sc.a.as<float>(ocal + dim * 4), scale, dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
      }
      prepare_into(p, sc.a.as<float>(oq), shift ? sc.a.as<float>(ocal) : nullptr,
                   shift ? sc.a.as<float>(ocal + dim * 4) : nullptr, sc.s);
      hip_check(hipStreamSynchronize(sc.s), "sync");
    } catch (...) {
      (void)hipStreamSynchronize(sc.s);
      (void)hipEventDestroy(p->ready);
      p->buf.release();
      delete p;
      throw;
    }
    *out = p;
  });
}

vc_status vc_prepare_queries_device(const float* queries_dev, size_t nq, size_t dim, size_t bits,
                                    const float* shift_dev, const float* scale_dev, void* stream, vc_prepared** out) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    if (out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (nq > 0 && queries_dev == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if ((shift_dev == nullptr) != (scale_dev == nullptr))
      fail(VC_ERR_INVALID_ARGUMENT, "shift and scale go together");
    vc_prepared* p = new_prepared(nq, dim, bits);
    try {
      prepare_into(p, queries_dev, shift_dev, scale_dev, static_cast<hipStream_t>(stream));
    } catch (...) {
      (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
      (void)hipEventDestroy(p->ready);
      p->buf.release();
      delete p;
      throw;
    }
    *out = p;
  });
}

vc_status vc_prepared_destroy(vc_prepared* p) {
  return guarded([&] {
    if (p == nullptr) return;
    (void)hipEventSynchronize(p->ready);
    (void)hipEventDestroy(p->ready);
    p->buf.release();
    delete p;
  });
}

vc_status vc_prepared_shape(const vc_prepared* p, size_t* nq, size_t* dim, size_t* bits, size_t* n_byte_groups) {
  return guarded([&] {
    if (p == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null prepared queries");
    if (nq) *nq = p->nq;
    if (dim) *dim = p->dim;
    if (bits) *bits = p->bits;
    if (n_byte_groups) *n_byte_groups = p->G;
  });
}

vc_status vc_prepared_export(const vc_prepared* p, uint8_t* luts, float* lut_scale, float* lut_bias,
                             float* bias_corr) {
  return guarded([&] {
    check_prepared(p);
    if (p->nq == 0) return;
    hip_check(hipEventSynchronize(p->ready), "hipEventSynchronize");
    auto down = [&](void* dst, size_t off, size_t bytes) {
      if (dst != nullptr) hip_check(hipMemcpy(dst, p->buf.as<uint8_t>(off), bytes, hipMemcpyDeviceToHost), "copy");
    };
    down(luts, p->o_luts, p->nq * p->G * 32);
    down(lut_scale, p->o_ls, p->nq * 4);
    down(lut_bias, p->o_lb, p->nq * 4);
    down(bias_corr, p->o_bc, p->nq * 4);
  });
}

vc_status vc_prepared_query_at(const vc_prepared* p, size_t qi, vc_prepared** out) {
  return guarded([&] {
    check_prepared(p);
    if (out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (qi >= p->nq) fail(VC_ERR_OUT_OF_RANGE, "prepared_query_at: query index out of range");
    Scoped sc;
    vc_prepared* q = new_prepared(1, p->dim, p->bits);
    try {
      hip_check(hipStreamWaitEvent(sc.s, p->ready, 0), "hipStreamWaitEvent");
      auto cp = [&](size_t dst_off, size_t src_off, size_t bytes) {
        hip_check(hipMemcpyAsync(q->buf.as<uint8_t>(dst_off), p->buf.as<uint8_t>(src_off), bytes,
                                 hipMemcpyDeviceToDevice, sc.s),
                  "copy");
      };
      cp(q->o_luts, p->o_luts + qi * p->G * 32, p->G * 32);
      cp(q->o_ls, p->o_ls + qi * 4, 4);
      cp(q->o_lb, p->o_lb + qi * 4, 4);
      cp(q->o_bc, p->o_bc + qi * 4, 4);
      cp(q->o_bt, p->o_bt + qi * 4, 4);
      hip_check(hipEventRecord(q->ready, sc.s), "hipEventRecord");
      hip_check(hipStreamSynchronize(sc.s), "sync");
    } catch (...) {
      (void)hipEventDestroy(q->ready);
      q->buf.release();
      delete q;
      throw;
    }
    *out = q;
  });
}

vc_status vc_score_prepared(const vc_prepared* p, size_t k, const uint8_t* perm0, const float* alpha, size_t n,
                            const uint64_t* id_map, const size_t* query_indices, size_t n_indices,
                            const float* residual_norms, const float* residual_scales, float* scores, uint64_t* ids,
                            size_t* k_out) {
  return guarded([&] {
    check_prepared(p);
    if (query_indices != nullptr)
      for (size_t i = 0; i < n_indices; ++i)
        if (query_indices[i] >= p->nq) fail(VC_ERR_OUT_OF_RANGE, "score_prepared: query_indices out of range");
    const size_t nq_out = query_indices != nullptr ? n_indices : p->nq;
    if (k_out) *k_out = 0;
    if (n == 0 || nq_out == 0) return;
    const size_t ke = std::min(k, n);
    if (k_out) *k_out = ke;
    if (ke == 0) return;
    if (perm0 == nullptr || alpha == nullptr || scores == nullptr || ids == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    const bool resid = residual_norms != nullptr && residual_scales != nullptr;
    std::vector<float> rs_sel;
    if (resid) {
      rs_sel.resize(nq_out);
      for (size_t l = 0; l < nq_out; ++l) rs_sel[l] = residual_scales[query_indices ? query_indices[l] : l];
    }
    Scoped sc;
    hip_check(hipStreamWaitEvent(sc.s, p->ready, 0), "hipStreamWaitEvent");
    const vchip::CodeTables t = make_tables(p->dim, p->bits);
    const size_t G = p->G;
    const size_t pbytes = (n + 191) / 192 * G * 192;
    Carve c;
    const size_t op = c.take(pbytes), of = c.take(n * G), ot = c.take(vchip::tiled_bytes(n, t.groups16)),
                 oa = c.take(n * 4), om = c.take(id_map ? n * 8 : 0), os = c.take(nq_out * ke * 4),
                 oi = c.take(nq_out * ke * 8), orn = c.take(resid ? n * 4 : 0), ors = c.take(resid ? nq_out * 4 : 0),
                 oidx = c.take(query_indices ? nq_out * 8 : 0), orows = c.take(query_indices ? nq_out * G * 32 : 0),
                 ols = c.take(query_indices ? nq_out * 4 : 0), obt = c.take(query_indices ? nq_out * 4 : 0);
    sc.a.reserve(c.off, sc.s);
    upload_perm0_tiled(t, perm0, n, sc.a, op, of, ot, sc.s);
    hip_check(hipMemcpyAsync(sc.a.as<float>(oa), alpha, n * 4, hipMemcpyHostToDevice, sc.s), "copy");
    if (id_map) hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(om), id_map, n * 8, hipMemcpyHostToDevice, sc.s), "copy");
    if (resid) {
      hip_check(hipMemcpyAsync(sc.a.as<float>(orn), residual_norms, n * 4, hipMemcpyHostToDevice, sc.s), "copy");
      hip_check(hipMemcpyAsync(sc.a.as<float>(ors), rs_sel.data(), nq_out * 4, hipMemcpyHostToDevice, sc.s), "copy");
    }
    const PreLuts pl = gather(p, query_indices, nq_out, sc.a, oidx, orows, ols, obt, sc.s);
    // with the residual correction the rows are ranked on positions first and
    // mapped afterwards (remap_ids follows the correction)
    const uint64_t* im_dev = id_map ? sc.a.as<uint64_t>(om) : nullptr;
    search_tiled(t, sc.a.as<uint8_t>(ot), sc.a.as<float>(oa), n, nullptr, nq_out, k, resid ? nullptr : im_dev,
                 sc.a.as<float>(os), sc.a.as<uint64_t>(oi), sc.scratch, sc.s, nullptr, nullptr, nullptr, &pl);
    if (resid)
      hip_check(vchip::launch_residual_post(nq_out, int(ke), sc.a.as<float>(os), sc.a.as<uint64_t>(oi),
                                            sc.a.as<float>(orn), sc.a.as<float>(ors), n, im_dev, sc.s),
                "residual_post");
    hip_check(hipMemcpyAsync(scores, sc.a.as<float>(os), nq_out * ke * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipMemcpyAsync(ids, sc.a.as<uint64_t>(oi), nq_out * ke * 8, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_score_prepared_into(const vc_prepared* p, size_t k, const uint8_t* perm0, const float* alpha, size_t n,
                                 const uint64_t* id_map, const size_t* query_indices, size_t n_indices, float* heap_s,
                                 uint64_t* heap_i, size_t* heap_sz, float* heap_min, size_t* heap_mi,
                                 const float* score_offsets) {
  return guarded([&] {
    check_prepared(p);
    const size_t nq = p->nq;
    if (n_indices == 0 || query_indices == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "score_prepared_into requires non-empty query_indices");
    for (size_t i = 0; i < n_indices; ++i)
      if (query_indices[i] >= nq) fail(VC_ERR_OUT_OF_RANGE, "score_prepared_into: query_indices out of range");
    if (n == 0 || k == 0) return;
    if (perm0 == nullptr || alpha == nullptr || heap_s == nullptr || heap_i == nullptr || heap_sz == nullptr ||
        heap_min == nullptr || heap_mi == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (k > size_t(vchip::kMaxHeapK)) fail(VC_ERR_INVALID_ARGUMENT, "k exceeds the GPU heap capacity of 4096");
    if (n > 0xFFFFFFFFull) fail(VC_ERR_INVALID_ARGUMENT, "bucket exceeds 2^32 vectors");
    {
      std::vector<size_t> sorted(query_indices, query_indices + n_indices);
      std::sort(sorted.begin(), sorted.end());
      if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        fail(VC_ERR_INVALID_ARGUMENT, "score_prepared_into: duplicate query index");
    }
    for (size_t i = 0; i < n_indices; ++i) {
      const size_t qi = query_indices[i];
      if (heap_mi[qi] == SIZE_MAX) fail(VC_ERR_INVALID_ARGUMENT, "buffered top-k collectors are not supported");
      if (heap_sz[qi] > k) fail(VC_ERR_INVALID_ARGUMENT, "heap_sz exceeds k");
      if (heap_sz[qi] == k && heap_mi[qi] >= k) fail(VC_ERR_INVALID_ARGUMENT, "heap_mi out of range");
    }
    const size_t m = n_indices;
    std::vector<float> hoff(m), hs(m * k), hmin(m);
    std::vector<uint64_t> hi(m * k), hsz(m), hmi(m);
    for (size_t i = 0; i < m; ++i) {
      const size_t qi = query_indices[i];
      hoff[i] = score_offsets ? score_offsets[qi] : 0.f;
      std::memcpy(&hs[i * k], heap_s + qi * k, k * 4);
      std::memcpy(&hi[i * k], heap_i + qi * k, k * 8);
      hsz[i] = heap_sz[qi];
      hmin[i] = heap_min[qi];
      hmi[i] = heap_mi[qi];
    }
    Scoped sc;
    hip_check(hipStreamWaitEvent(sc.s, p->ready, 0), "hipStreamWaitEvent");
    const vchip::CodeTables t = make_tables(p->dim, p->bits);
    const size_t G = p->G;
    const size_t pbytes = (n + 191) / 192 * G * 192;
    Carve c;
    const size_t op = c.take(pbytes), of = c.take(n * G), ot = c.take(vchip::tiled_bytes(n, t.groups16)),
                 oa = c.take(n * 4), om = c.take(id_map ? n * 8 : 0), opv = c.take(m * n * 4), ooff = c.take(m * 4),
                 ohs = c.take(m * k * 4), ohi = c.take(m * k * 8), ohsz = c.take(m * 8), ohmin = c.take(m * 4),
                 ohmi = c.take(m * 8), oidx = c.take(m * 8), orows = c.take(m * G * 32), ols = c.take(m * 4),
                 obt = c.take(m * 4);
    sc.a.reserve(c.off, sc.s);
    auto up = [&](size_t off, const void* src, size_t bytes) {
      hip_check(hipMemcpyAsync(sc.a.as<uint8_t>(off), src, bytes, hipMemcpyHostToDevice, sc.s), "copy");
    };
    upload_perm0_tiled(t, perm0, n, sc.a, op, of, ot, sc.s);
    up(oa, alpha, n * 4);
    if (id_map) up(om, id_map, n * 8);
    up(ooff, hoff.data(), m * 4);
    up(ohs, hs.data(), m * k * 4);
    up(ohi, hi.data(), m * k * 8);
    up(ohsz, hsz.data(), m * 8);
    up(ohmin, hmin.data(), m * 4);
    up(ohmi, hmi.data(), m * 8);
    const PreLuts pl = gather(p, query_indices, m, sc.a, oidx, orows, ols, obt, sc.s);
    hip_check(vchip::launch_bucket_scores(sc.a.as<uint8_t>(ot), sc.a.as<float>(oa), n, int(G), t.groups16, pl.luts,
                                          pl.scale, pl.bias_total, m, sc.a.as<float>(opv), sc.s),
              "bucket_scores");
    hip_check(vchip::launch_heap_replay(sc.a.as<float>(opv), n, int(k), id_map ? sc.a.as<uint64_t>(om) : nullptr,
                                        sc.a.as<float>(ooff), m, sc.a.as<float>(ohs), sc.a.as<uint64_t>(ohi),
                                        sc.a.as<uint64_t>(ohsz), sc.a.as<float>(ohmin), sc.a.as<uint64_t>(ohmi), sc.s),
              "heap_replay");
    auto down = [&](void* dst, size_t off, size_t bytes) {
      hip_check(hipMemcpyAsync(dst, sc.a.as<uint8_t>(off), bytes, hipMemcpyDeviceToHost, sc.s), "copy");
    };
    down(hs.data(), ohs, m * k * 4);
    down(hi.data(), ohi, m * k * 8);
    down(hsz.data(), ohsz, m * 8);
    down(hmin.data(), ohmin, m * 4);
    down(hmi.data(), ohmi, m * 8);
    hip_check(hipStreamSynchronize(sc.s), "sync");
    for (size_t i = 0; i < m; ++i) {
      const size_t qi = query_indices[i];
      std::memcpy(heap_s + qi * k, &hs[i * k], k * 4);
      std::memcpy(heap_i + qi * k, &hi[i * k], k * 8);
      heap_sz[qi] = size_t(hsz[i]);
      heap_min[qi] = hmin[i];
      heap_mi[qi] = size_t(hmi[i]);
    }
  });
}

// ---- repack free functions (src/pack/pack.cpp:352-401) -----------------------
namespace {
enum class Repack { Flat, Perm0, Seq };

void repack(const uint8_t* packed, size_t n, size_t dim, size_t bits, uint8_t* out, Repack kind) {
  check_dim_bits(dim, bits);
  if (n == 0) return;
  if (packed == nullptr || out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
  const size_t G = size_t(vchip::byte_groups(bits, dim));
  const size_t pk = n * bits * dim / 8, blocked = (n + 191) / 192 * G * 192;
  const size_t out_bytes = kind == Repack::Flat ? n * G : blocked;
  Scoped sc;
  Carve c;
  const size_t opk = c.take(pk), of = c.take(n * G), oo = c.take(kind == Repack::Flat ? 0 : blocked);
  sc.a.reserve(c.off, sc.s);
  hip_check(hipMemcpyAsync(sc.a.as<uint8_t>(opk), packed, pk, hipMemcpyHostToDevice, sc.s), "copy");
  hip_check(vchip::launch_packed_to_flat(sc.a.as<uint8_t>(opk), n, int(dim), int(bits), sc.a.as<uint8_t>(of), sc.s),
            "packed_to_flat");
  size_t src = of;
  if (kind == Repack::Perm0) {
    hip_check(vchip::launch_flat_to_perm0(sc.a.as<uint8_t>(of), n, int(G), sc.a.as<uint8_t>(oo), sc.s),
              "flat_to_perm0");
    src = oo;
  } else if (kind == Repack::Seq) {
    hip_check(vchip::launch_flat_to_seq(sc.a.as<uint8_t>(of), n, int(G), sc.a.as<uint8_t>(oo), sc.s), "flat_to_seq");
    src = oo;
  }
  hip_check(hipMemcpyAsync(out, sc.a.as<uint8_t>(src), out_bytes, hipMemcpyDeviceToHost, sc.s), "copy");
  hip_check(hipStreamSynchronize(sc.s), "sync");
}
}  // namespace

vc_status vc_codes_flat_from_packed(const uint8_t* packed, size_t n, size_t dim, size_t bits, uint8_t* flat_out) {
  return guarded([&] { repack(packed, n, dim, bits, flat_out, Repack::Flat); });
}

vc_status vc_repack_perm0(const uint8_t* packed, size_t n, size_t dim, size_t bits, uint8_t* perm0_out) {
  return guarded([&] { repack(packed, n, dim, bits, perm0_out, Repack::Perm0); });
}

vc_status vc_repack_seq(const uint8_t* packed, size_t n, size_t dim, size_t bits, uint8_t* seq_out) {
  return guarded([&] { repack(packed, n, dim, bits, seq_out, Repack::Seq); });
}

}  // extern "C"
