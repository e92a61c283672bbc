// C ABI of libvectorcache_hip.so (include/vectorcache_hip.h).
//
// Owns device memory and streams; the arithmetic lives in vc_kernels.hip and
// the host precompute in vc_host.cpp. There is no CPU fallback: every entry
// that computes runs on the GPU and reports VC_ERR_DEVICE when it cannot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "vc_host.hpp"
#include "vc_io.hpp"
#include "vc_internal.hpp"
#include "vectorcache_hip.h"
#include "vc_common.hpp"
#include "vc_index.hpp"

namespace vcapi {

thread_local std::string g_last_error;
thread_local const vchip::ScanPlan* g_plan_override = nullptr;
std::atomic<int> g_forced_tile{0};
std::atomic<int> g_debug_flags{0};

ScanProbe g_probe;

// Device copies of the rotation tables, per (device, dim).
struct DevRotation {
  uint32_t* inv_perm = nullptr;
  float* signs = nullptr;
  uint32_t* perm_sign = nullptr;  // inv_perm[j] | sign-of-destination << 31
  uint32_t* perm_gather = nullptr;  // perm[i] | sign[i] << 31 (by destination)
};
std::mutex g_rot_mu;
std::map<std::pair<int, size_t>, DevRotation> g_dev_rot;

DevRotation device_rotation(size_t dim) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard lock(g_rot_mu);
  auto& r = g_dev_rot[{dev, dim}];
  if (r.inv_perm == nullptr) {
    const auto& rt = vchip::rotation_tables(dim);
    hip_check(hipMalloc(&r.inv_perm, dim * 4), "hipMalloc");
    hip_check(hipMalloc(&r.signs, dim * 4), "hipMalloc");
    hip_check(hipMemcpy(r.inv_perm, rt.inv_perm.data(), dim * 4, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(r.signs, rt.signs.data(), dim * 4, hipMemcpyHostToDevice), "copy");
    std::vector<uint32_t> ps(dim);
    for (size_t j = 0; j < dim; ++j) {
      const uint32_t i = rt.inv_perm[j];
      ps[j] = i | (rt.signs[i] < 0.f ? 0x80000000u : 0u);
    }
    hip_check(hipMalloc(&r.perm_sign, dim * 4), "hipMalloc");
    hip_check(hipMemcpy(r.perm_sign, ps.data(), dim * 4, hipMemcpyHostToDevice), "copy");
    for (size_t i = 0; i < dim; ++i) ps[i] = rt.perm[i] | (rt.signs[i] < 0.f ? 0x80000000u : 0u);
    hip_check(hipMalloc(&r.perm_gather, dim * 4), "hipMalloc");
    hip_check(hipMemcpy(r.perm_gather, ps.data(), dim * 4, hipMemcpyHostToDevice), "copy");
  }
  return r;
}

// Device copies of the quantizer cell tables, per (device, bits, dim).
std::map<std::tuple<int, size_t, size_t>, void*> g_dev_cells;

const void* device_cells(size_t bits, size_t dim) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  const auto& ct = vchip::cell_table(bits, dim);
  if (!ct.ok) return nullptr;
  std::lock_guard lock(g_rot_mu);
  void*& p = g_dev_cells[{dev, bits, dim}];
  if (p == nullptr) {
    static_assert(sizeof(vchip::CellEntry) == 16, "cell entry layout");
    hip_check(hipMalloc(&p, sizeof(ct.cells)), "hipMalloc");
    hip_check(hipMemcpy(p, ct.cells, sizeof(ct.cells), hipMemcpyHostToDevice), "copy");
  }
  return p;
}

vchip::CodeTables make_tables(size_t dim, size_t bits) {
  vchip::CodeTables t{};
  const auto& rt = vchip::rotation_tables(dim);
  const auto& cb = vchip::codebook(bits, dim);
  std::fill(std::begin(t.boundaries), std::end(t.boundaries), 0.f);
  std::fill(std::begin(t.centroids), std::end(t.centroids), 0.f);
  std::copy(cb.boundaries.begin(), cb.boundaries.end(), t.boundaries);
  std::copy(cb.centroids.begin(), cb.centroids.end(), t.centroids);
  t.dim = int(dim);
  t.bits = int(bits);
  t.block = int(rt.block);
  t.groups = int(vchip::byte_groups(bits, dim));
  t.groups16 = (t.groups + 15) / 16;
  t.inv_sqrt_block = rt.inv_sqrt_block;
  const DevRotation dr = device_rotation(dim);
  t.inv_perm = dr.inv_perm;
  t.signs = dr.signs;
  t.perm_sign = dr.perm_sign;
  t.perm_gather = dr.perm_gather;
  const auto& ct = vchip::cell_table(bits, dim);
  t.cells = device_cells(bits, dim);
  t.use_cells = t.cells != nullptr ? 1 : 0;
  t.cell_inv_w = ct.inv_w;
  t.cell_off = ct.off;
  t.debug_flags = g_debug_flags.load();
  return t;
}

// Calibration from device-resident sample rows into device shift/scale/inv.
void calibrate_device(const vchip::CodeTables& t, const float* sample_dev, size_t n, float* shift,
                      float* scale, float* inv_scale, DevBuf& scratch, hipStream_t s) {
  if (n < 2) fail(VC_ERR_INVALID_ARGUMENT, "calibration needs at least 2 rows");
  const size_t D = size_t(t.dim);
  Carve c;
  const size_t o_rot = c.take(n * D * 4);
  const size_t o_keys = c.take(n * D * 4);
  scratch.reserve(c.off, s);
  vchip::CodeTables rot_only = t;
  rot_only.shift = rot_only.scale = rot_only.inv_scale = nullptr;
  hip_check(vchip::launch_encode(rot_only, sample_dev, n, 0, nullptr, nullptr, scratch.as<float>(o_rot), nullptr, s),
            "encode(rotate)");
  const auto anchors = vchip::tq_anchors(size_t(t.bits), D, n);
  hip_check(vchip::launch_calibrate(scratch.as<float>(o_rot), n, t.dim, anchors.lo_idx, anchors.hi_idx,
                                    anchors.qc_lo, anchors.qc_span, scratch.as<uint32_t>(o_keys), shift, scale,
                                    inv_scale, s),
            "calibrate");
}

SearchLayout search_layout(const vchip::CodeTables& t, size_t n, size_t nq, size_t k) {
  SearchLayout L;
  L.k_eff = std::min(k, n);
  if (nq == 0 || L.k_eff == 0) return L;
  if (L.k_eff > size_t(vchip::kMaxTopK))
    fail(VC_ERR_INVALID_ARGUMENT, "k exceeds the GPU top-k capacity of 2048");
  if (n > 0xFFFFFFFFull) fail(VC_ERR_INVALID_ARGUMENT, "index exceeds 2^32 vectors");
  const int ke = int(L.k_eff);
  vchip::ScanPlan& plan = L.plan;
  plan = g_plan_override != nullptr
             ? *g_plan_override
             : vchip::plan_scan(t.groups16, n, nq, ke, g_forced_tile.load(), (g_debug_flags.load() & 2) != 0);
  if (plan.tile == 0) fail(VC_ERR_INVALID_ARGUMENT, "no scan configuration fits LDS for this k / tile");
  if ((g_debug_flags.load() & 8) != 0 && plan.pair) plan.vec_per_thread = 3 - plan.vec_per_thread;  // test hook
  if (const size_t fr = std::min<size_t>(unsigned(g_debug_flags.load()) >> 16, 512); fr != 0 && n <= (size_t(1) << 30)) {
    // test hook: flags >> 16 = vector ranges per query tile (at most 512)
    size_t len = (n + fr - 1) / fr;
    len = (len + 63) / 64 * 64;
    plan.range_len = int(len);
    plan.n_ranges = int((n + len - 1) / len);
  }
  const size_t G = size_t(t.groups);
  Carve c;
  L.o_luts = c.take(nq * G * 32);
  L.o_ls = c.take(nq * 4);
  L.o_lb = c.take(nq * 4);
  L.o_bc = c.take(nq * 4);
  L.o_bt = c.take(nq * 4);
  const size_t parts = nq * size_t(plan.n_ranges);
  L.o_ps = c.take(parts * ke * 4);
  L.o_pi = c.take(parts * ke * 4);
  L.o_pc = c.take(parts * 4);
  L.o_lt = c.take(vchip::scan_lut_bytes(plan, t.groups16, nq));
  L.bytes = c.off;
  return L;
}

SearchOut search_tiled(const vchip::CodeTables& t, const uint8_t* tiled, const float* alpha, size_t n,
                       const float* q_dev, size_t nq, size_t k, const uint64_t* id_map_dev, float* out_s,
                       uint64_t* out_i, DevBuf& scratch, hipStream_t s, uint8_t** luts_out, size_t* luts_off,
                       ScratchArena* arena, const PreLuts* pre) {
  const SearchLayout L = search_layout(t, n, nq, k);
  SearchOut r{L.k_eff};
  if (L.bytes == 0) return r;
  const int ke = int(L.k_eff);
  const vchip::ScanPlan& plan = L.plan;
  const size_t o_luts = L.o_luts, o_ls = L.o_ls, o_lb = L.o_lb, o_bc = L.o_bc, o_bt = L.o_bt;
  const size_t o_ps = L.o_ps, o_pi = L.o_pi, o_pc = L.o_pc, o_lt = L.o_lt;
  Carve c;
  c.off = L.bytes;
  ScratchArena own{&scratch};
  uint8_t* base = (arena != nullptr ? arena : &own)->take(c.off, s);
  auto at = [base](size_t off) { return base + off; };
  auto atf = [base](size_t off) { return reinterpret_cast<float*>(base + off); };
  const uint8_t* luts = at(o_luts);
  const float *ls = atf(o_ls), *bt = atf(o_bt);
  if (pre != nullptr) {
    luts = pre->luts;
    ls = pre->scale;
    bt = pre->bias_total;
  } else {
    hip_check(vchip::launch_prepare_queries(t, q_dev, nq, at(o_luts), atf(o_ls), atf(o_lb), atf(o_bc), atf(o_bt), s),
              "prepare_queries");
  }
  // the table sources first, so that the profiling events bracket the scan kernel alone
  hip_check(vchip::launch_scan(plan, tiled, alpha, n, t.groups, t.groups16, luts, at(o_lt), ls,
                               bt, nq, ke, atf(o_ps), reinterpret_cast<uint32_t*>(at(o_pi)),
                               reinterpret_cast<int*>(at(o_pc)), s, 1),
            "scan sources");
  std::unique_lock probe_lock(g_probe.mu);
  const bool probe = g_probe.on;
  if (!probe) probe_lock.unlock();
  if (probe) hip_check(hipEventRecord(g_probe.start, s), "hipEventRecord");
  hip_check(vchip::launch_scan(plan, tiled, alpha, n, t.groups, t.groups16, luts, at(o_lt), ls,
                               bt, nq, ke, atf(o_ps), reinterpret_cast<uint32_t*>(at(o_pi)),
                               reinterpret_cast<int*>(at(o_pc)), s, 2),
            "scan");
  if (probe) {
    hip_check(hipEventRecord(g_probe.stop, s), "hipEventRecord");
    g_probe.recorded = true;
    g_probe.tile = plan.tile;
    g_probe.ranges = plan.n_ranges;
    probe_lock.unlock();
  }
  hip_check(vchip::launch_merge_ranges(atf(o_ps), reinterpret_cast<uint32_t*>(at(o_pi)),
                                       reinterpret_cast<int*>(at(o_pc)), nq, plan.n_ranges, ke, ke, id_map_dev, out_s,
                                       out_i, s),
            "merge_ranges");
  if (luts_out != nullptr) {
    *luts_out = base;
    *luts_off = o_luts;
  }
  return r;
}

}  // namespace vcapi

using namespace vcapi;

// ---------------------------------------------------------------------------
extern "C" {

const char* vc_last_error(void) { return g_last_error.c_str(); }

const char* vc_version(void) { return "vectorcache-hip 0.1.0 (gfx950)"; }

vc_status vc_index_create(size_t dim, size_t bits, vc_index** out) {
  return guarded([&] {
    if (out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "out is null");
    check_dim_bits(dim, bits);
    auto* ix = new vc_index();
    try {
      ix->dim = dim;
      ix->bits = bits;
      hip_check(hipGetDevice(&ix->device), "hipGetDevice");
      hip_check(hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking), "hipStreamCreate");
      ix->t = make_tables(dim, bits);
      ix->h_shift.assign(dim, 0.f);
      ix->h_scale.assign(dim, 1.f);
    } catch (...) {
      if (ix->stream) (void)hipStreamDestroy(ix->stream);
      delete ix;
      throw;
    }
    *out = ix;
  });
}

vc_status vc_index_destroy(vc_index* ix) {
  return guarded([&] {
    if (ix == nullptr) return;
    {
      std::lock_guard lock(ix->mu);
      refuse_if_served(ix);
      refuse_if_in_set(ix);
    }
    (void)hipStreamSynchronize(ix->stream);
    ix->calib.release();
    ix->tiled.release();
    ix->alpha.release();
    ix->pending.release();
    ix->scratch.release();
    ix->io.release();
    (void)hipStreamDestroy(ix->stream);
    delete ix;
  });
}

static void index_calibrate(vc_index* ix, const float* sample_dev, size_t n) {
  refuse_if_served(ix);
  refuse_if_in_set(ix);
  if (ix->n_encoded > 0) fail(VC_ERR_STATE, "calibrate() must precede prepare() of any vector");
  const size_t D = ix->dim;
  ix->calib.reserve(3 * D * 4, ix->stream);
  float* sh = ix->calib.as<float>(0);
  float* sc = ix->calib.as<float>(D * 4);
  float* is = ix->calib.as<float>(2 * D * 4);
  calibrate_device(ix->t, sample_dev, n, sh, sc, is, ix->scratch, ix->stream);
  hip_check(hipMemcpyAsync(ix->h_shift.data(), sh, D * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
  hip_check(hipMemcpyAsync(ix->h_scale.data(), sc, D * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
  hip_check(hipStreamSynchronize(ix->stream), "sync");
  ix->t.shift = sh;
  ix->t.scale = sc;
  ix->t.inv_scale = is;
  ix->calibrated = true;
  ix->prepared = false;
}

vc_status vc_index_calibrate(vc_index* ix, const float* sample, size_t n) {
  return guarded([&] {
    if (ix == nullptr || (sample == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (n < 2) fail(VC_ERR_INVALID_ARGUMENT, "calibration needs at least 2 rows");
    ix->io.reserve(n * ix->dim * 4, ix->stream);
    hip_check(hipMemcpyAsync(ix->io.p, sample, n * ix->dim * 4, hipMemcpyHostToDevice, ix->stream), "copy");
    index_calibrate(ix, ix->io.as<float>(), n);
  });
}

vc_status vc_index_calibrate_device(vc_index* ix, const float* sample_dev, size_t n) {
  return guarded([&] {
    if (ix == nullptr || (sample_dev == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    index_calibrate(ix, sample_dev, n);
  });
}

// Stages n rows for the next prepare(). Device rows are read in order after
// the work enqueued so far on `producer`, and later work on `producer` is
// ordered after the copy (the caller may reuse the rows' memory at once).
static void index_add(vc_index* ix, const float* v, size_t n, hipMemcpyKind kind, hipStream_t producer = nullptr) {
  refuse_if_served(ix);
  refuse_if_in_set(ix);
  if (n == 0) return;
  const size_t row = ix->dim * 4;
  ix->pending.reserve((ix->n_pending + n) * row, ix->stream, /*keep=*/true);
  if (kind == hipMemcpyDeviceToDevice) stream_after(ix->stream, producer);
  hip_check(hipMemcpyAsync(ix->pending.as<char>(ix->n_pending * row), v, n * row, kind, ix->stream), "copy");
  if (kind == hipMemcpyHostToDevice) hip_check(hipStreamSynchronize(ix->stream), "sync");
  else stream_after(producer, ix->stream);
  ix->n_pending += n;
  ix->prepared = false;
}

vc_status vc_index_add(vc_index* ix, const float* vectors, size_t n) {
  return guarded([&] {
    if (ix == nullptr || (vectors == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    index_add(ix, vectors, n, hipMemcpyHostToDevice);
  });
}

vc_status vc_index_add_device(vc_index* ix, const float* vectors_dev, size_t n, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (vectors_dev == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    index_add(ix, vectors_dev, n, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream));  // NULL: null stream
  });
}

vc_status vc_index_prepare(vc_index* ix) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    refuse_if_served(ix);
    if (ix->n_pending > 0) refuse_if_in_set(ix);
    if (ix->n_pending > 0) {
      const size_t total = ix->n_encoded + ix->n_pending;
      if (total > 0xFFFFFFFFull) fail(VC_ERR_INVALID_ARGUMENT, "index exceeds 2^32 vectors");
      index_reserve(ix, total, ix->stream);
      ix->t.debug_flags = g_debug_flags.load();
      hip_check(vchip::launch_encode(ix->t, ix->pending.as<float>(), ix->n_pending, ix->n_encoded,
                                     ix->tiled.as<uint8_t>(), nullptr, nullptr, ix->alpha.as<float>(), ix->stream),
                "encode");
      hip_check(hipStreamSynchronize(ix->stream), "sync");
      ix->n_encoded = total;
      ix->n_pending = 0;
    }
    ix->prepared = true;
  });
}

static size_t index_search_dev(vc_index* ix, const float* q_dev, size_t nq, size_t k, float* s_dev,
                               uint64_t* i_dev, hipStream_t s, size_t n_prefix = SIZE_MAX) {
  refuse_if_served(ix);
  if (!ix->prepared) fail(VC_ERR_STATE, "search() before prepare()");
  const size_t n = std::min(n_prefix, ix->n_encoded);
  const auto r = search_tiled(ix->t, ix->tiled.as<uint8_t>(), ix->alpha.as<float>(), n, q_dev, nq, k, nullptr, s_dev,
                              i_dev, ix->scratch, s);
  return r.k_eff;
}

vc_status vc_index_search(vc_index* ix, const float* queries, size_t nq, size_t k, float* scores, uint64_t* ids,
                          size_t* k_out) {
  return guarded([&] {
    if (ix == nullptr || (queries == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (!ix->prepared) fail(VC_ERR_STATE, "search() before prepare()");
    const size_t ke = std::min(k, ix->n_encoded);
    if (k_out) *k_out = ke;
    if (nq == 0 || ke == 0) return;
    if (scores == nullptr || ids == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null output");
    Carve c;
    const size_t oq = c.take(nq * ix->dim * 4), os = c.take(nq * ke * 4), oi = c.take(nq * ke * 8);
    ix->io.reserve(c.off, ix->stream);
    hip_check(hipMemcpyAsync(ix->io.as<float>(oq), queries, nq * ix->dim * 4, hipMemcpyHostToDevice, ix->stream),
              "copy");
    index_search_dev(ix, ix->io.as<float>(oq), nq, k, ix->io.as<float>(os), ix->io.as<uint64_t>(oi), ix->stream);
    hip_check(hipMemcpyAsync(scores, ix->io.as<float>(os), nq * ke * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipMemcpyAsync(ids, ix->io.as<uint64_t>(oi), nq * ke * 8, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_index_search_device(vc_index* ix, const float* queries_dev, size_t nq, size_t k, float* scores_dev,
                                 uint64_t* ids_dev, size_t* k_out, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (queries_dev == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
    const size_t ke = index_search_dev(ix, queries_dev, nq, k, scores_dev, ids_dev, s);
    if (k_out) *k_out = ke;
  });
}

vc_status vc_index_search_prefix_device(vc_index* ix, const float* queries_dev, size_t nq, size_t k, size_t n_prefix,
                                        float* scores_dev, uint64_t* ids_dev, size_t* k_out, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (queries_dev == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (n_prefix > ix->n_encoded) fail(VC_ERR_OUT_OF_RANGE, "n_prefix exceeds the encoded vectors");
    const size_t ke = index_search_dev(ix, queries_dev, nq, k, scores_dev, ids_dev, static_cast<hipStream_t>(stream),
                                       n_prefix);
    if (k_out) *k_out = ke;
  });
}

}  // extern "C"

// ---- search without allocation ----------------------------------------------
// Scratch of vc_index_search_device_noalloc: one array per GPU, held until the
// process ends. Host code never passes `scratch` itself (its host-side name is
// not a device address); it passes the device address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device on the entry point's first call there.
static constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_noalloc_mu;
std::map<int, uint8_t*> g_noalloc_ptr;

uint8_t* noalloc_scratch(int device) {
  std::lock_guard lock(g_noalloc_mu);
  auto it = g_noalloc_ptr.find(device);
  if (it != g_noalloc_ptr.end()) return it->second;
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (p == nullptr) fail(VC_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
  g_noalloc_ptr[device] = static_cast<uint8_t*>(p);
  return static_cast<uint8_t*>(p);
}
}  // namespace

extern "C" {

size_t vc_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

vc_status vc_index_search_device_noalloc(vc_index* ix, const float* queries_dev, size_t nq, size_t k,
                                         float* scores_dev, uint64_t* ids_dev, size_t* k_out, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (queries_dev == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    refuse_if_served(ix);
    if (!ix->prepared) fail(VC_ERR_STATE, "search() before prepare()");
    int dev = -1;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != ix->device) fail(VC_ERR_INVALID_ARGUMENT, "the index's device must be current");
    ScratchArena arena;
    arena.fixed = noalloc_scratch(dev);
    arena.fixed_cap = kNoAllocScratchBytes;
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
    const auto r = search_tiled(ix->t, ix->tiled.as<uint8_t>(), ix->alpha.as<float>(), ix->n_encoded, queries_dev,
                                nq, k, nullptr, scores_dev, ids_dev, ix->scratch, s, nullptr, nullptr, &arena);
    if (k_out) *k_out = r.k_eff;
  });
}

size_t vc_index_dim(const vc_index* ix) { return ix ? ix->dim : 0; }
size_t vc_index_size(const vc_index* ix) { return ix ? ix->n_encoded + ix->n_pending : 0; }
size_t vc_index_bits(const vc_index* ix) { return ix ? ix->bits : 0; }
int vc_index_is_prepared(const vc_index* ix) { return ix && ix->prepared ? 1 : 0; }
int vc_index_is_calibrated(const vc_index* ix) { return ix && ix->calibrated ? 1 : 0; }
size_t vc_index_codes_bytes(const vc_index* ix) {
  return ix ? ix->n_encoded * size_t(ix->t.groups) : 0;
}
size_t vc_index_index_bytes(const vc_index* ix) {
  return ix ? vc_index_codes_bytes(ix) + ix->n_encoded * 4 : 0;
}

// ---- exports --------------------------------------------------------------
static void export_flat_dev(vc_index* ix, uint8_t* flat_dev) {
  hip_check(vchip::launch_tiled_to_flat(ix->tiled.as<uint8_t>(), ix->n_encoded, ix->t.groups, ix->t.groups16,
                                        flat_dev, ix->stream),
            "tiled_to_flat");
}

static void require_prepared(vc_index* ix) {
  refuse_if_served(ix);
  if (ix->n_pending > 0 || !ix->prepared) fail(VC_ERR_STATE, "export before prepare()");
}

vc_status vc_index_export_flat(vc_index* ix, uint8_t* flat_host) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    require_prepared(ix);
    const size_t bytes = ix->n_encoded * size_t(ix->t.groups);
    if (bytes == 0) return;
    ix->io.reserve(bytes, ix->stream);
    export_flat_dev(ix, ix->io.as<uint8_t>());
    hip_check(hipMemcpyAsync(flat_host, ix->io.p, bytes, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_index_export_perm0(vc_index* ix, uint8_t* perm0_host) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    require_prepared(ix);
    const size_t n = ix->n_encoded, G = size_t(ix->t.groups);
    if (n == 0) return;
    const size_t pbytes = (n + 191) / 192 * G * 192;
    Carve c;
    const size_t of = c.take(n * G), op = c.take(pbytes);
    ix->io.reserve(c.off, ix->stream);
    export_flat_dev(ix, ix->io.as<uint8_t>(of));
    hip_check(vchip::launch_flat_to_perm0(ix->io.as<uint8_t>(of), n, int(G), ix->io.as<uint8_t>(op), ix->stream),
              "flat_to_perm0");
    hip_check(hipMemcpyAsync(perm0_host, ix->io.as<uint8_t>(op), pbytes, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_index_export_seq(vc_index* ix, uint8_t* seq_host) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    require_prepared(ix);
    const size_t n = ix->n_encoded, G = size_t(ix->t.groups);
    if (n == 0) return;
    const size_t sbytes = (n + 191) / 192 * G * 192;
    Carve c;
    const size_t of = c.take(n * G), os = c.take(sbytes);
    ix->io.reserve(c.off, ix->stream);
    export_flat_dev(ix, ix->io.as<uint8_t>(of));
    hip_check(vchip::launch_flat_to_seq(ix->io.as<uint8_t>(of), n, int(G), ix->io.as<uint8_t>(os), ix->stream),
              "flat_to_seq");
    hip_check(hipMemcpyAsync(seq_host, ix->io.as<uint8_t>(os), sbytes, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_index_export_gpu32(vc_index* ix, uint32_t* codes_host) {
  return guarded([&] {
    if (ix == nullptr || codes_host == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    require_prepared(ix);
    const size_t n = ix->n_encoded, G = size_t(ix->t.groups);
    if (n == 0) return;
    const size_t words = vchip::gpu32_words(n, int(G));
    Carve c;
    const size_t of = c.take(n * G), ow = c.take(words * 4);
    ix->io.reserve(c.off, ix->stream);
    export_flat_dev(ix, ix->io.as<uint8_t>(of));
    hip_check(vchip::launch_flat_to_gpu32(ix->io.as<uint8_t>(of), n, int(G), ix->io.as<uint32_t>(ow), ix->stream),
              "flat_to_gpu32");
    hip_check(hipMemcpyAsync(codes_host, ix->io.as<uint32_t>(ow), words * 4, hipMemcpyDeviceToHost, ix->stream),
              "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_index_export_packed(vc_index* ix, uint8_t* packed_host) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    require_prepared(ix);
    const size_t n = ix->n_encoded, G = size_t(ix->t.groups), bpr = ix->bits * ix->dim / 8;
    if (n == 0) return;
    Carve c;
    const size_t of = c.take(n * G), op = c.take(n * bpr);
    ix->io.reserve(c.off, ix->stream);
    export_flat_dev(ix, ix->io.as<uint8_t>(of));
    hip_check(vchip::launch_flat_to_packed(ix->io.as<uint8_t>(of), n, int(ix->dim), int(ix->bits),
                                           ix->io.as<uint8_t>(op), ix->stream),
              "flat_to_packed");
    hip_check(hipMemcpyAsync(packed_host, ix->io.as<uint8_t>(op), n * bpr, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_index_export_alpha(vc_index* ix, float* alpha_host) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    require_prepared(ix);
    if (ix->n_encoded == 0) return;
    hip_check(hipMemcpyAsync(alpha_host, ix->alpha.p, ix->n_encoded * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_index_get_calibration(const vc_index* ix, float* shift_host, float* scale_host) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    if (shift_host) std::memcpy(shift_host, ix->h_shift.data(), ix->dim * 4);
    if (scale_host) std::memcpy(scale_host, ix->h_scale.data(), ix->dim * 4);
  });
}

vc_status vc_index_prepare_queries(vc_index* ix, const float* queries, size_t nq, uint8_t* luts, float* lut_scale,
                                   float* lut_bias, float* bias_corr) {
  return guarded([&] {
    if (ix == nullptr || (queries == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (nq == 0) return;
    const size_t D = ix->dim, G = size_t(ix->t.groups);
    Carve c;
    const size_t oq = c.take(nq * D * 4), ol = c.take(nq * G * 32), os = c.take(nq * 4), ob = c.take(nq * 4),
                 oc = c.take(nq * 4), ot = c.take(nq * 4);
    ix->io.reserve(c.off, ix->stream);
    hip_check(hipMemcpyAsync(ix->io.as<float>(oq), queries, nq * D * 4, hipMemcpyHostToDevice, ix->stream), "copy");
    hip_check(vchip::launch_prepare_queries(ix->t, ix->io.as<float>(oq), nq, ix->io.as<uint8_t>(ol),
                                            ix->io.as<float>(os), ix->io.as<float>(ob), ix->io.as<float>(oc),
                                            ix->io.as<float>(ot), ix->stream),
              "prepare_queries");
    hip_check(hipMemcpyAsync(luts, ix->io.as<uint8_t>(ol), nq * G * 32, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipMemcpyAsync(lut_scale, ix->io.as<float>(os), nq * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipMemcpyAsync(lut_bias, ix->io.as<float>(ob), nq * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipMemcpyAsync(bias_corr, ix->io.as<float>(oc), nq * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

// ---- stateless entry points ------------------------------------------------
vc_status vc_rotation_tables(size_t dim, uint32_t* perm, float* signs, float* inv_sqrt_block, size_t* block) {
  return guarded([&] {
    const auto& rt = vchip::rotation_tables(dim);
    if (perm) std::memcpy(perm, rt.perm.data(), dim * 4);
    if (signs) std::memcpy(signs, rt.signs.data(), dim * 4);
    if (inv_sqrt_block) *inv_sqrt_block = rt.inv_sqrt_block;
    if (block) *block = rt.block;
  });
}

vc_status vc_codebook(size_t bits, size_t dim, float* boundaries, float* centroids) {
  return guarded([&] {
    const auto& cb = vchip::codebook(bits, dim);
    if (boundaries) std::memcpy(boundaries, cb.boundaries.data(), cb.boundaries.size() * 4);
    if (centroids) std::memcpy(centroids, cb.centroids.data(), cb.centroids.size() * 4);
  });
}

vc_status vc_landmarks_make(size_t dim, size_t m, uint64_t seed, float* L) {
  return guarded([&] {
    if (L == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "L is null");
    const auto& v = vchip::ortho_landmarks(dim, m, seed);
    std::memcpy(L, v.data(), v.size() * 4);
  });
}

vc_status vc_landmarks_quantize(const float* L, size_t dim, size_t m, uint8_t* codes, float* scales) {
  return guarded([&] {
    if (L == nullptr || codes == nullptr || scales == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    const auto q = vchip::quantize_landmarks(L, dim, m);
    std::memcpy(codes, q.codes.data(), q.codes.size());
    std::memcpy(scales, q.scales.data(), m * 4);
  });
}

vc_status vc_encode(const float* vectors, size_t n, size_t dim, size_t bits, const float* shift, const float* scale,
                    uint8_t* packed_out, float* alpha_out) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    if (n == 0) return;
    if (vectors == nullptr || packed_out == nullptr || alpha_out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if ((shift == nullptr) != (scale == nullptr)) fail(VC_ERR_INVALID_ARGUMENT, "shift and scale go together");
    Scoped sc;
    vchip::CodeTables t = make_tables(dim, bits);
    const size_t bpr = bits * dim / 8;
    Carve c;
    const size_t ox = c.take(n * dim * 4), op = c.take(n * bpr), oa = c.take(n * 4), ocal = c.take(3 * dim * 4);
    sc.a.reserve(c.off, sc.s);
    hip_check(hipMemcpyAsync(sc.a.as<float>(ox), vectors, n * dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
    if (shift != nullptr) {
      std::vector<float> inv(dim);
      for (size_t d = 0; d < dim; ++d) inv[d] = 1.f / scale[d];
      hip_check(hipMemcpyAsync(sc.a.as<float>(ocal), shift, dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
      hip_check(hipMemcpyAsync(sc.a.as<float>(ocal + dim * 4), scale, dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
      hip_check(hipMemcpyAsync(sc.a.as<float>(ocal + 2 * dim * 4), inv.data(), dim * 4, hipMemcpyHostToDevice, sc.s),
                "copy");
      hip_check(hipStreamSynchronize(sc.s), "sync");
      t.shift = sc.a.as<float>(ocal);
      t.scale = sc.a.as<float>(ocal + dim * 4);
      t.inv_scale = sc.a.as<float>(ocal + 2 * dim * 4);
    }
    hip_check(vchip::launch_encode(t, sc.a.as<float>(ox), n, 0, nullptr, sc.a.as<uint8_t>(op), nullptr,
                                   sc.a.as<float>(oa), sc.s),
              "encode");
    hip_check(hipMemcpyAsync(packed_out, sc.a.as<uint8_t>(op), n * bpr, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipMemcpyAsync(alpha_out, sc.a.as<float>(oa), n * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_rotate(const float* vectors, size_t n, size_t dim, float* rotated_out) {
  return guarded([&] {
    check_dim_bits(dim, 4);
    if (n == 0) return;
    if (vectors == nullptr || rotated_out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    Scoped sc;
    vchip::CodeTables t = make_tables(dim, 4);
    Carve c;
    const size_t ox = c.take(n * dim * 4), orot = c.take(n * dim * 4);
    sc.a.reserve(c.off, sc.s);
    hip_check(hipMemcpyAsync(sc.a.as<float>(ox), vectors, n * dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(vchip::launch_encode(t, sc.a.as<float>(ox), n, 0, nullptr, nullptr, sc.a.as<float>(orot), nullptr, sc.s),
              "encode(rotate)");
    hip_check(hipMemcpyAsync(rotated_out, sc.a.as<float>(orot), n * dim * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_fit_calibration(const float* sample, size_t n, size_t dim, size_t bits, float* shift_out,
                             float* scale_out) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    if (n < 2) fail(VC_ERR_INVALID_ARGUMENT, "calibration needs at least 2 rows");
    if (sample == nullptr || shift_out == nullptr || scale_out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    Scoped sc;
    vchip::CodeTables t = make_tables(dim, bits);
    Carve c;
    const size_t ox = c.take(n * dim * 4), ocal = c.take(3 * dim * 4);
    sc.a.reserve(c.off, sc.s);
    hip_check(hipMemcpyAsync(sc.a.as<float>(ox), sample, n * dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
    float* sh = sc.a.as<float>(ocal);
    calibrate_device(t, sc.a.as<float>(ox), n, sh, sh + dim, sh + 2 * dim, sc.scratch, sc.s);
    hip_check(hipMemcpyAsync(shift_out, sh, dim * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipMemcpyAsync(scale_out, sh + dim, dim * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_search_perm0(const float* queries, size_t nq, size_t dim, size_t bits, size_t k, const uint8_t* perm0,
                          const float* alpha, size_t n, const uint64_t* id_map, const float* shift, const float* scale,
                          float* scores, uint64_t* ids, size_t* k_out) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    const size_t ke = std::min(k, n);
    if (k_out) *k_out = ke;
    if (nq == 0 || ke == 0) return;
    if (queries == nullptr || perm0 == nullptr || alpha == nullptr || scores == nullptr || ids == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if ((shift == nullptr) != (scale == nullptr)) fail(VC_ERR_INVALID_ARGUMENT, "shift and scale go together");
    Scoped sc;
    vchip::CodeTables t = make_tables(dim, bits);
    const size_t G = size_t(t.groups);
    const size_t pbytes = (n + 191) / 192 * G * 192;
    Carve c;
    const size_t oq = c.take(nq * dim * 4), op = c.take(pbytes), of = c.take(n * G),
                 ot = c.take(vchip::tiled_bytes(n, t.groups16)), oa = c.take(n * 4), om = c.take(id_map ? n * 8 : 0),
                 os = c.take(nq * ke * 4), oi = c.take(nq * ke * 8), ocal = c.take(3 * dim * 4);
    sc.a.reserve(c.off, sc.s);
    hip_check(hipMemcpyAsync(sc.a.as<float>(oq), queries, nq * dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<uint8_t>(op), perm0, pbytes, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<float>(oa), alpha, n * 4, hipMemcpyHostToDevice, sc.s), "copy");
    if (id_map) hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(om), id_map, n * 8, hipMemcpyHostToDevice, sc.s), "copy");
    if (shift != nullptr) {
      hip_check(hipMemcpyAsync(sc.a.as<float>(ocal), shift, dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
      hip_check(hipMemcpyAsync(sc.a.as<float>(ocal + dim * 4), scale, dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
      t.shift = sc.a.as<float>(ocal);
      t.scale = sc.a.as<float>(ocal + dim * 4);
      t.inv_scale = nullptr;  // not used by query preparation
    }
    hip_check(vchip::launch_perm0_to_flat(sc.a.as<uint8_t>(op), n, int(G), sc.a.as<uint8_t>(of), sc.s), "perm0_to_flat");
    hip_check(vchip::launch_flat_to_tiled(sc.a.as<uint8_t>(of), n, int(G), t.groups16, sc.a.as<uint8_t>(ot), sc.s),
              "flat_to_tiled");
    search_tiled(t, sc.a.as<uint8_t>(ot), sc.a.as<float>(oa), n, sc.a.as<float>(oq), nq, k,
                 id_map ? sc.a.as<uint64_t>(om) : nullptr, sc.a.as<float>(os), sc.a.as<uint64_t>(oi), sc.scratch, sc.s);
    hip_check(hipMemcpyAsync(scores, sc.a.as<float>(os), nq * ke * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipMemcpyAsync(ids, sc.a.as<uint64_t>(oi), nq * ke * 8, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_score_perm0(const float* queries, size_t nq, size_t dim, size_t bits, size_t k, const uint8_t* perm0,
                         const float* alpha, size_t n, const uint64_t* id_map, const float* shift,
                         const float* scale, const size_t* query_indices, size_t n_indices,
                         const float* residual_norms, const float* residual_scales, float* scores, uint64_t* ids,
                         size_t* k_out) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    if (query_indices != nullptr)
      for (size_t i = 0; i < n_indices; ++i)
        if (query_indices[i] >= nq) fail(VC_ERR_OUT_OF_RANGE, "score_prepared: query_indices out of range");
    const size_t nq_out = query_indices != nullptr ? n_indices : nq;
    if (k_out) *k_out = 0;
    if (n == 0 || nq_out == 0) return;
    const size_t ke = std::min(k, n);
    if (k_out) *k_out = ke;
    if (ke == 0) return;
    if (queries == nullptr || perm0 == nullptr || alpha == nullptr || scores == nullptr || ids == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if ((shift == nullptr) != (scale == nullptr)) fail(VC_ERR_INVALID_ARGUMENT, "shift and scale go together");
    const bool resid = residual_norms != nullptr && residual_scales != nullptr;
    // the listed queries (repeats allowed) and their residual scales, in order
    std::vector<float> qsel, rs_sel;
    const float* qh = queries;
    if (query_indices != nullptr) {
      qsel.resize(nq_out * dim);
      for (size_t l = 0; l < nq_out; ++l)
        std::memcpy(qsel.data() + l * dim, queries + query_indices[l] * dim, dim * 4);
      qh = qsel.data();
    }
    if (resid) {
      rs_sel.resize(nq_out);
      for (size_t l = 0; l < nq_out; ++l) rs_sel[l] = residual_scales[query_indices ? query_indices[l] : l];
    }
    Scoped sc;
    vchip::CodeTables t = make_tables(dim, bits);
    const size_t G = size_t(t.groups);
    const size_t pbytes = (n + 191) / 192 * G * 192;
    Carve c;
    const size_t oq = c.take(nq_out * dim * 4), op = c.take(pbytes), of = c.take(n * G),
                 ot = c.take(vchip::tiled_bytes(n, t.groups16)), oa = c.take(n * 4), om = c.take(id_map ? n * 8 : 0),
                 os = c.take(nq_out * ke * 4), oi = c.take(nq_out * ke * 8), ocal = c.take(3 * dim * 4),
                 orn = c.take(resid ? n * 4 : 0), ors = c.take(resid ? nq_out * 4 : 0);
    sc.a.reserve(c.off, sc.s);
    hip_check(hipMemcpyAsync(sc.a.as<float>(oq), qh, nq_out * dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<uint8_t>(op), perm0, pbytes, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<float>(oa), alpha, n * 4, hipMemcpyHostToDevice, sc.s), "copy");
    if (id_map) hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(om), id_map, n * 8, hipMemcpyHostToDevice, sc.s), "copy");
    if (resid) {
      hip_check(hipMemcpyAsync(sc.a.as<float>(orn), residual_norms, n * 4, hipMemcpyHostToDevice, sc.s), "copy");
      hip_check(hipMemcpyAsync(sc.a.as<float>(ors), rs_sel.data(), nq_out * 4, hipMemcpyHostToDevice, sc.s), "copy");
    }
    if (shift != nullptr) {
      hip_check(hipMemcpyAsync(sc.a.as<float>(ocal), shift, dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
      hip_check(hipMemcpyAsync(sc.a.as<float>(ocal + dim * 4), scale, dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
      t.shift = sc.a.as<float>(ocal);
      t.scale = sc.a.as<float>(ocal + dim * 4);
      t.inv_scale = nullptr;  // not used by query preparation
    }
    hip_check(vchip::launch_perm0_to_flat(sc.a.as<uint8_t>(op), n, int(G), sc.a.as<uint8_t>(of), sc.s), "perm0_to_flat");
    hip_check(vchip::launch_flat_to_tiled(sc.a.as<uint8_t>(of), n, int(G), t.groups16, sc.a.as<uint8_t>(ot), sc.s),
              "flat_to_tiled");
    // with the residual correction the rows are ranked on positions first and
    // mapped afterwards (remap_ids follows the correction)
    const uint64_t* im_dev = id_map ? sc.a.as<uint64_t>(om) : nullptr;
    search_tiled(t, sc.a.as<uint8_t>(ot), sc.a.as<float>(oa), n, sc.a.as<float>(oq), nq_out, k,
                 resid ? nullptr : im_dev, sc.a.as<float>(os), sc.a.as<uint64_t>(oi), sc.scratch, sc.s);
    if (resid)
      hip_check(vchip::launch_residual_post(nq_out, int(ke), sc.a.as<float>(os), sc.a.as<uint64_t>(oi),
                                            sc.a.as<float>(orn), sc.a.as<float>(ors), n, im_dev, sc.s),
                "residual_post");
    hip_check(hipMemcpyAsync(scores, sc.a.as<float>(os), nq_out * ke * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipMemcpyAsync(ids, sc.a.as<uint64_t>(oi), nq_out * ke * 8, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

size_t vc_gpu32_words(size_t n, size_t dim, size_t bits) {
  if (dim == 0 || dim % 8 != 0 || dim > 16384 || (bits != 2 && bits != 3 && bits != 4)) return 0;
  const size_t G = bits == 2 ? dim / 4 : dim / 2;
  return vchip::gpu32_words(n, int(G));
}

vc_status vc_gpu32_to_perm0(const uint32_t* codes, size_t n, size_t dim, size_t bits, uint8_t* perm0_out) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    if (n == 0) return;
    if (codes == nullptr || perm0_out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    const vchip::CodeTables t = make_tables(dim, bits);
    const size_t G = size_t(t.groups);
    const size_t words = vchip::gpu32_words(n, int(G)), pbytes = (n + 191) / 192 * G * 192;
    Scoped sc;
    Carve c;
    const size_t ow = c.take(words * 4), of = c.take(n * G), op = c.take(pbytes);
    sc.a.reserve(c.off, sc.s);
    hip_check(hipMemcpyAsync(sc.a.as<uint32_t>(ow), codes, words * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(vchip::launch_gpu32_to_flat(sc.a.as<uint32_t>(ow), n, int(G), sc.a.as<uint8_t>(of), sc.s),
              "gpu32_to_flat");
    hip_check(vchip::launch_flat_to_perm0(sc.a.as<uint8_t>(of), n, int(G), sc.a.as<uint8_t>(op), sc.s),
              "flat_to_perm0");
    hip_check(hipMemcpyAsync(perm0_out, sc.a.as<uint8_t>(op), pbytes, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_score_perm0_into(const float* queries, size_t nq, size_t dim, size_t bits, size_t k,
                              const uint8_t* perm0, const float* alpha, size_t n, const uint64_t* id_map,
                              const float* shift, const float* scale, const size_t* query_indices,
                              size_t n_indices, float* heap_s, uint64_t* heap_i, size_t* heap_sz,
                              float* heap_min, size_t* heap_mi, const float* score_offsets) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    if (n_indices == 0 || query_indices == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "score_prepared_into requires non-empty query_indices");
    for (size_t i = 0; i < n_indices; ++i)
      if (query_indices[i] >= nq) fail(VC_ERR_OUT_OF_RANGE, "score_prepared_into: query_indices out of range");
    if (n == 0 || k == 0) return;
    if (queries == nullptr || perm0 == nullptr || alpha == nullptr || heap_s == nullptr || heap_i == nullptr ||
        heap_sz == nullptr || heap_min == nullptr || heap_mi == nullptr)
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if ((shift == nullptr) != (scale == nullptr)) fail(VC_ERR_INVALID_ARGUMENT, "shift and scale go together");
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
    // the listed queries, their offsets and heaps, gathered in list order
    std::vector<float> hq(m * dim), hoff(m), hs(m * k), hmin(m);
    std::vector<uint64_t> hi(m * k), hsz(m), hmi(m);
    for (size_t i = 0; i < m; ++i) {
      const size_t qi = query_indices[i];
      std::memcpy(&hq[i * dim], queries + qi * dim, dim * 4);
      hoff[i] = score_offsets ? score_offsets[qi] : 0.f;
      std::memcpy(&hs[i * k], heap_s + qi * k, k * 4);
      std::memcpy(&hi[i * k], heap_i + qi * k, k * 8);
      hsz[i] = heap_sz[qi];
      hmin[i] = heap_min[qi];
      hmi[i] = heap_mi[qi];
    }
    Scoped sc;
    vchip::CodeTables t = make_tables(dim, bits);
    const size_t G = size_t(t.groups);
    const size_t pbytes = (n + 191) / 192 * G * 192;
    Carve c;
    const size_t oq = c.take(m * dim * 4), op = c.take(pbytes), of = c.take(n * G),
                 ot = c.take(vchip::tiled_bytes(n, t.groups16)), oa = c.take(n * 4), om = c.take(id_map ? n * 8 : 0),
                 ocal = c.take(3 * dim * 4), oluts = c.take(m * G * 32), ols = c.take(m * 4), olb = c.take(m * 4),
                 obc = c.take(m * 4), obt = c.take(m * 4), opv = c.take(m * n * 4), ooff = c.take(m * 4),
                 ohs = c.take(m * k * 4), ohi = c.take(m * k * 8), ohsz = c.take(m * 8), ohmin = c.take(m * 4),
                 ohmi = c.take(m * 8);
    sc.a.reserve(c.off, sc.s);
    auto up = [&](size_t off, const void* src, size_t bytes) {
      hip_check(hipMemcpyAsync(sc.a.as<uint8_t>(off), src, bytes, hipMemcpyHostToDevice, sc.s), "copy");
    };
    up(oq, hq.data(), m * dim * 4);
    up(op, perm0, pbytes);
    up(oa, alpha, n * 4);
    if (id_map) up(om, id_map, n * 8);
    up(ooff, hoff.data(), m * 4);
    up(ohs, hs.data(), m * k * 4);
    up(ohi, hi.data(), m * k * 8);
    up(ohsz, hsz.data(), m * 8);
    up(ohmin, hmin.data(), m * 4);
    up(ohmi, hmi.data(), m * 8);
    if (shift != nullptr) {
      up(ocal, shift, dim * 4);
      up(ocal + dim * 4, scale, dim * 4);
      t.shift = sc.a.as<float>(ocal);
      t.scale = sc.a.as<float>(ocal + dim * 4);
      t.inv_scale = nullptr;  // not used by query preparation
    }
    hip_check(vchip::launch_perm0_to_flat(sc.a.as<uint8_t>(op), n, int(G), sc.a.as<uint8_t>(of), sc.s), "perm0_to_flat");
    hip_check(vchip::launch_flat_to_tiled(sc.a.as<uint8_t>(of), n, int(G), t.groups16, sc.a.as<uint8_t>(ot), sc.s),
              "flat_to_tiled");
    hip_check(vchip::launch_prepare_queries(t, sc.a.as<float>(oq), m, sc.a.as<uint8_t>(oluts), sc.a.as<float>(ols),
                                            sc.a.as<float>(olb), sc.a.as<float>(obc), sc.a.as<float>(obt), sc.s),
              "prepare_queries");
    hip_check(vchip::launch_bucket_scores(sc.a.as<uint8_t>(ot), sc.a.as<float>(oa), n, int(G), t.groups16,
                                          sc.a.as<uint8_t>(oluts), sc.a.as<float>(ols), sc.a.as<float>(obt), m,
                                          sc.a.as<float>(opv), sc.s),
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

vc_status vc_heaps_to_results(size_t nq, size_t k, const float* heap_s, const uint64_t* heap_i,
                              const size_t* heap_sz, float* scores, uint64_t* ids) {
  return guarded([&] {
    if (nq == 0 || k == 0) return;
    if (!heap_s || !heap_i || !heap_sz || !scores || !ids) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (k > size_t(vchip::kMaxHeapK)) fail(VC_ERR_INVALID_ARGUMENT, "k exceeds the GPU heap capacity of 4096");
    Scoped sc;
    Carve c;
    const size_t ohs = c.take(nq * k * 4), ohi = c.take(nq * k * 8), ohsz = c.take(nq * 8), oos = c.take(nq * k * 4),
                 ooi = c.take(nq * k * 8);
    sc.a.reserve(c.off, sc.s);
    std::vector<uint64_t> hsz(heap_sz, heap_sz + nq);
    hip_check(hipMemcpyAsync(sc.a.as<float>(ohs), heap_s, nq * k * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(ohi), heap_i, nq * k * 8, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(ohsz), hsz.data(), nq * 8, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(vchip::launch_heaps_to_results(nq, int(k), sc.a.as<float>(ohs), sc.a.as<uint64_t>(ohi),
                                             sc.a.as<uint64_t>(ohsz), sc.a.as<float>(oos), sc.a.as<uint64_t>(ooi),
                                             sc.s),
              "heaps_to_results");
    hip_check(hipMemcpyAsync(scores, sc.a.as<float>(oos), nq * k * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipMemcpyAsync(ids, sc.a.as<uint64_t>(ooi), nq * k * 8, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_float_rescore(size_t nq, size_t k, float* scores, uint64_t* ids, const size_t* counts,
                           const float* queries, const float* corpus, size_t n_corpus, size_t dim) {
  return guarded([&] {
    if (nq == 0 || k == 0 || dim == 0) return;
    if (!scores || !ids || !queries || (!corpus && n_corpus != 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (k > size_t(vchip::kMaxHeapK)) fail(VC_ERR_INVALID_ARGUMENT, "k exceeds the GPU capacity of 4096");
    Scoped sc;
    Carve c;
    const size_t os = c.take(nq * k * 4), oi = c.take(nq * k * 8), oc = c.take(counts ? nq * 8 : 0),
                 oq = c.take(nq * dim * 4), ox = c.take(n_corpus * dim * 4);
    sc.a.reserve(c.off, sc.s);
    std::vector<uint64_t> cnt;
    if (counts) cnt.assign(counts, counts + nq);
    hip_check(hipMemcpyAsync(sc.a.as<float>(os), scores, nq * k * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(oi), ids, nq * k * 8, hipMemcpyHostToDevice, sc.s), "copy");
    if (counts) hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(oc), cnt.data(), nq * 8, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<float>(oq), queries, nq * dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
    if (n_corpus) hip_check(hipMemcpyAsync(sc.a.as<float>(ox), corpus, n_corpus * dim * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(vchip::launch_float_rescore(nq, int(k), sc.a.as<float>(os), sc.a.as<uint64_t>(oi),
                                          counts ? sc.a.as<uint64_t>(oc) : nullptr, sc.a.as<float>(oq),
                                          sc.a.as<float>(ox), n_corpus, int(dim), sc.s),
              "float_rescore");
    hip_check(hipMemcpyAsync(scores, sc.a.as<float>(os), nq * k * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipMemcpyAsync(ids, sc.a.as<uint64_t>(oi), nq * k * 8, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_float_rescore_device(size_t nq, size_t k, float* scores_dev, uint64_t* ids_dev,
                                  const uint64_t* counts_dev, const float* queries_dev, const float* corpus_dev,
                                  size_t n_corpus, size_t dim, void* stream) {
  return guarded([&] {
    if (nq == 0 || k == 0 || dim == 0) return;
    if (!scores_dev || !ids_dev || !queries_dev || (!corpus_dev && n_corpus != 0))
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (k > size_t(vchip::kMaxHeapK)) fail(VC_ERR_INVALID_ARGUMENT, "k exceeds the GPU capacity of 4096");
    hip_check(vchip::launch_float_rescore(nq, int(k), scores_dev, ids_dev, counts_dev, queries_dev, corpus_dev,
                                          n_corpus, int(dim), static_cast<hipStream_t>(stream)),
              "float_rescore");
  });
}

vc_status vc_merge_results(size_t nq, size_t k_dst, float* dst_scores, uint64_t* dst_ids, size_t k_src,
                           const float* src_scores, const uint64_t* src_ids) {
  return guarded([&] {
    if (nq == 0 || k_src == 0 || k_dst == 0) return;  // k_dst == 0: the caller takes src as is
    if (!dst_scores || !dst_ids || !src_scores || !src_ids) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (k_dst + k_src > 8192) fail(VC_ERR_INVALID_ARGUMENT, "k_dst + k_src exceeds 8192");
    Scoped sc;
    Carve c;
    const size_t ods = c.take(nq * k_dst * 4), odi = c.take(nq * k_dst * 8), oss = c.take(nq * k_src * 4),
                 osi = c.take(nq * k_src * 8);
    sc.a.reserve(c.off, sc.s);
    hip_check(hipMemcpyAsync(sc.a.as<float>(ods), dst_scores, nq * k_dst * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(odi), dst_ids, nq * k_dst * 8, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<float>(oss), src_scores, nq * k_src * 4, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(hipMemcpyAsync(sc.a.as<uint64_t>(osi), src_ids, nq * k_src * 8, hipMemcpyHostToDevice, sc.s), "copy");
    hip_check(vchip::launch_merge_results(nq, int(k_dst), sc.a.as<float>(ods), sc.a.as<uint64_t>(odi), int(k_src),
                                          sc.a.as<float>(oss), sc.a.as<uint64_t>(osi), sc.s),
              "merge_results");
    hip_check(hipMemcpyAsync(dst_scores, sc.a.as<float>(ods), nq * k_dst * 4, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipMemcpyAsync(dst_ids, sc.a.as<uint64_t>(odi), nq * k_dst * 8, hipMemcpyDeviceToHost, sc.s), "copy");
    hip_check(hipStreamSynchronize(sc.s), "sync");
  });
}

vc_status vc_set_scan_tile(size_t tile) {
  return guarded([&] {
    if (tile != 0 && tile != 1 && tile != 2 && tile != 4 && tile != 8)
      fail(VC_ERR_INVALID_ARGUMENT, "tile must be 0, 1, 2, 4 or 8");
    g_forced_tile.store(int(tile));
  });
}

vc_status vc_set_debug(unsigned flags) {
  return guarded([&] { g_debug_flags.store(int(flags)); });
}

vc_status vc_set_profiling(int on) {
  return guarded([&] {
    std::lock_guard lock(g_probe.mu);
    if (on && g_probe.start == nullptr) {
      hip_check(hipEventCreate(&g_probe.start), "hipEventCreate");
      hip_check(hipEventCreate(&g_probe.stop), "hipEventCreate");
    }
    g_probe.on = on != 0;
    g_probe.recorded = false;
  });
}

vc_status vc_last_scan_stats(double* ms, size_t* tile, size_t* n_ranges) {
  return guarded([&] {
    std::lock_guard lock(g_probe.mu);
    if (!g_probe.recorded) fail(VC_ERR_STATE, "no profiled scan recorded");
    hip_check(hipEventSynchronize(g_probe.stop), "hipEventSynchronize");
    float t = 0.f;
    hip_check(hipEventElapsedTime(&t, g_probe.start, g_probe.stop), "hipEventElapsedTime");
    if (ms) *ms = double(t);
    if (tile) *tile = size_t(g_probe.tile);
    if (n_ranges) *n_ranges = size_t(g_probe.ranges);
  });
}

size_t vc_scan_tile_for(size_t dim, size_t bits, size_t nq) {
  try {
    const size_t G = vchip::byte_groups(bits, dim);
    const auto p = vchip::plan_scan(int((G + 15) / 16), 1u << 20, nq, 10, g_forced_tile.load(),
                                    (g_debug_flags.load() & 2) != 0);
    return size_t(p.tile);
  } catch (...) {
    return 0;
  }
}

}  // extern "C"

// ---- on-disk formats (host) -------------------------------------------------
struct vc_npy {
  std::unique_ptr<vcio::NpyMatrix> m;
};

template <typename F>
vc_status guarded_io(F&& f) {
  return guarded([&] {
    try {
      f();
    } catch (const vcio::IoError& e) {
      fail(VC_ERR_STATE, e.what());
    }
  });
}

static vcio::ExactTopkKey vcex_key(const vc_vcex_key* k) {
  if (k == nullptr || k->source_label == nullptr || k->split == nullptr || k->query_split == nullptr)
    fail(VC_ERR_INVALID_ARGUMENT, "null argument");
  vcio::ExactTopkKey key;
  key.source_label = k->source_label;
  key.split = k->split;
  key.query_split = k->query_split;
  key.index_n = k->index_n;
  key.query_n = k->query_n;
  key.k = k->k;
  key.seed = k->seed;
  return key;
}

static void put_string(const std::string& s, char* out, size_t cap, size_t* len) {
  if (len == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
  *len = s.size();
  if (out != nullptr && cap > s.size()) std::memcpy(out, s.c_str(), s.size() + 1);
}

vc_status vc_npy_open(const char* path, vc_npy** out, size_t* count, size_t* dim) {
  return guarded_io([&] {
    if (path == nullptr || out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    auto h = std::make_unique<vc_npy>();
    h->m = std::make_unique<vcio::NpyMatrix>(path);
    if (count) *count = h->m->count();
    if (dim) *dim = h->m->dim();
    *out = h.release();
  });
}

const float* vc_npy_data(const vc_npy* npy) { return npy ? npy->m->data() : nullptr; }

vc_status vc_npy_read_rows(const vc_npy* npy, size_t first, size_t n, float* out, size_t* n_read) {
  return guarded_io([&] {
    if (npy == nullptr || (out == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    const size_t got = npy->m->copy_rows(first, n, out);
    if (n_read) *n_read = got;
  });
}

vc_status vc_npy_close(vc_npy* npy) {
  delete npy;
  return VC_OK;
}

vc_status vc_npy_write_f32(const char* path, size_t rows, size_t cols, const float* data) {
  return guarded_io([&] {
    if (path == nullptr || (data == nullptr && rows * cols > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    vcio::write_npy_f32(path, rows, cols, data);
  });
}

vc_status vc_vcex_save(const char* path, const vc_vcex_key* key, const uint64_t* ids) {
  return guarded_io([&] {
    const vcio::ExactTopkKey k = vcex_key(key);
    if (path == nullptr || (ids == nullptr && k.query_n * k.k > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (!vcio::save_exact_topk(path, k, ids)) fail(VC_ERR_STATE, std::string("failed to write exact top-k cache ") + path);
  });
}

vc_status vc_vcex_load(const char* path, const vc_vcex_key* key, uint64_t* ids, int* loaded) {
  return guarded_io([&] {
    const vcio::ExactTopkKey k = vcex_key(key);
    if (path == nullptr || loaded == nullptr || (ids == nullptr && k.query_n * k.k > 0))
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<uint64_t> v;
    *loaded = vcio::load_exact_topk(path, k, v) ? 1 : 0;
    if (*loaded && !v.empty()) std::memcpy(ids, v.data(), v.size() * 8);
  });
}

vc_status vc_vcex_source_label(const char* npy_path, const char* dataset, char* out, size_t cap, size_t* len) {
  return guarded_io([&] {
    put_string(vcio::cache_source_label(npy_path ? npy_path : "", dataset ? dataset : ""), out, cap, len);
  });
}

vc_status vc_vcex_cache_path(const vc_vcex_key* key, char* out, size_t cap, size_t* len) {
  return guarded_io([&] { put_string(vcio::exact_topk_cache_path(vcex_key(key)), out, cap, len); });
}

// ---------------------------------------------------------------------------
// Landmark index (reference LandmarkIndex, src/index.cpp:120-289, on the AVX2
// path). HBM state: the landmark basis L (dim x m floats, column-major) and
// its 4-bit form (m columns of dim_pad / 2 bytes + m scales); landmark rows
// (n x scan_dim floats) from add() until prepare(); residual norms (n);
// after prepare() the scan codes (tiled, scan_dim / (8 / scan_bits) groups)
// and alpha (n). prepare() encodes every row and releases the rows, so rows
// added after a prepare() find the earlier ones zeroed, as in the reference
// (its landmark_rows_ is released and resized with zeros).
// ---------------------------------------------------------------------------
struct vc_lm_index {
  size_t dim = 0, scan_bits = 0, rerank_bits = 0, m = 0, scan_dim = 0, dim_pad = 0;
  uint64_t seed = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  vchip::CodeTables t_add{};   // (dim, rerank_bits): in-order norm + rotation, y-hat codebook
  vchip::CodeTables t_query{}; // (dim): query norm + rotation
  vchip::CodeTables t_scan{};  // (scan_dim, scan_bits): landmark rows, prerotated
  DevBuf basis;                // L | codes as words [dim_pad / 8][m] | scales
  size_t o_codes = 0, o_scales = 0;
  DevBuf rows;                 // n x scan_dim while live
  bool rows_live = true;
  DevBuf rn;                   // residual norms
  DevBuf tiled, alpha;
  size_t cap = 0;              // vector capacity of tiled / alpha
  size_t n = 0, n_encoded = 0;
  bool prepared = false;
  DevBuf work, qwork, scratch, io;
  // staged sign-plane search (the reference's AVX-512 VNNI host behaviour):
  // sign bytes of every encoded vector in the tiled layout, and 1.0 per vector
  // (the sign scan's scores carry no alpha)
  int planes = 0;
  bool has_sign = false;
  DevBuf sign_tiled, ones, stage;
  std::mutex mu;
};

namespace {

// rows per add() pass (rotated rows staged in HBM: 768 MiB at dim 1536); the
// y-hat kernel runs one thread per row, so a pass needs many rows to keep
// enough wavefronts in flight
constexpr size_t kLmAddChunk = 131072;

void lm_add(vc_lm_index* ix, const float* v, size_t n_new, bool host) {
  if (n_new == 0) return;
  const size_t D = ix->dim, S = ix->scan_dim, base = ix->n, total = ix->n + n_new;
  if (total > 0xFFFFFFFFull) fail(VC_ERR_INVALID_ARGUMENT, "index exceeds 2^32 vectors");
  hipStream_t s = ix->stream;
  if (!ix->rows_live) {
    ix->rows.reserve(total * S * 4, s);
    hip_check(hipMemsetAsync(ix->rows.p, 0, base * S * 4, s), "memset");
    ix->rows_live = true;
  } else {
    ix->rows.reserve(total * S * 4, s, /*keep=*/true);
  }
  ix->rn.reserve(total * 4, s, /*keep=*/true);
  const size_t chunk = std::min(n_new, kLmAddChunk);
  Carve c;
  const size_t ox = c.take(host ? chunk * D * 4 : 0), orot = c.take(chunk * D * 4), ovn = c.take(chunk * 4);
  ix->work.reserve(c.off, s);
  for (size_t off = 0; off < n_new; off += chunk) {
    const size_t cn = std::min(chunk, n_new - off);
    const float* x = v + off * D;
    if (host) {
      hip_check(hipMemcpyAsync(ix->work.as<float>(ox), x, cn * D * 4, hipMemcpyHostToDevice, s), "copy");
      x = ix->work.as<float>(ox);
    }
    float* rot = ix->work.as<float>(orot);
    float* vn = ix->work.as<float>(ovn);
    float* dst = ix->rows.as<float>((base + off) * S * 4);
    hip_check(vchip::launch_encode(ix->t_add, x, cn, 0, nullptr, nullptr, rot, nullptr, s), "encode(rotate)");
    hip_check(vchip::launch_lm_yhat(ix->t_add, rot, cn, vn, s), "lm_yhat");
    hip_check(vchip::launch_lm_project(rot, cn, int(D), ix->basis.as<float>(), int(ix->m), dst, int(S), s),
              "lm_project");
    hip_check(vchip::launch_lm_resid(dst, cn, int(ix->m), int(S), vn, ix->rn.as<float>((base + off) * 4), s),
              "lm_resid");
    if (host) hip_check(hipStreamSynchronize(s), "sync");  // the staging buffer is reused
  }
  hip_check(hipStreamSynchronize(s), "sync");
  ix->n = total;
  ix->prepared = false;
}

// Rotated query rows -> landmark rows (nq x scan_dim) and |q_perp| (nq), in qwork.
void lm_query_rows(vc_lm_index* ix, const float* q_dev, size_t nq, float** q_land, float** q_perp, hipStream_t s) {
  Carve c;
  const size_t orot = c.take(nq * ix->dim * 4), oland = c.take(nq * ix->scan_dim * 4), operp = c.take(nq * 4);
  ix->qwork.reserve(c.off, s);
  float* rot = ix->qwork.as<float>(orot);
  *q_land = ix->qwork.as<float>(oland);
  *q_perp = ix->qwork.as<float>(operp);
  const bool fused = (g_debug_flags.load() & 32) == 0 &&  // flags bit 5: separate rotation (test hook)
                     vchip::lm_query_rotates(ix->t_query, int(ix->dim), int(ix->dim_pad), int(ix->scan_dim));
  if (!fused)
    hip_check(vchip::launch_encode(ix->t_query, q_dev, nq, 0, nullptr, nullptr, rot, nullptr, s), "encode(rotate)");
  hip_check(vchip::launch_lm_query(fused ? &ix->t_query : nullptr, fused ? q_dev : rot, nq, int(ix->dim),
                                   int(ix->dim_pad), ix->basis.as<uint32_t>(ix->o_codes),
                                   ix->basis.as<float>(ix->o_scales), int(ix->m), *q_land, int(ix->scan_dim),
                                   *q_perp, s),
            "lm_query");
}

// Sign planes kept at prepare on a VNNI host (src/pack/pack.cpp:428-449): 4-bit
// scan codes and at least 32768 prepared vectors (VECTORCACHE_PLANES_MIN's
// default). The staged branch also needs scan_dim % 32 == 0: below that the
// reference's sign scan has no kernel for its 32-wide sign blocks and falls
// through to the 192-wide 4-bit kernel (src/search/search_vnni.cpp:1466-1469,
// 803-843), which reads past the sign codes, so no defined result exists; such
// an index keeps the exact scan, the AVX2 host's result.
constexpr size_t kPlanesMinVectors = 32768;

void lm_build_sign(vc_lm_index* ix) {
  ix->has_sign = false;
  const size_t n = ix->n_encoded, S = ix->scan_dim, Gs = S / 8;
  if (ix->planes == 0 || ix->scan_bits != 4 || n < kPlanesMinVectors || S % 32 != 0) {
    ix->sign_tiled.release();
    ix->ones.release();
    return;
  }
  const int g16s = int((Gs + 15) / 16);
  const size_t bytes = vchip::tiled_bytes(n, g16s);
  ix->sign_tiled.reserve(bytes, ix->stream);
  ix->ones.reserve(n * 4, ix->stream);
  hip_check(hipMemsetAsync(ix->sign_tiled.p, 0, bytes, ix->stream), "memset");
  hip_check(vchip::launch_lm_sign_codes(ix->tiled.as<uint8_t>(), n, ix->t_scan.groups16, int(Gs), g16s,
                                        ix->sign_tiled.as<uint8_t>(), ix->stream),
            "sign codes");
  hip_check(vchip::launch_fill_value(ix->ones.as<float>(), n, 1.f, ix->stream), "fill");
  hip_check(hipStreamSynchronize(ix->stream), "sync");
  ix->has_sign = true;
}

// The reference's shortlist lengths (src/pack/pack.cpp:451-463) for 4-bit scans.
size_t planes_shortlist_len4(size_t k) { return std::max<size_t>(256, k >= 64 ? k * 16 : k * 20); }
size_t planes_exact_len4(size_t k) { return std::min(planes_shortlist_len4(k), std::max<size_t>(256, 6 * k)); }

// The staged branch of search_landmark_rerank (src/search/search.cpp:
// 1379-1510): the E best sign-plane scores (exact, ties to the lower
// position), then their 4-bit permute-dot rescore with the residual term into
// a k-heap in shortlist order.
void lm_search_staged(vc_lm_index* ix, const float* q_land, const float* q_perp, size_t nq, size_t ke, float* s_dev,
                      uint64_t* i_dev, hipStream_t s) {
  const size_t S = ix->scan_dim, Gs = S / 8, n = ix->n_encoded, E = planes_exact_len4(ke);
  if (Gs > 256) fail(VC_ERR_INVALID_ARGUMENT, "staged sign search supports at most 2048 landmarks");
  if (E > size_t(vchip::kMaxTopK) || vchip::staged_lds_bytes(int(S), int(E), int(ke)) > size_t(160 * 1024))
    fail(VC_ERR_INVALID_ARGUMENT, "k too large for the staged sign search (shortlist above 2048)");
  vchip::CodeTables t_sign = ix->t_scan;
  t_sign.groups = int(Gs);
  t_sign.groups16 = int((Gs + 15) / 16);
  Carve c;
  const size_t ol = c.take(nq * Gs * 32), ols = c.take(nq * 4), obt = c.take(nq * 4), oss = c.take(nq * E * 4),
               osi = c.take(nq * E * 8);
  ix->stage.reserve(c.off, s);
  hip_check(vchip::launch_sign_luts(q_land, nq, int(S), ix->stage.as<uint8_t>(ol), ix->stage.as<float>(ols),
                                    ix->stage.as<float>(obt), s),
            "sign luts");
  const PreLuts pre{ix->stage.as<uint8_t>(ol), ix->stage.as<float>(ols), ix->stage.as<float>(obt)};
  search_tiled(t_sign, ix->sign_tiled.as<uint8_t>(), ix->ones.as<float>(), n, nullptr, nq, E, nullptr,
               ix->stage.as<float>(oss), ix->stage.as<uint64_t>(osi), ix->scratch, s, nullptr, nullptr, nullptr, &pre);
  hip_check(vchip::launch_lm_staged(q_land, nq, int(S), q_perp, ix->t_scan.centroids, ix->tiled.as<uint8_t>(),
                                    ix->t_scan.groups16, ix->alpha.as<float>(), ix->rn.as<float>(),
                                    ix->stage.as<uint64_t>(osi), int(E), int(ke), s_dev, i_dev, s),
            "staged rescore");
}

size_t lm_search_dev(vc_lm_index* ix, const float* q_dev, size_t nq, size_t k, const float* corpus_dev,
                     float* s_dev, uint64_t* i_dev, hipStream_t s) {
  if (!ix->prepared) fail(VC_ERR_STATE, "call prepare() before search");
  const size_t ke = std::min(k, ix->n_encoded);
  if (nq == 0 || ke == 0) return ke;
  if (ke > size_t(vchip::kMaxTopK)) fail(VC_ERR_INVALID_ARGUMENT, "k exceeds the GPU top-k capacity of 2048");
  float *q_land = nullptr, *q_perp = nullptr;
  lm_query_rows(ix, q_dev, nq, &q_land, &q_perp, s);
  if (ix->has_sign && planes_shortlist_len4(ke) < ix->n_encoded) {
    lm_search_staged(ix, q_land, q_perp, nq, ke, s_dev, i_dev, s);
  } else {
    search_tiled(ix->t_scan, ix->tiled.as<uint8_t>(), ix->alpha.as<float>(), ix->n_encoded, q_land, nq, ke, nullptr,
                 s_dev, i_dev, ix->scratch, s);
    hip_check(vchip::launch_residual_post(nq, int(ke), s_dev, i_dev, ix->rn.as<float>(), q_perp, ix->n_encoded,
                                          nullptr, s),
              "residual_post");
  }
  if (corpus_dev != nullptr)
    hip_check(vchip::launch_float_rescore(nq, int(ke), s_dev, i_dev, nullptr, q_dev, corpus_dev, ix->n, int(ix->dim), s),
              "float_rescore");
  return ke;
}

}  // namespace

extern "C" {

vc_status vc_lm_index_create(size_t dim, size_t scan_bits, size_t n_landmarks, uint64_t landmark_seed,
                             size_t rerank_bits, vc_lm_index** out) {
  return guarded([&] {
    if (out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "out is null");
    const size_t rb = rerank_bits == 0 ? scan_bits : rerank_bits;
    if (dim == 0 || dim % 8 != 0 || dim > vchip::kMaxDim)
      fail(VC_ERR_INVALID_ARGUMENT, "dim must be a multiple of 8 in [8, MAX_DIM]");
    if (scan_bits < 2 || scan_bits > 4) fail(VC_ERR_INVALID_ARGUMENT, "scan_bits must be 2, 3, or 4");
    if (rb < 2 || rb > 4) fail(VC_ERR_INVALID_ARGUMENT, "rerank_bits must be 2, 3, or 4");
    if (n_landmarks == 0 || n_landmarks > dim) fail(VC_ERR_INVALID_ARGUMENT, "n_landmarks must be in 1..dim");
    auto* ix = new vc_lm_index();
    try {
      ix->dim = dim;
      ix->scan_bits = scan_bits;
      ix->rerank_bits = rb;
      ix->m = n_landmarks;
      ix->scan_dim = (n_landmarks + 7) & ~size_t(7);
      ix->dim_pad = (dim + 15) & ~size_t(15);
      ix->seed = landmark_seed;
      hip_check(hipGetDevice(&ix->device), "hipGetDevice");
      hip_check(hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking), "hipStreamCreate");
      ix->t_add = make_tables(dim, rb);
      ix->t_add.norm_seq = 1;
      ix->t_query = make_tables(dim, rb);
      ix->t_scan = make_tables(ix->scan_dim, scan_bits);
      ix->t_scan.prerotated = 1;
      ix->t_scan.norm_seq = 1;
      const auto& L = vchip::ortho_landmarks(dim, n_landmarks, landmark_seed);
      const auto q = vchip::quantize_landmarks(L.data(), dim, n_landmarks);
      // column j's u32 word w (dims 8w .. 8w + 7) at w * m + j
      const size_t W = ix->dim_pad / 8, bpc = ix->dim_pad / 2;
      std::vector<uint32_t> codes_t(W * n_landmarks);
      for (size_t j = 0; j < n_landmarks; ++j)
        for (size_t w = 0; w < W; ++w) std::memcpy(&codes_t[w * n_landmarks + j], &q.codes[j * bpc + 4 * w], 4);
      Carve c;
      const size_t oL = c.take(L.size() * 4);
      ix->o_codes = c.take(codes_t.size() * 4);
      ix->o_scales = c.take(q.scales.size() * 4);
      ix->basis.reserve(c.off, ix->stream);
      hip_check(hipMemcpyAsync(ix->basis.as<float>(oL), L.data(), L.size() * 4, hipMemcpyHostToDevice, ix->stream), "copy");
      hip_check(hipMemcpyAsync(ix->basis.as<uint32_t>(ix->o_codes), codes_t.data(), codes_t.size() * 4,
                               hipMemcpyHostToDevice, ix->stream), "copy");
      hip_check(hipMemcpyAsync(ix->basis.as<float>(ix->o_scales), q.scales.data(), q.scales.size() * 4,
                               hipMemcpyHostToDevice, ix->stream), "copy");
      hip_check(hipStreamSynchronize(ix->stream), "sync");
    } catch (...) {
      if (ix->stream) (void)hipStreamDestroy(ix->stream);
      ix->basis.release();
      delete ix;
      throw;
    }
    *out = ix;
  });
}

vc_status vc_lm_index_destroy(vc_lm_index* ix) {
  return guarded([&] {
    if (ix == nullptr) return;
    (void)hipStreamSynchronize(ix->stream);
    for (DevBuf* b : {&ix->basis, &ix->rows, &ix->rn, &ix->tiled, &ix->alpha, &ix->work, &ix->qwork, &ix->scratch,
                      &ix->io})
      b->release();
    (void)hipStreamDestroy(ix->stream);
    delete ix;
  });
}

vc_status vc_lm_index_add(vc_lm_index* ix, const float* vectors, size_t n) {
  return guarded([&] {
    if (ix == nullptr || (vectors == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    lm_add(ix, vectors, n, true);
  });
}

vc_status vc_lm_index_add_device(vc_lm_index* ix, const float* vectors_dev, size_t n, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (vectors_dev == nullptr && n > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    stream_after(ix->stream, static_cast<hipStream_t>(stream));  // rows are read after the producer's work
    lm_add(ix, vectors_dev, n, false);  // ends with the index stream drained
  });
}

vc_status vc_lm_index_prepare(vc_lm_index* ix) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    if (ix->rows_live) {
      // every row (zeros for rows added before an earlier prepare)
      const size_t n = ix->n;
      if (n > ix->cap) {
        size_t ncap = std::max(n, ix->cap + ix->cap / 2);
        ncap = (ncap + 63) / 64 * 64;
        ix->tiled.reserve(vchip::tiled_bytes(ncap, ix->t_scan.groups16), ix->stream);
        ix->alpha.reserve(ncap * 4, ix->stream);
        ix->cap = ncap;
      }
      ix->t_scan.debug_flags = g_debug_flags.load();
      if (n > 0)
        hip_check(vchip::launch_encode(ix->t_scan, ix->rows.as<float>(), n, 0, ix->tiled.as<uint8_t>(), nullptr,
                                       nullptr, ix->alpha.as<float>(), ix->stream),
                  "encode(landmark rows)");
      hip_check(hipStreamSynchronize(ix->stream), "sync");
      ix->rows.release();
      ix->rows_live = false;
      ix->n_encoded = n;
      lm_build_sign(ix);
    }
    ix->prepared = true;
  });
}

vc_status vc_lm_index_search(vc_lm_index* ix, const float* queries, size_t nq, size_t k, const float* float_corpus,
                             float* scores, uint64_t* ids, size_t* k_out) {
  return guarded([&] {
    if (ix == nullptr || (queries == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (!ix->prepared) fail(VC_ERR_STATE, "call prepare() before search");
    const size_t ke = std::min(k, ix->n_encoded);
    if (k_out) *k_out = ke;
    if (nq == 0 || ke == 0) return;
    if (scores == nullptr || ids == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null output");
    const size_t D = ix->dim, m = nq * ke;
    Carve c;
    const size_t oq = c.take(nq * D * 4), os = c.take(m * 4), oi = c.take(m * 8),
                 ox = c.take(float_corpus ? m * D * 4 : 0);
    ix->io.reserve(c.off, ix->stream);
    hip_check(hipMemcpyAsync(ix->io.as<float>(oq), queries, nq * D * 4, hipMemcpyHostToDevice, ix->stream), "copy");
    lm_search_dev(ix, ix->io.as<float>(oq), nq, k, nullptr, ix->io.as<float>(os), ix->io.as<uint64_t>(oi),
                  ix->stream);
    if (float_corpus) {
      // Re-rank through the host corpus: only the nq x k candidate rows go up.
      // Row r of the upload holds the candidate of local id r; within a query
      // the local ids follow the global ids, so the rescore's tie order (id
      // asc) is the global one, and the ids come back through `global`.
      std::vector<uint64_t> got(m), global(m), local(m);
      hip_check(hipMemcpyAsync(got.data(), ix->io.as<uint64_t>(oi), m * 8, hipMemcpyDeviceToHost, ix->stream), "copy");
      hip_check(hipStreamSynchronize(ix->stream), "sync");
      std::vector<size_t> order(ke);
      for (size_t i = 0; i < nq; ++i) {
        const uint64_t* g = got.data() + i * ke;
        for (size_t j = 0; j < ke; ++j) order[j] = j;
        std::sort(order.begin(), order.end(), [g](size_t a, size_t b) { return g[a] < g[b]; });
        for (size_t r = 0; r < ke; ++r) {
          global[i * ke + r] = g[order[r]];
          local[i * ke + order[r]] = i * ke + r;
        }
      }
      std::vector<float> rows(m * D, 0.f);
      for (size_t r = 0; r < m; ++r)
        if (global[r] < ix->n) std::memcpy(rows.data() + r * D, float_corpus + global[r] * D, D * 4);
      hip_check(hipMemcpyAsync(ix->io.as<float>(ox), rows.data(), m * D * 4, hipMemcpyHostToDevice, ix->stream),
                "copy");
      hip_check(hipMemcpyAsync(ix->io.as<uint64_t>(oi), local.data(), m * 8, hipMemcpyHostToDevice, ix->stream),
                "copy");
      hip_check(vchip::launch_float_rescore(nq, int(ke), ix->io.as<float>(os), ix->io.as<uint64_t>(oi), nullptr,
                                            ix->io.as<float>(oq), ix->io.as<float>(ox), m, int(D), ix->stream),
                "float_rescore");
      hip_check(hipMemcpyAsync(scores, ix->io.as<float>(os), m * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
      hip_check(hipMemcpyAsync(ids, ix->io.as<uint64_t>(oi), m * 8, hipMemcpyDeviceToHost, ix->stream), "copy");
      hip_check(hipStreamSynchronize(ix->stream), "sync");
      for (size_t r = 0; r < m; ++r) ids[r] = ids[r] < m ? global[ids[r]] : ids[r];
      return;
    }
    hip_check(hipMemcpyAsync(scores, ix->io.as<float>(os), m * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipMemcpyAsync(ids, ix->io.as<uint64_t>(oi), m * 8, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_lm_index_search_device(vc_lm_index* ix, const float* queries_dev, size_t nq, size_t k,
                                    const float* float_corpus_dev, float* scores_dev, uint64_t* ids_dev,
                                    size_t* k_out, void* stream) {
  return guarded([&] {
    if (ix == nullptr || (queries_dev == nullptr && nq > 0)) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
    const size_t ke = lm_search_dev(ix, queries_dev, nq, k, float_corpus_dev, scores_dev, ids_dev, s);
    if (k_out) *k_out = ke;
  });
}

vc_status vc_lm_index_query_rows(vc_lm_index* ix, const float* queries, size_t nq, float* q_land, float* q_perp) {
  return guarded([&] {
    if (ix == nullptr || (nq > 0 && (queries == nullptr || q_land == nullptr || q_perp == nullptr)))
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(ix->mu);
    if (nq == 0) return;
    ix->io.reserve(nq * ix->dim * 4, ix->stream);
    hip_check(hipMemcpyAsync(ix->io.p, queries, nq * ix->dim * 4, hipMemcpyHostToDevice, ix->stream), "copy");
    float *ql = nullptr, *qp = nullptr;
    lm_query_rows(ix, ix->io.as<float>(), nq, &ql, &qp, ix->stream);
    hip_check(hipMemcpyAsync(q_land, ql, nq * ix->scan_dim * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipMemcpyAsync(q_perp, qp, nq * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

vc_status vc_lm_index_export(vc_lm_index* ix, float* residual_norms, float* scales, uint8_t* codes_flat) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    if (residual_norms && ix->n)
      hip_check(hipMemcpyAsync(residual_norms, ix->rn.p, ix->n * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    if ((scales || codes_flat) && !ix->prepared) fail(VC_ERR_STATE, "call prepare() before exporting codes");
    const size_t ne = ix->n_encoded, G = size_t(ix->t_scan.groups);
    if (scales && ne)
      hip_check(hipMemcpyAsync(scales, ix->alpha.p, ne * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
    if (codes_flat && ne) {
      ix->io.reserve(ne * G, ix->stream);
      hip_check(vchip::launch_tiled_to_flat(ix->tiled.as<uint8_t>(), ne, int(G), ix->t_scan.groups16,
                                            ix->io.as<uint8_t>(), ix->stream),
                "tiled_to_flat");
      hip_check(hipMemcpyAsync(codes_flat, ix->io.p, ne * G, hipMemcpyDeviceToHost, ix->stream), "copy");
    }
    hip_check(hipStreamSynchronize(ix->stream), "sync");
  });
}

size_t vc_lm_index_dim(const vc_lm_index* ix) { return ix ? ix->dim : 0; }
size_t vc_lm_index_size(const vc_lm_index* ix) { return ix ? ix->n : 0; }
size_t vc_lm_index_scan_bits(const vc_lm_index* ix) { return ix ? ix->scan_bits : 0; }
size_t vc_lm_index_rerank_bits(const vc_lm_index* ix) { return ix ? ix->rerank_bits : 0; }
size_t vc_lm_index_n_landmarks(const vc_lm_index* ix) { return ix ? ix->m : 0; }
size_t vc_lm_index_scan_dim(const vc_lm_index* ix) { return ix ? ix->scan_dim : 0; }
vc_status vc_lm_index_set_planes(vc_lm_index* ix, int on) {
  return guarded([&] {
    if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
    std::lock_guard lock(ix->mu);
    ix->planes = on ? 1 : 0;
    if (ix->prepared) lm_build_sign(ix);
  });
}
int vc_lm_index_planes_active(const vc_lm_index* ix) { return ix && ix->has_sign ? 1 : 0; }
int vc_lm_index_is_prepared(const vc_lm_index* ix) { return ix && ix->prepared ? 1 : 0; }
size_t vc_lm_index_codes_bytes(const vc_lm_index* ix) {
  return ix ? (ix->n_encoded + 191) / 192 * 192 * size_t(ix->t_scan.groups) : 0;
}
size_t vc_lm_index_index_bytes(const vc_lm_index* ix) {
  return ix ? vc_lm_index_codes_bytes(ix) + 4 * ix->n + 12 * ix->n_encoded : 0;
}

}  // extern "C"
