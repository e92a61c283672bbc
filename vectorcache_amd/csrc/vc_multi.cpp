// Many collections searched together (vc_multi_*), and the scan-plan
// autotuner (vc_tune_*).
//
// No counterpart in the reference, which searches one index per call
// (src/index.cpp:267-289). A vc_multi holds a fixed set of prepared flat
// indexes of any dim / bits / size, with a query batch size and k per
// collection. Its chain is: per collection the query preparation (rotation,
// TQ+, LUTs, prepare_queries) and the pair-source rows of the launch's tile,
// then ONE persistent scan launch over work items (collection, query tile,
// vector range) of every collection, then per collection the exact merge of
// its lists into its top-k rows. Each item is the single-index scan's
// (tile, range) workgroup, so every collection's rows are bit-identical to
// its own search. The chain runs eagerly or as a captured HIP graph
// replayed on the caller's stream; it reads queries from, and writes results
// to, device rows the set owns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "vc_common.hpp"
#include "vc_index.hpp"
#include "vc_internal.hpp"
#include "vectorcache_hip.h"

using namespace vcapi;

namespace {

constexpr int kCus = 256;  // MI355X compute units

struct Coll {
  vc_index* ix = nullptr;
  vchip::CodeTables t{};
  const uint8_t* tiled = nullptr;
  const float* alpha = nullptr;
  size_t dim = 0, G = 0, G16 = 0, n = 0, nq = 0, k = 0, k_eff = 0;
  size_t o_q = 0, o_s = 0, o_i = 0;  // rows / out (stable)
  size_t o_luts = 0, o_ls = 0, o_lb = 0, o_bc = 0, o_bt = 0, o_src = 0, o_ps = 0, o_pi = 0, o_pc = 0;  // work
  int cap = 0, n_lists = 0, tile = 0;
  size_t len = 0;
};

struct Plan {
  int tile = 8, vpt = 2;  // largest query tile; vectors per thread (tiles >= 2)
  float item_scale = 1.f;
  int wg_per_cu = 2;      // 2: tiles shrink until two workgroups share a CU; 1: one per CU
  size_t dyn = 0;
  int grid = 0;
  size_t n_items = 0;
  double ms = 0.0;  // measured by the tuner (0: not measured)
};

// ---- measured multi-set plans, keyed by a signature of the set --------------
std::mutex g_multi_mu;
std::map<uint64_t, Plan> g_multi_tuned;

int log2_floor(size_t v) {
  int r = 0;
  while (v > 1) {
    v >>= 1;
    ++r;
  }
  return r;
}

}  // namespace

struct vc_multi {
  int device = 0;
  std::vector<Coll> c;
  DevBuf rows, out, work, meta;
  Plan plan;
  uint64_t sig = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::mutex mu;

  void drop_graph() {
    if (exec != nullptr) (void)hipGraphExecDestroy(exec);
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void layout(const Plan& want, hipStream_t s);
  void enqueue(hipStream_t s, bool capturing = false);
};

// Signature of a set for the measured-plan table: per collection its
// (groups16, floor log2 n, floor log2 nq, k).
static uint64_t set_signature(const std::vector<Coll>& cs) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    for (int b = 0; b < 8; ++b) {
      h ^= (v >> (8 * b)) & 0xFF;
      h *= 1099511628211ull;
    }
  };
  mix(cs.size());
  for (const Coll& c : cs) {
    mix(c.G16);
    mix(uint64_t(log2_floor(std::max<size_t>(c.n, 1))));
    mix(uint64_t(log2_floor(std::max<size_t>(c.nq, 1))));
    mix(c.k_eff);
  }
  return h;
}

// Work buffers, per-collection tiles and candidate caps, LDS and work items
// of a plan.
void vc_multi::layout(const Plan& want, hipStream_t s) {
  const int V = want.vpt;
  if ((want.tile != 1 && want.tile != 2 && want.tile != 4 && want.tile != 8) || (V != 1 && V != 2) ||
      !(want.item_scale > 0.f) || (want.wg_per_cu != 1 && want.wg_per_cu != 2))
    fail(VC_ERR_INVALID_ARGUMENT, "tile must be 1, 2, 4 or 8, vec_per_thread 1 or 2, item_scale > 0, "
                                  "workgroups per CU 1 or 2");
  const size_t fixed = vchip::pair_fixed_lds();
  const size_t lim = size_t(vchip::kLdsPerCu) / size_t(want.wg_per_cu) - fixed;  // dynamic LDS per workgroup
  auto cap_for = [&](int T, size_t k) {
    const long room = long(lim) - long(vchip::pair_topk_lds(T, int(k), 0));
    return std::min<long>(vchip::kMaxCandCap, room / (8L * T) / 32 * 32);
  };
  Plan p = want;
  p.dyn = 0;
  double total = 0.0;
  for (Coll& x : c) {
    if (x.nq == 0 || x.k_eff == 0) continue;
    // the smallest tile that holds the batch, at most the plan's, shrunk
    // until its top-k state keeps >= 32 candidates in the LDS share
    int T = 1;
    while (T < want.tile && size_t(T) < x.nq) T *= 2;
    while (T > 1 && cap_for(T, x.k_eff) < 32) T /= 2;
    const long cap = cap_for(T, x.k_eff);
    if (cap < 32) fail(VC_ERR_INVALID_ARGUMENT, "k too large for the LDS share of this plan");
    x.tile = T;
    x.cap = int(cap);
    p.dyn = std::max(p.dyn, vchip::pair_topk_lds(T, int(x.k_eff), x.cap));
    total += double((x.nq + T - 1) / T) * double(x.n) * double(x.G16);
  }
  const int slots = kCus * want.wg_per_cu;
  // about 4 items per workgroup slot, scaled; each collection's ranges hold
  // whole steps of the scan and at most 512 lists per query
  const double target = std::max(1.0, total / (slots * 4.0) * p.item_scale);
  const size_t step = size_t(1024) * size_t(V);
  std::vector<std::pair<double, size_t>> order;
  for (size_t j = 0; j < c.size(); ++j) {
    Coll& x = c[j];
    if (x.nq == 0 || x.k_eff == 0) continue;
    size_t len = size_t(std::ceil(target / double(x.G16)));
    len = std::max(step, (len + 63) / 64 * 64);
    if ((x.n + len - 1) / len > 512) len = ((x.n + 511) / 512 + 63) / 64 * 64;
    x.len = len;
    x.n_lists = int((x.n + len - 1) / len);
    order.emplace_back(double(len) * double(x.G16), j);
  }
  // largest items first (a static round-robin then ends balanced); within a
  // collection range-major, so neighbouring workgroups read the same codes
  std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  std::vector<vchip::MultiItem> items;
  for (const auto& [cost, j] : order) {
    const Coll& x = c[j];
    const long n_qt = long((x.nq + x.tile - 1) / x.tile);
    for (int l = 0; l < x.n_lists; ++l)
      for (long qt = 0; qt < n_qt; ++qt) {
        vchip::MultiItem it{};
        it.coll = int(j);
        it.list = l;
        it.qt = qt;
        it.v_begin = long(l) * long(x.len);
        it.v_end = std::min<long>(long(x.n), it.v_begin + long(x.len));
        items.push_back(it);
      }
  }
  p.n_items = items.size();
  p.grid = int(std::min<size_t>(items.size(), size_t(slots)));
  // work buffer
  Carve cv;
  for (Coll& x : c) {
    if (x.nq == 0) continue;
    x.o_luts = cv.take(x.nq * x.G * 32);
    x.o_ls = cv.take(x.nq * 4);
    x.o_lb = cv.take(x.nq * 4);
    x.o_bc = cv.take(x.nq * 4);
    x.o_bt = cv.take(x.nq * 4);
    x.o_src = cv.take(x.k_eff > 0 ? vchip::pair_source_bytes(x.tile, int(x.G16), x.nq) : 0);
    const size_t lists = x.nq * size_t(std::max(x.n_lists, 0));
    x.o_ps = cv.take(lists * x.k_eff * 4);
    x.o_pi = cv.take(lists * x.k_eff * 4);
    x.o_pc = cv.take(lists * 4);
  }
  hip_check(hipStreamSynchronize(s), "sync");
  work.reserve(std::max<size_t>(cv.off, 256), s);
  std::vector<vchip::MultiColl> mc(c.size());
  for (size_t j = 0; j < c.size(); ++j) {
    const Coll& x = c[j];
    vchip::MultiColl& m = mc[j];
    m = vchip::MultiColl{};
    if (x.nq == 0 || x.k_eff == 0) continue;
    m.tiled = x.tiled;
    m.alpha = x.alpha;
    m.src = work.as<uint8_t>(x.o_src);
    m.lut_scale = work.as<float>(x.o_ls);
    m.bias_total = work.as<float>(x.o_bt);
    m.part_s = work.as<float>(x.o_ps);
    m.part_i = work.as<uint32_t>(x.o_pi);
    m.part_cnt = work.as<int>(x.o_pc);
    m.n = long(x.n);
    m.nq = long(x.nq);
    m.G16 = int(x.G16);
    m.k = int(x.k_eff);
    m.cap = x.cap;
    m.n_lists = x.n_lists;
    m.tile = x.tile;
  }
  Carve mv;
  const size_t o_c = mv.take(mc.size() * sizeof(vchip::MultiColl));
  const size_t o_it = mv.take(std::max<size_t>(items.size(), 1) * sizeof(vchip::MultiItem));
  meta.reserve(mv.off, s);
  hip_check(hipMemcpyAsync(meta.as<char>(o_c), mc.data(), mc.size() * sizeof(vchip::MultiColl),
                           hipMemcpyHostToDevice, s),
            "copy sets");
  if (!items.empty())
    hip_check(hipMemcpyAsync(meta.as<char>(o_it), items.data(), items.size() * sizeof(vchip::MultiItem),
                             hipMemcpyHostToDevice, s),
              "copy items");
  hip_check(hipStreamSynchronize(s), "sync");
  hip_check(vchip::prepare_multi_scan(V, p.dyn), "multi scan LDS limit");
  plan = p;
  drop_graph();  // captured buffers and items may have moved
}

void vc_multi::enqueue(hipStream_t s, bool capturing) {
  for (Coll& x : c) {
    if (x.nq == 0 || x.k_eff == 0) continue;
    hip_check(vchip::launch_prepare_queries(x.t, rows.as<float>(x.o_q), x.nq, work.as<uint8_t>(x.o_luts),
                                            work.as<float>(x.o_ls), work.as<float>(x.o_lb), work.as<float>(x.o_bc),
                                            work.as<float>(x.o_bt), s),
              "prepare_queries");
    hip_check(vchip::launch_pair_source(x.tile, work.as<uint8_t>(x.o_luts), x.nq, int(x.G), int(x.G16),
                                        work.as<uint8_t>(x.o_src), s),
              "pair_source");
  }
  Carve mv;
  const size_t o_c = mv.take(c.size() * sizeof(vchip::MultiColl));
  const size_t o_it = mv.take(std::max<size_t>(plan.n_items, 1) * sizeof(vchip::MultiItem));
  // profiling (vc_set_profiling): the scan launch between HIP events on its
  // stream, outside graph capture; vc_last_scan_stats reports tile and items
  std::unique_lock probe_lock(g_probe.mu);
  const bool probe = g_probe.on && !capturing;
  if (!probe) probe_lock.unlock();
  if (probe) hip_check(hipEventRecord(g_probe.start, s), "hipEventRecord");
  hip_check(vchip::launch_multi_scan(plan.vpt, meta.as<vchip::MultiColl>(o_c), meta.as<vchip::MultiItem>(o_it),
                                     long(plan.n_items), plan.dyn, plan.grid, s),
            "multi scan");
  if (probe) {
    hip_check(hipEventRecord(g_probe.stop, s), "hipEventRecord");
    g_probe.recorded = true;
    g_probe.tile = plan.tile;
    g_probe.ranges = int(plan.n_items);
    probe_lock.unlock();
  }
  for (Coll& x : c) {
    if (x.nq == 0 || x.k_eff == 0) continue;
    hip_check(vchip::launch_merge_ranges(work.as<float>(x.o_ps), work.as<uint32_t>(x.o_pi), work.as<int>(x.o_pc),
                                         x.nq, x.n_lists, int(x.k_eff), int(x.k_eff), nullptr,
                                         out.as<float>(x.o_s), out.as<uint64_t>(x.o_i), s),
              "merge");
  }
}

namespace {

// Median time (ms) of `reps` runs of f on stream s, after one warm-up run.
template <typename F>
double time_ms(hipStream_t s, int reps, F&& f) {
  f();
  hip_check(hipStreamSynchronize(s), "sync");
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  std::vector<float> ts;
  for (int r = 0; r < reps; ++r) {
    hip_check(hipEventRecord(e0, s), "hipEventRecord");
    f();
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    ts.push_back(ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::sort(ts.begin(), ts.end());
  return ts[ts.size() / 2];
}

struct StreamGuard {
  hipStream_t s = nullptr;
  StreamGuard() { hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
  ~StreamGuard() {
    if (s != nullptr) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

}  // namespace

extern "C" {

vc_status vc_multi_create(vc_index* const* indexes, size_t n_sets, const size_t* nq, const size_t* k,
                          vc_multi** out) {
  return guarded([&] {
    if (out == nullptr || (n_sets > 0 && (indexes == nullptr || nq == nullptr || k == nullptr)))
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (n_sets == 0) fail(VC_ERR_INVALID_ARGUMENT, "a set needs at least one collection");
    int dev = -1;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    auto m = std::make_unique<vc_multi>();
    m->device = dev;
    m->c.resize(n_sets);
    Carve rq, ro;
    for (size_t j = 0; j < n_sets; ++j) {
      vc_index* ix = indexes[j];
      if (ix == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null index");
      std::lock_guard lock(ix->mu);
      refuse_if_served(ix);
      if (!ix->prepared || ix->n_pending > 0) fail(VC_ERR_STATE, "every collection must be prepared");
      for (size_t i = 0; i < j; ++i)
        if (indexes[i] == ix) fail(VC_ERR_INVALID_ARGUMENT, "a collection appears twice in the set");
      if (ix->device != dev) fail(VC_ERR_INVALID_ARGUMENT, "every collection must live on the current device");
      if (k[j] < 1 || k[j] > size_t(vchip::kMaxTopK)) fail(VC_ERR_INVALID_ARGUMENT, "k must be in 1..2048");
      Coll& x = m->c[j];
      x.ix = ix;
      x.t = ix->t;
      x.tiled = ix->tiled.as<uint8_t>();
      x.alpha = ix->alpha.as<float>();
      x.dim = ix->dim;
      x.G = size_t(ix->t.groups);
      x.G16 = size_t(ix->t.groups16);
      x.n = ix->n_encoded;
      x.nq = nq[j];
      x.k = k[j];
      x.k_eff = std::min(k[j], x.n);
      if (x.n > 0xFFFFFFFFull) fail(VC_ERR_INVALID_ARGUMENT, "index exceeds 2^32 vectors");
      x.o_q = rq.take(x.nq * x.dim * 4);
      x.o_s = ro.take(x.nq * x.k_eff * 4);
      x.o_i = ro.take(x.nq * x.k_eff * 8);
    }
    StreamGuard sg;
    m->rows.reserve(std::max<size_t>(rq.off, 256), sg.s);
    m->out.reserve(std::max<size_t>(ro.off, 256), sg.s);
    m->sig = set_signature(m->c);
    Plan want;
    {
      std::lock_guard lock(g_multi_mu);
      auto it = g_multi_tuned.find(m->sig);
      if (it != g_multi_tuned.end()) want = it->second;
    }
    try {
      m->layout(want, sg.s);
    } catch (const VcError&) {
      m->layout(Plan{}, sg.s);  // a measured plan that no longer fits: the default
    }
    // the codes must outlive the set: the indexes refuse to change or go away while it exists
    for (Coll& x : m->c) {
      std::lock_guard lock(x.ix->mu);
      ++x.ix->sets;
    }
    *out = m.release();
  });
}

vc_status vc_multi_destroy(vc_multi* m) {
  return guarded([&] {
    if (m == nullptr) return;
    (void)hipDeviceSynchronize();
    m->drop_graph();
    for (Coll& x : m->c) {
      std::lock_guard lock(x.ix->mu);
      --x.ix->sets;
    }
    m->rows.release();
    m->out.release();
    m->work.release();
    m->meta.release();
    delete m;
  });
}

vc_status vc_multi_rows(vc_multi* m, size_t set, float** queries_dev, float** scores_dev, uint64_t** ids_dev,
                        size_t* k_eff) {
  return guarded([&] {
    if (m == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null set");
    if (set >= m->c.size()) fail(VC_ERR_OUT_OF_RANGE, "collection index out of range");
    const Coll& x = m->c[set];
    if (queries_dev) *queries_dev = m->rows.as<float>(x.o_q);
    if (scores_dev) *scores_dev = m->out.as<float>(x.o_s);
    if (ids_dev) *ids_dev = m->out.as<uint64_t>(x.o_i);
    if (k_eff) *k_eff = x.k_eff;
  });
}

vc_status vc_multi_collection_plan(vc_multi* m, size_t set, size_t* tile, size_t* lists) {
  return guarded([&] {
    if (m == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null set");
    std::lock_guard lock(m->mu);
    if (set >= m->c.size()) fail(VC_ERR_OUT_OF_RANGE, "collection index out of range");
    if (tile) *tile = size_t(m->c[set].tile);
    if (lists) *lists = size_t(std::max(m->c[set].n_lists, 0));
  });
}

vc_status vc_multi_set_queries(vc_multi* m, size_t set, const float* queries_dev, void* stream) {
  return guarded([&] {
    if (m == nullptr || queries_dev == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (set >= m->c.size()) fail(VC_ERR_OUT_OF_RANGE, "collection index out of range");
    const Coll& x = m->c[set];
    if (x.nq > 0)
      hip_check(hipMemcpyAsync(m->rows.as<float>(x.o_q), queries_dev, x.nq * x.dim * 4, hipMemcpyDeviceToDevice,
                               static_cast<hipStream_t>(stream)),
                "copy queries");
  });
}

vc_status vc_multi_get_results(vc_multi* m, size_t set, float* scores_dev, uint64_t* ids_dev, void* stream) {
  return guarded([&] {
    if (m == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null set");
    if (set >= m->c.size()) fail(VC_ERR_OUT_OF_RANGE, "collection index out of range");
    const Coll& x = m->c[set];
    const size_t cnt = x.nq * x.k_eff;
    if (cnt == 0) return;
    if (scores_dev == nullptr || ids_dev == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hip_check(hipMemcpyAsync(scores_dev, m->out.as<float>(x.o_s), cnt * 4, hipMemcpyDeviceToDevice, s), "copy");
    hip_check(hipMemcpyAsync(ids_dev, m->out.as<uint64_t>(x.o_i), cnt * 8, hipMemcpyDeviceToDevice, s), "copy");
  });
}

vc_status vc_multi_search(vc_multi* m, void* stream) {
  return guarded([&] {
    if (m == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null set");
    std::lock_guard lock(m->mu);
    m->enqueue(static_cast<hipStream_t>(stream));
  });
}

vc_status vc_multi_capture(vc_multi* m, void* stream) {
  return guarded([&] {
    if (m == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null set");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (s == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "capture needs a stream other than the null stream");
    std::lock_guard lock(m->mu);
    m->drop_graph();
    hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      m->enqueue(s, /*capturing=*/true);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g != nullptr) (void)hipGraphDestroy(g);
      throw;
    }
    hip_check(hipStreamEndCapture(s, &m->graph), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  });
}

vc_status vc_multi_replay(vc_multi* m, void* stream) {
  return guarded([&] {
    if (m == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null set");
    std::lock_guard lock(m->mu);
    if (m->exec == nullptr) fail(VC_ERR_STATE, "no captured graph (capture first; a new plan drops it)");
    hip_check(hipGraphLaunch(m->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
  });
}

static void plan_info(const Plan& p, vc_multi_plan_info* out) {
  out->tile = size_t(p.tile);
  out->vec_per_thread = size_t(p.vpt);
  out->item_scale = p.item_scale;
  out->workgroups_per_cu = size_t(p.wg_per_cu);
  out->items = p.n_items;
  out->grid = size_t(p.grid);
  out->lds_bytes = vchip::pair_fixed_lds() + p.dyn;
  out->measured_ms = p.ms;
}

vc_status vc_multi_set_plan(vc_multi* m, size_t tile, size_t vec_per_thread, float item_scale,
                            size_t workgroups_per_cu) {
  return guarded([&] {
    if (m == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null set");
    std::lock_guard lock(m->mu);
    Plan want;
    want.tile = int(tile);
    want.vpt = int(vec_per_thread);
    want.item_scale = item_scale;
    want.wg_per_cu = int(workgroups_per_cu);
    StreamGuard sg;
    m->layout(want, sg.s);
  });
}

vc_status vc_multi_get_plan(vc_multi* m, vc_multi_plan_info* out) {
  return guarded([&] {
    if (m == nullptr || out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lock(m->mu);
    plan_info(m->plan, out);
  });
}

vc_status vc_multi_tune(vc_multi* m, size_t reps, vc_multi_plan_info* best) {
  return guarded([&] {
    if (m == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null set");
    std::lock_guard lock(m->mu);
    StreamGuard sg;
    const int r = int(std::max<size_t>(reps, 1));
    Plan best_p;
    double best_ms = 1e30;
    for (int wg : {2, 1})
      for (int T : {8, 4})
        for (int V : {2, 1})
          for (float sc : {0.5f, 1.f, 2.f}) {
          Plan want;
          want.tile = T;
          want.vpt = V;
          want.item_scale = sc;
          want.wg_per_cu = wg;
          try {
            m->layout(want, sg.s);
          } catch (const VcError&) {
            continue;  // does not fit (k too large for this tile)
          }
          const double ms = time_ms(sg.s, r, [&] { m->enqueue(sg.s); });
          if (ms < best_ms) {
            best_ms = ms;
            best_p = m->plan;
          }
        }
    if (best_ms >= 1e30) fail(VC_ERR_INVALID_ARGUMENT, "no plan fits this set");
    best_p.ms = best_ms;
    m->layout(best_p, sg.s);
    m->plan.ms = best_ms;
    {
      std::lock_guard g(g_multi_mu);
      g_multi_tuned[m->sig] = m->plan;
    }
    if (best != nullptr) plan_info(m->plan, best);
  });
}

// ---- single-index scan plans ---------------------------------------------------
vc_status vc_tune_scan(size_t dim, size_t bits, size_t n, size_t nq, size_t k, size_t reps,
                       vc_scan_choice* out) {
  return guarded([&] {
    check_dim_bits(dim, bits);
    if (n == 0 || nq == 0 || k == 0) fail(VC_ERR_INVALID_ARGUMENT, "n, nq and k must be positive");
    if (n > 0xFFFFFFFFull) fail(VC_ERR_INVALID_ARGUMENT, "n exceeds 2^32 vectors");
    const size_t ke = std::min(k, n);
    if (ke > size_t(vchip::kMaxTopK)) fail(VC_ERR_INVALID_ARGUMENT, "k exceeds 2048");
    StreamGuard sg;
    hipStream_t s = sg.s;
    vchip::CodeTables t = make_tables(dim, bits);
    const size_t chunk = std::min<size_t>(n, 65536);
    DevBuf codes, alpha, rowsb, q, res, scratch;
    codes.reserve(vchip::tiled_bytes(n, t.groups16), s);
    alpha.reserve(std::max<size_t>(n, 1) * 4, s);
    rowsb.reserve(chunk * dim * 4, s);
    for (size_t off = 0; off < n; off += chunk) {
      const size_t cn = std::min(chunk, n - off);
      hip_check(vchip::launch_fill_rows(rowsb.as<float>(), cn, int(dim), 0x5EED0000ull + off, s), "fill rows");
      hip_check(vchip::launch_encode(t, rowsb.as<float>(), cn, off, codes.as<uint8_t>(), nullptr, nullptr,
                                     alpha.as<float>(), s),
                "encode");
    }
    q.reserve(nq * dim * 4, s);
    hip_check(vchip::launch_fill_rows(q.as<float>(), nq, int(dim), 0xC0FFEEull, s), "fill queries");
    Carve rc;
    const size_t o_rs = rc.take(nq * ke * 4), o_ri = rc.take(nq * ke * 8);
    res.reserve(rc.off, s);
    hip_check(hipStreamSynchronize(s), "sync");
    const int r = int(std::max<size_t>(reps, 1));
    auto run = [&](const vchip::ScanPlan* plan) {
      g_plan_override = plan;
      try {
        search_tiled(t, codes.as<uint8_t>(), alpha.as<float>(), n, q.as<float>(), nq, ke, nullptr,
                     res.as<float>(o_rs), res.as<uint64_t>(o_ri), scratch, s);
      } catch (...) {
        g_plan_override = nullptr;
        throw;
      }
      g_plan_override = nullptr;
    };
    const vchip::ScanPlan rule = vchip::plan_scan_rule(t.groups16, n, nq, int(ke), 0, false);
    const double rule_ms = time_ms(s, r, [&] { run(&rule); });
    double best_ms = rule_ms;
    vchip::TunedScan best{rule.tile, rule.vec_per_thread, 1.f, rule_ms};
    for (int T : {1, 2, 4, 8}) {
      if (T > 1 && size_t(T / 2) >= nq) continue;  // a tile at least twice the batch only pads
      for (int V : {1, 2})
        for (float sc : {0.5f, 1.f, 2.f}) {
          const vchip::ScanPlan p = vchip::plan_scan_with(t.groups16, n, nq, int(ke), T, V, sc);
          if (p.tile == 0) continue;
          const double ms = time_ms(s, r, [&] { run(&p); });
          if (ms < best_ms) {
            best_ms = ms;
            best = vchip::TunedScan{T, V, sc, ms};
          }
        }
    }
    best.ms = best_ms;
    vchip::tune_put(vchip::tune_key(t.groups16, n, nq, int(ke)), best);
    if (out != nullptr) {
      out->tile = size_t(best.tile);
      out->vec_per_thread = size_t(best.vec_per_thread);
      out->range_scale = best.range_scale;
      out->ms = best_ms;
      out->rule_ms = rule_ms;
      out->rule_tile = size_t(rule.tile);
    }
  });
}

vc_status vc_tune_save(const char* path) {
  return guarded([&] {
    if (path == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null path");
    std::string text = vchip::tune_dump();
    {
      std::lock_guard lock(g_multi_mu);
      char line[256];
      for (const auto& [sig, p] : g_multi_tuned) {
        std::snprintf(line, sizeof line, "multi sig=%016llx tile=%d vpt=%d iscale=%.4g wg=%d ms=%.6g\n",
                      (unsigned long long)sig, p.tile, p.vpt, (double)p.item_scale, p.wg_per_cu, p.ms);
        text += line;
      }
    }
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) fail(VC_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    f << text;
    if (!f) fail(VC_ERR_DEVICE, std::string("write failed: ") + path);
  });
}

vc_status vc_tune_load(const char* path, size_t* n_entries) {
  return guarded([&] {
    if (path == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null path");
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(VC_ERR_STATE, std::string("cannot read ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    // multi entries first (validated), then the scan entries
    std::vector<std::pair<uint64_t, Plan>> multi;
    std::istringstream lines(text);
    std::string line;
    while (std::getline(lines, line)) {
      if (line.rfind("multi ", 0) != 0) continue;
      unsigned long long sig = 0;
      Plan p;
      double sc = 1.0;
      if (std::sscanf(line.c_str(), "multi sig=%llx tile=%d vpt=%d iscale=%lf wg=%d ms=%lf", &sig, &p.tile, &p.vpt,
                      &sc, &p.wg_per_cu, &p.ms) != 6 ||
          (p.tile != 1 && p.tile != 2 && p.tile != 4 && p.tile != 8) || (p.vpt != 1 && p.vpt != 2) ||
          (p.wg_per_cu != 1 && p.wg_per_cu != 2) ||
          !(sc > 0.0 && sc <= 64.0))
        fail(VC_ERR_STATE, std::string("malformed tuning entry: ") + line);
      p.item_scale = float(sc);
      multi.emplace_back(uint64_t(sig), p);
    }
    const int n_scan = vchip::tune_parse(text);
    if (n_scan < 0) fail(VC_ERR_STATE, std::string("malformed scan entry in ") + path);
    {
      std::lock_guard lock(g_multi_mu);
      for (const auto& [sig, p] : multi) g_multi_tuned[sig] = p;
    }
    if (n_entries) *n_entries = size_t(n_scan) + multi.size();
  });
}

vc_status vc_tune_clear(void) {
  return guarded([&] {
    vchip::tune_clear();
    std::lock_guard lock(g_multi_mu);
    g_multi_tuned.clear();
  });
}

size_t vc_tune_count(void) {
  std::lock_guard lock(g_multi_mu);
  return vchip::tune_count() + g_multi_tuned.size();
}

}  // extern "C"
