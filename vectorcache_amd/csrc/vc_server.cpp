// Request server over one prepared flat index (vc_server_* in
// include/vectorcache_hip.h): continuous batching of search requests on
// several HIP streams, admission of each batch by its device memory against a
// fixed pool, and add batches encoded concurrently on a stream of their own.
//
// No counterpart in the reference, whose GPU search mutates shared device
// buffers and is not safe to call concurrently (src/gpu/kernels.cu:602-627).
// Each request's results are those of TurboQuantIndex::search
// (src/index.cpp:267-289) over the vectors its batch saw.
//
// Threads: one worker per slot, one adder. All queue and pool state is under
// vc_server::mu; the index's codes are written only by the adder, at rows no
// running batch reads (a batch scans the first n_snapshot rows, the adder
// writes rows from n_visible >= n_snapshot on), and are reallocated only
// while no batch is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vc_common.hpp"
#include "vc_index.hpp"
#include "vc_internal.hpp"
#include "vectorcache_hip.h"

using namespace vcapi;

namespace {

double now_s() {
  // steady_clock is CLOCK_MONOTONIC on Linux: comparable with a caller's
  // monotonic timestamps (time.monotonic() in Python)
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// k class of a request: the smallest c with k <= 2^c
int k_class(size_t k) {
  int c = 0;
  while ((size_t(1) << c) < k) ++c;
  return c;
}

constexpr size_t kGranule = 256;
size_t granules(size_t bytes) { return (bytes + kGranule - 1) / kGranule * kGranule; }

// First-fit allocator over one device allocation; freed blocks coalesce, so
// an empty pool is one block again.
class Pool {
 public:
  void init(size_t bytes) {
    cap_ = bytes / kGranule * kGranule;
    free_.clear();
    free_[0] = cap_;
  }
  size_t cap() const { return cap_; }
  size_t used() const { return used_; }
  bool alloc(size_t bytes, size_t* off) {
    bytes = granules(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      free_.erase(it);
      if (rest > 0) free_[*off + bytes] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = granules(std::max<size_t>(bytes, 1));
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
    used_ -= bytes;
  }

 private:
  std::map<size_t, size_t> free_;  // offset -> bytes
  size_t cap_ = 0, used_ = 0;
};

struct Request {
  bool is_add = false;
  std::vector<float> rows;  // queries (search) or vectors (add), copied at submit
  size_t n = 0;             // rows
  size_t k = 0;
  std::vector<float> scores;  // n x info.k_eff
  std::vector<uint64_t> ids;
  vc_request_info info{};
  vc_status status = VC_OK;
  std::string err;
  bool done = false;
};

// Device bytes one batch of nq queries at top-k over n vectors holds: the
// search's scratch, the query rows and the results.
size_t batch_device_bytes(const vchip::CodeTables& t, size_t dim, size_t n, size_t nq, size_t k) {
  const SearchLayout L = search_layout(t, n, nq, k);
  return granules(L.bytes) + granules(nq * dim * 4) + granules(nq * L.k_eff * 4) + granules(nq * L.k_eff * 8);
}

struct PinnedBuf {
  void* p = nullptr;
  void alloc(size_t bytes) { hip_check(hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault), "hipHostMalloc"); }
  void release() {
    if (p != nullptr) (void)hipHostFree(p);
    p = nullptr;
  }
};

}  // namespace

struct vc_server {
  vc_index* ix = nullptr;
  vc_server_params p{};
  int device = 0;
  size_t dim = 0;
  vchip::CodeTables t{};

  DevBuf pool_mem;
  Pool pool;
  DevBuf add_stage;
  std::vector<hipStream_t> streams;
  hipStream_t add_stream = nullptr;
  std::vector<PinnedBuf> stage;  // per slot: queries, then scores, then ids
  PinnedBuf add_host;

  std::mutex mu;
  std::condition_variable cv_work, cv_add, cv_done;
  std::deque<Request*> search_q, add_q;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_ticket = 1, next_batch = 0;
  size_t n_visible = 0;
  size_t in_flight = 0;
  size_t not_done = 0;
  bool growing = false;
  bool stop = false;
  vc_server_stats stats{};

  std::vector<std::thread> workers;
  std::thread adder;

  size_t batch_bytes(size_t n, size_t nq, size_t k) const { return batch_device_bytes(t, dim, n, nq, k); }

  void finish(Request* r) {  // under mu
    r->info.t_done = now_s();
    r->done = true;
    --not_done;
    cv_done.notify_all();
  }

  void worker(int slot);
  void add_loop();
  void run_batch(int slot, std::vector<Request*>& batch, size_t nq, size_t kb, size_t n, const uint8_t* tiled,
                 const float* alpha, uint8_t* mem);
};

void vc_server::run_batch(int slot, std::vector<Request*>& batch, size_t nq, size_t kb, size_t n,
                          const uint8_t* tiled, const float* alpha, uint8_t* mem) {
  hipStream_t s = streams[size_t(slot)];
  const SearchLayout L = search_layout(t, n, nq, kb);
  const size_t ke = L.k_eff;
  Carve c;
  const size_t o_scr = c.take(L.bytes), o_q = c.take(nq * dim * 4), o_s = c.take(nq * ke * 4),
               o_i = c.take(nq * ke * 8);
  char* h = static_cast<char*>(stage[size_t(slot)].p);
  const size_t h_s = nq * dim * 4, h_i = h_s + nq * ke * 4;
  size_t q0 = 0;
  for (Request* r : batch) {
    std::memcpy(h + q0 * dim * 4, r->rows.data(), r->n * dim * 4);
    q0 += r->n;
  }
  hip_check(hipMemcpyAsync(mem + o_q, h, nq * dim * 4, hipMemcpyHostToDevice, s), "copy queries");
  ScratchArena arena;
  arena.fixed = mem + o_scr;
  arena.fixed_cap = L.bytes;
  DevBuf unused;
  search_tiled(t, tiled, alpha, n, reinterpret_cast<const float*>(mem + o_q), nq, kb, nullptr,
               reinterpret_cast<float*>(mem + o_s), reinterpret_cast<uint64_t*>(mem + o_i), unused, s, nullptr,
               nullptr, &arena);
  hip_check(hipMemcpyAsync(h + h_s, mem + o_s, nq * ke * 4, hipMemcpyDeviceToHost, s), "copy scores");
  hip_check(hipMemcpyAsync(h + h_i, mem + o_i, nq * ke * 8, hipMemcpyDeviceToHost, s), "copy ids");
  hip_check(hipStreamSynchronize(s), "sync batch");
  const float* hs = reinterpret_cast<const float*>(h + h_s);
  const uint64_t* hi = reinterpret_cast<const uint64_t*>(h + h_i);
  q0 = 0;
  for (Request* r : batch) {
    const size_t kr = std::min(r->k, n);  // a prefix of the batch's rows: the request's own top-k
    r->info.k_eff = kr;
    r->scores.resize(r->n * kr);
    r->ids.resize(r->n * kr);
    for (size_t q = 0; q < r->n; ++q) {
      std::memcpy(&r->scores[q * kr], hs + (q0 + q) * ke, kr * 4);
      std::memcpy(&r->ids[q * kr], hi + (q0 + q) * ke, kr * 8);
    }
    q0 += r->n;
  }
}

void vc_server::worker(int slot) {
  (void)hipSetDevice(device);
  std::unique_lock lk(mu);
  for (;;) {
    cv_work.wait(lk, [&] { return (stop && search_q.empty()) || (!search_q.empty() && !growing); });
    if (search_q.empty()) break;  // stopping, nothing left
    // the batch: the waiting requests of the oldest request's k class (k in
    // (2^(c-1), 2^c]: a batch runs at its largest k, so requests of a much
    // smaller k go to batches of their own), in arrival order, up to
    // max_batch_queries
    std::vector<Request*> batch;
    size_t nq = 0;
    const int cls = k_class(search_q.front()->k);
    for (Request* r : search_q) {
      if (k_class(r->k) != cls) continue;
      if (nq + r->n > p.max_batch_queries) break;
      batch.push_back(r);
      nq += r->n;
    }
    const size_t n = n_visible;
    auto k_of = [&] {
      size_t k = 0;
      for (Request* r : batch) k = std::max(k, r->k);
      return k;
    };
    size_t kb = k_of(), bytes = 0, off = 0;
    bool placed = false, wait_memory = false;
    try {
      for (;;) {
        bytes = batch_bytes(n, nq, kb);
        if (pool.alloc(bytes, &off)) {
          placed = true;
          break;
        }
        if (batch.size() > 1) {  // admit fewer requests
          nq -= batch.back()->n;
          batch.pop_back();
          kb = k_of();
          continue;
        }
        if (bytes > pool.cap())
          fail(VC_ERR_OOM, "the request needs " + std::to_string(bytes) + " bytes of device memory; the server's "
                           "pool holds " + std::to_string(pool.cap()));
        wait_memory = true;  // running batches hold the memory: wait for one to finish
        break;
      }
    } catch (const VcError& e) {
      Request* r = search_q.front();
      search_q.pop_front();
      r->status = e.status;
      r->err = e.msg;
      finish(r);
      continue;
    }
    if (wait_memory) {
      ++stats.admission_waits;
      cv_work.wait(lk);
      continue;
    }
    (void)placed;
    const uint64_t bid = next_batch++;
    const double t0 = now_s();
    {
      size_t taken = 0;  // the batch's requests leave the queue, the others keep their order
      search_q.erase(std::remove_if(search_q.begin(), search_q.end(),
                                    [&](Request* r) {
                                      if (taken < batch.size() && r == batch[taken]) {
                                        ++taken;
                                        return true;
                                      }
                                      return false;
                                    }),
                     search_q.end());
    }
    for (Request* r : batch) {
      r->info.t_start = t0;
      r->info.n_snapshot = n;
      r->info.batch = bid;
      r->info.batch_queries = nq;
      r->info.slot = slot;
    }
    ++in_flight;
    stats.in_flight_peak = std::max(stats.in_flight_peak, in_flight);
    stats.pool_peak_bytes = std::max(stats.pool_peak_bytes, pool.used());
    ++stats.batches;
    const uint8_t* tiled = ix->tiled.as<uint8_t>();
    const float* alpha = ix->alpha.as<float>();
    uint8_t* mem = pool_mem.as<uint8_t>(off);
    if (!search_q.empty()) cv_work.notify_all();  // another slot may take the rest
    lk.unlock();
    vc_status st = VC_OK;
    std::string err;
    try {
      run_batch(slot, batch, nq, kb, n, tiled, alpha, mem);
    } catch (const VcError& e) {
      st = e.status;
      err = e.msg;
    } catch (const std::exception& e) {
      st = VC_ERR_DEVICE;
      err = e.what();
    }
    lk.lock();
    pool.release(off, bytes);
    --in_flight;
    for (Request* r : batch) {
      r->status = st;
      r->err = err;
      stats.queries += r->n;
      finish(r);
    }
    cv_work.notify_all();
    cv_add.notify_all();
  }
}

void vc_server::add_loop() {
  (void)hipSetDevice(device);
  std::unique_lock lk(mu);
  for (;;) {
    cv_add.wait(lk, [&] { return stop || !add_q.empty(); });
    if (add_q.empty()) break;  // stopping, nothing left
    Request* r = add_q.front();
    r->info.t_start = now_s();
    size_t done_rows = 0;
    try {
      while (done_rows < r->n) {
        const size_t cn = std::min(p.add_chunk, r->n - done_rows);
        const size_t total = n_visible + cn;
        if (p.max_vectors != 0 && total > p.max_vectors)
          fail(VC_ERR_OOM, "add refused: the index would hold " + std::to_string(total) + " vectors, above the "
                           "server's max_vectors of " + std::to_string(p.max_vectors));
        if (total > 0xFFFFFFFFull) fail(VC_ERR_INVALID_ARGUMENT, "index exceeds 2^32 vectors");
        if (total > ix->cap) {
          // reallocation: no batch may hold the old buffers
          growing = true;
          cv_add.wait(lk, [&] { return in_flight == 0; });
          try {
            index_reserve(ix, total, add_stream);
          } catch (...) {
            growing = false;
            cv_work.notify_all();
            throw;
          }
          growing = false;
          cv_work.notify_all();
        }
        const size_t at = n_visible;
        uint8_t* tiled = ix->tiled.as<uint8_t>();
        float* alpha = ix->alpha.as<float>();
        const float* src = r->rows.data() + done_rows * dim;
        lk.unlock();
        try {
          std::memcpy(add_host.p, src, cn * dim * 4);
          hip_check(hipMemcpyAsync(add_stage.p, add_host.p, cn * dim * 4, hipMemcpyHostToDevice, add_stream),
                    "copy rows");
          hip_check(vchip::launch_encode(t, add_stage.as<float>(), cn, at, tiled, nullptr, nullptr, alpha,
                                         add_stream),
                    "encode");
          hip_check(hipStreamSynchronize(add_stream), "sync add");
        } catch (...) {
          lk.lock();
          throw;
        }
        lk.lock();
        n_visible = at + cn;  // batches started from now on see these rows
        ix->n_encoded = n_visible;
        stats.vectors_added += cn;
        stats.size = n_visible;
        done_rows += cn;
        cv_work.notify_all();
      }
    } catch (const VcError& e) {
      r->status = e.status;
      r->err = e.msg + " (" + std::to_string(done_rows) + " of " + std::to_string(r->n) + " rows were added)";
    } catch (const std::exception& e) {
      r->status = VC_ERR_DEVICE;
      r->err = e.what();
    }
    add_q.pop_front();
    r->info.n_snapshot = n_visible;
    r->info.k_eff = done_rows;
    ++stats.adds;
    finish(r);
  }
}

extern "C" {

vc_status vc_server_params_default(vc_server_params* p) {
  return guarded([&] {
    if (p == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null params");
    p->slots = 4;
    p->max_batch_queries = 512;
    p->max_k = 100;
    p->hbm_budget_bytes = size_t(1) << 30;
    p->reserve_vectors = 0;
    p->max_vectors = 0;
    p->add_chunk = 65536;
  });
}

size_t vc_server_batch_bytes(const vc_index* ix, size_t n, size_t nq, size_t k) {
  if (ix == nullptr || nq == 0) return 0;
  try {
    return batch_device_bytes(ix->t, ix->dim, n, nq, k);
  } catch (...) {
    return 0;
  }
}

vc_status vc_server_create(vc_index* ix, const vc_server_params* params, vc_server** out) {
  return guarded([&] {
    if (ix == nullptr || out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    vc_server_params p{};
    vc_server_params_default(&p);
    if (params != nullptr) p = *params;
    if (p.slots < 1 || p.slots > 64) fail(VC_ERR_INVALID_ARGUMENT, "slots must be in 1..64");
    if (p.max_batch_queries < 1) fail(VC_ERR_INVALID_ARGUMENT, "max_batch_queries must be >= 1");
    if (p.max_k < 1 || p.max_k > size_t(vchip::kMaxTopK)) fail(VC_ERR_INVALID_ARGUMENT, "max_k must be in 1..2048");
    if (p.add_chunk < 1) fail(VC_ERR_INVALID_ARGUMENT, "add_chunk must be >= 1");
    std::lock_guard ilk(ix->mu);
    refuse_if_served(ix);
    refuse_if_in_set(ix);
    if (!ix->prepared || ix->n_pending > 0) fail(VC_ERR_STATE, "the server needs a prepared index");
    int dev = -1;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != ix->device) fail(VC_ERR_INVALID_ARGUMENT, "the index's device must be current");
    if (p.reserve_vectors == 0) p.reserve_vectors = ix->n_encoded + 1000000;
    if (p.max_vectors != 0) p.reserve_vectors = std::min(p.reserve_vectors, p.max_vectors);
    auto sv = std::make_unique<vc_server>();
    sv->ix = ix;
    sv->p = p;
    sv->device = dev;
    sv->dim = ix->dim;
    sv->t = ix->t;
    sv->n_visible = ix->n_encoded;
    sv->stats.size = sv->n_visible;
    // a request of max_batch_queries queries at max_k must fit the pool alone
    const size_t n_hi = std::max(ix->n_encoded, p.reserve_vectors);
    const size_t worst = std::max(sv->batch_bytes(std::max<size_t>(ix->n_encoded, 1), p.max_batch_queries, p.max_k),
                                  sv->batch_bytes(std::max<size_t>(n_hi, 1), p.max_batch_queries, p.max_k));
    if (worst > p.hbm_budget_bytes)
      fail(VC_ERR_INVALID_ARGUMENT, "hbm_budget_bytes holds " + std::to_string(p.hbm_budget_bytes) +
                                    " bytes; one batch of max_batch_queries at max_k needs " + std::to_string(worst));
    try {
      index_reserve(ix, p.reserve_vectors, ix->stream);
      hip_check(hipStreamSynchronize(ix->stream), "sync");
      sv->pool_mem.reserve(p.hbm_budget_bytes, ix->stream);
      sv->pool.init(p.hbm_budget_bytes);
      sv->add_stage.reserve(p.add_chunk * sv->dim * 4, ix->stream);
      sv->add_host.alloc(p.add_chunk * sv->dim * 4);
      sv->streams.assign(p.slots, nullptr);
      sv->stage.resize(p.slots);
      for (size_t i = 0; i < p.slots; ++i) {
        hip_check(hipStreamCreateWithFlags(&sv->streams[i], hipStreamNonBlocking), "hipStreamCreate");
        sv->stage[i].alloc(p.max_batch_queries * (sv->dim * 4 + p.max_k * 12));
      }
      hip_check(hipStreamCreateWithFlags(&sv->add_stream, hipStreamNonBlocking), "hipStreamCreate");
    } catch (...) {
      for (hipStream_t s : sv->streams)
        if (s != nullptr) (void)hipStreamDestroy(s);
      if (sv->add_stream != nullptr) (void)hipStreamDestroy(sv->add_stream);
      for (auto& b : sv->stage) b.release();
      sv->add_host.release();
      sv->pool_mem.release();
      sv->add_stage.release();
      throw;
    }
    vc_server* raw = sv.release();
    ix->server = raw;
    for (size_t i = 0; i < p.slots; ++i) raw->workers.emplace_back(&vc_server::worker, raw, int(i));
    raw->adder = std::thread(&vc_server::add_loop, raw);
    *out = raw;
  });
}

vc_status vc_server_destroy(vc_server* sv) {
  return guarded([&] {
    if (sv == nullptr) return;
    {
      std::lock_guard lk(sv->mu);
      sv->stop = true;
    }
    sv->cv_work.notify_all();
    sv->cv_add.notify_all();
    for (auto& w : sv->workers) w.join();
    sv->adder.join();
    for (hipStream_t s : sv->streams) (void)hipStreamDestroy(s);
    (void)hipStreamDestroy(sv->add_stream);
    for (auto& b : sv->stage) b.release();
    sv->add_host.release();
    sv->pool_mem.release();
    sv->add_stage.release();
    {
      std::lock_guard ilk(sv->ix->mu);
      sv->ix->n_encoded = sv->n_visible;
      sv->ix->server = nullptr;
    }
    delete sv;
  });
}

vc_status vc_server_submit_search(vc_server* sv, const float* queries, size_t nq, size_t k, uint64_t* ticket) {
  return guarded([&] {
    if (sv == nullptr || ticket == nullptr || (queries == nullptr && nq > 0))
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    if (nq < 1) fail(VC_ERR_INVALID_ARGUMENT, "a search request needs at least one query");
    if (nq > sv->p.max_batch_queries)
      fail(VC_ERR_INVALID_ARGUMENT, "the request holds more queries than max_batch_queries");
    if (k < 1 || k > sv->p.max_k) fail(VC_ERR_INVALID_ARGUMENT, "k must be in 1..max_k");
    auto r = std::make_unique<Request>();
    r->rows.assign(queries, queries + nq * sv->dim);
    r->n = nq;
    r->k = k;
    r->info.t_submit = now_s();
    r->info.slot = -1;
    {
      std::lock_guard lk(sv->mu);
      if (sv->stop) fail(VC_ERR_STATE, "the server is stopping");
      *ticket = sv->next_ticket++;
      sv->search_q.push_back(r.get());
      sv->reqs.emplace(*ticket, std::move(r));
      ++sv->not_done;
      ++sv->stats.requests;
    }
    sv->cv_work.notify_all();
  });
}

vc_status vc_server_submit_add(vc_server* sv, const float* vectors, size_t n, uint64_t* ticket) {
  return guarded([&] {
    if (sv == nullptr || ticket == nullptr || (vectors == nullptr && n > 0))
      fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    auto r = std::make_unique<Request>();
    r->is_add = true;
    r->rows.assign(vectors, vectors + n * sv->dim);
    r->n = n;
    r->info.t_submit = now_s();
    r->info.slot = -1;
    {
      std::lock_guard lk(sv->mu);
      if (sv->stop) fail(VC_ERR_STATE, "the server is stopping");
      *ticket = sv->next_ticket++;
      sv->add_q.push_back(r.get());
      sv->reqs.emplace(*ticket, std::move(r));
      ++sv->not_done;
    }
    sv->cv_add.notify_all();
  });
}

vc_status vc_server_wait(vc_server* sv, uint64_t ticket, float* scores, uint64_t* ids, vc_request_info* info) {
  return guarded([&] {
    if (sv == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null server");
    std::unique_ptr<Request> r;
    {
      std::unique_lock lk(sv->mu);
      auto it = sv->reqs.find(ticket);
      if (it == sv->reqs.end()) fail(VC_ERR_INVALID_ARGUMENT, "unknown ticket (never issued, or already waited)");
      Request* rp = it->second.get();
      sv->cv_done.wait(lk, [&] { return rp->done; });
      r = std::move(it->second);
      sv->reqs.erase(it);
    }
    if (info != nullptr) *info = r->info;
    if (r->status != VC_OK) fail(r->status, r->err);
    if (!r->is_add) {
      if ((scores == nullptr || ids == nullptr) && !r->scores.empty()) fail(VC_ERR_INVALID_ARGUMENT, "null output");
      if (!r->scores.empty()) {
        std::memcpy(scores, r->scores.data(), r->scores.size() * 4);
        std::memcpy(ids, r->ids.data(), r->ids.size() * 8);
      }
    }
  });
}

vc_status vc_server_flush(vc_server* sv) {
  return guarded([&] {
    if (sv == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock lk(sv->mu);
    sv->cv_done.wait(lk, [&] { return sv->not_done == 0; });
  });
}

vc_status vc_server_get_stats(vc_server* sv, vc_server_stats* out) {
  return guarded([&] {
    if (sv == nullptr || out == nullptr) fail(VC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard lk(sv->mu);
    *out = sv->stats;
    out->size = sv->n_visible;
    out->pool_bytes = sv->pool.cap();
  });
}

}  // extern "C"
