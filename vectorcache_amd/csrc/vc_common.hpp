// Host-side helpers shared by the C ABI translation units (vc_capi.cpp,
// vc_bucketed.cpp): error mapping, device buffers, and the search entry
// points other index types build on. Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

#include "vc_host.hpp"
#include "vc_internal.hpp"
#include "vectorcache_hip.h"

namespace vcapi {

// scan-kernel instrumentation (vc_set_profiling / vc_last_scan_stats): when
// on, the dominant scan kernel of a search is bracketed by HIP events on its
// launch stream (tile = query tile, ranges = work split; for the bucketed
// scan, tile 1 and ranges = 0).
struct ScanProbe {
  std::mutex mu;
  bool on = false;
  hipEvent_t start = nullptr, stop = nullptr;
  bool recorded = false;
  int tile = 0, ranges = 0;
};
extern ScanProbe g_probe;

extern thread_local std::string g_last_error;
// When set, the searches of this thread use this scan plan instead of
// plan_scan's (the autotuner's candidates).
extern thread_local const vchip::ScanPlan* g_plan_override;
extern std::atomic<int> g_forced_tile;
extern std::atomic<int> g_debug_flags;

struct VcError {
  vc_status status;
  std::string msg;
};

[[noreturn]] inline void fail(vc_status st, const std::string& msg) { throw VcError{st, msg}; }

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const vc_status st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? VC_ERR_OOM : VC_ERR_DEVICE;
  fail(st, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
vc_status guarded(F&& f) {
  try {
    f();
    return VC_OK;
  } catch (const VcError& e) {
    g_last_error = e.msg;
    return e.status;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return VC_ERR_INVALID_ARGUMENT;
  } catch (const std::out_of_range& e) {
    g_last_error = e.what();
    return VC_ERR_OUT_OF_RANGE;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return VC_ERR_OOM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return VC_ERR_DEVICE;
  }
}

inline void check_dim_bits(size_t dim, size_t bits) {
  if (dim == 0 || dim % 8 != 0) fail(VC_ERR_INVALID_ARGUMENT, "dim must be a positive multiple of 8");
  if (dim > vchip::kMaxDim) fail(VC_ERR_INVALID_ARGUMENT, "dim exceeds the maximum of 16384");
  if (bits < 2 || bits > 4) fail(VC_ERR_INVALID_ARGUMENT, "bits must be 2, 3, or 4");
}

// Growable device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void reserve(size_t need, hipStream_t s, bool keep = false) {
    if (need <= bytes) return;
    size_t nb = std::max(need, bytes + bytes / 2);
    void* q = nullptr;
    hip_check(hipMalloc(&q, nb), "hipMalloc");
    if (keep && p != nullptr && bytes > 0) hip_check(hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, s), "copy");
    if (p != nullptr) {
      hip_check(hipStreamSynchronize(s), "sync");
      hip_check(hipFree(p), "hipFree");
    }
    p = q;
    bytes = nb;
  }
  void release() {
    if (p != nullptr) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as(size_t byte_offset = 0) const {
    return reinterpret_cast<T*>(static_cast<char*>(p) + byte_offset);
  }
};

// Orders the work enqueued next on `consumer` after the work enqueued so far
// on `producer` (NULL: the null stream), with no host wait.
inline void stream_after(hipStream_t consumer, hipStream_t producer) {
  if (consumer == producer) return;
  hipEvent_t ev;
  hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
  const hipError_t e1 = hipEventRecord(ev, producer);
  const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(consumer, ev, 0) : e1;
  (void)hipEventDestroy(ev);
  hip_check(e2, "stream ordering");
}

// Carves one scratch allocation into aligned sub-buffers.
struct Carve {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t at = off;
    off = (off + bytes + 255) & ~size_t(255);
    return at;
  }
};

// Constant tables of one (dim, bits) configuration on the current device.
vchip::CodeTables make_tables(size_t dim, size_t bits);

// TQ+ calibration of device sample rows into device shift / scale / 1/scale.
void calibrate_device(const vchip::CodeTables& t, const float* sample_dev, size_t n, float* shift, float* scale,
                      float* inv_scale, DevBuf& scratch, hipStream_t s);

// Scan a tiled code set: queries (device) -> results (device).
struct SearchOut {
  size_t k_eff;
};

// Scratch of a search: a growable device buffer, or (when `fixed` is set) a
// region the caller owns of `fixed_cap` bytes, which is never grown.
struct ScratchArena {
  DevBuf* grow = nullptr;
  uint8_t* fixed = nullptr;
  size_t fixed_cap = 0;
  uint8_t* take(size_t bytes, hipStream_t s) {
    if (fixed != nullptr) {
      if (bytes > fixed_cap)
        fail(VC_ERR_INVALID_ARGUMENT, "search needs " + std::to_string(bytes) + " bytes of scratch; the fixed scratch "
                                      "holds " + std::to_string(fixed_cap) + " (fewer queries or a smaller k)");
      return fixed;
    }
    grow->reserve(bytes, s);
    return grow->as<uint8_t>();
  }
};

// Scan plan and scratch carve of one search of nq queries over n vectors:
// `bytes` of scratch (0 when there is nothing to scan), offsets into it.
struct SearchLayout {
  vchip::ScanPlan plan{};
  size_t k_eff = 0;
  size_t o_luts = 0, o_ls = 0, o_lb = 0, o_bc = 0, o_bt = 0, o_ps = 0, o_pi = 0, o_pc = 0, o_lt = 0;
  size_t bytes = 0;
};
SearchLayout search_layout(const vchip::CodeTables& t, size_t n, size_t nq, size_t k);

// Query state already on the device (a prepared-queries handle): nq rows of
// G x 32 LUT bytes, their scales and total biases (lut_bias + bias_corr).
struct PreLuts {
  const uint8_t* luts = nullptr;
  const float* scale = nullptr;
  const float* bias_total = nullptr;
};

// `pre` (not NULL): scan with those LUTs; q_dev is not read.
SearchOut search_tiled(const vchip::CodeTables& t, const uint8_t* tiled, const float* alpha, size_t n,
                       const float* q_dev, size_t nq, size_t k, const uint64_t* id_map_dev, float* out_s,
                       uint64_t* out_i, DevBuf& scratch, hipStream_t s, uint8_t** luts_out = nullptr,
                       size_t* luts_off = nullptr, ScratchArena* arena = nullptr, const PreLuts* pre = nullptr);

// A private non-blocking stream with its own device buffers, synchronized and
// released at scope exit (the stateless host-buffer entry points).
struct Scoped {
  hipStream_t s = nullptr;
  DevBuf a, scratch;
  Scoped() { hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
  ~Scoped() {
    (void)hipStreamSynchronize(s);
    a.release();
    scratch.release();
    (void)hipStreamDestroy(s);
  }
  Scoped(const Scoped&) = delete;
  Scoped& operator=(const Scoped&) = delete;
};

}  // namespace vcapi
