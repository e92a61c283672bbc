// The flat index object behind vc_index_* (vc_capi.cpp) and the request
// server built on it (vc_server.cpp). Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <vector>

#include "vc_common.hpp"

struct vc_index {
  size_t dim = 0, bits = 0;
  vchip::CodeTables t{};
  int device = 0;
  hipStream_t stream = nullptr;
  vcapi::DevBuf calib;  // shift | scale | inv_scale (3 x dim floats)
  bool calibrated = false;
  std::vector<float> h_shift, h_scale;
  vcapi::DevBuf tiled;  // codes, tiled layout
  vcapi::DevBuf alpha;
  size_t cap = 0;  // vector capacity of tiled/alpha (multiple of 64)
  size_t n_encoded = 0;
  vcapi::DevBuf pending;  // staged raw vectors
  size_t n_pending = 0;
  bool prepared = false;
  vcapi::DevBuf scratch;
  vcapi::DevBuf io;  // host-call staging
  std::mutex mu;
  // set while a vc_server serves the index: the server owns the codes, and
  // the index's own entry points refuse to add, prepare, search or destroy
  vc_server* server = nullptr;
  // vc_multi sets that scan the index's codes: while any exists the index
  // refuses to add, prepare, calibrate or be destroyed
  int sets = 0;
};

namespace vcapi {

// Grows the code capacity of `ix` to hold `total` vectors (at least 1.5x the
// old capacity, a multiple of 64), keeping the encoded codes. The old buffers
// are freed after `s` drains: no other stream may still read them.
inline void index_reserve(vc_index* ix, size_t total, hipStream_t s) {
  if (total <= ix->cap) return;
  size_t ncap = std::max(total, ix->cap + ix->cap / 2);
  ncap = (ncap + 63) / 64 * 64;
  ix->tiled.reserve(vchip::tiled_bytes(ncap, ix->t.groups16), s, /*keep=*/true);
  ix->alpha.reserve(ncap * 4, s, /*keep=*/true);
  ix->cap = ncap;
}

inline void refuse_if_served(const vc_index* ix) {
  if (ix->server != nullptr)
    fail(VC_ERR_STATE, "the index is being served: submit through its vc_server, or destroy the server first");
}

inline void refuse_if_in_set(const vc_index* ix) {
  if (ix->sets > 0) fail(VC_ERR_STATE, "the index belongs to a vc_multi set: destroy the set first");
}

}  // namespace vcapi
