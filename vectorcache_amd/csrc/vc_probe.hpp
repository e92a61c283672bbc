// The learned probe-budget model of the bucketed index: an MLP
// 96 -> 256 -> 256 -> 1 (ReLU) that reads a query's top-64 routing scores and
// the sizes of its top-32 buckets and predicts log2 of the number of buckets
// it needs. bf16 weights on the matrix cores with fp32 accumulation; fp32
// master weights and Adam state. Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace vchip {

constexpr int kProbeIn = 96, kProbeHid = 256, kProbeBatch = 64;
constexpr int kProbeMaxP = 256;  // probe counts the model chooses from
constexpr int kProbeMaxBins = 64;
// fp32 parameter layout
constexpr int kPoW1 = 0;                               // [256][96]
constexpr int kPob1 = kPoW1 + kProbeHid * kProbeIn;    // [256]
constexpr int kPoW2 = kPob1 + kProbeHid;               // [256][256]
constexpr int kPob2 = kPoW2 + kProbeHid * kProbeHid;   // [256]
constexpr int kPow3 = kPob2 + kProbeHid;               // [256]
constexpr int kPob3 = kPow3 + kProbeHid;               // [1] (+3 pad)
constexpr int kProbeParams = kPob3 + 4;

struct ProbeDev {
  float* w;       // kProbeParams fp32 master weights
  float* m;       // Adam first moments
  float* v;       // Adam second moments
  __bf16* w1b;    // bf16 copy of W1 [256][96]
  __bf16* w2b;    // bf16 copy of W2 [256][256]
  float* grad;    // kProbeParams, written by the gradient-only step
  int* table;     // kProbeMaxBins probe counts
};

struct ProbeModel {
  void* mem = nullptr;
  ProbeDev d{};
  bool trained = false;
  float shift = 0.f;  // log2 offset applied to the prediction
  // probe table (when n_bins > 0): prediction y falls in bin
  // clamp(floor((y - y0) / y_step), 0, n_bins - 1), which opens table[bin]
  // buckets (table in device memory, kProbeMaxBins entries at mem's end)
  int n_bins = 0;
  float y0 = 0.f, y_step = 1.f;
  long step = 0;      // Adam steps taken
  void release() {
    if (mem != nullptr) (void)hipFree(mem);
    mem = nullptr;
    d = ProbeDev{};
    trained = false;
  }
};

// Allocates the model and initializes it from `seed` (He-uniform weights,
// zero biases and moments).
hipError_t probe_model_init(ProbeModel& pm, std::uint64_t seed, hipStream_t s);
// Uploads / downloads the fp32 master weights (bf16 copies refreshed).
hipError_t probe_model_set(ProbeModel& pm, const float* host_w, hipStream_t s);
hipError_t probe_model_get(const ProbeModel& pm, float* host_w, hipStream_t s);
hipError_t probe_model_set_table(ProbeModel& pm, const std::int32_t* host_counts, int n_bins, float y0, float y_step,
                                 hipStream_t s);

// features [nq][96] of routing lists (probe_s sorted desc, P >= 1 entries,
// -inf past the valid ones) and bucket sizes.
hipError_t launch_probe_features(const float* probe_s, const int* probe_b, const int* bsize, std::size_t nq, int P,
                                 float* feats, hipStream_t s);
// Training: n_steps minibatches of 64 rows (row ids order[step * 64 + i],
// clamped to [0, n_rows)) of feats / targets, Adam (lr, beta1, beta2, eps) from step pm.step; loss_out
// [n_steps] the minibatch mean squared errors. grads_only: one minibatch,
// gradients into pm.d.grad, no update.
hipError_t launch_probe_train(ProbeModel& pm, const float* feats, const float* targets, std::size_t n_rows,
                              const int* order, int n_steps,
                              float lr, float beta1, float beta2, float eps, bool grads_only, float* loss_out,
                              hipStream_t s);
// y[i] = model(feats[i]).
hipError_t launch_probe_predict(const ProbeModel& pm, const float* feats, std::size_t n, float* y, hipStream_t s);
// Routing lists -> features -> model -> probe_n[q] = clamp(table[bin(y)] or,
// without a table, ceil(2^(y + shift)), 1, valid entries of q's list), in one
// kernel.
hipError_t probe_counts_from_model(const ProbeModel& pm, std::size_t nq, int P, const float* probe_s,
                                   const int* probe_b, const int* bsize, int* probe_n, hipStream_t s);
// probe_n[q] = clamp(counts[q], 1, valid entries of q's list), at most P.
hipError_t launch_clamp_counts(std::size_t nq, int P, const std::int32_t* counts, const int* probe_b, int* probe_n,
                               hipStream_t s);

}  // namespace vchip
