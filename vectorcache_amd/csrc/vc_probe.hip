// gfx950 kernels of the learned probe-budget model (MLP 96 -> 256 -> 256 -> 1).
//
// Every matrix product runs on the matrix cores: v_mfma_f32_16x16x32_bf16
// with bf16 operands and fp32 accumulation (lane i holds row i % 16 of a
// fragment and eight consecutive k of group i / 16; result element j of lane
// i is row 4 (i / 16) + j, column i % 16). The rounding points are fixed so a
// CPU fp64 restatement can follow them: the features, W1, W2, the first
// hidden layer's activations and the backward signals that enter a product
// are rounded to bf16; biases, w3, the second hidden layer, the output, the
// loss, Adam and the master weights are fp32.
//
// Training runs in one workgroup (4 wavefronts), one launch for many
// minibatches of 64 rows: forward, backward and an Adam update per
// minibatch, with the small parameters (b1, b2, w3, b3) held in LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "vc_probe.hpp"

namespace vchip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kI = kProbeIn, kH = kProbeHid, kB = kProbeBatch;

// Fragment of rows r0 .. r0 + 15 and k0 .. k0 + 31 of a bf16 matrix whose
// element (r, k) sits at base[r * rs + k * ks].
__device__ __forceinline__ bf16x8 frag(const __bf16* base, int rs, int ks, int r0, int k0, int lane) {
  const __bf16* p = base + (long)(r0 + (lane & 15)) * rs + (long)(k0 + 8 * (lane >> 4)) * ks;
  bf16x8 f;
  if (ks == 1) {
    f = *reinterpret_cast<const bf16x8*>(p);
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) f[e] = p[(long)e * ks];
  }
  return f;
}

// acc[a][b] = sum over k < K of A(ar0 + 16 a + ., k) B(br0 + 16 b + ., k).
template <int NA, int NB>
__device__ __forceinline__ void mma(const __bf16* A, int ars, int aks, int ar0, const __bf16* B, int brs, int bks,
                                    int br0, int K, int lane, f32x4 (&acc)[NA][NB]) {
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 fa[NA], fb[NB];
#pragma unroll
    for (int a = 0; a < NA; ++a) fa[a] = frag(A, ars, aks, ar0 + 16 * a, k0, lane);
#pragma unroll
    for (int b = 0; b < NB; ++b) fb[b] = frag(B, brs, bks, br0 + 16 * b, k0, lane);
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
  }
}

// Row / column of element j of lane's result in tile (a, b) of a 64 x 64 block.
__device__ __forceinline__ int res_row(int a, int j, int lane) { return 16 * a + 4 * (lane >> 4) + j; }
__device__ __forceinline__ int res_col(int b, int lane) { return 16 * b + (lane & 15); }

// Features of one query from its routing list: f0 = best score; f1..63 =
// 16 (s0 - s_j) (the last valid entry repeats past the list); f64..95 =
// log2(1 + size of bucket j) / 10 (0 past the list).
__device__ __forceinline__ float feature(const float* ps, const int* pb, const int* bsize, int P, int f) {
  int nv = 0;
  while (nv < P && nv < 64 && pb[nv] >= 0) ++nv;
  if (nv == 0) return 0.f;
  const float s0 = ps[0];
  if (f == 0) return s0;
  if (f < 64) return 16.f * (s0 - ps[min(f, nv - 1)]);
  const int j = f - 64;
  return j < nv ? log2f(1.f + (float)bsize[pb[j]]) * 0.1f : 0.f;
}

// Forward of 64 rows held in X (LDS, bf16 [64][96]) by 4 wavefronts:
// H1 (LDS, bf16 [64][256]) = relu(X W1^T + b1); h2 (registers; wavefront w
// owns columns 64 w .. 64 w + 63) = relu(H1 W2^T + b2); yw[w][row] = the
// wavefront's part of h2 . w3. b1, b2, w3 are read from `small` (b1 | b2 | w3).
__device__ __forceinline__ void forward64(const ProbeDev& p, const float* small, const __bf16* X, __bf16* H1, float* yw,
                                          f32x4 (&h2)[4][4], int lane, int w) {
  const float* b1 = small;
  const float* b2 = small + kH;
  const float* w3 = small + 2 * kH;
  {
    f32x4 acc[4][4];
    mma<4, 4>(X, kI, 1, 0, p.w1b, kI, 1, 64 * w, kI, lane, acc);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int n = 64 * w + res_col(b, lane);
#pragma unroll
        for (int j = 0; j < 4; ++j) H1[res_row(a, j, lane) * kH + n] = (__bf16)fmaxf(acc[a][b][j] + b1[n], 0.f);
      }
  }
  __syncthreads();
  mma<4, 4>(H1, kH, 1, 0, p.w2b, kH, 1, 64 * w, kH, lane, h2);
  float yp[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) yp[r] = 0.f;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int n = 64 * w + res_col(b, lane);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        h2[a][b][j] = fmaxf(h2[a][b][j] + b2[n], 0.f);
        yp[4 * a + j] = fmaf(h2[a][b][j], w3[n], yp[4 * a + j]);
      }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r)
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) yp[r] += __shfl_xor(yp[r], o);
  if ((lane & 15) == 0)
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int j = 0; j < 4; ++j) yw[w * 64 + res_row(a, j, lane)] = yp[4 * a + j];
  __syncthreads();
}

__device__ __forceinline__ float out_of(const float* yw, float b3, int row) {
  return ((yw[row] + yw[64 + row]) + (yw[128 + row] + yw[192 + row])) + b3;
}

__device__ __forceinline__ void adam(float* w, float* m, float* v, int i, float g, float lr, float b1, float b2,
                                     float eps, float c1, float c2) {
  const float mi = fmaf(b1, m[i], (1.f - b1) * g);
  const float vi = fmaf(b2, v[i], (1.f - b2) * g * g);
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  w[i] = w[i] - lr * mh / (sqrtf(vh) + eps);
}

__global__ __launch_bounds__(256) void probe_train_kernel(ProbeDev p, const float* __restrict__ feats,
                                                          const float* __restrict__ targets,
                                                          const int* __restrict__ order, long n_rows, int n_steps,
                                                          float lr,
                                                          float beta1, float beta2, float eps, float pow1,
                                                          float pow2, int grads_only, float* __restrict__ loss_out) {
  __shared__ __align__(16) __bf16 X[kB * kI];
  __shared__ __align__(16) __bf16 H1[kB * kH];
  __shared__ __align__(16) __bf16 dZ2[kB * kH];
  __shared__ __align__(16) __bf16 dZ1[kB * kH];
  __shared__ float small[3 * kH + 4];  // b1 | b2 | w3 | b3
  __shared__ float yw[4 * 64];
  __shared__ float dy[kB];
  __shared__ float g_small[3 * kH + 4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int i = tid; i < kH; i += 256) {
    small[i] = p.w[kPob1 + i];
    small[kH + i] = p.w[kPob2 + i];
    small[2 * kH + i] = p.w[kPow3 + i];
  }
  if (tid == 0) small[3 * kH] = p.w[kPob3];
  __syncthreads();
  float c1 = pow1, c2 = pow2;  // beta^t before this launch's first step
  for (int st = 0; st < n_steps; ++st) {
    c1 *= beta1;
    c2 *= beta2;
    const float bc1 = 1.f - c1, bc2 = 1.f - c2;
    const int* rows = order + (long)st * kB;
    // row ids outside [0, n_rows) are clamped into it
    auto row_at = [&](int i) -> long { return min(max((long)rows[i], 0L), n_rows - 1); };
    for (int t = tid; t < kB * kI; t += 256) X[t] = (__bf16)feats[row_at(t / kI) * kI + t % kI];
    __syncthreads();
    f32x4 h2[4][4];
    forward64(p, small, X, H1, yw, h2, lane, w);
    if (tid < kB) {
      const float y = out_of(yw, small[3 * kH], tid);
      const float e = y - targets[row_at(tid)];
      dy[tid] = 2.f * e / (float)kB;
      yw[tid] = e * e;  // the forward's partial sums are consumed
    }
    __syncthreads();
    if (tid == 0) {
      float l = 0.f, gb3 = 0.f;
      for (int i = 0; i < kB; ++i) {
        l += yw[i];
        gb3 += dy[i];
      }
      if (loss_out != nullptr) loss_out[st] = l / (float)kB;
      g_small[3 * kH] = gb3;
    }
    // dw3, db2 and dZ2 = dy w3 relu'(z2)
    {
      float gw3[4] = {0.f, 0.f, 0.f, 0.f}, gb2[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int n = 64 * w + res_col(b, lane);
        const float w3n = small[2 * kH + n];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int m = res_row(a, j, lane);
            const float g = h2[a][b][j] > 0.f ? dy[m] * w3n : 0.f;
            dZ2[m * kH + n] = (__bf16)g;
            gw3[b] = fmaf(dy[m], h2[a][b][j], gw3[b]);
            gb2[b] += g;
          }
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        gw3[b] += __shfl_xor(gw3[b], 16);
        gw3[b] += __shfl_xor(gw3[b], 32);
        gb2[b] += __shfl_xor(gb2[b], 16);
        gb2[b] += __shfl_xor(gb2[b], 32);
        if (lane < 16) {
          g_small[2 * kH + 64 * w + res_col(b, lane)] = gw3[b];
          g_small[kH + 64 * w + res_col(b, lane)] = gb2[b];
        }
      }
    }
    __syncthreads();
    // dH1 = dZ2 W2 (the weights before this step's update), dZ1 = dH1 relu'(z1), db1
    {
      f32x4 acc[4][4];
      mma<4, 4>(dZ2, kH, 1, 0, p.w2b, 1, kH, 64 * w, kH, lane, acc);
      float gb1[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int kcol = 64 * w + res_col(b, lane);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int m = res_row(a, j, lane);
            const float g = (float)H1[m * kH + kcol] > 0.f ? acc[a][b][j] : 0.f;
            dZ1[m * kH + kcol] = (__bf16)g;
            gb1[b] += g;
          }
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        gb1[b] += __shfl_xor(gb1[b], 16);
        gb1[b] += __shfl_xor(gb1[b], 32);
        if (lane < 16) g_small[64 * w + res_col(b, lane)] = gb1[b];
      }
    }
    __syncthreads();
    // dW2[n][k] = sum_m dZ2[m][n] H1[m][k]: wavefront w owns rows n of block w
    for (int kb = 0; kb < 4; ++kb) {
      f32x4 acc[4][4];
      mma<4, 4>(dZ2, 1, kH, 64 * w, H1, 1, kH, 64 * kb, kB, lane, acc);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int n = 64 * w + res_row(a, j, lane), kk = 64 * kb + res_col(b, lane);
            const int i = kPoW2 + n * kH + kk;
            if (grads_only) {
              p.grad[i] = acc[a][b][j];
            } else {
              adam(p.w, p.m, p.v, i, acc[a][b][j], lr, beta1, beta2, eps, bc1, bc2);
              p.w2b[n * kH + kk] = (__bf16)p.w[i];
            }
          }
    }
    // dW1[k][j] = sum_m dZ1[m][k] X[m][j]: wavefront w owns rows k of block w
    {
      f32x4 acc[4][6];
      mma<4, 6>(dZ1, 1, kH, 64 * w, X, 1, kI, 0, kB, lane, acc);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 6; ++b)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int kk = 64 * w + res_row(a, j, lane), jj = res_col(b, lane);
            const int i = kPoW1 + kk * kI + jj;
            if (grads_only) {
              p.grad[i] = acc[a][b][j];
            } else {
              adam(p.w, p.m, p.v, i, acc[a][b][j], lr, beta1, beta2, eps, bc1, bc2);
              p.w1b[kk * kI + jj] = (__bf16)p.w[i];
            }
          }
    }
    // the small parameters: b1, b2, w3, b3
    for (int i = tid; i < 3 * kH + 1; i += 256) {
      const int gi = i < kH ? kPob1 + i : i < 2 * kH ? kPob2 + (i - kH) : i < 3 * kH ? kPow3 + (i - 2 * kH) : kPob3;
      if (grads_only) {
        p.grad[gi] = g_small[i];
      } else {
        adam(p.w, p.m, p.v, gi, g_small[i], lr, beta1, beta2, eps, bc1, bc2);
        small[i] = p.w[gi];
      }
    }
    __syncthreads();  // every wavefront's weight writes are complete before the next forward reads them
  }
}

__global__ __launch_bounds__(256) void probe_predict_kernel(ProbeDev p, const float* __restrict__ feats, long n,
                                                            float* __restrict__ y,
                                                            const float* __restrict__ probe_s,
                                                            const int* __restrict__ probe_b,
                                                            const int* __restrict__ bsize, int P, float shift,
                                                            int n_bins, float y0, float y_step,
                                                            int* __restrict__ probe_n) {
  __shared__ __align__(16) __bf16 X[kB * kI];
  __shared__ __align__(16) __bf16 H1[kB * kH];
  __shared__ float small[3 * kH + 4];
  __shared__ float yw[4 * 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int i = tid; i < kH; i += 256) {
    small[i] = p.w[kPob1 + i];
    small[kH + i] = p.w[kPob2 + i];
    small[2 * kH + i] = p.w[kPow3 + i];
  }
  if (tid == 0) small[3 * kH] = p.w[kPob3];
  for (long r0 = (long)blockIdx.x * kB; r0 < n; r0 += (long)gridDim.x * kB) {
    __syncthreads();
    for (int t = tid; t < kB * kI; t += 256) {
      const long r = min(r0 + t / kI, n - 1);
      const int f = t % kI;
      const float v = feats != nullptr ? feats[r * kI + f]
                                       : feature(probe_s + r * P, probe_b + r * P, bsize, P, f);
      X[t] = (__bf16)v;
    }
    __syncthreads();
    f32x4 h2[4][4];
    forward64(p, small, X, H1, yw, h2, lane, w);
    if (tid < kB && r0 + tid < n) {
      const long r = r0 + tid;
      const float out = out_of(yw, small[3 * kH], tid);
      if (y != nullptr) y[r] = out;
      if (probe_n != nullptr) {
        const int* pb = probe_b + r * P;
        int nv = 0;
        while (nv < P && pb[nv] >= 0) ++nv;
        float want;
        if (n_bins > 0) {
          const float fb = floorf((out - y0) / y_step);
          want = (float)p.table[(int)fminf(fmaxf(fb, 0.f), (float)(n_bins - 1))];
        } else {
          want = ceilf(exp2f(fminf(out + shift, 30.f)));
        }
        probe_n[r] = nv == 0 ? 0 : (int)fminf(fmaxf(want, 1.f), (float)nv);
      }
    }
  }
}

__global__ void probe_features_kernel(const float* __restrict__ probe_s, const int* __restrict__ probe_b,
                                      const int* __restrict__ bsize, long nq, int P, float* __restrict__ feats) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nq * kI; t += (long)gridDim.x * blockDim.x) {
    const long q = t / kI;
    feats[t] = feature(probe_s + q * P, probe_b + q * P, bsize, P, (int)(t - q * kI));
  }
}

__global__ void clamp_counts_kernel(long nq, int P, const int32_t* __restrict__ counts,
                                    const int* __restrict__ probe_b, int* __restrict__ probe_n) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nq; q += (long)gridDim.x * blockDim.x) {
    int nv = 0;
    while (nv < P && probe_b[q * P + nv] >= 0) ++nv;
    probe_n[q] = nv == 0 ? 0 : max(1, min(counts[q], nv));
  }
}

__global__ void probe_bf16_kernel(ProbeDev p) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kH * kH; i += gridDim.x * blockDim.x) {
    if (i < kH * kI) p.w1b[i] = (__bf16)p.w[kPoW1 + i];
    p.w2b[i] = (__bf16)p.w[kPoW2 + i];
  }
}

int grid_n(long work, int threads, int cap = 65536) {
  long g = (work + threads - 1) / threads;
  return (int)(g < 1 ? 1 : g > cap ? cap : g);
}

// splitmix64: the initial weights' generator
uint64_t mix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace

hipError_t probe_model_init(ProbeModel& pm, std::uint64_t seed, hipStream_t s) {
  if (pm.mem == nullptr) {
    const size_t bytes = (size_t)kProbeParams * 4 * 4 + (size_t)(kH * kI + kH * kH) * 2 + kProbeMaxBins * 4;
    hipError_t e = hipMalloc(&pm.mem, bytes);
    if (e != hipSuccess) return e;
    char* b = static_cast<char*>(pm.mem);
    pm.d.w = reinterpret_cast<float*>(b);
    pm.d.m = pm.d.w + kProbeParams;
    pm.d.v = pm.d.m + kProbeParams;
    pm.d.grad = pm.d.v + kProbeParams;
    pm.d.w1b = reinterpret_cast<__bf16*>(pm.d.grad + kProbeParams);
    pm.d.w2b = pm.d.w1b + kH * kI;
    pm.d.table = reinterpret_cast<int*>(pm.d.w2b + kH * kH);
  }
  pm.n_bins = 0;
  std::vector<float> w(kProbeParams, 0.f);
  uint64_t st = seed;
  auto uni = [&](float lim) { return (float)(((double)(mix64(st) >> 11) * 0x1.0p-53) * 2.0 - 1.0) * lim; };
  for (int i = 0; i < kH * kI; ++i) w[kPoW1 + i] = uni(sqrtf(6.f / kI));
  for (int i = 0; i < kH * kH; ++i) w[kPoW2 + i] = uni(sqrtf(6.f / kH));
  for (int i = 0; i < kH; ++i) w[kPow3 + i] = uni(sqrtf(6.f / kH));
  hipError_t e = hipMemsetAsync(pm.d.m, 0, (size_t)kProbeParams * 4 * 3, s);
  if (e != hipSuccess) return e;
  pm.step = 0;
  pm.trained = false;
  return probe_model_set(pm, w.data(), s);
}

hipError_t probe_model_set(ProbeModel& pm, const float* host_w, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(pm.d.w, host_w, (size_t)kProbeParams * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(probe_bf16_kernel, dim3(grid_n((long)kH * kH, 256)), dim3(256), 0, s, pm.d);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t probe_model_set_table(ProbeModel& pm, const std::int32_t* host_counts, int n_bins, float y0, float y_step,
                                 hipStream_t s) {
  if (n_bins < 0 || n_bins > kProbeMaxBins || !(y_step > 0.f)) return hipErrorInvalidValue;
  if (n_bins > 0) {
    hipError_t e = hipMemcpyAsync(pm.d.table, host_counts, (size_t)n_bins * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
  }
  pm.n_bins = n_bins;
  pm.y0 = y0;
  pm.y_step = y_step;
  return hipSuccess;
}

hipError_t probe_model_get(const ProbeModel& pm, float* host_w, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(host_w, pm.d.w, (size_t)kProbeParams * 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t launch_probe_features(const float* probe_s, const int* probe_b, const int* bsize, std::size_t nq, int P,
                                 float* feats, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(probe_features_kernel, dim3(grid_n((long)nq * kI, 256)), dim3(256), 0, s, probe_s, probe_b, bsize,
                     (long)nq, P, feats);
  return hipGetLastError();
}

hipError_t launch_probe_train(ProbeModel& pm, const float* feats, const float* targets, std::size_t n_rows,
                              const int* order, int n_steps, float lr, float beta1, float beta2, float eps, bool grads_only, float* loss_out,
                              hipStream_t s) {
  if (n_steps <= 0) return hipSuccess;
  const float pow1 = (float)std::pow((double)beta1, (double)pm.step);
  const float pow2 = (float)std::pow((double)beta2, (double)pm.step);
  if (n_rows == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(probe_train_kernel, dim3(1), dim3(256), 0, s, pm.d, feats, targets, order, (long)n_rows,
                     grads_only ? 1 : n_steps, lr, beta1, beta2, eps, pow1, pow2, grads_only ? 1 : 0, loss_out);
  if (!grads_only) pm.step += n_steps;
  return hipGetLastError();
}

hipError_t launch_probe_predict(const ProbeModel& pm, const float* feats, std::size_t n, float* y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(probe_predict_kernel, dim3(grid_n((long)n, kB, 4096)), dim3(256), 0, s, pm.d, feats, (long)n, y,
                     nullptr, nullptr, nullptr, 0, 0.f, 0, 0.f, 1.f, nullptr);
  return hipGetLastError();
}

hipError_t probe_counts_from_model(const ProbeModel& pm, std::size_t nq, int P, const float* probe_s,
                                   const int* probe_b, const int* bsize, int* probe_n, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(probe_predict_kernel, dim3(grid_n((long)nq, kB, 4096)), dim3(256), 0, s, pm.d, nullptr, (long)nq,
                     nullptr, probe_s, probe_b, bsize, P, pm.shift, pm.n_bins, pm.y0, pm.y_step, probe_n);
  return hipGetLastError();
}

hipError_t launch_clamp_counts(std::size_t nq, int P, const std::int32_t* counts, const int* probe_b, int* probe_n,
                               hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(clamp_counts_kernel, dim3(grid_n((long)nq, 256)), dim3(256), 0, s, (long)nq, P, counts, probe_b,
                     probe_n);
  return hipGetLastError();
}

}  // namespace vchip
