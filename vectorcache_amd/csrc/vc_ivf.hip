// gfx950 kernels of the bucketed (IVF) index: coarse-quantizer build,
// routing and probe planning.
//
// Bucket assignment and routing only choose which buckets a query opens;
// they never enter a score. So their dot products run on the matrix cores
// (bf16 inputs, fp32 accumulation). What does enter a score is exact and
// order-fixed: the residual each vector is encoded from (its normalized row
// minus its bucket's encode centroid, element by element) and the per-bucket
// score offset (a sequential double chain over the dimensions).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>

#include "vc_ivf.hpp"

namespace vchip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool better_ci(float s, int c, float ts, int tc) { return s > ts || (s == ts && c < tc); }

// ---------------------------------------------------------------------------
// C = A B^T on the matrix cores. A: M rows of K bf16 (row stride lda), B: N
// rows of K bf16 (row stride ldb), K a multiple of 32, both 16-byte aligned
// rows. Workgroup = 4 wavefronts in 2 x 2 over a 128 x 128 tile; a wavefront
// owns 64 x 64 = 4 x 4 v_mfma_f32_16x16x32_bf16 tiles. Operand layout of that
// instruction: lane i holds row i % 16 of A (column i % 16 of B^T) and eight
// consecutive k of group i / 16; result element j of lane i is row
// 4 (i / 16) + j, column i % 16.
// gemm_argmax_kernel keeps, per row, the best column over all N (score desc,
// column asc) and writes it to best_col.
// ---------------------------------------------------------------------------
constexpr int kGemmTile = 128;

__device__ __forceinline__ void load_frags(const __bf16* __restrict__ base, long rows, long r0, int ld, int k0,
                                           int lane, bf16x8 (&f)[4]) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const long r = min(r0 + 16 * t + (lane & 15), rows - 1);  // clamped: in bounds, masked at the output
    f[t] = *reinterpret_cast<const bf16x8*>(base + r * ld + k0 + 8 * (lane >> 4));
  }
}

__device__ __forceinline__ void tile_mma(const __bf16* __restrict__ A, long M, int lda, long m0,
                                         const __bf16* __restrict__ B, long N, int ldb, long n0, int K, int lane,
                                         f32x4 (&acc)[4][4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  bf16x8 fa[4], fb[4], na[4], nb[4];
  load_frags(A, M, m0, lda, 0, lane, fa);
  load_frags(B, N, n0, ldb, 0, lane, fb);
  for (int k0 = 0; k0 < K; k0 += 32) {
    const int kn = k0 + 32 < K ? k0 + 32 : k0;  // the last step reloads its own block (unused)
    load_frags(A, M, m0, lda, kn, lane, na);
    load_frags(B, N, n0, ldb, kn, lane, nb);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      fa[t] = na[t];
      fb[t] = nb[t];
    }
  }
}

__global__ __launch_bounds__(256) void gemm_argmax_kernel(const __bf16* __restrict__ A, long M, int lda,
                                                          const __bf16* __restrict__ B, long N, int ldb, int K,
                                                          int* __restrict__ best_col) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long m0 = (long)blockIdx.x * kGemmTile + (w >> 1) * 64;
  float bs[16];
  int bc[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    bs[r] = -INFINITY;
    bc[r] = 0x7fffffff;
  }
  for (long nt = 0; nt < N; nt += kGemmTile) {
    const long n0 = nt + (w & 1) * 64;
    f32x4 acc[4][4];
    tile_mma(A, M, lda, m0, B, N, ldb, n0, K, lane, acc);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const long col = n0 + 16 * b + (lane & 15);
      if (col < N) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (acc[a][b][j] > bs[4 * a + j]) {  // columns arrive in increasing order: ties keep the lower
              bs[4 * a + j] = acc[a][b][j];
              bc[4 * a + j] = (int)col;
            }
      }
    }
  }
  // the 16 lanes of a row group hold the same rows (different columns)
#pragma unroll
  for (int r = 0; r < 16; ++r)
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const float os = __shfl_xor(bs[r], o);
      const int oc = __shfl_xor(bc[r], o);
      if (better_ci(os, oc, bs[r], bc[r])) {
        bs[r] = os;
        bc[r] = oc;
      }
    }
  __shared__ float s_s[2][64];
  __shared__ int s_c[2][64];
  // wavefronts (wr, 1) hand their rows' best to (wr, 0)
  if ((w & 1) == 1 && (lane & 15) == 0) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 16 * (r >> 2) + 4 * (lane >> 4) + (r & 3);
      s_s[w >> 1][row] = bs[r];
      s_c[w >> 1][row] = bc[r];
    }
  }
  __syncthreads();
  if ((w & 1) == 0 && (lane & 15) == 0) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 16 * (r >> 2) + 4 * (lane >> 4) + (r & 3);
      float s = bs[r];
      int c = bc[r];
      if (better_ci(s_s[w >> 1][row], s_c[w >> 1][row], s, c)) c = s_c[w >> 1][row];
      if (m0 + row < M) best_col[m0 + row] = c;
    }
  }
}

// ---------------------------------------------------------------------------
// Routing top-P per row: the P best columns of a score row by (score desc,
// column asc), sorted. One workgroup per row. A threshold is read off 1024
// evenly spaced samples of the row (sorted), so that about 1.25 P + 4 N /
// 1024 entries lie at or above it; those entries are compacted into LDS (if
// fewer than P or more than the buffer, the threshold moves and the pass
// repeats) and bitonic-sorted. Every entry above the threshold is a
// candidate, so the result is the exact top P.
// ---------------------------------------------------------------------------
constexpr int kTopThreads = 256;
constexpr int kTopMaxP = 512;
constexpr int kTopCand = 2048;   // candidate buffer
constexpr int kTopSamples = 512;

__device__ __forceinline__ void bitonic_desc(float* s, int* c, int size, int tid) {
  for (int kk = 2; kk <= size; kk <<= 1)
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < size; i += kTopThreads) {
        const int p = i ^ j;
        if (p > i) {
          const bool desc = (i & kk) == 0;
          const bool pb = better_ci(s[p], c[p], s[i], c[i]);
          if (desc ? pb : !pb) {
            const float ts = s[i];
            const int tc = c[i];
            s[i] = s[p];
            c[i] = c[p];
            s[p] = ts;
            c[p] = tc;
          }
        }
      }
      __syncthreads();
    }
}

__device__ __forceinline__ uint32_t desc_key(float f) {
  // larger float -> larger key; NaN never wins (mapped to 0)
  const uint32_t b = __float_as_uint(f);
  if (f != f) return 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ __launch_bounds__(kTopThreads) void topk_rows_kernel(const float* __restrict__ scores, long M, int N,
                                                                long lds_, int P, int* __restrict__ out_c,
                                                                float* __restrict__ out_s) {
  __shared__ float cs[kTopCand];
  __shared__ int cc[kTopCand];
  __shared__ float smp[kTopSamples];
  __shared__ int smp_c[kTopSamples];
  __shared__ int n_cand, hist[256], sh_digit, sh_need, scan_buf[kTopThreads];
  const int tid = threadIdx.x;
  const int total = min(P, N);
  for (long row = blockIdx.x; row < M; row += gridDim.x) {
    const float* sr = scores + row * lds_;
    // ---- sampled threshold: at most 12 moves of the sample rank ----
    bool done = false;
    if (N <= kTopCand) {
      __syncthreads();
      if (tid == 0) n_cand = 0;
      __syncthreads();
      for (int c = tid; c < N; c += kTopThreads) {
        const float v = sr[c];
        if (v == v) {
          const int at = atomicAdd(&n_cand, 1);
          cs[at] = v;
          cc[at] = c;
        }
      }
      __syncthreads();
      done = true;  // every finite entry is a candidate
    } else {
      __syncthreads();
      for (int i = tid; i < kTopSamples; i += kTopThreads) {
        const float v = sr[(long)i * N / kTopSamples];
        smp[i] = v == v ? v : -INFINITY;
        smp_c[i] = i;
      }
      __syncthreads();
      bitonic_desc(smp, smp_c, kTopSamples, tid);
      int lo = -1, hi = kTopSamples;  // ranks known too high (too few above) / too low (too many)
      int r = (int)min((long)kTopSamples - 1, ((long)5 * total * kTopSamples) / (4L * N) + 4);
      for (int it = 0; it < 12 && r > lo && r < hi; ++it) {
        const float thr = smp[r];
        __syncthreads();
        if (tid == 0) n_cand = 0;
        __syncthreads();
        // eight loads in flight per thread (one wavefront load covers 256
        // consecutive scores), then their compares and appends
        constexpr int U = 8;
        for (int c0 = tid; c0 < N; c0 += kTopThreads * U) {
          float v[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int c = c0 + u * kTopThreads;
            v[u] = c < N ? sr[c] : 0.f;
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int c = c0 + u * kTopThreads;
            if (c < N && v[u] >= thr) {  // false for NaN
              const int at = atomicAdd(&n_cand, 1);
              if (at < kTopCand) {
                cs[at] = v[u];
                cc[at] = c;
              }
            }
          }
        }
        __syncthreads();
        const int nc = n_cand;
        if (nc >= total && nc <= kTopCand) {
          done = true;
          break;
        }
        if (nc < total) lo = r; else hi = r;
        r = hi == kTopSamples ? min(kTopSamples - 1, 2 * r + 8) : (lo + hi) / 2;
      }
    }
    if (!done) {
      // ---- exact fallback (degenerate rows: runs of equal scores): radix
      // select of the P-th key, the keys above it, the lowest-column ties ----
      uint32_t prefix = 0u, mask = 0u;
      int need = total;
      for (int shift = 24; shift >= 0; shift -= 8) {
        for (int d = tid; d < 256; d += kTopThreads) hist[d] = 0;
        __syncthreads();
        for (int c = tid; c < N; c += kTopThreads) {
          const uint32_t k = desc_key(sr[c]);
          if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255], 1);
        }
        __syncthreads();
        if (tid == 0) {
          int cum = 0, d = 255;
          for (; d > 0; --d) {
            if (cum + hist[d] >= need) break;
            cum += hist[d];
          }
          sh_digit = d;
          sh_need = need - cum;
        }
        __syncthreads();
        prefix |= (uint32_t)sh_digit << shift;
        mask |= 255u << shift;
        need = sh_need;
        __syncthreads();
      }
      if (tid == 0) n_cand = 0;
      __syncthreads();
      for (int c = tid; c < N; c += kTopThreads)
        if (desc_key(sr[c]) > prefix) {
          const int at = atomicAdd(&n_cand, 1);  // fewer than total <= kTopMaxP
          cs[at] = sr[c];
          cc[at] = c;
        }
      const int per = (N + kTopThreads - 1) / kTopThreads;
      const int c0 = tid * per, c1 = min(N, c0 + per);
      int mine = 0;
      for (int c = c0; c < c1; ++c) mine += desc_key(sr[c]) == prefix ? 1 : 0;
      scan_buf[tid] = mine;
      __syncthreads();
      for (int o = 1; o < kTopThreads; o <<= 1) {
        const int v = tid >= o ? scan_buf[tid - o] : 0;
        __syncthreads();
        scan_buf[tid] += v;
        __syncthreads();
      }
      int rank = scan_buf[tid] - mine;
      const int base = n_cand;
      __syncthreads();
      for (int c = c0; c < c1 && rank < need; ++c)
        if (desc_key(sr[c]) == prefix) {
          cs[base + rank] = sr[c];
          cc[base + rank] = c;
          ++rank;
        }
      __syncthreads();
      if (tid == 0) n_cand = min(total, base + need);
      __syncthreads();
    }
    const int nc = min(n_cand, kTopCand);
    int size = 1;
    while (size < nc) size <<= 1;
    for (int i = nc + tid; i < size; i += kTopThreads) {
      cs[i] = -INFINITY;
      cc[i] = 0x7fffffff;
    }
    __syncthreads();
    bitonic_desc(cs, cc, size, tid);
    for (int i = tid; i < P; i += kTopThreads) {
      const bool ok = i < total && i < nc;
      out_c[row * P + i] = ok ? cc[i] : -1;
      out_s[row * P + i] = ok ? cs[i] : -INFINITY;
    }
  }
}

// ---------------------------------------------------------------------------
// Probe planning: for each query the first probe_n[q] entries of its probe
// list, their running vector counts, and the work items they are cut into.
// One wavefront per query.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void plan_probes_kernel(long nq, int P, const int* __restrict__ probe_b,
                                                         const int* __restrict__ probe_n,
                                                         const int* __restrict__ bsize, int min_item,
                                                         int max_items, int* __restrict__ pre,
                                                         int* __restrict__ item_len, int* __restrict__ n_items) {
  const int lane = threadIdx.x;
  for (long q = blockIdx.x; q < nq; q += gridDim.x) {
    const int np = probe_n[q];
    int run = 0;
    for (int j0 = 0; j0 < np; j0 += 64) {
      const int j = j0 + lane;
      const int b = j < np ? probe_b[q * P + j] : -1;
      int v = b >= 0 ? bsize[b] : 0;
      // inclusive wave scan
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o) v += u;
      }
      if (j < np) pre[q * (P + 1) + j + 1] = run + v;
      run += __shfl(v, 63);
    }
    if (lane == 0) {
      pre[q * (P + 1)] = 0;
      // items of min_item vectors, lengthened (in whole min_item units) so a
      // query has at most max_items of them
      const long need = ((long)run + max_items - 1) / max_items;
      const int len = (int)(need <= min_item ? min_item : (need + min_item - 1) / min_item * min_item);
      item_len[q] = len;
      n_items[q] = (run + len - 1) / len;
    }
  }
}

// Exclusive prefix of per-query item counts into item_off [nq + 1] (one
// workgroup; nq is a request batch).
__global__ __launch_bounds__(1024) void scan_items_kernel(const int* __restrict__ n_items, long nq,
                                                          int* __restrict__ item_off) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const long per = (nq + 1023) / 1024;
  const long a = tid * per, b = min(nq, a + per);
  int s = 0;
  for (long q = a; q < b; ++q) s += n_items[q];
  part[tid] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - s;
  for (long q = a; q < b; ++q) {
    item_off[q] = run;
    run += n_items[q];
  }
  if (tid == 1023) item_off[nq] = part[1023];
}

// Score offsets of the probed buckets: (float) of the sequential double chain
// sum_d q_hat[d] * c[b][d] (each product exact in double), one thread per
// (query, entry).
__global__ __launch_bounds__(256) void probe_offsets_kernel(const float* __restrict__ qhat, long nq, int D,
                                                            const float* __restrict__ cent, int P,
                                                            const int* __restrict__ probe_b,
                                                            const int* __restrict__ probe_n,
                                                            float* __restrict__ off) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nq * P; t += (long)gridDim.x * blockDim.x) {
    const long q = t / P;
    const int j = (int)(t - q * P);
    if (j >= probe_n[q]) continue;
    if (cent == nullptr) {  // codes of the rows themselves: no offset
      off[t] = 0.f;
      continue;
    }
    const float* qr = qhat + q * D;
    const float* cr = cent + (long)probe_b[t] * D;
    double acc = 0.0;
    for (int d = 0; d < D; ++d) acc = fma((double)qr[d], (double)cr[d], acc);
    off[t] = (float)acc;
  }
}

// Fixed probe count: probe_n[q] = min(n, number of listed buckets).
__global__ void fixed_probes_kernel(long nq, int P, int n, const int* __restrict__ probe_b,
                                    int* __restrict__ probe_n) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nq; q += (long)gridDim.x * blockDim.x) {
    int c = 0;
    while (c < n && c < P && probe_b[q * P + c] >= 0) ++c;
    probe_n[q] = c;
  }
}

// ---------------------------------------------------------------------------
// Build-side kernels.
// ---------------------------------------------------------------------------

// Rows of a float matrix as bf16 rows of Kp (zero past D), scaled by inv[r].
__global__ void rows_to_bf16_kernel(const float* __restrict__ x, long n, int D, const float* __restrict__ inv,
                                    int Kp, __bf16* __restrict__ out) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n * Kp; t += (long)gridDim.x * blockDim.x) {
    const long r = t / Kp;
    const int d = (int)(t - r * Kp);
    const float v = d < D ? (inv != nullptr ? x[r * D + d] * inv[r] : x[r * D + d]) : 0.f;
    out[t] = (__bf16)v;
  }
}

// Residual rows of positions [p0, p0 + n): out[p - p0][d] = x[row][d] * inv[row]
// - cent[bucket][d], row = perm[p], bucket = bucket_of_pos[p] (each product
// and difference rounded: the index's definition of a residual).
__global__ void residual_rows_kernel(const float* __restrict__ x, int D, const float* __restrict__ inv,
                                     const uint32_t* __restrict__ perm, const int* __restrict__ pos_bucket,
                                     const float* __restrict__ cent, long p0, long n, float* __restrict__ out) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n * D; t += (long)gridDim.x * blockDim.x) {
    const long i = t / D;
    const int d = (int)(t - i * D);
    const long p = p0 + i;
    const long row = perm[p];
    const float v = x[row * D + d] * inv[row];
    out[t] = cent != nullptr ? v - cent[(long)pos_bucket[p] * D + d] : v;
  }
}

// Bucket sizes from assignments.
__global__ void count_buckets_kernel(const int* __restrict__ assign, long n, int* __restrict__ counts) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    atomicAdd(&counts[assign[t]], 1);
}

// Bucket of each position from the bucket offsets (one thread per bucket).
__global__ void fill_pos_bucket_kernel(const long* __restrict__ boff, const int* __restrict__ bsize, int L,
                                       int* __restrict__ pos_bucket) {
  for (int b = blockIdx.x; b < L; b += gridDim.x)
    for (int i = threadIdx.x; i < bsize[b]; i += blockDim.x) pos_bucket[boff[b] + i] = b;
}

// Centroids as normalized sums of their members (members in row order, sums
// in that order: deterministic). One workgroup per bucket; buckets with no
// member keep their centroid. rows: member rows sorted by bucket, starting at
// boff[b]; x * inv[row] is the normalized row.
__global__ __launch_bounds__(256) void centroid_sums_kernel(const float* __restrict__ x, int D,
                                                            const float* __restrict__ inv,
                                                            const uint32_t* __restrict__ rows,
                                                            const long* __restrict__ boff,
                                                            const int* __restrict__ bsize, int L,
                                                            float* __restrict__ cent) {
  extern __shared__ float red[];
  for (int b = blockIdx.x; b < L; b += gridDim.x) {
    const int n = bsize[b];
    if (n == 0) continue;
    const uint32_t* rb = rows + boff[b];
    float ss = 0.f;
    for (int d0 = 0; d0 < D; d0 += blockDim.x) {
      const int d = d0 + threadIdx.x;
      float s = 0.f;
      if (d < D) {
        int i = 0;
        for (; i + 8 <= n; i += 8) {
          float v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const uint32_t r = rb[i + u];
            v[u] = x[(long)r * D + d] * inv[r];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; i < n; ++i) {
          const uint32_t r = rb[i];
          s += x[(long)r * D + d] * inv[r];
        }
        red[d] = s;
        ss = fmaf(s, s, ss);
      }
    }
    // norm of the sum
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
    __shared__ float wsum[4];
    __syncthreads();
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = ss;
    __syncthreads();
    const float nrm = sqrtf(wsum[0] + wsum[1] + wsum[2] + wsum[3]);
    if (nrm > 0.f)
      for (int d = threadIdx.x; d < D; d += blockDim.x) cent[(long)b * D + d] = red[d] / nrm;
    __syncthreads();
  }
}

// Dot product of one normalized row with a vector held in LDS, by one
// wavefront (lanes over dimensions, then a wave sum).
__device__ __forceinline__ float wave_dot(const float* __restrict__ xr, float inv, const float* __restrict__ v,
                                          int D, int lane) {
  float s = 0.f;
  for (int d = lane; d < D; d += 64) s = fmaf(xr[d] * inv, v[d], s);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  return s;
}

// One split of every listed oversized bucket by one iteration of spherical
// 2-means: seeds = the first and the middle member (members in row order);
// members go to the nearer seed, each side's normalized mean becomes a child
// centroid, members go to the nearer child. side[i] (over the bucket's sorted
// members) = 1 for the second child; a split that leaves a side with fewer
// than n / 8 members cuts the member list in half instead. One workgroup per
// bucket; cent_a / cent_b receive the children.
__global__ __launch_bounds__(256) void split_buckets_kernel(const float* __restrict__ x, int D,
                                                            const float* __restrict__ inv,
                                                            const uint32_t* __restrict__ rows,
                                                            const long* __restrict__ boff,
                                                            const int* __restrict__ bsize,
                                                            const int* __restrict__ list, int n_list,
                                                            uint8_t* __restrict__ side, float* __restrict__ cent_a,
                                                            float* __restrict__ cent_b) {
  extern __shared__ float sm[];  // s1[D], s2[D], ca[D], cb[D]
  float* s1 = sm;
  float* s2 = sm + D;
  float* ca = sm + 2 * D;
  float* cb = sm + 3 * D;
  __shared__ int cnt[2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int li = blockIdx.x; li < n_list; li += gridDim.x) {
    const int b = list[li];
    const int n = bsize[b];
    const uint32_t* rb = rows + boff[b];
    uint8_t* sd = side + boff[b];
    __syncthreads();
    const uint32_t r1 = rb[0];
    for (int d = threadIdx.x; d < D; d += blockDim.x) s1[d] = x[(long)r1 * D + d] * inv[r1];
    __syncthreads();
    // second seed: the middle member (members are in row order, so two
    // members far apart in it; the member least similar to the first would
    // make an outlier a seed and split off a handful of rows per round)
    const uint32_t r2 = rb[n / 2];
    for (int d = threadIdx.x; d < D; d += blockDim.x) s2[d] = x[(long)r2 * D + d] * inv[r2];
    __syncthreads();
    for (int i = w; i < n; i += 4) {
      const uint32_t r = rb[i];
      const float a = wave_dot(x + (long)r * D, inv[r], s1, D, lane);
      const float c = wave_dot(x + (long)r * D, inv[r], s2, D, lane);
      if (lane == 0) sd[i] = c > a ? 1 : 0;
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      // side means (members in order), normalized
      for (int d = threadIdx.x; d < D; d += blockDim.x) {
        float sa = 0.f, sb = 0.f;
        for (int i = 0; i < n; ++i) {
          const uint32_t r = rb[i];
          const float v = x[(long)r * D + d] * inv[r];
          if (sd[i]) sb += v; else sa += v;
        }
        ca[d] = sa;
        cb[d] = sb;
      }
      __syncthreads();
      float na = 0.f, nb = 0.f;
      for (int d = lane; d < D; d += 64) {
        na = fmaf(ca[d], ca[d], na);
        nb = fmaf(cb[d], cb[d], nb);
      }
      for (int o = 32; o > 0; o >>= 1) {
        na += __shfl_xor(na, o);
        nb += __shfl_xor(nb, o);
      }
      na = sqrtf(na);
      nb = sqrtf(nb);
      __syncthreads();
      for (int d = threadIdx.x; d < D; d += blockDim.x) {
        ca[d] = na > 0.f ? ca[d] / na : 0.f;
        cb[d] = nb > 0.f ? cb[d] / nb : 0.f;
      }
      __syncthreads();
      if (pass == 1) break;
      if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
      __syncthreads();
      for (int i = w; i < n; i += 4) {
        const uint32_t r = rb[i];
        const float a = wave_dot(x + (long)r * D, inv[r], ca, D, lane);
        const float c = wave_dot(x + (long)r * D, inv[r], cb, D, lane);
        if (lane == 0) {
          const int s = c > a ? 1 : 0;
          sd[i] = (uint8_t)s;
          atomicAdd(&cnt[s], 1);
        }
      }
      __syncthreads();
      if (8 * min(cnt[0], cnt[1]) < n) {  // a side below n / 8: halves instead, so every round halves
        for (int i = threadIdx.x; i < n; i += blockDim.x) sd[i] = i >= n / 2 ? 1 : 0;
        __syncthreads();
      }
    }
    for (int d = threadIdx.x; d < D; d += blockDim.x) {
      cent_a[(long)li * D + d] = ca[d];
      cent_b[(long)li * D + d] = cb[d];
    }
  }
}

// Members of split buckets on side 1 move to their bucket's new id.
__global__ void apply_split_kernel(const uint32_t* __restrict__ rows, const long* __restrict__ boff,
                                   const int* __restrict__ bsize, const int* __restrict__ list, int n_list,
                                   const int* __restrict__ new_id, const uint8_t* __restrict__ side,
                                   int* __restrict__ assign) {
  for (int li = blockIdx.x; li < n_list; li += gridDim.x) {
    const int b = list[li];
    for (int i = threadIdx.x; i < bsize[b]; i += blockDim.x)
      if (side[boff[b] + i]) assign[rows[boff[b] + i]] = new_id[li];
  }
}


// dst[i] = src[rows[i]] (bf16 rows of Kp).
__global__ void gather_rows_bf16_kernel(const __bf16* __restrict__ src, int Kp, const uint32_t* __restrict__ rows,
                                        long n, __bf16* __restrict__ dst) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n * Kp; t += (long)gridDim.x * blockDim.x) {
    const long i = t / Kp;
    dst[t] = src[(long)rows[i] * Kp + (t - i * Kp)];
  }
}

// out[i] = x[rows[i]] * inv[rows[i]] (normalized float rows).
__global__ void normalized_rows_kernel(const float* __restrict__ x, int D, const float* __restrict__ inv,
                                       const uint32_t* __restrict__ rows, long n, float* __restrict__ out) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n * D; t += (long)gridDim.x * blockDim.x) {
    const long i = t / D;
    const long r = rows[i];
    out[t] = x[r * D + (t - i * D)] * inv[r];
  }
}

// Split children into the centroid table: bucket list[li] takes cent_a[li],
// the new bucket base + li takes cent_b[li].
__global__ void scatter_children_kernel(float* __restrict__ cent, int D, const int* __restrict__ list, int n_list,
                                        int base, const float* __restrict__ cent_a, const float* __restrict__ cent_b) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)n_list * D; t += (long)gridDim.x * blockDim.x) {
    const int li = (int)(t / D);
    const int d = (int)(t - (long)li * D);
    cent[(long)list[li] * D + d] = cent_a[t];
    cent[(long)(base + li) * D + d] = cent_b[t];
  }
}

// ids[p] = base + rows[p].
__global__ void widen_ids_kernel(const uint32_t* __restrict__ rows, long n, uint64_t base, uint64_t* __restrict__ ids) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    ids[t] = base + rows[t];
}


// dst[i] = src[rows[i]] (float rows of D).
__global__ void gather_rows_f32_kernel(const float* __restrict__ src, int D, const int* __restrict__ rows, long n,
                                       float* __restrict__ dst) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n * D; t += (long)gridDim.x * blockDim.x) {
    const long i = t / D;
    dst[t] = src[(long)rows[i] * D + (t - i * D)];
  }
}

// v[i] = map[v[i]].
__global__ void remap_kernel(int* __restrict__ v, long n, const int* __restrict__ map) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) v[t] = map[v[t]];
}

int grid_of(long work, int threads, int cap = 65536) {
  long g = (work + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace

// ---- launchers ---------------------------------------------------------------
// The routing scores are a plain GEMM (bf16 in, fp32 accumulate and out):
// rocBLAS's gemm_ex, which runs it about 3x faster than this file's MFMA tile
// kernel (whose fused argmax form still serves the k-means assignment). One
// handle per device, used under a lock (its stream is set per call).
hipError_t launch_gemm_scores(const __bf16* A, std::size_t M, int lda, const __bf16* B, std::size_t N, int ldb, int K,
                              float* C, std::size_t ldc, hipStream_t s) {
  if (M == 0 || N == 0) return hipSuccess;
  static std::mutex mu;
  static std::map<int, rocblas_handle> handles;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return hipErrorInvalidDevice;
  std::lock_guard lock(mu);
  rocblas_handle& h = handles[dev];
  if (h == nullptr && rocblas_create_handle(&h) != rocblas_status_success) {
    h = nullptr;
    return hipErrorNotInitialized;
  }
  if (rocblas_set_stream(h, s) != rocblas_status_success) return hipErrorInvalidHandle;
  // row-major C (M x N) = A (M x K) B^T is column-major C^T (N x M) = op(B^T memory) (A^T memory)
  const float alpha = 1.f, beta = 0.f;
  const rocblas_status st = rocblas_gemm_ex(
      h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)N, (rocblas_int)M, K, &alpha, B,
      rocblas_datatype_bf16_r, ldb, A, rocblas_datatype_bf16_r, lda, &beta, C, rocblas_datatype_f32_r,
      (rocblas_int)ldc, C, rocblas_datatype_f32_r, (rocblas_int)ldc, rocblas_datatype_f32_r, rocblas_gemm_algo_standard,
      0, 0);
  return st == rocblas_status_success ? hipSuccess : hipErrorLaunchFailure;
}

hipError_t launch_gemm_argmax(const __bf16* A, std::size_t M, int lda, const __bf16* B, std::size_t N, int ldb, int K,
                              int* best_col, hipStream_t s) {
  if (M == 0 || N == 0) return hipSuccess;
  hipLaunchKernelGGL(gemm_argmax_kernel, dim3((unsigned)((M + kGemmTile - 1) / kGemmTile)), dim3(256), 0, s, A,
                     (long)M, lda, B, (long)N, ldb, K, best_col);
  return hipGetLastError();
}

hipError_t launch_topk_rows(const float* scores, std::size_t M, int N, std::size_t ld, int P, int* out_c,
                            float* out_s, hipStream_t s) {
  if (M == 0) return hipSuccess;
  if (P < 1 || P > kTopMaxP) return hipErrorInvalidValue;
  hipLaunchKernelGGL(topk_rows_kernel, dim3(grid_of((long)M, 1, 65536)), dim3(kTopThreads), 0, s, scores, (long)M, N,
                     (long)ld, P, out_c, out_s);
  return hipGetLastError();
}

hipError_t launch_fixed_probes(std::size_t nq, int P, int n, const int* probe_b, int* probe_n, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(fixed_probes_kernel, dim3(grid_of((long)nq, 256)), dim3(256), 0, s, (long)nq, P, n, probe_b,
                     probe_n);
  return hipGetLastError();
}

hipError_t launch_plan_probes(std::size_t nq, int P, const int* probe_b, const int* probe_n, const int* bsize,
                              int min_item, int max_items, int* pre, int* item_len, int* n_items, int* item_off,
                              hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(plan_probes_kernel, dim3(grid_of((long)nq, 1, 65536)), dim3(64), 0, s, (long)nq, P, probe_b,
                     probe_n, bsize, min_item, max_items, pre, item_len, n_items);
  hipLaunchKernelGGL(scan_items_kernel, dim3(1), dim3(1024), 0, s, n_items, (long)nq, item_off);
  return hipGetLastError();
}

hipError_t launch_probe_offsets(const float* qhat, std::size_t nq, int D, const float* cent, int P,
                                const int* probe_b, const int* probe_n, float* off, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(probe_offsets_kernel, dim3(grid_of((long)nq * P, 256, 16384)), dim3(256), 0, s, qhat, (long)nq, D,
                     cent, P, probe_b, probe_n, off);
  return hipGetLastError();
}

hipError_t launch_rows_to_bf16(const float* x, std::size_t n, int D, const float* inv, int Kp, __bf16* out,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(rows_to_bf16_kernel, dim3(grid_of((long)n * Kp, 256, 65536)), dim3(256), 0, s, x, (long)n, D, inv,
                     Kp, out);
  return hipGetLastError();
}

hipError_t launch_residual_rows(const float* x, int D, const float* inv, const std::uint32_t* perm,
                                const int* pos_bucket, const float* cent, std::size_t p0, std::size_t n, float* out,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(residual_rows_kernel, dim3(grid_of((long)n * D, 256, 65536)), dim3(256), 0, s, x, D, inv, perm,
                     pos_bucket, cent, (long)p0, (long)n, out);
  return hipGetLastError();
}

hipError_t launch_count_buckets(const int* assign, std::size_t n, int* counts, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(count_buckets_kernel, dim3(grid_of((long)n, 256, 16384)), dim3(256), 0, s, assign, (long)n,
                     counts);
  return hipGetLastError();
}

hipError_t launch_fill_pos_bucket(const long* boff, const int* bsize, int L, int* pos_bucket, hipStream_t s) {
  if (L == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_pos_bucket_kernel, dim3(grid_of(L, 1, 65536)), dim3(256), 0, s, boff, bsize, L, pos_bucket);
  return hipGetLastError();
}

hipError_t launch_centroid_sums(const float* x, int D, const float* inv, const std::uint32_t* rows, const long* boff,
                                const int* bsize, int L, float* cent, hipStream_t s) {
  if (L == 0) return hipSuccess;
  hipLaunchKernelGGL(centroid_sums_kernel, dim3(grid_of(L, 1, 65536)), dim3(256), (size_t)D * 4, s, x, D, inv, rows,
                     boff, bsize, L, cent);
  return hipGetLastError();
}

hipError_t launch_split_buckets(const float* x, int D, const float* inv, const std::uint32_t* rows, const long* boff,
                                const int* bsize, const int* list, int n_list, std::uint8_t* side, float* cent_a,
                                float* cent_b, hipStream_t s) {
  if (n_list == 0) return hipSuccess;
  const size_t lds = (size_t)D * 16;
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)split_buckets_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(split_buckets_kernel, dim3(grid_of(n_list, 1, 65536)), dim3(256), lds, s, x, D, inv, rows, boff,
                     bsize, list, n_list, side, cent_a, cent_b);
  return hipGetLastError();
}

hipError_t launch_apply_split(const std::uint32_t* rows, const long* boff, const int* bsize, const int* list,
                              int n_list, const int* new_id, const std::uint8_t* side, int* assign, hipStream_t s) {
  if (n_list == 0) return hipSuccess;
  hipLaunchKernelGGL(apply_split_kernel, dim3(grid_of(n_list, 1, 65536)), dim3(256), 0, s, rows, boff, bsize, list,
                     n_list, new_id, side, assign);
  return hipGetLastError();
}


hipError_t launch_gather_rows_bf16(const __bf16* src, int Kp, const std::uint32_t* rows, std::size_t n, __bf16* dst,
                                   hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_rows_bf16_kernel, dim3(grid_of((long)n * Kp, 256, 65536)), dim3(256), 0, s, src, Kp, rows,
                     (long)n, dst);
  return hipGetLastError();
}

hipError_t launch_normalized_rows(const float* x, int D, const float* inv, const std::uint32_t* rows, std::size_t n,
                                  float* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(normalized_rows_kernel, dim3(grid_of((long)n * D, 256, 65536)), dim3(256), 0, s, x, D, inv, rows,
                     (long)n, out);
  return hipGetLastError();
}

hipError_t launch_scatter_children(float* cent, int D, const int* list, int n_list, int base, const float* cent_a,
                                   const float* cent_b, hipStream_t s) {
  if (n_list == 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_children_kernel, dim3(grid_of((long)n_list * D, 256, 65536)), dim3(256), 0, s, cent, D,
                     list, n_list, base, cent_a, cent_b);
  return hipGetLastError();
}

hipError_t launch_widen_ids(const std::uint32_t* rows, std::size_t n, std::uint64_t base, std::uint64_t* ids,
                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(widen_ids_kernel, dim3(grid_of((long)n, 256, 65536)), dim3(256), 0, s, rows, (long)n, base, ids);
  return hipGetLastError();
}


hipError_t launch_gather_rows_f32(const float* src, int D, const int* rows, std::size_t n, float* dst, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_rows_f32_kernel, dim3(grid_of((long)n * D, 256, 65536)), dim3(256), 0, s, src, D, rows,
                     (long)n, dst);
  return hipGetLastError();
}

hipError_t launch_remap(int* v, std::size_t n, const int* map, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(remap_kernel, dim3(grid_of((long)n, 256, 65536)), dim3(256), 0, s, v, (long)n, map);
  return hipGetLastError();
}

}  // namespace vchip
