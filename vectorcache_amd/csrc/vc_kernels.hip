// gfx950 kernels of the TurboQuant encode + search path.
//
// Compiled with -ffp-contract=off: every fused multiply-add that the
// reference CPU path performs is an explicit __fmaf_rn / fma here, every
// other multiply and add is rounded on its own, so scores, alpha and LUTs are
// bit-identical to the reference's AVX2 PERM0 path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "vc_internal.hpp"
#include "vc_ivf.hpp"

namespace vchip {
namespace {

constexpr float kMinNorm = 1e-10f;

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Correctly rounded sqrt / division (IEEE, as the reference's sqrtf and '/').
// Note: HIP's __fsqrt_rn maps to the approximate native sqrt unless
// OCML_BASIC_ROUNDED_OPERATIONS is defined, so it is not used here.
__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float div_rn(float a, float b) { return a / b; }
constexpr int kCandCap = 512;   // per-query candidate buffer of the scan
constexpr int kScanThreads = 1024;
constexpr int kLdsBudget = 160 * 1024;

// ---------------------------------------------------------------------------
// Shared device helpers
// ---------------------------------------------------------------------------

// Rotation of one row held in LDS by a single wavefront (one-round SRHT of
// the reference, src/transform/rotation.cpp:38-121,185-205):
//   y[i] = (x[perm[i]] * inv) * sign[i]      (scattered through inv_perm)
//   per block of B: butterflies len = 1 .. B/2, (a + b, a - b)
//   y *= 1 / sqrt(B)
// The butterfly pairs and their order are the reference's, so the float
// result is identical however lanes share the work.
__device__ void rotate_row_lds(const CodeTables& t, const float* __restrict__ xr, float inv,
                               float* __restrict__ y, int lane) {
  const int D = t.dim;
  for (int j = lane; j < D; j += 64) {
    const int i = (int)t.inv_perm[j];
    y[i] = (xr[j] * inv) * t.signs[i];
  }
  __syncthreads();
  const int B = t.block, npairs = D >> 1;
  const int lhb = __ffs(B) - 2;  // log2(B / 2)
  for (int lg = 0; (1 << lg) < B; ++lg) {
    const int len = 1 << lg;
    for (int p = lane; p < npairs; p += 64) {
      const int blk = p >> lhb, q = p & ((B >> 1) - 1);
      const int i = blk * B + ((q >> lg) << (lg + 1)) + (q & (len - 1));
      const float u = y[i], v = y[i + len];
      y[i] = u + v;
      y[i + len] = u - v;
    }
    __syncthreads();
  }
  for (int i = lane; i < D; i += 64) y[i] = y[i] * t.inv_sqrt_block;
  __syncthreads();
}

// Norm of a row staged in LDS: lanes 0..7 run the reference's chains.
__device__ float norm8_lds(const float* __restrict__ x, int D, int lane) {
  float c = 0.f;
  if (lane < 8) {
    constexpr int kAhead = 32;
    const int steps = D / 8;
    for (int t0 = 0; t0 < steps; t0 += kAhead) {
      float v[kAhead];
#pragma unroll
      for (int u = 0; u < kAhead; ++u) v[u] = x[lane + 8 * min(t0 + u, steps - 1)];  // in bounds, no branch
#pragma unroll
      for (int u = 0; u < kAhead; ++u)
        if (t0 + u < steps) c = __fmaf_rn(v[u], v[u], c);
    }
  }
  c = c + __shfl_xor(c, 1);
  c = c + __shfl_xor(c, 2);
  c = c + __shfl_xor(c, 4);
  return __shfl(sqrt_rn(c), 0);
}

// Row norm of the reference (src/encode/encode.cpp:16-26): 8 interleaved
// chains c[j % 8] += x_j^2 (fused), tree ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)).
// seq: the chains summed in order c0 + c1 + ... + c7 instead (the landmark
// index's l2-normalize, src/index.cpp:21-50, whose double sqrt of a float
// sum rounds to the same float as sqrtf).
// Called with lanes grouped by 8: lane & 7 is the chain. Returns sqrtf.
__device__ float norm8(const float* __restrict__ xr, int D, int chain, bool active, bool seq = false) {
  // the chain is sequential; its loads are not: fetch 32 ahead
  constexpr int kAhead = 32;
  float c = 0.f;
  const int steps = D / 8;
  if (active) {
    for (int t0 = 0; t0 < steps; t0 += kAhead) {
      float v[kAhead];
#pragma unroll
      for (int u = 0; u < kAhead; ++u) v[u] = t0 + u < steps ? xr[chain + 8 * (t0 + u)] : 0.f;
#pragma unroll
      for (int u = 0; u < kAhead; ++u)
        if (t0 + u < steps) c = __fmaf_rn(v[u], v[u], c);
    }
  }
  if (seq) {
    const int base = (int)(threadIdx.x & 63) & ~7;
    float sum = __shfl(c, base);
#pragma unroll
    for (int j = 1; j < 8; ++j) sum = sum + __shfl(c, base + j);
    return sqrt_rn(sum);
  }
  c = c + __shfl_xor(c, 1);
  c = c + __shfl_xor(c, 2);
  c = c + __shfl_xor(c, 4);
  return sqrt_rn(c);
}

// Inverse row norms as the encode computes them (8 fused chains, tree,
// correctly rounded sqrt; 0 below 1e-10), and optionally the normalized rows
// x * inv. Eight lanes per row.
__global__ __launch_bounds__(256) void row_inv_norms_kernel(const float* __restrict__ x, long n, int D,
                                                            float* __restrict__ inv_out, float* __restrict__ qhat) {
  const int chain = threadIdx.x & 7;
  for (long r0 = (blockIdx.x * (long)blockDim.x + threadIdx.x) >> 3; r0 < (n + 7) / 8 * 8;
       r0 += ((long)gridDim.x * blockDim.x) >> 3) {
    const bool active = r0 < n;  // whole groups of eight lanes share a row
    const long r = active ? r0 : 0;
    const float nrm = norm8(x + r * D, D, chain, active);
    const float inv = nrm > kMinNorm ? div_rn(1.f, nrm) : 0.f;
    if (active) {
      if (chain == 0) inv_out[r] = inv;
      if (qhat != nullptr)
        for (int d = chain; d < D; d += 8) qhat[r * D + d] = x[r * D + d] * inv;
    }
  }
}

// ---------------------------------------------------------------------------
// Encode (src/encode/encode.cpp:112-231): one wavefront per workgroup, rows
// in batches of 8 (the norm phase keeps all 64 lanes busy as 8 rows x 8
// chains). Alpha uses the reference's 4 double chains in element order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void encode_kernel(CodeTables t, const float* __restrict__ x,
                                                    long n, long out_offset,
                                                    uint8_t* __restrict__ tiled,
                                                    uint8_t* __restrict__ packed,
                                                    float* __restrict__ rotated,
                                                    float* __restrict__ alpha) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int D = t.dim;
  float* y = reinterpret_cast<float*>(smem);
  uint8_t* code = smem + (size_t)D * 4;
  const int lane = threadIdx.x;
  const bool quant = tiled != nullptr || packed != nullptr || alpha != nullptr;
  const int nbounds = (1 << t.bits) - 1;

  for (long batch = blockIdx.x; batch * 8 < n; batch += gridDim.x) {
    const long r8 = batch * 8 + (lane >> 3);
    const float nrm_mine = norm8(x + (r8 < n ? r8 : 0) * (long)D, D, lane & 7, r8 < n, t.norm_seq != 0);
    const int rows = (int)((n - batch * 8) < 8 ? (n - batch * 8) : 8);
    for (int rr = 0; rr < rows; ++rr) {
      const long row = batch * 8 + rr;
      const float nv = __shfl(nrm_mine, rr * 8);
      float inv = nv > kMinNorm ? div_rn(1.f, nv) : 0.f;
      if (t.norm_seq) inv = nv < kMinNorm ? 0.f : div_rn(1.f, nv);
      __syncthreads();
      if (t.prerotated) {
        // rows already in the rotated basis (landmark rows): unit = x * inv
        for (int i = lane; i < D; i += 64) y[i] = x[row * (long)D + i] * inv;
        __syncthreads();
      } else {
        rotate_row_lds(t, x + row * (long)D, inv, y, lane);
      }
      if (rotated != nullptr)
        for (int i = lane; i < D; i += 64) rotated[row * (long)D + i] = y[i];
      if (!quant) continue;
      // quantize: code = #{boundaries b : value > b} (strict)
      for (int i = lane; i < D; i += 64) {
        const float v = y[i];
        const float c = t.shift ? (v + t.shift[i]) * t.scale[i] : v;
        unsigned k = 0;
        for (int b = 0; b < nbounds; ++b) k += c > t.boundaries[b] ? 1u : 0u;
        code[i] = (uint8_t)k;
      }
      __syncthreads();
      // alpha = norm / (float) inner, inner = (d0 + d1) + (d2 + d3) over 4
      // sequential double chains d[j % 4] (fused multiply-add).
      double d = 0.0;
      if (lane < 4) {
        for (int j = lane; j < D; j += 4) {
          const double yv = (double)y[j];
          const int k = code[j];
          if (t.shift == nullptr) {
            d = fma(yv, (double)t.centroids[k], d);
          } else {
            const double rec = fma((double)t.centroids[k], (double)t.inv_scale[j], -(double)t.shift[j]);
            d = fma(yv, rec, d);
          }
        }
      }
      const double pair = d + __shfl_xor(d, 1);
      const double inner = pair + __shfl_xor(pair, 2);
      const long v = out_offset + row;
      if (lane == 0 && alpha != nullptr) alpha[v] = inner > 1e-12 ? div_rn(nv, (float)inner) : 0.f;
      if (tiled != nullptr) {
        const int G = t.groups, G16 = t.groups16;
        for (int c16 = lane; c16 < G16; c16 += 64) {
          uint32_t w[4] = {0u, 0u, 0u, 0u};
          for (int e = 0; e < 16; ++e) {
            const int g = c16 * 16 + e;
            uint32_t byte = 0;
            if (g < G) {
              if (t.bits == 2) {
                const uint8_t* c4 = code + g * 4;
                byte = (uint32_t(c4[0]) << 6) | (uint32_t(c4[1]) << 4) | (uint32_t(c4[2]) << 2) | c4[3];
              } else {
                byte = (uint32_t(code[2 * g]) << 4) | code[2 * g + 1];
              }
            }
            w[e >> 2] |= byte << ((e & 3) * 8);
          }
          uint4* dst = reinterpret_cast<uint4*>(tiled + ((v >> 6) * G16 + c16) * 1024 + (v & 63) * 16);
          *dst = make_uint4(w[0], w[1], w[2], w[3]);
        }
      }
      if (packed != nullptr) {
        const int bpp = D / 8;
        for (int c8 = lane; c8 < bpp; c8 += 64) {
          for (int p = 0; p < t.bits; ++p) {
            uint32_t byte = 0;
            for (int kk = 0; kk < 8; ++kk) byte |= ((uint32_t(code[c8 * 8 + kk]) >> p) & 1u) << (7 - kk);
            packed[row * (long)(t.bits * bpp) + p * bpp + c8] = (uint8_t)byte;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Encode, tiled fast path: blocks of B = 64 R elements, one wavefront per row,
// 16 rows (a quarter of a 64-vector code tile) per wavefront iteration.
//   * norms: 8 rows x 8 chains at a time (all 64 lanes, the reference order);
//   * rotation: scatter through inv_perm into LDS, then per block lane l holds
//     elements l R .. l R + R - 1: butterflies with len < R in registers, the
//     rest across lanes (partner l ^ len / R) in the reference's stage order;
//   * quantize through the cell table (one LDS read per element);
//   * alpha: the reference's 4 sequential double chains give a value within
//     (D/4 + 3) u sum|t| of the exact sum; a plain-double wavefront sum with
//     its own bound brackets it. When both ends of the bracket round to the
//     same float, that float IS the reference's (float) inner; otherwise the
//     wavefront runs the reference's chains exactly (rare);
//   * codes are staged per row in LDS and written as whole 16-byte pieces.
// ---------------------------------------------------------------------------
struct CellDev {
  uint32_t base;
  float bound, c_lo, c_hi;
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ unsigned quant_code(const CodeTables& t, const CellDev* __restrict__ cells, float v,
                                               float& cen) {
  if (t.use_cells) {
    const float f = __fmaf_rn(v, t.cell_inv_w, t.cell_off);
    int ci = 0;
    if (f >= 255.f) ci = 255;
    else if (f >= 0.f) ci = (int)f;
    const CellDev e = cells[ci];
    const bool up = v > e.bound;
    cen = up ? e.c_hi : e.c_lo;
    return e.base + (up ? 1u : 0u);
  }
  unsigned k = 0;
  for (int b = 0; b < (1 << t.bits) - 1; ++b) k += v > t.boundaries[b] ? 1u : 0u;
  cen = t.centroids[k];
  return k;
}

constexpr int kEncodeWaves = 4;
constexpr int kEncodeRows = 8;

__host__ __device__ inline size_t enc_align(size_t v) { return (v + 15) & ~size_t(15); }
__host__ __device__ inline size_t enc_wave_bytes(int D, int G16) {
  return enc_align((size_t)D * 4) + enc_align((size_t)kEncodeRows * G16 * 16) + 2 * kEncodeRows * 4;
}
__host__ __device__ inline size_t enc_lds_bytes(int D, int G16) {
  return enc_align((size_t)D * 4) + 256 * 16 + kEncodeWaves * enc_wave_bytes(D, G16);
}

template <int R>
__global__ __launch_bounds__(64 * kEncodeWaves) void encode_tile_kernel(CodeTables t, const float* __restrict__ x,
                                                                        long n, long out_offset,
                                                                        uint8_t* __restrict__ tiled,
                                                                        float* __restrict__ alpha_out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int D = t.dim, G16 = t.groups16, B = t.block, nblk = D / B;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t* ps = reinterpret_cast<uint32_t*>(smem);
  CellDev* cells = reinterpret_cast<CellDev*>(smem + enc_align((size_t)D * 4));
  unsigned char* wb = smem + enc_align((size_t)D * 4) + 256 * 16 + wave * enc_wave_bytes(D, G16);
  float* y = reinterpret_cast<float*>(wb);
  uint8_t* codes = wb + enc_align((size_t)D * 4);
  float* nrm = reinterpret_cast<float*>(codes + enc_align((size_t)kEncodeRows * G16 * 16));
  float* alp = nrm + kEncodeRows;

  for (int j = threadIdx.x; j < D; j += blockDim.x) ps[j] = t.perm_sign[j];
  if (t.use_cells)
    for (int j = threadIdx.x; j < 256 * 4; j += blockDim.x)
      reinterpret_cast<uint32_t*>(cells)[j] = reinterpret_cast<const uint32_t*>(t.cells)[j];
  __syncthreads();

  const int row_bytes = G16 * 16;
  const int cpb = t.bits == 2 ? 4 : 2;  // codes per byte
  const long s_begin = out_offset / kEncodeRows, s_end = (out_offset + n + kEncodeRows - 1) / kEncodeRows;
  const long stride = (long)gridDim.x * kEncodeWaves;
  for (long st = s_begin + (long)blockIdx.x * kEncodeWaves + wave; st < s_end; st += stride) {
    const long v0 = st * kEncodeRows;
    for (int i = lane; i < kEncodeRows * row_bytes / 4; i += 64) reinterpret_cast<uint32_t*>(codes)[i] = 0u;
    for (int rnd = 0; rnd < kEncodeRows / 8; ++rnd) {
      const int q = rnd * 8 + (lane >> 3);
      const long row = v0 + q - out_offset;
      const bool ok = row >= 0 && row < n;
      const float nv = norm8(x + (ok ? row : 0) * (long)D, D, lane & 7, ok);
      if ((lane & 7) == 0) nrm[q] = nv;
    }
    wave_lds_sync();
    for (int q = 0; q < kEncodeRows; ++q) {
      const long row = v0 + q - out_offset;
      if (row < 0 || row >= n) continue;  // uniform across the wavefront
      const float nv = nrm[q];
      const float inv = nv > kMinNorm ? div_rn(1.f, nv) : 0.f;
      const float* xr = x + row * (long)D;
#pragma unroll 8
      for (int j = lane; j < D; j += 64) {
        const uint32_t w = ps[j];
        y[w & 0x7fffffffu] = (xr[j] * inv) * ((w >> 31) ? -1.f : 1.f);
      }
      wave_lds_sync();
      double sum = 0.0, asum = 0.0;
      for (int b = 0; b < nblk; ++b) {
        float* yb = y + b * B + lane * R;
        float v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = yb[r];
#pragma unroll
        for (int len = 1; len < R; len <<= 1) {
#pragma unroll
          for (int i = 0; i < R; i += 2 * len) {
#pragma unroll
            for (int j = i; j < i + len; ++j) {
              const float a = v[j], c = v[j + len];
              v[j] = a + c;
              v[j + len] = a - c;
            }
          }
        }
#pragma unroll
        for (int lm = 1; lm < 64; lm <<= 1) {
          // lower lane: a + b; upper lane: a - b = partner - self; fma with +-1 is exact
          const float sg = (lane & lm) ? -1.f : 1.f;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const float p = __shfl_xor(v[r], lm);
            v[r] = __fmaf_rn(sg, v[r], p);
          }
        }
        uint64_t cw = 0;  // this lane's codes of the block, packed as flat bytes
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float yv = v[r] * t.inv_sqrt_block;
          yb[r] = yv;  // post-rotation values, read only by the exact alpha path
          const int j = b * B + lane * R + r;
          float cen;
          const unsigned k = quant_code(t, cells, t.shift ? (yv + t.shift[j]) * t.scale[j] : yv, cen);
          double term;
          if (t.shift == nullptr)
            term = (double)yv * (double)cen;
          else
            term = (double)yv * fma((double)cen, (double)t.inv_scale[j], -(double)t.shift[j]);
          sum += term;
          asum += fabs(term);
          const int slot = r % cpb;  // position inside its byte, first code highest
          const int shift = (cpb - 1 - slot) * (8 / cpb);
          cw |= (uint64_t)k << (8 * (r / cpb) + shift);
        }
        uint8_t* dst = codes + q * row_bytes + (b * B + lane * R) / cpb;
#pragma unroll
        for (int i = 0; i < R / cpb; ++i) dst[i] = (uint8_t)(cw >> (8 * i));
      }
      for (int o = 32; o > 0; o >>= 1) {
        sum += __shfl_xor(sum, o);
        asum += __shfl_xor(asum, o);
      }
      const double u = 1.1102230246251565e-16;  // 2^-53
      double bound = (double)(D / 4 + D / 64 + 16) * u * asum * 1.01;
      bound += 2.0 * u * (fabs(sum) + bound);
      const double lo = sum - bound, hi = sum + bound;
      float a = 0.f;
      bool exact = (t.debug_flags & 1) == 0;
      if (exact && lo > 1e-12 && __double2float_rn(lo) == __double2float_rn(hi))
        a = div_rn(nv, __double2float_rn(hi));
      else if (exact && hi <= 1e-12)
        a = 0.f;
      else
        exact = false;
      if (!exact) {
        // the reference's chains, in order (src/encode/encode.cpp:112-157)
        wave_lds_sync();
        double d = 0.0;
        if (lane < 4) {
          for (int j = lane; j < D; j += 4) {
            const float yv = y[j];
            float cen;
            (void)quant_code(t, cells, t.shift ? (yv + t.shift[j]) * t.scale[j] : yv, cen);
            if (t.shift == nullptr)
              d = fma((double)yv, (double)cen, d);
            else
              d = fma((double)yv, fma((double)cen, (double)t.inv_scale[j], -(double)t.shift[j]), d);
          }
        }
        const double pr = d + __shfl_xor(d, 1);
        const double inner = pr + __shfl_xor(pr, 2);
        a = inner > 1e-12 ? div_rn(nv, (float)inner) : 0.f;
      }
      if (lane == 0) alp[q] = a;
      wave_lds_sync();
    }
    // write the 16 rows' codes as 16-byte pieces and their alphas
    for (int c16 = lane / kEncodeRows; c16 < G16; c16 += 64 / kEncodeRows) {
      const int q = lane % kEncodeRows;
      const long v = v0 + q;
      const long row = v - out_offset;
      if (row >= 0 && row < n) {
        const uint4 val = *reinterpret_cast<const uint4*>(codes + q * row_bytes + c16 * 16);
        *reinterpret_cast<uint4*>(tiled + ((v >> 6) * G16 + c16) * 1024 + (v & 63) * 16) = val;
      }
    }
    if (lane < kEncodeRows) {
      const long v = v0 + lane;
      const long row = v - out_offset;
      if (row >= 0 && row < n) alpha_out[v] = alp[lane];
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Streaming encode in two passes over the rows.
//
// norm_rows_kernel: the reference's row norm (8 fused chains c[j % 8] and the
// tree ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)), src/encode/encode.cpp:16-26) for
// 8 rows per wavefront, lane = (row, chain); the chains are sequential, their
// loads are issued kNormAhead at a time.
//
// encode_row_kernel: one wavefront per row. The row is prefetched into
// registers one row ahead, staged in the wavefront's LDS buffer, gathered into
// rotation order (y[i] = (x[perm[i]] * inv) * sign[i]), transformed and
// quantized in registers, and its codes leave as 16-byte pieces.
// ---------------------------------------------------------------------------
constexpr int kNormAhead = 48;

__global__ __launch_bounds__(256) void norm_rows_kernel(const float* __restrict__ x, long n, int D,
                                                        float* __restrict__ norms) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 3, c = lane & 7;
  const long n_waves = (long)gridDim.x * (blockDim.x >> 6);
  const int steps = D / 8;
  for (long g = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g * 8 < n; g += n_waves) {
    const long row = g * 8 + q;
    const bool ok = row < n;
    const float* xr = x + (ok ? row : 0) * (long)D + c;
    float acc = 0.f;
    for (int t0 = 0; t0 < steps; t0 += kNormAhead) {
      float v[kNormAhead];
#pragma unroll
      for (int u = 0; u < kNormAhead; ++u) v[u] = xr[8 * min(t0 + u, steps - 1)];  // clamped: always in bounds
#pragma unroll
      for (int u = 0; u < kNormAhead; ++u)
        if (t0 + u < steps) acc = __fmaf_rn(v[u], v[u], acc);
    }
    acc = acc + __shfl_xor(acc, 1);
    acc = acc + __shfl_xor(acc, 2);
    acc = acc + __shfl_xor(acc, 4);
    if (ok && c == 0) norms[row] = sqrt_rn(acc);
  }
}

constexpr int kRowMaxWaves = 16;

// per wavefront: the row (D floats) and a staging area that holds the row's
// codes, later (exact alpha path only) one rotated block at a time
__host__ __device__ inline size_t row_stage_bytes(int D, int G16, int B) {
  const size_t c = (size_t)G16 * 16, y = (size_t)B * 4;
  return enc_align(c > y ? c : y);
}
__host__ __device__ inline size_t row_wave_bytes(int D, int G16, int B) {
  return enc_align((size_t)D * 4) + row_stage_bytes(D, G16, B);
}
// shared by the workgroup: gather table, cell table, TQ+ shift | scale | 1/scale
__host__ __device__ inline size_t row_shared_bytes(int D, bool tq) {
  return enc_align((size_t)D * 4) + 256 * 16 + (tq ? 3 * enc_align((size_t)D * 4) : 0);
}

// Value of lane ^ LM: DPP quad permutes for 1 and 2, ds_swizzle (xor within
// 32 lanes) for 4, 8 and 16, a lane permute for 32.
template <int LM>
__device__ __forceinline__ float lane_xor(float v) {
  const int b = __builtin_bit_cast(int, v);
  if constexpr (LM == 1) return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(b, 0xB1, 0xF, 0xF, false));
  else if constexpr (LM == 2) return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(b, 0x4E, 0xF, 0xF, false));
  else if constexpr (LM < 32) return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(b, (LM << 10) | 0x1F));
  else return __shfl_xor(v, 32);
}

// One cross-lane butterfly stage: lower lane a + b, upper lane a - b =
// partner - self (fma with +-1 is exact).
template <int R, int LM>
__device__ __forceinline__ void lane_stage(float (&v)[R], int lane) {
  const float sg = (lane & LM) ? -1.f : 1.f;
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = __fmaf_rn(sg, v[r], lane_xor<LM>(v[r]));
}

// Block b of the rotation for this lane: v[r] = y[b * B + lane * R + r]
// (gathered from the row in X, butterflies in the lane then across lanes,
// scaled by 1 / sqrt(B)); the reference's pairs and order
// (src/transform/rotation.cpp:38-121,185-205).
template <int R>
__device__ __forceinline__ void rotate_block(const CodeTables& t, const uint32_t* __restrict__ pg,
                                             const float* __restrict__ X, float inv, int b, int lane, float (&v)[R]) {
  const int B = t.block;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t w = pg[b * B + lane * R + r];
    v[r] = (X[w & 0x7fffffffu] * inv) * ((w >> 31) ? -1.f : 1.f);
  }
#pragma unroll
  for (int len = 1; len < R; len <<= 1) {
#pragma unroll
    for (int i = 0; i < R; i += 2 * len) {
#pragma unroll
      for (int j = i; j < i + len; ++j) {
        const float a = v[j], c = v[j + len];
        v[j] = a + c;
        v[j + len] = a - c;
      }
    }
  }
  lane_stage<R, 1>(v, lane);
  lane_stage<R, 2>(v, lane);
  lane_stage<R, 4>(v, lane);
  lane_stage<R, 8>(v, lane);
  lane_stage<R, 16>(v, lane);
  lane_stage<R, 32>(v, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = v[r] * t.inv_sqrt_block;
}

// rotate_block over a lane-major gather table: pgl[b * B + r * 64 + lane] =
// perm_gather[b * B + lane * R + r], so the 64 lanes of one (block, r) read 64
// consecutive words (no bank conflicts; the by-destination table read with a
// lane stride of R words costs an R-way conflict per read). The sign is
// applied by flipping the product's sign bit, which equals the product with
// -1 exactly.
template <int R>
__device__ __forceinline__ void rotate_block_lm(const CodeTables& t, const uint32_t* __restrict__ pgl,
                                                const float* __restrict__ X, float inv, int b, int lane,
                                                float (&v)[R]) {
  const int B = t.block;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t w = pgl[b * B + r * 64 + lane];
    v[r] = __uint_as_float(__float_as_uint(X[w & 0x7fffffffu] * inv) ^ (w & 0x80000000u));
  }
#pragma unroll
  for (int len = 1; len < R; len <<= 1) {
#pragma unroll
    for (int i = 0; i < R; i += 2 * len) {
#pragma unroll
      for (int j = i; j < i + len; ++j) {
        const float a = v[j], c = v[j + len];
        v[j] = a + c;
        v[j + len] = a - c;
      }
    }
  }
  lane_stage<R, 1>(v, lane);
  lane_stage<R, 2>(v, lane);
  lane_stage<R, 4>(v, lane);
  lane_stage<R, 8>(v, lane);
  lane_stage<R, 16>(v, lane);
  lane_stage<R, 32>(v, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = v[r] * t.inv_sqrt_block;
}

// The butterfly stage across lanes 16 apart (X32: 32 apart) for pairs of
// registers with the gfx950 row swaps: v_permlane16_swap exchanges the odd
// 16-lane rows of its first operand with the even rows of its second
// (v_permlane32_swap: the upper 32 lanes of the first with the lower 32 of the
// second). After one swap each lane holds both members of a butterfly pair of
// one register (register r in the lower row, r + 1 in the upper); the sum and
// difference are formed in the lane, and a second swap puts every result back
// in its own lane and register. Same operations and operands as lane_stage
// (a + c, a - c), with no LDS traffic.
template <int R, bool X32>
__device__ __forceinline__ void swap_stage(float (&v)[R]) {
#pragma unroll
  for (int r = 0; r < R; r += 2) {
    const auto p = X32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(v[r]), __float_as_uint(v[r + 1]), false, false)
                       : __builtin_amdgcn_permlane16_swap(__float_as_uint(v[r]), __float_as_uint(v[r + 1]), false, false);
    const float a = __uint_as_float(p[0]), c = __uint_as_float(p[1]);
    const float sum = a + c, dif = a - c;
    const auto q = X32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(sum), __float_as_uint(dif), false, false)
                       : __builtin_amdgcn_permlane16_swap(__float_as_uint(sum), __float_as_uint(dif), false, false);
    v[r] = __uint_as_float(q[0]);
    v[r + 1] = __uint_as_float(q[1]);
  }
}

// The stage across lanes 8 apart by a DPP row rotation by 8 (lane ^ 8 inside
// each 16-lane row).
template <int R>
__device__ __forceinline__ void dpp8_stage(float (&v)[R], int lane) {
  const float sg = (lane & 8) ? -1.f : 1.f;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const float pv = __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v[r]), 0x128, 0xF, 0xF, false));
    v[r] = __fmaf_rn(sg, v[r], pv);
  }
}

// Lane-major slot of rotated position j (block B = 64 R): the position
// b * B + lane * R + r is kept at b * B + r * 64 + lane.
__device__ __forceinline__ int lane_major_slot(int j, int B, int R) {
  const int b = j / B, w = j - b * B;
  return b * B + (w % R) * 64 + w / R;
}

// The NB code bytes of one lane's block (little-endian in cw) with one store.
template <int NB>
__device__ __forceinline__ void store_code_bytes(uint8_t* dst, uint64_t cw) {
  if constexpr (NB == 8) *reinterpret_cast<uint64_t*>(dst) = cw;
  else if constexpr (NB == 4) *reinterpret_cast<uint32_t*>(dst) = (uint32_t)cw;
  else if constexpr (NB == 2) *reinterpret_cast<uint16_t*>(dst) = (uint16_t)cw;
  else *dst = (uint8_t)cw;
}

// R = block / 64 values per lane per rotation block; NXM >= D / 64 values per
// lane per row. alpha_io[out_offset + row] holds the row norm on entry and
// alpha on exit.
template <int R, int NXM>
__global__ __launch_bounds__(64 * kRowMaxWaves) void encode_row_kernel(CodeTables t, const float* __restrict__ x,
                                                                    long n, long out_offset,
                                                                    uint8_t* __restrict__ tiled,
                                                                    float* __restrict__ alpha_io) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int D = t.dim, G16 = t.groups16, B = t.block;
  const int NX = D / 64, nblk = D / B;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  const bool tq = t.shift != nullptr;
  uint32_t* pg = reinterpret_cast<uint32_t*>(smem);
  CellDev* cells = reinterpret_cast<CellDev*>(smem + enc_align((size_t)D * 4));
  float* tsh = reinterpret_cast<float*>(smem + enc_align((size_t)D * 4) + 256 * 16);
  float* tsc = tsh + enc_align((size_t)D * 4) / 4;
  float* tis = tsc + enc_align((size_t)D * 4) / 4;
  unsigned char* wb = smem + row_shared_bytes(D, tq) + wave * row_wave_bytes(D, G16, B);
  float* X = reinterpret_cast<float*>(wb);
  unsigned char* stage = wb + enc_align((size_t)D * 4);
  uint8_t* codes = stage;
  float* ys = reinterpret_cast<float*>(stage);

  // gather and TQ+ tables in lane-major order (lane_major_slot)
  for (int j = threadIdx.x; j < D; j += blockDim.x) pg[lane_major_slot(j, B, R)] = t.perm_gather[j];
  if (t.use_cells)
    for (int j = threadIdx.x; j < 256 * 4; j += blockDim.x)
      reinterpret_cast<uint32_t*>(cells)[j] = reinterpret_cast<const uint32_t*>(t.cells)[j];
  if (tq)
    for (int j = threadIdx.x; j < D; j += blockDim.x) {
      const int sl = lane_major_slot(j, B, R);
      tsh[sl] = t.shift[j];
      tsc[sl] = t.scale[j];
      tis[sl] = t.inv_scale[j];
    }
  __syncthreads();

  const int cpb = t.bits == 2 ? 4 : 2;  // codes per byte
  // A workgroup encodes whole 64-row output tiles, so every byte of a tile's
  // 1 KiB code lines is written through one L2 (one XCD) before eviction.
  const long tile0 = out_offset >> 6;
  const long ntile = ((out_offset + n + 63) >> 6) - tile0;
  for (long ti = blockIdx.x; ti < ntile; ti += gridDim.x)
  for (int tr = wave; tr < 64; tr += waves) {
    const long row = ((tile0 + ti) << 6) + tr - out_offset;
    if (row < 0 || row >= n) continue;  // wave-uniform
    {
      // the row, lane + 64 * i (clamped loads: always in bounds), into X
      const float* xr = x + row * (long)D + lane;
      float xv[NXM];
#pragma unroll
      for (int i = 0; i < NXM; ++i) xv[i] = xr[64 * min(i, NX - 1)];
#pragma unroll
      for (int i = 0; i < NXM; ++i)
        if (i < NX) X[lane + 64 * i] = xv[i];
    }
    // bytes past the last group of the final 16-byte piece stay zero
    for (int i = (t.groups & ~15) + lane; i < G16 * 16; i += 64) codes[i] = 0;
    const float nv = alpha_io[out_offset + row];
    wave_lds_sync();
    const float inv = nv > kMinNorm ? div_rn(1.f, nv) : 0.f;

    double sum = 0.0, asum = 0.0;
#pragma unroll 1
    for (int b = 0; b < nblk; ++b) {
      float v[R];
      rotate_block_lm<R>(t, pg, X, inv, b, lane, v);
      uint64_t cw = 0;  // this lane's codes of the block, packed as flat bytes
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float yv = v[r];
        const int jl = b * B + r * 64 + lane;  // lane-major slot of position b * B + lane * R + r
        float cen;
        const unsigned k = quant_code(t, cells, tq ? (yv + tsh[jl]) * tsc[jl] : yv, cen);
        double term;
        if (!tq)
          term = (double)yv * (double)cen;
        else
          term = (double)yv * fma((double)cen, (double)tis[jl], -(double)tsh[jl]);
        sum += term;
        asum += fabs(term);
        const int slot = r % cpb;  // position inside its byte, first code highest
        const int shift = (cpb - 1 - slot) * (8 / cpb);
        cw |= (uint64_t)k << (8 * (r / cpb) + shift);
      }
      uint8_t* dst = codes + (b * B + lane * R) / cpb;
      if (cpb == 2) store_code_bytes<R / 2>(dst, cw);
      else store_code_bytes<(R >= 4 ? R / 4 : 1)>(dst, cw);
    }
    for (int o = 32; o > 0; o >>= 1) {
      sum += __shfl_xor(sum, o);
      asum += __shfl_xor(asum, o);
    }
    wave_lds_sync();  // the row's codes are complete
    const long v = out_offset + row;
    for (int c16 = lane; c16 < G16; c16 += 64)
      *reinterpret_cast<uint4*>(tiled + ((v >> 6) * G16 + c16) * 1024 + (v & 63) * 16) =
          *reinterpret_cast<const uint4*>(codes + c16 * 16);

    // certified alpha (the bound of encode_tile_kernel), else the reference's
    // sequential chains over the rotated values, one block at a time
    const double u = 1.1102230246251565e-16;  // 2^-53
    double bound = (double)(D / 4 + D / 64 + 16) * u * asum * 1.01;
    bound += 2.0 * u * (fabs(sum) + bound);
    const double lo = sum - bound, hi = sum + bound;
    float a = 0.f;
    bool exact = (t.debug_flags & 1) == 0;
    if (exact && lo > 1e-12 && __double2float_rn(lo) == __double2float_rn(hi))
      a = div_rn(nv, __double2float_rn(hi));
    else if (exact && hi <= 1e-12)
      a = 0.f;
    else
      exact = false;
    if (!exact) {
      wave_lds_sync();  // codes written out: the staging area takes rotated blocks
      double d = 0.0;   // lane m < 4: chain m
#pragma unroll 1
      for (int b = 0; b < nblk; ++b) {
        float yb[R];
        rotate_block_lm<R>(t, pg, X, inv, b, lane, yb);
#pragma unroll
        for (int r = 0; r < R; ++r) ys[lane * R + r] = yb[r];
        wave_lds_sync();
        if (lane < 4) {
          for (int jj = lane; jj < B; jj += 4) {
            const int jl = b * B + (jj % R) * 64 + jj / R;  // lane-major slot of position b * B + jj
            const float yv = ys[jj];
            float cen;
            (void)quant_code(t, cells, tq ? (yv + tsh[jl]) * tsc[jl] : yv, cen);
            if (!tq)
              d = fma((double)yv, (double)cen, d);
            else
              d = fma((double)yv, fma((double)cen, (double)tis[jl], -(double)tsh[jl]), d);
          }
        }
        wave_lds_sync();
      }
      const double pr = d + __shfl_xor(d, 1);
      const double inner = pr + __shfl_xor(pr, 2);
      a = inner > 1e-12 ? div_rn(nv, (float)inner) : 0.f;
    }
    if (lane == 0) alpha_io[v] = a;
    wave_lds_sync();  // X and the staging area reused by the next row
  }
}

// ---------------------------------------------------------------------------
// One-pass encode for the common shapes: D = 64 NX with NX in
// {2, 4, 6, 8, 12, 24} (rotation block B = 64 R, R = NX & -NX), cell
// quantization. A wavefront takes eight consecutive rows of a 64-row output
// tile (a workgroup owns whole tiles, so a tile's code lines are written
// through one L2):
//   1. their norms: the reference's 8 fused chains c[j % 8] and the tree
//      ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)) (src/encode/encode.cpp:16-26), lane
//      = (row, chain), loads straight from global memory;
//   2. each row in turn: a second read of the row (from L2 / MALL: step 1 has
//      just read it; the next row's loads are in flight while a row is
//      encoded), staged in the wavefront's LDS slot, gathered through the
//      lane-major table (address (w << 2) + slot: the sign bit shifts out;
//      the sign rides on the multiplier, x * (-inv) = (x * inv) * -1 exactly),
//      butterflies, cell quantization, alpha and 16-byte code pieces.
// The separate norm pass of the two-pass kernels (a whole extra read of the
// rows from HBM) is gone.
// ---------------------------------------------------------------------------
constexpr int kFusedWaves = 16;

template <int NX, int CPB>
struct FusedShape {
  static constexpr int D = 64 * NX;
  static constexpr int R = NX & -NX;
  static constexpr int B = 64 * R;
  static constexpr int NBLK = NX / R;
  static constexpr int G = D / CPB;
  static constexpr int G16 = (G + 15) / 16;
  static constexpr int STAGE = (G16 * 16 > B * 4) ? G16 * 16 : B * 4;  // codes, or one rotated block
  static constexpr int SLOT = ((D * 4 + 15) & ~15) + ((STAGE + 15) & ~15);
  static constexpr int NT = D / 8;  // steps of each norm chain
  static constexpr int NA = NT <= 48 ? NT : (NT % 48 == 0 ? 48 : 32);
  static constexpr int VW = NX % 4 == 0 ? 4 : 2;  // floats per lane per row load
  static constexpr int NV = NX / VW;
  static constexpr size_t shared_bytes(bool tq) {
    return (size_t)D * 4 + 256 * 16 + (tq ? 3 * (size_t)D * 4 : 0);
  }
};

// Row pieces as native vectors (HIP's float4 is a struct, and an array of
// them crossing the row loop lands in scratch memory).
typedef float rowv4 __attribute__((ext_vector_type(4)));
typedef float rowv2 __attribute__((ext_vector_type(2)));
template <int VW>
struct RowVec;
template <>
struct RowVec<4> {
  using T = rowv4;
};
template <>
struct RowVec<2> {
  using T = rowv2;
};

template <int NX, int CPB, bool TQ>
__global__ __launch_bounds__(64 * kFusedWaves) void encode_fused_kernel(CodeTables t, const float* __restrict__ x,
                                                                       long n, long out_offset,
                                                                       uint8_t* __restrict__ tiled,
                                                                       float* __restrict__ alpha) {
  using S = FusedShape<NX, CPB>;
  constexpr int D = S::D, R = S::R, B = S::B;
  using V = typename RowVec<S::VW>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: row loops stay scalar
  uint32_t* pgl = reinterpret_cast<uint32_t*>(smem);
  CellDev* cells = reinterpret_cast<CellDev*>(smem + D * 4);
  float* tsh = reinterpret_cast<float*>(smem + D * 4 + 256 * 16);
  float* tsc = tsh + D;
  float* tis = tsc + D;
  unsigned char* slot = smem + S::shared_bytes(TQ) + (size_t)wave * S::SLOT;
  float* X = reinterpret_cast<float*>(slot);
  unsigned char* stage = slot + ((D * 4 + 15) & ~15);
  uint8_t* codes = stage;
  float* ys = reinterpret_cast<float*>(stage);

  for (int j = threadIdx.x; j < D; j += blockDim.x) pgl[lane_major_slot(j, B, R)] = t.perm_gather[j];
  for (int j = threadIdx.x; j < 256 * 4; j += blockDim.x)
    reinterpret_cast<uint32_t*>(cells)[j] = reinterpret_cast<const uint32_t*>(t.cells)[j];
  if constexpr (TQ)
    for (int j = threadIdx.x; j < D; j += blockDim.x) {
      const int sl = lane_major_slot(j, B, R);
      tsh[sl] = t.shift[j];
      tsc[sl] = t.scale[j];
      tis[sl] = t.inv_scale[j];
    }
  __syncthreads();

  const float cell_off = t.cell_off, cell_inv_w = t.cell_inv_w, isb = t.inv_sqrt_block;
  const bool certify = (t.debug_flags & 1) == 0;
  // octets (eight rows) of whole 64-row tiles; with a multiple of 8 wavefronts
  // per workgroup, a workgroup's octets of one round form whole tiles
  const long tile0 = out_offset >> 6;
  const long oct_total = (((out_offset + n + 63) >> 6) - tile0) * 8;
  for (long oi = (long)blockIdx.x * waves + wave; oi < oct_total; oi += (long)gridDim.x * waves) {
    const long v0 = (tile0 << 6) + oi * 8;  // output vector of the octet's first row
    const long r0 = v0 - out_offset;        // its input row (may be < 0 or past n at the edges)

    // 1. norms of the eight rows
    float nrm, invn;
    {
      const int q = lane >> 3, c = lane & 7;
      const long row = r0 + q;
      const long rc = (row >= 0 && row < n) ? row : 0;  // rows outside [0, n): any valid row, result unused
      const float* xr = x + rc * (long)D + c;
      float acc = 0.f;
#pragma unroll 1
      for (int t0 = 0; t0 < S::NT; t0 += S::NA) {
        float vv[S::NA];
#pragma unroll
        for (int u = 0; u < S::NA; ++u) vv[u] = xr[8 * (t0 + u)];
#pragma unroll
        for (int u = 0; u < S::NA; ++u) acc = __fmaf_rn(vv[u], vv[u], acc);
      }
      acc = acc + __shfl_xor(acc, 1);
      acc = acc + __shfl_xor(acc, 2);
      acc = acc + __shfl_xor(acc, 4);
      nrm = sqrt_rn(acc);
      invn = nrm > kMinNorm ? div_rn(1.f, nrm) : 0.f;
    }

    // 2. the rows, one at a time; the next row's loads fly while a row is encoded
    int q = 0;
    while (q < 8 && (r0 + q < 0 || r0 + q >= n)) ++q;  // first row of the octet inside [0, n)
    if (q == 8) continue;
    V cur[S::NV];
    {
      const V* xr = reinterpret_cast<const V*>(x + (r0 + q) * (long)D) + lane;
#pragma unroll
      for (int i = 0; i < S::NV; ++i) cur[i] = xr[64 * i];
    }
    for (; q < 8; ++q) {
      const long row = r0 + q;
      if (row >= n) break;  // rows past the end: none after this one either
      const float nv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nrm), 8 * q));
      const float inv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(invn), 8 * q));
      wave_lds_sync();  // the previous row is done with X and the stage
#pragma unroll
      for (int i = 0; i < S::NV; ++i) reinterpret_cast<V*>(X)[lane + 64 * i] = cur[i];
      if (q + 1 < 8 && row + 1 < n) {
        const V* xr = reinterpret_cast<const V*>(x + (row + 1) * (long)D) + lane;
#pragma unroll
        for (int i = 0; i < S::NV; ++i) cur[i] = xr[64 * i];
      }
      wave_lds_sync();  // the row is in X

      const uint32_t inv_bits = __float_as_uint(inv);
      double sum = 0.0, asum = 0.0;
#pragma unroll
      for (int b = 0; b < S::NBLK; ++b) {
        float v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t w = pgl[b * B + r * 64 + lane];
          const float xv = *reinterpret_cast<const float*>(reinterpret_cast<const unsigned char*>(X) + (w << 2));
          v[r] = xv * __uint_as_float((w & 0x80000000u) | (inv_bits & 0x7fffffffu));
        }
#pragma unroll
        for (int len = 1; len < R; len <<= 1) {
#pragma unroll
          for (int i = 0; i < R; i += 2 * len) {
#pragma unroll
            for (int j = i; j < i + len; ++j) {
              const float a = v[j], c = v[j + len];
              v[j] = a + c;
              v[j + len] = a - c;
            }
          }
        }
        lane_stage<R, 1>(v, lane);
        lane_stage<R, 2>(v, lane);
        lane_stage<R, 4>(v, lane);
        dpp8_stage<R>(v, lane);
        swap_stage<R, false>(v);
        swap_stage<R, true>(v);
        uint64_t cw = 0;  // this lane's codes of the block, packed as flat bytes
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float yv = v[r] * isb;
          const int jl = b * B + r * 64 + lane;
          float z = yv;
          if constexpr (TQ) z = (yv + tsh[jl]) * tsc[jl];
          const float f = __fmaf_rn(z, cell_inv_w, cell_off);
          const int ci = (int)__builtin_fminf(__builtin_fmaxf(f, 0.f), 255.f);
          const CellDev e = cells[ci];
          const bool up = z > e.bound;
          const float cen = up ? e.c_hi : e.c_lo;
          const unsigned k = e.base + (up ? 1u : 0u);
          // the term's rounding folds into the running sum's (fused multiply-add):
          // the error bound below covers one rounding per term either way
          const double rc = TQ ? fma((double)cen, (double)tis[jl], -(double)tsh[jl]) : (double)cen;
          sum = fma((double)yv, rc, sum);
          asum = fma(fabs((double)yv), fabs(rc), asum);
          constexpr int bits_per = 8 / CPB;
          const int shift = (CPB - 1 - r % CPB) * bits_per;  // first code highest in its byte
          cw |= (uint64_t)k << (8 * (r / CPB) + shift);
        }
        uint8_t* dst = codes + (b * B + lane * R) / CPB;
        store_code_bytes<(R / CPB > 0 ? R / CPB : 1)>(dst, cw);
      }
      for (int o = 32; o > 0; o >>= 1) {
        sum += __shfl_xor(sum, o);
        asum += __shfl_xor(asum, o);
      }
      wave_lds_sync();  // the row's codes are complete
      const long vo = out_offset + row;
      for (int c16 = lane; c16 < S::G16; c16 += 64)
        *reinterpret_cast<uint4*>(tiled + ((vo >> 6) * S::G16 + c16) * 1024 + (vo & 63) * 16) =
            *reinterpret_cast<const uint4*>(codes + c16 * 16);

      // certified alpha (the bound of encode_tile_kernel), else the reference's
      // sequential chains over the rotated values, one block at a time
      const double u = 1.1102230246251565e-16;  // 2^-53
      double bound = (double)(D / 4 + D / 64 + 16) * u * asum * 1.01;
      bound += 2.0 * u * (fabs(sum) + bound);
      const double lo = sum - bound, hi = sum + bound;
      float a = 0.f;
      bool exact = certify;
      if (exact && lo > 1e-12 && __double2float_rn(lo) == __double2float_rn(hi))
        a = div_rn(nv, __double2float_rn(hi));
      else if (exact && hi <= 1e-12)
        a = 0.f;
      else
        exact = false;
      if (!exact) {
        wave_lds_sync();  // codes written out: the stage takes rotated blocks
        double d = 0.0;   // lane m < 4: chain m
#pragma unroll 1
        for (int b = 0; b < S::NBLK; ++b) {
          float yb[R];
          rotate_block_lm<R>(t, pgl, X, inv, b, lane, yb);
#pragma unroll
          for (int r = 0; r < R; ++r) ys[lane * R + r] = yb[r];
          wave_lds_sync();
          if (lane < 4) {
            for (int jj = lane; jj < B; jj += 4) {
              const int jl = b * B + (jj % R) * 64 + jj / R;
              const float yv = ys[jj];
              float z = yv;
              if constexpr (TQ) z = (yv + tsh[jl]) * tsc[jl];
              const float f = __fmaf_rn(z, cell_inv_w, cell_off);
              const int ci = (int)__builtin_fminf(__builtin_fmaxf(f, 0.f), 255.f);
              const CellDev e = cells[ci];
              const float cen = z > e.bound ? e.c_hi : e.c_lo;
              if constexpr (TQ)
                d = fma((double)yv, fma((double)cen, (double)tis[jl], -(double)tsh[jl]), d);
              else
                d = fma((double)yv, (double)cen, d);
            }
          }
          wave_lds_sync();
        }
        const double pr = d + __shfl_xor(d, 1);
        const double inner = pr + __shfl_xor(pr, 2);
        a = inner > 1e-12 ? div_rn(nv, (float)inner) : 0.f;
      }
      if (lane == 0) alpha[vo] = a;
    }
  }
}

// ---------------------------------------------------------------------------
// TQ+ calibration (src/encode/encode.cpp:142-150,282-326)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f32_key(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float f32_unkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// rotated (n x D row-major) -> keys (D x n), 32x32 tiles through LDS.
__global__ __launch_bounds__(256) void keys_transpose_kernel(const float* __restrict__ rot, long n,
                                                             int D, uint32_t* __restrict__ keys) {
  __shared__ uint32_t tile[32][33];
  const long r0 = (long)blockIdx.y * 32;
  const int d0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int i = ty; i < 32; i += 8) {
    const long r = r0 + i;
    const int d = d0 + tx;
    if (r < n && d < D) tile[i][tx] = f32_key(rot[r * D + d]);
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int d = d0 + i;
    const long r = r0 + tx;
    if (r < n && d < D) keys[(long)d * n + r] = tile[tx][i];
  }
}

// Per dimension: the lo_idx-th and hi_idx-th smallest keys by 4-pass radix
// select (8 bits per pass, both ranks per pass), then the TQ+ shift/scale.
__global__ __launch_bounds__(256) void radix_select_kernel(const uint32_t* __restrict__ keys, long n,
                                                           long lo_idx, long hi_idx, float qc_lo,
                                                           float qc_span, float* __restrict__ shift,
                                                           float* __restrict__ scale,
                                                           float* __restrict__ inv_scale) {
  __shared__ uint32_t hist[2][256];
  __shared__ uint32_t prefix[2];
  __shared__ unsigned long long rank[2];
  const int d = blockIdx.x;
  const uint32_t* col = keys + (long)d * n;
  if (threadIdx.x == 0) {
    prefix[0] = prefix[1] = 0u;
    rank[0] = (unsigned long long)lo_idx;
    rank[1] = (unsigned long long)hi_idx;
  }
  uint32_t mask = 0u;
  for (int sh = 24; sh >= 0; sh -= 8) {
    for (int i = threadIdx.x; i < 512; i += 256) (&hist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t p0 = prefix[0], p1 = prefix[1];
    for (long i = threadIdx.x; i < n; i += 256) {
      const uint32_t k = col[i];
      const uint32_t b = (k >> sh) & 255u;
      if ((k & mask) == p0) atomicAdd(&hist[0][b], 1u);
      if ((k & mask) == p1) atomicAdd(&hist[1][b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      const int w = threadIdx.x;
      unsigned long long r = rank[w], below = 0;
      for (int b = 0; b < 256; ++b) {
        const unsigned long long h = hist[w][b];
        if (below + h > r) {
          prefix[w] |= uint32_t(b) << sh;
          rank[w] = r - below;
          break;
        }
        below += h;
      }
    }
    __syncthreads();
    mask |= 255u << sh;
  }
  if (threadIdx.x == 0) {
    const float q_lo = f32_unkey(prefix[0]);
    const float q_hi = f32_unkey(prefix[1]);
    const float span = q_hi - q_lo;
    float s = 1.f, sft = 0.f;
    if (span > 1e-12f && qc_span > 0.f) {
      s = div_rn(qc_span, span);
      sft = div_rn(qc_lo, s) - q_lo;
    }
    shift[d] = sft;
    scale[d] = s;
    inv_scale[d] = div_rn(1.f, s);
  }
}

// ---------------------------------------------------------------------------
// Query preparation (src/search/search.cpp:149-219,762-860): one wavefront
// per query.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lut_value(const float* __restrict__ q, const float* __restrict__ cent, int bits,
                                           int base, int cpn, unsigned nib) {
  const unsigned mask = (1u << bits) - 1u;
  float s = 0.f;
  for (int c = 0; c < cpn; ++c) {
    const unsigned code = (nib >> ((cpn - 1 - c) * bits)) & mask;
    s = __fmaf_rn(q[base + c], cent[code], s);
  }
  return s;
}

// One query per 256-thread workgroup (src/search/search.cpp:149-219,762-860):
// row norm (8 chains) -> rotation -> TQ+ (bias_corr, y / scale) -> per group
// the 16 + 16 nibble sums, their mins and the global span -> scale = span /
// 127, u8 entries, bias = sequential sum of the group mins. The chains and the
// two sequential sums run on one lane each in the reference's order; every
// other step is spread over the workgroup.
constexpr int kPrepThreads = 256;

// dst[j] = src[j] for j < n over the workgroup, 8 loads in flight per thread
__device__ __forceinline__ void stage_to_lds(float* __restrict__ dst, const float* __restrict__ src, int n) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int base = 0; base < n; base += 8 * nt) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = base + u * nt + tid;
      v[u] = j < n ? src[j] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = base + u * nt + tid;
      if (j < n) dst[j] = v[u];
    }
  }
}

// The 16 + 16 LUT values of group g (lut_value order: the fused products of
// the group's rotated coordinates with the codebook, src/search/search.cpp:
// 149-219), the codebook in registers.
template <int CPN>
__device__ __forceinline__ void group_values(const float* __restrict__ y, const float (&cr)[16], int bits, int g,
                                             float (&sv)[2][16]) {
  const unsigned mask = (1u << bits) - 1u;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int base = g * (2 * CPN) + half * CPN;
    float yv[CPN];
#pragma unroll
    for (int c = 0; c < CPN; ++c) yv[c] = y[base + c];
#pragma unroll
    for (int nib = 0; nib < 16; ++nib) {
      float acc = 0.f;
#pragma unroll
      for (int c = 0; c < CPN; ++c) acc = __fmaf_rn(yv[c], cr[(nib >> ((CPN - 1 - c) * bits)) & mask], acc);
      sv[half][nib] = acc;
    }
  }
}

// Query rotation by one wavefront per rotation block (the encode kernel's
// rotate_block: gather, in-lane and cross-lane butterflies, 1/sqrt(B)).
template <int R>
__device__ __forceinline__ void rotate_query_blocks(const CodeTables& t, const float* __restrict__ xs, float inv,
                                                    float* __restrict__ yout, int wave, int waves, int lane) {
  const int nblk = t.dim / t.block;
  for (int b = wave; b < nblk; b += waves) {
    float v[R];
    rotate_block<R>(t, t.perm_gather, xs, inv, b, lane, v);
#pragma unroll
    for (int r = 0; r < R; ++r) yout[b * t.block + lane * R + r] = v[r];
  }
}

__global__ __launch_bounds__(kPrepThreads) void prepare_queries_kernel(CodeTables t, const float* __restrict__ q,
                                                                      long nq, uint8_t* __restrict__ luts,
                                                                      float* __restrict__ lut_scale,
                                                                      float* __restrict__ lut_bias,
                                                                      float* __restrict__ bias_corr,
                                                                      float* __restrict__ bias_total) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int D = t.dim, G = t.groups, B = t.block;
  const bool tq = t.shift != nullptr;
  float* y = reinterpret_cast<float*>(smem);   // D: rotated query (divided by the TQ+ scale)
  float* xs = y + D;                            // D: the staged query row, then y before the TQ+ division
  float* mins = xs + D;                         // 2 * G
  float* tsh = mins + 2 * G;                    // D (TQ+)
  float* tsc = tsh + D;                         // D (TQ+)
  __shared__ float s_norm, s_bsum, s_bc, s_span[kPrepThreads / 64];
  __shared__ double s_red[2][kPrepThreads / 64];
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, waves = nt >> 6;
  const int cpb = 8 / t.bits, cpn = cpb / 2;
  float cr[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) cr[i] = t.centroids[i & ((1 << t.bits) - 1)];
  if (tq) {
    stage_to_lds(tsh, t.shift, D);
    stage_to_lds(tsc, t.scale, D);
  }
  const int R = B / 64;
  const bool wave_rot = t.perm_gather != nullptr && (R == 2 || R == 4 || R == 8 || R == 16);
  for (long qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    const float* xr = q + qi * (long)D;
    __syncthreads();
    if (t.prerotated) {
      // LUTs of the rows as given (landmark query rows, src/search/search.cpp:973-1050)
      stage_to_lds(y, xr, D);
      __syncthreads();
    } else {
    stage_to_lds(xs, xr, D);
    __syncthreads();
    if (wave == 0) {
      const float nrm = norm8_lds(xs, D, lane);
      if (lane == 0) s_norm = nrm;
    }
    __syncthreads();
    const float nrm = s_norm;
    const float inv = nrm > kMinNorm ? div_rn(1.f, nrm) : 0.f;
    if (wave_rot) {
      // y[i] = (x[perm[i]] * inv) * sign[i], butterflies, 1/sqrt(B): into y
      if (R == 2) rotate_query_blocks<2>(t, xs, inv, y, wave, waves, lane);
      else if (R == 4) rotate_query_blocks<4>(t, xs, inv, y, wave, waves, lane);
      else if (R == 8) rotate_query_blocks<8>(t, xs, inv, y, wave, waves, lane);
      else rotate_query_blocks<16>(t, xs, inv, y, wave, waves, lane);
      __syncthreads();
    } else {
      // y[i] = (x[perm[i]] * inv) * sign[i], scattered through inv_perm
      for (int base = 0; base < D; base += 8 * nt) {
        uint32_t w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int j = base + u * nt + tid;
          w[u] = j < D ? t.perm_sign[j] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int j = base + u * nt + tid;
          if (j < D) y[w[u] & 0x7fffffffu] = (xs[j] * inv) * ((w[u] >> 31) ? -1.f : 1.f);
        }
      }
      __syncthreads();
      const int npairs = D >> 1, lhb = __ffs(B) - 2;  // log2(B / 2)
      for (int lg = 0; (1 << lg) < B; ++lg) {
        const int len = 1 << lg;
        for (int p = tid; p < npairs; p += nt) {
          const int blk = p >> lhb, qq = p & ((B >> 1) - 1);
          const int i = blk * B + ((qq >> lg) << (lg + 1)) + (qq & (len - 1));
          const float u = y[i], v = y[i + len];
          y[i] = u + v;
          y[i + len] = u - v;
        }
        __syncthreads();
      }
      for (int i = tid; i < D; i += nt) y[i] = y[i] * t.inv_sqrt_block;
      __syncthreads();
    }
    }  // rotated rows
    if (tq) {
      // keep y for bias_corr, divide by the TQ+ scale; bias_corr =
      // (float) -sum_d (double) y_d * shift_d, sequential in d in the
      // reference: the products are exact in double, so a tree sum plus the
      // error bound of the sequential one certifies the float result
      double sum = 0.0, asum = 0.0;
      for (int i = tid; i < D; i += nt) {
        const float r = y[i];
        xs[i] = r;
        y[i] = div_rn(r, tsc[i]);
        const double term = -(double)r * (double)tsh[i];
        sum += term;
        asum += fabs(term);
      }
      for (int o = 32; o > 0; o >>= 1) {
        sum += __shfl_xor(sum, o);
        asum += __shfl_xor(asum, o);
      }
      if (lane == 0) {
        s_red[0][wave] = sum;
        s_red[1][wave] = asum;
      }
      __syncthreads();
      if (tid == 0) {
        double ts = 0.0, ta = 0.0;
        for (int w = 0; w < waves; ++w) {
          ts += s_red[0][w];
          ta += s_red[1][w];
        }
        const double u = 1.1102230246251565e-16;  // 2^-53
        double bound = (double)(D + 64) * u * ta * 1.01;
        bound += 2.0 * u * (fabs(ts) + bound);
        const float lo = (float)(ts - bound), hi = (float)(ts + bound);
        if (lo == hi && (t.debug_flags & 16) == 0) {  // flags bit 4: always the chain (test hook)
          s_bc = hi;
        } else {
          // not certified: the reference's sequential chain
          double acc = 0.0;
          for (int d = 0; d < D; ++d) acc = fma(-(double)xs[d], (double)tsh[d], acc);
          s_bc = (float)acc;
        }
      }
    } else if (tid == 0) {
      s_bc = 0.f;
    }
    __syncthreads();
    // pass 1: per group the min/max of both 16-entry halves; global span
    float span = 0.f;
    for (int g = tid; g < G; g += nt) {
      float sv[2][16];
      if (cpn == 1) group_values<1>(y, cr, t.bits, g, sv);
      else group_values<2>(y, cr, t.bits, g, sv);
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        float mn = FLT_MAX, mx = -FLT_MAX;
#pragma unroll
        for (int nib = 0; nib < 16; ++nib) {
          mn = sv[half][nib] < mn ? sv[half][nib] : mn;
          mx = sv[half][nib] > mx ? sv[half][nib] : mx;
        }
        mins[2 * g + half] = mn;
        span = fmaxf(span, mx - mn);
      }
    }
    for (int off = 32; off > 0; off >>= 1) span = fmaxf(span, __shfl_xor(span, off));
    if (lane == 0) s_span[wave] = span;
    __syncthreads();
    span = s_span[0];
    for (int w = 1; w < waves; ++w) span = fmaxf(span, s_span[w]);
    float scale = span > 0.f ? div_rn(span, 127.f) : 1.f;
    float iscale = 1.f;
    if (scale >= FLT_MIN)
      iscale = div_rn(1.f, scale);
    else
      scale = 1.f;
    // pass 2 (all but the last wavefront): quantized entries; meanwhile lane 0
    // of the last wavefront runs the bias, sequential over groups as in the
    // reference
    uint8_t* out = luts + qi * (long)G * 32;
    if (wave < waves - 1) {
      const int pt = nt - 64;
      for (int g = tid; g < G; g += pt) {
        float sv[2][16];
        if (cpn == 1) group_values<1>(y, cr, t.bits, g, sv);
        else group_values<2>(y, cr, t.bits, g, sv);
        uint32_t w[8];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          const float mn = mins[2 * g + half];
#pragma unroll
          for (int e = 0; e < 16; e += 4) {
            uint32_t word = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
              float r = roundf((sv[half][e + b] - mn) * iscale);
              r = r < 0.f ? 0.f : (r > 127.f ? 127.f : r);
              word |= uint32_t(r) << (8 * b);
            }
            w[half * 4 + e / 4] = word;
          }
        }
        uint4* dst = reinterpret_cast<uint4*>(out + (long)g * 32);
        dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
        dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
      }
    } else if (lane == 0) {
      float bsum = 0.f;
      int g = 0;
      for (; g + 8 <= G; g += 8) {
        float m[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) m[u] = mins[2 * g + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) bsum = bsum + (m[2 * u] + m[2 * u + 1]);
      }
      for (; g < G; ++g) bsum = bsum + (mins[2 * g] + mins[2 * g + 1]);
      lut_scale[qi] = scale;
      lut_bias[qi] = bsum;
      s_bsum = bsum;
    }
    __syncthreads();
    if (tid == 0) {
      bias_corr[qi] = s_bc;
      bias_total[qi] = s_bsum + s_bc;
    }
  }
}

// ---------------------------------------------------------------------------
// LUT scan with exact per-range top-k (src/search/search_avx2.cpp:95-222,
// src/search/search.cpp:600-758).
//
// Workgroup = 1024 threads, one vector per thread per step, T queries per
// workgroup. The T queries' LUTs sit in LDS interleaved by query:
// entry (g, e, t) at byte (g * 32 + e) * T + t, so one 32-bit (T = 4) or
// 64-bit (T = 8) LDS read fetches a nibble's entry for all T queries. Byte
// sums are accumulated SWAR in packed u16 lanes (exact: a 256-group window sum
// is <= 256 * 254 = 65024), flushed per window into the float chain
// fa = fmaf(scale, (float) A_w, fa), score = fa * alpha.
// Top-k per query: candidates that beat the current k-th best under
// (score desc, id asc) go to an LDS buffer, merged by rank into the sorted
// list when it fills. Exact; ties resolve to the lower id.
// ---------------------------------------------------------------------------
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ bool better(float s, uint32_t id, float ts, uint32_t ti) {
  return s > ts || (s == ts && id < ti);
}

struct TopkShared {
  float* top_s;     // [T][k]
  uint32_t* top_i;  // [T][k]
  float* cand_s;    // [T][cap]
  uint32_t* cand_i; // [T][cap]
  int* cnt;         // [T]
  int* m;           // [T]
  float* thr_s;     // [T]
  uint32_t* thr_i;  // [T]
  int* bound;       // [T] score bound (order-preserving int image of a float)
};

// Order-preserving map of floats to ints (for LDS atomicMax) and back.
__device__ __forceinline__ int float_order(float f) {
  const int b = __float_as_int(f);
  return b >= 0 ? b : (b ^ 0x7fffffff);
}
__device__ __forceinline__ float order_float(int o) { return __int_as_float(o >= 0 ? o : (o ^ 0x7fffffff)); }

// Merge query t's candidates into its sorted list (all threads participate).
__device__ void topk_merge(const TopkShared& sh, int t, int k, int cap) {
  const int tid = threadIdx.x;
  const int c = __builtin_amdgcn_readfirstlane(min(sh.cnt[t], cap));
  const int m = __builtin_amdgcn_readfirstlane(sh.m[t]);
  float* ts = sh.top_s + t * k;
  uint32_t* ti = sh.top_i + t * k;
  const float* cs = sh.cand_s + t * cap;
  const uint32_t* ci = sh.cand_i + t * cap;
  // rank = start + number of candidates better than (s, id) (16-byte LDS
  // reads; the buffer is 16-byte aligned and cap is a multiple of 32, so
  // reads past c stay inside it and are masked)
  // Eight candidates per step, their reads issued together: the early exit
  // (rank >= k) makes each step wait for its reads, and the LDS is shared with
  // another workgroup's lookups, so a chain of single reads is slow.
  auto rank_of = [&](float s, uint32_t id, int rank) {
    for (int f = 0; f < c && rank < k; f += 8) {
      const float4 sa = *reinterpret_cast<const float4*>(cs + f);
      const float4 sb = *reinterpret_cast<const float4*>(cs + f + 4);
      const uint4 ia = *reinterpret_cast<const uint4*>(ci + f);
      const uint4 ib = *reinterpret_cast<const uint4*>(ci + f + 4);
      rank += (better(sa.x, ia.x, s, id) ? 1 : 0) + (f + 1 < c && better(sa.y, ia.y, s, id) ? 1 : 0) +
              (f + 2 < c && better(sa.z, ia.z, s, id) ? 1 : 0) + (f + 3 < c && better(sa.w, ia.w, s, id) ? 1 : 0) +
              (f + 4 < c && better(sb.x, ib.x, s, id) ? 1 : 0) + (f + 5 < c && better(sb.y, ib.y, s, id) ? 1 : 0) +
              (f + 6 < c && better(sb.z, ib.z, s, id) ? 1 : 0) + (f + 7 < c && better(sb.w, ib.w, s, id) ? 1 : 0);
    }
    return rank;
  };
  // Every rank first, the writes after a barrier, so the merged list is
  // written in place: a thread holds at most one candidate (cap <= the
  // workgroup size) and two list entries (k <= twice the workgroup size).
  const int nt = blockDim.x;
  float es = 0.f, ls0 = 0.f, ls1 = 0.f;
  uint32_t ei = 0u, li0 = 0u, li1 = 0u;
  int er = k, lr0 = k, lr1 = k;
  if (tid < c) {
    es = cs[tid];
    ei = ci[tid];
    int lo = 0, hi = m;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (better(ts[mid], ti[mid], es, ei)) lo = mid + 1; else hi = mid;
    }
    er = rank_of(es, ei, lo);
  }
  if (tid < m) {
    ls0 = ts[tid];
    li0 = ti[tid];
    lr0 = rank_of(ls0, li0, tid);
  }
  if (tid + nt < m) {
    ls1 = ts[tid + nt];
    li1 = ti[tid + nt];
    lr1 = rank_of(ls1, li1, tid + nt);
  }
  __syncthreads();
  if (er < k) { ts[er] = es; ti[er] = ei; }
  if (lr0 < k) { ts[lr0] = ls0; ti[lr0] = li0; }
  if (lr1 < k) { ts[lr1] = ls1; ti[lr1] = li1; }
  const int m2 = min(k, m + c);
  __syncthreads();
  if (tid == 0) {
    sh.m[t] = m2;
    sh.cnt[t] = 0;
    if (m2 == k) { sh.thr_s[t] = ts[k - 1]; sh.thr_i[t] = ti[k - 1]; }
  }
  __syncthreads();
}

// topk_merge for every query of a tile at once: wavefront t takes query t's
// list and candidates into registers and writes the merged list back: up to
// 64 entries by one bitonic sort across the lanes, more (at most 256, four
// per lane) by repeated wavefront argmax, both under (score desc, id asc).
// One barrier in all; the queries merge in parallel.
// Needs T <= the wavefronts of the workgroup and k + cap <= 256. Defined
// after wave_best (below).
__device__ void topk_merge_all(const TopkShared& sh, int T, int k, int cap);

// Carve the top-k state out of LDS at `base` and reset it (threads < T write;
// the caller synchronizes before first use).
__device__ __forceinline__ void init_topk(TopkShared& sh, unsigned char* base, int T, int k, int cap) {
  size_t off = 0;
  sh.top_s = reinterpret_cast<float*>(base + off); off += (size_t)T * k * 4;
  sh.top_i = reinterpret_cast<uint32_t*>(base + off); off += (size_t)T * k * 4;
  sh.cand_s = reinterpret_cast<float*>(base + off); off += (size_t)T * cap * 4;
  sh.cand_i = reinterpret_cast<uint32_t*>(base + off); off += (size_t)T * cap * 4;
  sh.cnt = reinterpret_cast<int*>(base + off); off += T * 4;
  sh.m = reinterpret_cast<int*>(base + off); off += T * 4;
  sh.thr_s = reinterpret_cast<float*>(base + off); off += T * 4;
  sh.thr_i = reinterpret_cast<uint32_t*>(base + off); off += T * 4;
  sh.bound = reinterpret_cast<int*>(base + off);
  const int tid = threadIdx.x;
  if (tid < T) {
    sh.cnt[tid] = 0;
    sh.m[tid] = 0;
    sh.thr_s[tid] = -INFINITY;
    sh.thr_i[tid] = 0u;
    sh.bound[tid] = float_order(-INFINITY);
  }
}

// Bitonic sort of one key per lane across the wavefront into descending
// (score, then ascending id) order: lane i ends with the i-th best key.
// Value of lane ^ M of a wavefront, without LDS: DPP quad permutes for 1 and
// 2, two DPP moves for 4 (quad reversal = ^3, then the half-row mirror = ^7),
// a DPP row rotation by 8 (^8 inside each 16-lane row), and for 16 and 32 the
// gfx950 row swaps of the value with itself (v_permlane16_swap: the first
// result holds the even rows twice, the second the odd rows; the same for the
// halves with v_permlane32_swap), the partner's copy picked per lane.
template <int M>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, int lane) {
  const int b = (int)v;
  if constexpr (M == 1) return (uint32_t)__builtin_amdgcn_mov_dpp(b, 0xB1, 0xF, 0xF, false);
  else if constexpr (M == 2) return (uint32_t)__builtin_amdgcn_mov_dpp(b, 0x4E, 0xF, 0xF, false);
  else if constexpr (M == 4)
    return (uint32_t)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(b, 0x1B, 0xF, 0xF, false), 0x141, 0xF, 0xF,
                                              false);
  else if constexpr (M == 8) return (uint32_t)__builtin_amdgcn_mov_dpp(b, 0x128, 0xF, 0xF, false);
  else if constexpr (M == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? p[0] : p[1];
  } else {
    static_assert(M == 32, "xor distance");
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? p[0] : p[1];
  }
}

// One compare-exchange stage of a bitonic sort across the wavefront.
template <int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_stage(float& s, uint32_t& id, int lane) {
  const float os = __uint_as_float(xor_lane<STRIDE>(__float_as_uint(s), lane));
  const uint32_t oi = xor_lane<STRIDE>(id, lane);
  const bool desc = (lane & SIZE) == 0 || SIZE == 64;
  const bool lower = (lane & STRIDE) == 0;
  const bool o_better = better(os, oi, s, id);
  if ((lower == desc) ? o_better : !o_better) { s = os; id = oi; }
}

__device__ __forceinline__ void wave_sort_desc(float& s, uint32_t& id, int lane) {
  static_for<6>([&](auto lg) {
    constexpr int L = decltype(lg)::value;
    static_for<L + 1>([&](auto st) {
      constexpr int STRIDE = (1 << L) >> decltype(st)::value;
      bitonic_stage<2 << L, STRIDE>(s, id, lane);
    });
  });
}

// The kw-th largest u among the lanes with `cand` set (at least kw of them),
// bit by bit from the top with ballots. Not inlined: the scan kernels reach it
// only on a workgroup's first steps, and inlined into the admission of a
// two-vector tile it cost the scan half its occupancy.
__device__ __attribute__((noinline)) uint32_t wave_kth_largest(uint32_t u, bool cand, int kw) {
  uint32_t kth = 0u;
  for (int b = 31; b >= 0; --b) {
    const uint32_t trial = kth | (1u << b);
    if (__popcll(__ballot(cand && u >= trial)) >= kw) kth = trial;
  }
  return kth;
}

// Scores of V vectors x T queries (fa * alpha, -0.f reported as +0.f, as the
// reference's flush_batch_to_fa epilogue) and their admission into the
// workgroup's top-k lists: strict "better than the current k-th" test,
// candidate buffer, merge when half full, retry on overflow. Every thread of
// the workgroup calls it (it synchronizes).
// voff (when not null): a per-vector offset added after alpha, the reference's
// score_offsets of score_prepared_into (src/search/search_avx2.cpp:160-214:
// fa * alpha, then + offset, each rounded); toff (when voff is null): the
// same offset per query instead (a tile of queries over one bucket). fresh:
// the lists are empty (the first step of a work item), so the test against
// their k-th entries is skipped.
template <int T, int V>
__device__ void admit_scores(const TopkShared& sh, const float (&fa)[V][T], const long (&vv)[V],
                             const bool (&valid)[V], const float* __restrict__ alpha, long q0, long nq,
                             int k, int cap, const float (*voff)[V] = nullptr,
                             const float (*av_pre)[V] = nullptr, const float* toff = nullptr,
                             bool fresh = false) {
  const int lane = threadIdx.x & 63;
  const int kw = min(k, 64);
#pragma unroll
  for (int j = 0; j < V; ++j) {
    float s[T];
    uint32_t sid[T];
    const uint32_t vid = (uint32_t)vv[j];
    uint32_t pend = 0u, candm = 0u;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      s[t] = -INFINITY;
      sid[t] = vid;
    }
    // a wavefront without vectors (the tail of a range) goes straight to the barriers
    if (__ballot(valid[j]) != 0) {
    const float av = valid[j] ? (av_pre != nullptr ? (*av_pre)[j] : alpha[vv[j]]) : 0.f;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      s[t] = valid[j] ? fa[j][t] * av + (voff != nullptr ? (*voff)[j] : toff != nullptr ? toff[t] : 0.f)
                      : -INFINITY;
      // below the bound: worse than k vectors already offered (see below)
      // fresh: every list is still empty (its k-th is -inf), no need to read it
      const bool cand = valid[j] && q0 + t < nq && (fresh || better(s[t], vid, sh.thr_s[t], sh.thr_i[t])) &&
                        !(s[t] < order_float(sh.bound[t]));
      candm |= (cand ? 1u : 0u) << t;
    }
    // More candidates in this wavefront than can reach a query's list (a
    // workgroup's first steps admit almost every vector): offer only the
    // wavefront's best kw of that query. The wavefront then offers its k best;
    // every later vector scoring below the k-th of them is worse than k
    // offered vectors (each of which stays in the lists or is displaced by a
    // better one), so it can never enter the top k: that score bounds the
    // admission from below (an equal score may still win on its id). Without
    // this a tile of 8 queries pushed its whole first step through the
    // candidate buffer. (nc <= 64, so this happens for k < 64 only.)
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const bool cand = ((candm >> t) & 1u) != 0u;
      const int nc = __popcll(__ballot(cand));
      if (nc > kw) {
        if constexpr (T <= 2) {
          // tiles of one or two queries: sort the best kw into lanes 0 .. kw-1
          if (!cand) { s[t] = -INFINITY; sid[t] = 0xffffffffu; }
          wave_sort_desc(s[t], sid[t], lane);
          if (lane < kw) pend |= 1u << t;
          if (kw == k && lane == k - 1) atomicMax(&sh.bound[t], float_order(s[t]));
        } else {
          // larger tiles: the kw-th largest score by a radix select over the
          // candidates' order-preserving bits (ballots, no lane exchange: an
          // unrolled sort per query cost these tiles half their occupancy);
          // the lanes at or above it are offered (ties may add a few)
          const uint32_t u = (uint32_t)float_order(s[t]) ^ 0x80000000u;  // unsigned, same order
          const uint32_t kth = wave_kth_largest(u, cand, kw);
          if (cand && u >= kth) pend |= 1u << t;
          if (kw == k && lane == 0) atomicMax(&sh.bound[t], (int)(kth ^ 0x80000000u));
        }
      } else if (cand) {
        pend |= 1u << t;
      }
    }
    }
    while (true) {
      // the branches here are uniform (a ballot, or counts read by every
      // thread): no lane-dependent control flow around the barriers
      if (__ballot(pend != 0u) != 0)  // uniform: a wavefront with nothing to store skips the atomics
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const bool want = ((pend >> t) & 1u) != 0u;
        // one LDS atomic per wavefront: base + rank among the wanting lanes
        const uint64_t wm = __ballot(want);
        int base = 0;
        if ((threadIdx.x & 63) == 0) base = atomicAdd(&sh.cnt[t], __popcll(wm));
        base = __builtin_amdgcn_readfirstlane(base);  // lane 0's (the whole wavefront is active here)
        const int slot = base + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(wm >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)wm, 0u));
        const bool store = want && slot < cap;
        if (store) {
          sh.cand_s[t * cap + slot] = s[t];
          sh.cand_i[t * cap + slot] = sid[t];
        }
        pend &= ~((store ? 1u : 0u) << t);
      }
      __syncthreads();
      // Every thread reads the same counts; taking them through readfirstlane
      // makes the loop exit and the merges (which synchronize) wave-uniform
      // branches for the compiler as well.
      int overflow = 0, need = 0;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int c = sh.cnt[t];
        overflow |= c > cap ? 1 : 0;
        need |= (c > cap / 2 ? 1 : 0) << t;
      }
      overflow = __builtin_amdgcn_readfirstlane(overflow);
      need = __builtin_amdgcn_readfirstlane(need);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < T; ++t)
        if ((need >> t) & 1) topk_merge(sh, t, k, cap);
      if (!overflow) break;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const bool keep = better(s[t], sid[t], sh.thr_s[t], sh.thr_i[t]);
        pend &= ~((keep ? 0u : 1u) << t);
      }
    }
  }
}

// Final merge of the remaining candidates and output of this workgroup's
// sorted lists: part_s/part_i [q][n_ranges][k], part_cnt [q][n_ranges].
template <int T>
__device__ void write_partials(const TopkShared& sh, long q0, long nq, int n_ranges, int r, int k, int cap,
                               float* __restrict__ part_s, uint32_t* __restrict__ part_i,
                               int* __restrict__ part_cnt) {
  const int tid = threadIdx.x;
  __syncthreads();
#pragma unroll
  for (int t = 0; t < T; ++t)
    if (__builtin_amdgcn_readfirstlane(sh.cnt[t]) > 0) topk_merge(sh, t, k, cap);
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const long q = q0 + t;
    if (q >= nq) continue;
    const long slot = q * n_ranges + r;
    const int m = sh.m[t];
    for (int j = tid; j < m; j += blockDim.x) {
      part_s[slot * k + j] = sh.top_s[t * k + j];
      part_i[slot * k + j] = sh.top_i[t * k + j];
    }
    if (tid == 0) part_cnt[slot] = m;
  }
}

template <int T>
struct LutWord;
template <> struct LutWord<1> { typedef uint8_t type; };
template <> struct LutWord<4> { typedef uint32_t type; };
template <> struct LutWord<8> { typedef uint2 type; };

// Query-interleaved LUTs of one query tile: byte (g * 32 + e) * T + t holds
// entry e of group g of query q0 + t; padded groups / queries are zero.
__global__ void interleave_luts_kernel(const uint8_t* __restrict__ luts, long nq, int G, int G16, int T,
                                       uint8_t* __restrict__ out) {
  const long n_qt = (nq + T - 1) / T;
  const long per_tile = (long)G16 * 16 * 32 * T;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n_qt * per_tile; i += (long)gridDim.x * blockDim.x) {
    const long qt = i / per_tile;
    const long r = i - qt * per_tile;
    const int t = (int)(r % T);
    const int e = (int)((r / T) & 31);
    const int g = (int)(r / (T * 32));
    const long q = qt * T + t;
    out[i] = (g < G && q < nq) ? luts[q * (long)G * 32 + (long)g * 32 + e] : 0;
  }
}

__device__ __forceinline__ uint32_t nib(uint32_t w, int bitpos) { return __builtin_amdgcn_ubfe(w, bitpos, 4); }

template <int T, int V>
__global__ __launch_bounds__(kScanThreads) void scan_topk_kernel(
    const uint8_t* __restrict__ tiled, const float* __restrict__ alpha, long n, int G16, int chunk16,
    const uint8_t* __restrict__ lut_t, const float* __restrict__ lut_scale,
    const float* __restrict__ bias_total, long nq, int k, int cap, int n_ranges, long range_len,
    float* __restrict__ part_s, uint32_t* __restrict__ part_i, int* __restrict__ part_cnt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const long n_qt = (nq + T - 1) / T;
  const long L = blockIdx.x;
  const int r = (int)(L / n_qt);   // range-major: concurrent workgroups share codes
  const long qt = L % n_qt;
  const long q0 = qt * T;
  const long r_begin = (long)r * range_len;
  const long r_end = min(n, r_begin + range_len);

  // ---- LDS carve-up ----
  const size_t lut_bytes = (size_t)chunk16 * 16 * 32 * T;
  uint8_t* lut_s = smem;
  TopkShared sh;
  init_topk(sh, smem + ((lut_bytes + 15) & ~size_t(15)), T, k, cap);
  float scl[T], bias[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const long q = q0 + t < nq ? q0 + t : nq - 1;
    scl[t] = lut_scale[q];
    bias[t] = bias_total[q];
  }

  // LDS LUT for 16-group chunks [c0, c0 + chunk16): a straight copy of the
  // query tile's interleaved LUT.
  const uint4* tile_lut = reinterpret_cast<const uint4*>(lut_t + qt * ((long)G16 * 16 * 32 * T));
  auto load_lut = [&](int c0) {
    const int n16 = min(chunk16, G16 - c0) * 16 * 32 * T / 16;
    const uint4* src = tile_lut + (long)c0 * 16 * 32 * T / 16;
    uint4* dst = reinterpret_cast<uint4*>(lut_s);
    for (int w = tid; w < n16; w += blockDim.x) dst[w] = src[w];
  };
  const bool one_chunk = chunk16 >= G16;
  if (one_chunk) load_lut(0);
  __syncthreads();

  constexpr int H = T >= 4 ? T / 4 : 1;   // 32-bit LUT words per entry
  typedef typename LutWord<T>::type W;
  const long step = (long)blockDim.x * V;
  for (long base = r_begin; base < r_end; base += step) {
    long vv[V];
    bool valid[V];
    const uint4* cp[V];
    float fa[V][T];
#pragma unroll
    for (int j = 0; j < V; ++j) {
      vv[j] = base + tid + (long)j * blockDim.x;
      valid[j] = vv[j] < r_end;
      const long vs = valid[j] ? vv[j] : 0;
      cp[j] = reinterpret_cast<const uint4*>(tiled) + (vs >> 6) * (long)G16 * 64 + (vs & 63);
#pragma unroll
      for (int t = 0; t < T; ++t) fa[j][t] = bias[t];
    }

    for (int c0 = 0; c0 < G16; c0 += chunk16) {
      if (!one_chunk) {
        __syncthreads();
        load_lut(c0);
        __syncthreads();
      }
      const int c1 = min(G16, c0 + chunk16);
      uint32_t accA[V][H], accB[V][H];
#pragma unroll
      for (int j = 0; j < V; ++j)
#pragma unroll
        for (int h = 0; h < H; ++h) { accA[j][h] = 0u; accB[j][h] = 0u; }
      constexpr int NB = V == 1 ? 4 : 2;  // 16-byte code loads in flight per vector
      for (int gb = c0; gb < c1; gb += NB) {
        uint4 cb[V][NB];
#pragma unroll
        for (int j = 0; j < V; ++j)
#pragma unroll
          for (int u = 0; u < NB; ++u)
            // unconditional, clamped (in-bounds) load: no branch and no select
            // around it, so it stays in flight until its block is consumed.
            // Blocks past c1 are never consumed; lanes of invalid vectors
            // (pointing at vector 0) get -inf scores.
            cb[j][u] = cp[j][(long)min(gb + u, G16 - 1) * 64];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const int g16 = gb + u;
          if (g16 >= c1) break;
          const W* lw = reinterpret_cast<const W*>(lut_s) + (long)(g16 - c0) * 16 * 32;
#pragma unroll
          for (int j = 0; j < V; ++j) {
            const uint4 c = cb[j][u];
            const uint32_t words[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
            for (int e = 0; e < 16; ++e) {
              const uint32_t w = words[e >> 2];
              const int bp = (e & 3) * 8;
              const W rh = lw[e * 32 + nib(w, bp + 4)];
              const W rl = lw[e * 32 + 16 + nib(w, bp)];
              if constexpr (T == 1) {
                accA[j][0] += uint32_t(rh) + uint32_t(rl);
              } else if constexpr (T == 4) {
                const uint32_t rr = rh + rl;  // bytes <= 254: no carries
                accA[j][0] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, accA[j][0]) + __builtin_bit_cast(u16x2, rr));
                accB[j][0] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, accB[j][0]) +
                                                              (__builtin_bit_cast(u16x2, rr) >> (unsigned short)8));
              } else {
                const uint32_t r0 = rh.x + rl.x, r1 = rh.y + rl.y;
                accA[j][0] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, accA[j][0]) + __builtin_bit_cast(u16x2, r0));
                accB[j][0] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, accB[j][0]) +
                                                              (__builtin_bit_cast(u16x2, r0) >> (unsigned short)8));
                accA[j][1] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, accA[j][1]) + __builtin_bit_cast(u16x2, r1));
                accB[j][1] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, accB[j][1]) +
                                                              (__builtin_bit_cast(u16x2, r1) >> (unsigned short)8));
              }
            }
          }
          if (((g16 + 1) & 15) == 0 || g16 == G16 - 1) {
            // window flush: exact u16 sums per query -> the reference's float chain
#pragma unroll
            for (int j = 0; j < V; ++j) {
              if constexpr (T == 1) {
                fa[j][0] = __fmaf_rn(scl[0], (float)accA[j][0], fa[j][0]);
                accA[j][0] = 0u;
              } else {
#pragma unroll
                for (int h = 0; h < H; ++h) {
                  const u16x2 a = __builtin_bit_cast(u16x2, accA[j][h]);
                  const u16x2 b = __builtin_bit_cast(u16x2, accB[j][h]);
                  const uint32_t s0 = (unsigned short)(a.x - (unsigned short)(b.x << 8));
                  const uint32_t s1 = b.x;
                  const uint32_t s2 = (unsigned short)(a.y - (unsigned short)(b.y << 8));
                  const uint32_t s3 = b.y;
                  fa[j][4 * h + 0] = __fmaf_rn(scl[4 * h + 0], (float)s0, fa[j][4 * h + 0]);
                  fa[j][4 * h + 1] = __fmaf_rn(scl[4 * h + 1], (float)s1, fa[j][4 * h + 1]);
                  fa[j][4 * h + 2] = __fmaf_rn(scl[4 * h + 2], (float)s2, fa[j][4 * h + 2]);
                  fa[j][4 * h + 3] = __fmaf_rn(scl[4 * h + 3], (float)s3, fa[j][4 * h + 3]);
                  accA[j][h] = 0u;
                  accB[j][h] = 0u;
                }
              }
            }
          }
        }
      }
    }

    admit_scores<T, V>(sh, fa, vv, valid, alpha, q0, nq, k, cap);
  }
  write_partials<T>(sh, q0, nq, n_ranges, r, k, cap, part_s, part_i, part_cnt);
}

// ---------------------------------------------------------------------------
// Byte-pair scan. A code byte holds two nibbles of one group (hi = first
// code). Instead of two 16-entry lookups per byte, each group's two nibble
// tables are expanded into one 256-entry table over the whole byte:
//   pair[g][b] = lut[g][b >> 4] + lut[g][16 + (b & 15)]   (<= 254)
// so one LDS read serves both nibbles of T queries. The window sums (256
// groups, the reference's kFlushEvery) are the same exact integers, so the
// float chain fa = fmaf(scale, (float)A_w, fa) and the scores are unchanged.
//
// Entries hold T queries: u8 per query for T = 1 and T = 8 (T = 8 splits the
// bytes into u16 lanes after the read), u16 per query for T = 2 and 4. The
// table lives in a 64 KiB LDS chunk of CG groups, rebuilt per chunk from an
// 8 KiB slice of the query tile's source rows ([group][32][T] of u8 or u16),
// which is copied into an LDS stage by DMA one chunk ahead.
// ---------------------------------------------------------------------------
constexpr int kPairTableBytes = 65536;
constexpr int kPairStageBytes = 8192;

template <int T> struct PairFmt { static constexpr bool wide = (T == 2 || T == 4); };
template <int T> constexpr int pair_entry_bytes() { return T * (PairFmt<T>::wide ? 2 : 1); }
// 16-group code blocks per 64 KiB table chunk
__host__ __device__ constexpr int pair_chunk_blocks(int T) { return kPairTableBytes / (256 * T * ((T == 2 || T == 4) ? 2 : 1)) / 16; }
// 16-group blocks of source rows per query tile: padded to whole chunks
__host__ __device__ inline long pair_src_blocks(int T, int G16) {
  const int cb = pair_chunk_blocks(T);
  return (long)((G16 + cb - 1) / cb) * cb;
}

// Source rows of every query tile: element (tile, g, e, t) = lut[q][g][e] of
// query q = tile * T + t (zero for padded groups / queries), u16 when wide.
template <int T>
__global__ void pair_source_kernel(const uint8_t* __restrict__ luts, long nq, int G, int G16,
                                   uint8_t* __restrict__ out) {
  const long n_qt = (nq + T - 1) / T;
  const long per_tile = pair_src_blocks(T, G16) * 16 * 32 * T;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n_qt * per_tile; i += (long)gridDim.x * blockDim.x) {
    const long qt = i / per_tile;
    const long r = i - qt * per_tile;
    const int t = (int)(r % T);
    const int e = (int)((r / T) & 31);
    const int g = (int)(r / (T * 32));
    const long q = qt * T + t;
    const uint8_t v = (g < G && q < nq) ? luts[q * (long)G * 32 + (long)g * 32 + e] : 0;
    if constexpr (PairFmt<T>::wide) reinterpret_cast<uint16_t*>(out)[i] = v;
    else out[i] = v;
  }
}

// Table chunk from the staged source rows of its CG groups. The table is
// 4096 16-byte pieces; instruction v of thread t writes piece v * 1024 + t,
// so each wavefront store covers 1 KiB of consecutive bytes (no bank
// conflicts). Piece c holds entries (g, h, l0 .. l0 + 16 / EB - 1) of its
// first entry c * 16 / EB = (g, h, l0): lo[g][l0 ..] + hi[g][h].
template <int T>
__device__ __forceinline__ void build_pair_chunk(const uint8_t* __restrict__ stage, uint8_t* __restrict__ table) {
  constexpr int EB = pair_entry_bytes<T>();
  constexpr int EPC = 16 / EB;  // entries per piece
  // piece v * 1024 + t lies 4 * EPC groups after piece t: its source rows are
  // a fixed 4 * EPC * 32 * EB = 2 KiB further into the stage
  constexpr int kVStep = 4 * EPC * 32 * EB;
  static_assert(kScanThreads * EPC % 256 == 0 && kVStep == 2048, "pieces of one thread differ in g only");
  const int tid = threadIdx.x;
  const int i0 = tid * EPC;
  const int g = i0 >> 8, h = (i0 >> 4) & 15, l0 = i0 & 15;
  const uint8_t* lo_src = stage + (g * 32 + 16 + l0) * EB;
  const uint8_t* hi_src = stage + (g * 32 + h) * EB;
#pragma unroll
  for (int v = 0; v < kPairTableBytes / 16 / kScanThreads; ++v) {
    const int c = v * kScanThreads + tid;
    const uint4 lo = *reinterpret_cast<const uint4*>(lo_src + v * kVStep);
    uint32_t o[4] = {lo.x, lo.y, lo.z, lo.w};
    if constexpr (EB == 1) {
      const uint32_t bc = (uint32_t)hi_src[v * kVStep] * 0x01010101u;  // bytes <= 127: no carries
#pragma unroll
      for (int w = 0; w < 4; ++w) o[w] += bc;
    } else {
      constexpr int WPE = EB / 4;  // 32-bit words per entry
      uint32_t hw[WPE];
#pragma unroll
      for (int w = 0; w < WPE; ++w) hw[w] = reinterpret_cast<const uint32_t*>(hi_src + v * kVStep)[w];
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        if constexpr (PairFmt<T>::wide)
          o[w] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, o[w]) + __builtin_bit_cast(u16x2, hw[w % WPE]));
        else
          o[w] += hw[w % WPE];  // bytes <= 127: no carries
      }
    }
    reinterpret_cast<uint4*>(table)[c] = make_uint4(o[0], o[1], o[2], o[3]);
    // one piece at a time: the build runs with the code ring and the score
    // chains live, so it must not hold four pieces' sources in registers
    __builtin_amdgcn_sched_barrier(0);
  }
}

// f(integral_constant<int, I>) for I = 0 .. N-1, unrolled at compile time.

// Byte e of w times 2^SH in one VALU op: the SDWA operand select picks the
// byte, so the LDS address of a table entry costs one instruction.
template <int E, int SH>
__device__ __forceinline__ uint32_t byte_scaled(uint32_t w) {
  static_assert(E >= 0 && E < 4 && SH >= 0 && SH < 8, "byte and shift");
  uint32_t r;
  if constexpr (SH == 0) {
    r = __builtin_amdgcn_ubfe(w, 8 * E, 8);
  } else if constexpr (E == 0) {
    asm("v_lshlrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
        : "=v"(r) : "v"(w), "i"(SH));
  } else if constexpr (E == 1) {
    asm("v_lshlrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_1"
        : "=v"(r) : "v"(w), "i"(SH));
  } else if constexpr (E == 2) {
    asm("v_lshlrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_2"
        : "=v"(r) : "v"(w), "i"(SH));
  } else {
    asm("v_lshlrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_3"
        : "=v"(r) : "v"(w), "i"(SH));
  }
  return r;
}

// acc (two words, low first) += w as one 64-bit integer add. kAsm writes the
// instruction out: the compiler splits an add of two separately computed
// words into a 64-bit add and a 32-bit add.
template <bool kAsm = false>
__device__ __forceinline__ void acc64_add(uint32_t (&acc)[2], uint2 w) {
  uint64_t a = ((uint64_t)acc[1] << 32) | acc[0];
  const uint64_t b = ((uint64_t)w.y << 32) | w.x;
  if constexpr (kAsm)
    asm("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(a) : "v"(b), "v"(a));
  else
    a += b;
  acc[0] = (uint32_t)a;
  acc[1] = (uint32_t)(a >> 32);
}

// 32-bit accumulators per vector (u16 lanes; narrow T = 8 keeps a second set)
template <int T> constexpr int pair_acc_words() { return T == 1 ? 1 : (PairFmt<T>::wide ? T / 2 : T / 4); }

// One table chunk of the scan: CB code blocks of each of the thread's V
// vectors against the chunk's table, window flushes into the float chains,
// code-ring reloads; first, when dma_src is set, the DMA of the next chunk's
// source rows into the stage (global_load_lds: wavefront w copies bytes
// [1024 w, 1024 (w + 1)), 16 per lane, with no registers in flight).
// The DMA is complete at the barrier before the next build. (The compiler
// still retires it, with every load in flight, at the first wait for a code
// block after it: measured faster than staging the rows through registers,
// which costs the two registers that push the two-vector variant into spills.)
template <int T, int V, int NR>
__device__ __forceinline__ void pair_chunk(const uint8_t* __restrict__ table, uint8_t* __restrict__ stage,
                                           const uint8_t* __restrict__ dma_src, const float* __restrict__ scl,
                                           const uint4* const (&cp)[V], int c, int G16, uint4 (&ring)[NR][V],
                                           uint32_t (&acc)[V][pair_acc_words<T>()],
                                           uint32_t (&acc2)[V][pair_acc_words<T>()], float (&fa)[V][T]) {
  constexpr bool kWide = PairFmt<T>::wide;
  constexpr int EB = pair_entry_bytes<T>();
  constexpr int CB = kPairTableBytes / (256 * EB) / 16;  // 16-group code blocks per chunk
  constexpr int NA = pair_acc_words<T>();
  constexpr int SH = EB == 8 ? 3 : EB == 4 ? 2 : 0;  // log2 of the entry size
  const int tid = threadIdx.x;
  if (dma_src != nullptr) {
    const int w = tid >> 6;
    if (w < kPairStageBytes / 1024)
      __builtin_amdgcn_global_load_lds(dma_src + w * 1024 + (tid & 63) * 16, stage + w * 1024, 16, 0, 0);
  }
  auto load_block = [&](int j, int b) -> uint4 {
    // unconditional, clamped (in-bounds) load: no branch and no select around
    // it, so it stays in flight until its block is consumed. Blocks past G16
    // are never consumed; lanes of invalid vectors (pointing at vector 0)
    // get -inf scores.
    return cp[j][(long)min(b, G16 - 1) * 64];
  };
#pragma unroll
  for (int u = 0; u < CB; ++u) {
    const int bb = c * CB + u;
    const bool live = bb < G16;  // uniform
#pragma unroll
    for (int j = 0; j < V; ++j) {
      if (live) {
        const uint8_t* tb = table + u * 16 * 256 * EB;
        const uint32_t words[4] = {ring[u % NR][j].x, ring[u % NR][j].y, ring[u % NR][j].z, ring[u % NR][j].w};
        static_for<16>([&](auto ec) {
          constexpr int e = decltype(ec)::value;
          const uint8_t* p = tb + e * 256 * EB + byte_scaled<e & 3, SH>(words[e >> 2]);
          if constexpr (T == 1) {
            acc[j][0] += *p;
          } else if constexpr (kWide && NA == 2) {
            // the whole entry in one read (its offset folds into the
            // instruction); its four u16 lanes added by one 64-bit add (every
            // lane's window sum is < 2^16, so no carry crosses a lane)
            const uint2 w2 = *reinterpret_cast<const uint2*>(p);
            acc64_add(acc[j], w2);
          } else if constexpr (kWide) {
            const uint32_t e0 = *reinterpret_cast<const uint32_t*>(p);
            acc[j][0] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, acc[j][0]) + __builtin_bit_cast(u16x2, e0));
          } else {
            // Entry = bytes of queries 0 .. 7. acc adds it as it is, one
            // 64-bit add: sum over q of 2^(8q) S_q (mod 2^64); acc2 adds the
            // odd queries' bytes alone as u16 lanes (S_1, S_3, S_5, S_7; two
            // v_perm, 0x0c = zero). Every window sum S_q is < 2^16, so the
            // flush recovers the even queries exactly from acc - 2^8 acc2.
            const uint2 w2 = *reinterpret_cast<const uint2*>(p);
            acc64_add(acc[j], w2);
            acc64_add<true>(acc2[j], make_uint2(__builtin_amdgcn_perm(0u, w2.x, 0x0c030c01u),
                                          __builtin_amdgcn_perm(0u, w2.y, 0x0c030c01u)));
          }
        });
      }
      // the slot is reloaded only after its block is consumed: no second copy
      // of the block stays live (issued even past G16: clamped, unused)
      ring[u % NR][j] = load_block(j, bb + NR);
    }
    if (live && (((bb + 1) & 15) == 0 || bb == G16 - 1)) {
      // window flush: exact integer sums per query -> the reference's float chain
#pragma unroll
      for (int j = 0; j < V; ++j) {
        if constexpr (T == 1) {
          fa[j][0] = __fmaf_rn(scl[0], (float)acc[j][0], fa[j][0]);
        } else if constexpr (kWide) {
#pragma unroll
          for (int a = 0; a < NA; ++a) {
            const u16x2 sm = __builtin_bit_cast(u16x2, acc[j][a]);
            fa[j][2 * a] = __fmaf_rn(scl[2 * a], (float)sm.x, fa[j][2 * a]);
            fa[j][2 * a + 1] = __fmaf_rn(scl[2 * a + 1], (float)sm.y, fa[j][2 * a + 1]);
          }
        } else {
          const uint64_t a64 = ((uint64_t)acc[j][1] << 32) | acc[j][0];
          const uint64_t b64 = ((uint64_t)acc2[j][1] << 32) | acc2[j][0];
          const uint64_t e64 = a64 - (b64 << 8);  // sum over q of 2^(16q) S_2q
          const uint32_t evw[2] = {(uint32_t)e64, (uint32_t)(e64 >> 32)};
#pragma unroll
          for (int a = 0; a < NA; ++a) {
            const u16x2 od = __builtin_bit_cast(u16x2, acc2[j][a]);  // queries 4a + 1, 4a + 3
            const u16x2 ev = __builtin_bit_cast(u16x2, evw[a]);      // queries 4a, 4a + 2
            fa[j][4 * a + 0] = __fmaf_rn(scl[4 * a + 0], (float)ev.x, fa[j][4 * a + 0]);
            fa[j][4 * a + 1] = __fmaf_rn(scl[4 * a + 1], (float)od.x, fa[j][4 * a + 1]);
            fa[j][4 * a + 2] = __fmaf_rn(scl[4 * a + 2], (float)ev.y, fa[j][4 * a + 2]);
            fa[j][4 * a + 3] = __fmaf_rn(scl[4 * a + 3], (float)od.y, fa[j][4 * a + 3]);
          }
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) { acc[j][a] = 0u; acc2[j][a] = 0u; }
      }
    }
  }
}

#ifndef VC_PAIR_MIN_WAVES
#define VC_PAIR_MIN_WAVES 8  // waves per SIMD the register budget must allow
#endif
// Wavefront lists for one-query tiles with k <= 64: each wavefront keeps the
// best 64 of the vectors it scored, sorted across its lanes (lane r holds
// rank r by (score desc, id asc)), with no barrier per step: a step whose
// scores include one better than the list's k-th sorts them and merges them
// in (the best of the list against the reversed new ones, a bitonic
// sequence, then six half-cleaner stages). At the end of the work item the
// workgroup's lists meet in LDS and one wavefront merges them into the item's
// partial list. The workgroup-wide admission (candidate buffer, barriers,
// rank merges) costs the one-query scan ~9 % of its time.
struct WaveList {
  float s = -INFINITY;  // this lane's rank of the list
  uint32_t i = 0xffffffffu;
  float thr_s = -INFINITY;  // the list's k-th entry
  uint32_t thr_i = 0xffffffffu;
};

__device__ __forceinline__ void wave_list_merge_sorted(WaveList& w, float x, uint32_t xi, int lane) {
  // x, xi: 64 entries sorted descending across the lanes; reversed (lane ^ 63
  // = the row mirror ^15, then ^16 and ^32) they form a bitonic sequence with the list
  auto rev = [&](uint32_t v) {
    return xor_lane<32>(xor_lane<16>((uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false), lane), lane);
  };
  const float ys = __uint_as_float(rev(__float_as_uint(x)));
  const uint32_t yi = rev(xi);
  if (better(ys, yi, w.s, w.i)) { w.s = ys; w.i = yi; }
  static_for<6>([&](auto si) {
    constexpr int st = 32 >> decltype(si)::value;
    const float os = __uint_as_float(xor_lane<st>(__float_as_uint(w.s), lane));
    const uint32_t oi = xor_lane<st>(w.i, lane);
    const bool ob = better(os, oi, w.s, w.i);
    if (((lane & st) == 0) ? ob : !ob) { w.s = os; w.i = oi; }
  });
}

__device__ __forceinline__ void wave_list_admit(WaveList& w, float sc, uint32_t vid, bool valid, int k, int lane) {
  const bool cand = valid && better(sc, vid, w.thr_s, w.thr_i);
  if (__ballot(cand) == 0ull) return;  // uniform
  float x = cand ? sc : -INFINITY;
  uint32_t xi = cand ? vid : 0xffffffffu;
  wave_sort_desc(x, xi, lane);
  wave_list_merge_sorted(w, x, xi, lane);
  w.thr_s = __shfl(w.s, k - 1);
  w.thr_i = (uint32_t)__shfl((int)w.i, k - 1);
}

// The workgroup's wavefront lists into partial list `list` of query q
// (part_s/part_i [q][n_lists][k], part_cnt [q][n_lists]); scratch: 8 KiB of
// LDS no wavefront reads any more. Every thread calls it (it synchronizes).
__device__ void wave_lists_write(const WaveList& w, unsigned char* __restrict__ scratch, long q, int n_lists, int list,
                                 int k, float* __restrict__ part_s, uint32_t* __restrict__ part_i,
                                 int* __restrict__ part_cnt) {
  const int tid = threadIdx.x, lane = tid & 63;
  float* ws = reinterpret_cast<float*>(scratch);
  uint32_t* wi = reinterpret_cast<uint32_t*>(scratch) + kScanThreads;
  __syncthreads();  // the scratch is free
  ws[tid] = w.s;
  wi[tid] = w.i;
  __syncthreads();
  if ((tid >> 6) == 0) {
    WaveList m = w;
    for (int v = 1; v < (int)(blockDim.x >> 6); ++v)
      wave_list_merge_sorted(m, ws[v * 64 + lane], wi[v * 64 + lane], lane);
    const bool has = lane < k && m.i != 0xffffffffu;
    const long slot = q * n_lists + list;
    if (has) {
      part_s[slot * k + lane] = m.s;
      part_i[slot * k + lane] = m.i;
    }
    const int cnt = __popcll(__ballot(has));
    if (lane == 0) part_cnt[slot] = cnt;
  }
}

// One work item of the byte-pair scan: the T queries of one tile (first
// query q0; their source rows at tile_src) over vectors [r_begin, r_end) of
// a code set, ending in the item's sorted lists (list `list` of each query's
// n_lists). table / stage: the caller's static LDS; smem: dynamic LDS of at
// least topk_lds(T, k, cap) bytes.
template <int T, int V>
__device__ __forceinline__ void scan_pair_body(uint8_t* __restrict__ table, uint8_t* __restrict__ stage,
                                               unsigned char* __restrict__ smem, const uint8_t* __restrict__ tiled,
                                               const float* __restrict__ alpha, int G16,
                                               const uint8_t* __restrict__ tile_src,
                                               const float* __restrict__ lut_scale,
                                               const float* __restrict__ bias_total, long q0, long nq, int k, int cap,
                                               long r_begin, long r_end, int n_lists, int list,
                                               float* __restrict__ part_s, uint32_t* __restrict__ part_i,
                                               int* __restrict__ part_cnt) {
  constexpr int EB = pair_entry_bytes<T>();
  constexpr int CG = kPairTableBytes / (256 * EB);  // groups per table chunk
  constexpr int CB = CG / 16;                       // 16-group code blocks per chunk
  static_assert(CG * 32 * EB == kPairStageBytes && kPairStageBytes == kScanThreads * 8, "stage = 8 bytes per thread");
  constexpr int NR = (T == 1 && V == 1) ? 4 : 2;  // code blocks in flight per vector
  static_assert(CB % NR == 0, "a chunk holds whole rounds of the code-block ring");
  constexpr int NA = pair_acc_words<T>();
  const int tid = threadIdx.x;

  TopkShared sh;
  init_topk(sh, smem, T, k, cap);
  // one query and k <= 64: wavefront lists instead of the workgroup's admission
  const bool wave_lists = T == 1 && k <= 64;
  WaveList wl;
  // per-query LUT scale and bias in LDS (read at window flushes and step
  // starts only; keeps them out of the scalar register budget)
  float* scl = reinterpret_cast<float*>(smem + (size_t)T * (2 * (size_t)k + 2 * (size_t)cap) * 4 + 20 * T);
  float* bias = scl + T;
  if (tid < T) {
    const long q = q0 + tid < nq ? q0 + tid : nq - 1;
    scl[tid] = lut_scale[q];
    bias[tid] = bias_total[q];
  }

  // the source rows are padded to whole chunks, so every DMA is in bounds;
  // chunk 0's rows are staged by plain loads before the first build
  const int n_chunks = (G16 + CB - 1) / CB;
  reinterpret_cast<uint64_t*>(stage)[tid] = reinterpret_cast<const uint64_t*>(tile_src)[tid];
  __syncthreads();
  bool built = false;

  const long step = (long)kScanThreads * V;
  for (long base = r_begin; base < r_end; base += step) {
    long vv[V];
    bool valid[V];
    const uint4* cp[V];
    float fa[V][T], avp[V];
    uint32_t acc[V][NA], acc2[V][NA];
#pragma unroll
    for (int j = 0; j < V; ++j) {
      vv[j] = base + tid + (long)j * kScanThreads;
      valid[j] = vv[j] < r_end;
      const long vs = valid[j] ? vv[j] : 0;
      cp[j] = reinterpret_cast<const uint4*>(tiled) + (vs >> 6) * (long)G16 * 64 + (vs & 63);
      // tiles of one or two queries: alpha loaded now, so its latency is not
      // exposed at the admission (a workgroup of the bandwidth regime scans
      // only a few steps)
      if constexpr (T <= 2) avp[j] = alpha[vs];
#pragma unroll
      for (int t = 0; t < T; ++t) fa[j][t] = bias[t];
#pragma unroll
      for (int a = 0; a < NA; ++a) { acc[j][a] = 0u; acc2[j][a] = 0u; }
    }
    // Code blocks NR ahead. The block loop of a chunk is unrolled and CB is a
    // multiple of NR, so each block always uses the same ring slot: no
    // register rotation, and each wait covers exactly the block consumed.
    uint4 ring[NR][V];
#pragma unroll
    for (int u = 0; u < NR; ++u)
#pragma unroll
      for (int j = 0; j < V; ++j) ring[u][j] = cp[j][(long)min(u, G16 - 1) * 64];

    for (int c = 0; c < n_chunks; ++c) {
      const uint8_t* dma = nullptr;
      if (n_chunks > 1 || !built) {
        __syncthreads();  // table free (previous chunk's lookups done); stage DMA complete
        build_pair_chunk<T>(stage, table);
        __syncthreads();  // table ready; stage consumed
        // rows of the next chunk (of the next step after the last one); none
        // after the last chunk of the last step
        if (n_chunks > 1 && (c + 1 < n_chunks || base + step < r_end))
          dma = tile_src + (long)(c + 1 < n_chunks ? c + 1 : 0) * kPairStageBytes;
        built = true;
      }
      pair_chunk<T, V, NR>(table, stage, dma, scl, cp, c, G16, ring, acc, acc2, fa);
    }
    if constexpr (T == 1) {
      if (wave_lists) {
#pragma unroll
        for (int j = 0; j < V; ++j)
          wave_list_admit(wl, valid[j] ? fa[j][0] * avp[j] + 0.f : -INFINITY, (uint32_t)vv[j], valid[j], k,
                          tid & 63);
        continue;
      }
    }
    admit_scores<T, V>(sh, fa, vv, valid, alpha, q0, nq, k, cap, nullptr, T <= 2 ? &avp : nullptr);
  }
  if constexpr (T == 1) {
    if (wave_lists) {
      wave_lists_write(wl, table, q0, n_lists, list, k, part_s, part_i, part_cnt);
      return;
    }
  }
  write_partials<T>(sh, q0, nq, n_lists, list, k, cap, part_s, part_i, part_cnt);
}

template <int T, int V>
__global__ __launch_bounds__(kScanThreads, VC_PAIR_MIN_WAVES) void scan_pair_kernel(
    const uint8_t* __restrict__ tiled, const float* __restrict__ alpha, long n, int G16,
    const uint8_t* __restrict__ src, const float* __restrict__ lut_scale,
    const float* __restrict__ bias_total, long nq, int k, int cap, int n_ranges, long range_len,
    float* __restrict__ part_s, uint32_t* __restrict__ part_i, int* __restrict__ part_cnt) {
  constexpr int CB = kPairTableBytes / (256 * pair_entry_bytes<T>()) / 16;
  // the table at a fixed LDS address, so entry offsets fold into the reads
  __shared__ __align__(16) uint8_t table[kPairTableBytes];
  __shared__ __align__(16) uint8_t stage[kPairStageBytes];
  extern __shared__ __align__(16) unsigned char smem[];
  const long n_qt = (nq + T - 1) / T;
  const long L = blockIdx.x;
  const int r = (int)(L / n_qt);   // range-major: concurrent workgroups share codes
  const long qt = L % n_qt;
  const long r_begin = (long)r * range_len;
  const long r_end = min(n, r_begin + range_len);
  const int n_chunks = (G16 + CB - 1) / CB;
  scan_pair_body<T, V>(table, stage, smem, tiled, alpha, G16, src + qt * ((long)n_chunks * kPairStageBytes),
                       lut_scale, bias_total, qt * T, nq, k, cap, r_begin, r_end, n_ranges, r, part_s, part_i,
                       part_cnt);
}

// Many code sets in one persistent launch: each workgroup walks the work
// items blockIdx.x, blockIdx.x + gridDim.x, ... and runs each as the
// single-set scan runs its (tile, range) workgroup, at its code set's query
// tile (1, 2, 4 or 8; tile 1 always with one vector per thread), so every
// item's lists are bit-identical to what a search of its set alone merges.
template <int T, int V>
__device__ __forceinline__ void multi_item(uint8_t* table, uint8_t* stage, unsigned char* smem, const MultiColl& c,
                                           const MultiItem& w) {
  constexpr int CB = kPairTableBytes / (256 * pair_entry_bytes<T>()) / 16;
  const int n_chunks = (c.G16 + CB - 1) / CB;
  scan_pair_body<T, V>(table, stage, smem, c.tiled, c.alpha, c.G16, c.src + w.qt * ((long)n_chunks * kPairStageBytes),
                       c.lut_scale, c.bias_total, w.qt * T, c.nq, c.k, c.cap, w.v_begin, w.v_end, c.n_lists, w.list,
                       c.part_s, c.part_i, c.part_cnt);
}

template <int V>
__global__ __launch_bounds__(kScanThreads, VC_PAIR_MIN_WAVES) void multi_scan_kernel(
    const MultiColl* __restrict__ colls, const MultiItem* __restrict__ items, long n_items) {
  __shared__ __align__(16) uint8_t table[kPairTableBytes];
  __shared__ __align__(16) uint8_t stage[kPairStageBytes];
  extern __shared__ __align__(16) unsigned char smem[];
  for (long it = blockIdx.x; it < n_items; it += gridDim.x) {
    const MultiItem w = items[it];
    const MultiColl c = colls[w.coll];
    switch (c.tile) {
      case 8: multi_item<8, V>(table, stage, smem, c, w); break;
      case 4: multi_item<4, V>(table, stage, smem, c, w); break;
      case 2: multi_item<2, V>(table, stage, smem, c, w); break;
      default: multi_item<1, 1>(table, stage, smem, c, w); break;
    }
    __syncthreads();  // the next item reuses the LDS
  }
}

// Merge the per-range sorted lists of one query into its final row (one
// wavefront): each lane holds up to 8 lists and keeps their head entries in
// registers; every output is a wave argmax under (score desc, id asc).
// ps / pi: the query's lists [R][k]; pc: their lengths [R].
constexpr int kMaxListsPerLane = 8;  // n_ranges <= 512

// Best (score desc, id asc) entry of a full wavefront, the same on every lane:
// an all-reduce inside each row of 16 lanes through DPP (quad xor 1, quad xor
// 2, half-row mirror, row mirror), then the four row winners through
// readlane. The order is total, so any reduction tree gives the same winner.
template <int kCtrl>
__device__ __forceinline__ void dpp_best_step(bool& h, float& s, uint32_t& i) {
  const int oh = __builtin_amdgcn_update_dpp(0, (int)h, kCtrl, 0xF, 0xF, false);
  const float os = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s), kCtrl, 0xF, 0xF, false));
  const uint32_t oi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)i, kCtrl, 0xF, 0xF, false);
  if (oh && (!h || better(os, oi, s, i))) {
    h = true;
    s = os;
    i = oi;
  }
}

__device__ __forceinline__ void wave_best(bool& h, float& s, uint32_t& i) {
  dpp_best_step<0xB1>(h, s, i);   // quad_perm [1, 0, 3, 2]
  dpp_best_step<0x4E>(h, s, i);   // quad_perm [2, 3, 0, 1]
  dpp_best_step<0x141>(h, s, i);  // row_half_mirror
  dpp_best_step<0x140>(h, s, i);  // row_mirror
  bool bh = false;
  float bs = 0.f;
  uint32_t bi = 0u;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const bool rh = __builtin_amdgcn_readlane((int)h, 16 * r) != 0;
    const float rs = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 16 * r));
    const uint32_t ri = (uint32_t)__builtin_amdgcn_readlane((int)i, 16 * r);
    if (rh && (!bh || better(rs, ri, bs, bi))) {
      bh = true;
      bs = rs;
      bi = ri;
    }
  }
  h = bh;
  s = bs;
  i = bi;
}

__device__ void merge_query(const float* __restrict__ ps, const uint32_t* __restrict__ pi,
                            const int* __restrict__ pc, int R, int k, int k_out,
                            const uint64_t* __restrict__ id_map, float* __restrict__ out_s,
                            uint64_t* __restrict__ out_i, int lane) {
  int head[kMaxListsPerLane], len[kMaxListsPerLane];
  float hs[kMaxListsPerLane];
  uint32_t hid[kMaxListsPerLane];
#pragma unroll
  for (int j = 0; j < kMaxListsPerLane; ++j) {
    const int r = lane + 64 * j;
    head[j] = 0;
    len[j] = r < R ? pc[r] : 0;
    hs[j] = len[j] > 0 ? ps[(long)r * k] : 0.f;
    hid[j] = len[j] > 0 ? pi[(long)r * k] : 0u;
  }
  int j_out = 0;
  for (; j_out < k_out; ++j_out) {
    bool have = false;
    float bs = 0.f;
    uint32_t bi = 0u;
    int bj = -1;
#pragma unroll
    for (int j = 0; j < kMaxListsPerLane; ++j)
      if (head[j] < len[j] && (!have || better(hs[j], hid[j], bs, bi))) {
        have = true;
        bs = hs[j];
        bi = hid[j];
        bj = j;
      }
    bool wh = have;
    float ws = bs;
    uint32_t wi = bi;
    wave_best(wh, ws, wi);
    if (!wh) break;  // the same on every lane
    if (have && bs == ws && bi == wi) {
#pragma unroll
      for (int j = 0; j < kMaxListsPerLane; ++j)
        if (j == bj) {
          ++head[j];
          if (head[j] < len[j]) {
            const long p = (long)(lane + 64 * j) * k + head[j];
            hs[j] = ps[p];
            hid[j] = pi[p];
          }
        }
    }
    if (lane == 0) {
      out_s[j_out] = ws;
      out_i[j_out] = id_map ? id_map[wi] : (uint64_t)wi;
    }
  }
  for (int j = j_out + lane; j < k_out; j += 64) {
    out_s[j] = 0.f;
    out_i[j] = 0u;
  }
}

__global__ __launch_bounds__(64) void merge_ranges_kernel(const float* __restrict__ part_s,
                                                          const uint32_t* __restrict__ part_i,
                                                          const int* __restrict__ part_cnt, long nq,
                                                          int R, int k, int k_out,
                                                          const uint64_t* __restrict__ id_map,
                                                          float* __restrict__ out_s,
                                                          uint64_t* __restrict__ out_i) {
  for (long q = blockIdx.x; q < nq; q += gridDim.x)
    merge_query(part_s + q * R * (long)k, part_i + q * R * (long)k, part_cnt + q * R, R, k, k_out, id_map,
                out_s + q * k_out, out_i + q * k_out, threadIdx.x);
}

// The same with the query's lists first copied to LDS by the whole workgroup
// (one coalesced pass), when they fit.
constexpr int kMergeStageBytes = 48 * 1024;

// Partial lists staged in LDS (scores, ids and counts in one round of loads,
// eight per thread in flight: the lists come from every XCD's share of the
// scan, so each round is a trip to memory), then one wavefront merges. Small
// batches launch it with 1024 threads, large ones with 256.
__global__ __launch_bounds__(1024) void merge_ranges_lds_kernel(const float* __restrict__ part_s,
                                                               const uint32_t* __restrict__ part_i,
                                                               const int* __restrict__ part_cnt, long nq,
                                                               int R, int k, int k_out,
                                                               const uint64_t* __restrict__ id_map,
                                                               float* __restrict__ out_s,
                                                               uint64_t* __restrict__ out_i) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = R * k;
  float* ls = reinterpret_cast<float*>(smem);
  uint32_t* li = reinterpret_cast<uint32_t*>(ls + n);
  int* lc = reinterpret_cast<int*>(li + n);
  const int tid = threadIdx.x, nt = blockDim.x;
  for (long q = blockIdx.x; q < nq; q += gridDim.x) {
    __syncthreads();
    const float* ss = part_s + q * n;
    const uint32_t* si = part_i + q * n;
    const int* sc = part_cnt + q * R;
    for (int base = 0; base < n || base < R; base += 8 * nt) {
      float a[8];
      uint32_t b[8];
      int c[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = base + u * nt + tid;
        a[u] = j < n ? ss[j] : 0.f;
        b[u] = j < n ? si[j] : 0u;
        c[u] = j < R ? sc[j] : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = base + u * nt + tid;
        if (j < n) {
          ls[j] = a[u];
          li[j] = b[u];
        }
        if (j < R) lc[j] = c[u];
      }
    }
    __syncthreads();
    if (threadIdx.x < 64)
      merge_query(ls, li, lc, R, k, k_out, id_map, out_s + q * k_out, out_i + q * k_out, threadIdx.x);
  }
}

// ---------------------------------------------------------------------------
// Bucketed scan (the multi-bucket plumbing of score_prepared_into, src/search/
// search.cpp:1156-1267, driven over many buckets per query in one launch).
// Codes are ordered by bucket: bucket b holds positions [boff[b], boff[b] +
// bsz[b]). Query q probes the first probe_n[q] buckets of its probe list
// probe_b[q][0 .. P), each with the score offset probe_off[q][j] (the
// query's dot product with the bucket's encode centroid); probe_pre[q][j] is
// the number of vectors in the list before entry j. The probed vectors of q
// are cut into work items of item_len[q] consecutive list vectors; item_off
// [q] .. item_off[q + 1] are q's items (item_off[nq] = all items). Every
// vector scores fa * alpha + offset exactly as the reference's AVX2 path
// scores a bucket; each item keeps its exact top-k by (score desc, position
// asc), and merge_items_kernel merges a query's items and maps positions to
// ids.
// ---------------------------------------------------------------------------

constexpr int kBucketScanGrid = 512;  // persistent workgroups (two per CU)

template <int V>
__global__ __launch_bounds__(kScanThreads, VC_PAIR_MIN_WAVES) void bucket_scan_kernel(BucketScanParams a) {
  constexpr int T = 1;
  constexpr int CG = kPairTableBytes / 256;  // groups per table chunk
  constexpr int CB = CG / 16;
  constexpr int NR = V == 1 ? 4 : 2;
  static_assert(CB % NR == 0, "a chunk holds whole rounds of the code-block ring");
  __shared__ __align__(16) uint8_t table[kPairTableBytes];
  __shared__ __align__(16) uint8_t stage[kPairStageBytes];
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int G16 = a.G16;
  const int n_chunks = (G16 + CB - 1) / CB;
  const size_t tk = (size_t)T * (2 * (size_t)a.k + 2 * (size_t)a.cap) * 4 + 28 * T + 64;
  float* scl = reinterpret_cast<float*>(smem + (size_t)T * (2 * (size_t)a.k + 2 * (size_t)a.cap) * 4 + 20 * T);
  float* bias = scl + T;
  int* l_pre = reinterpret_cast<int*>(smem + ((tk + 15) & ~size_t(15)));  // [P + 1]
  long* l_pos = reinterpret_cast<long*>(l_pre + ((a.P + 2) & ~1));        // [P] first position of entry j
  float* l_off = reinterpret_cast<float*>(l_pos + a.P);                     // [P]
  const int n_items = __builtin_amdgcn_readfirstlane(a.item_off[a.nq]);
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    // query of the item: last q with item_off[q] <= it (uniform)
    int lo = 0, hi = (int)a.nq;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.item_off[mid] <= it) lo = mid; else hi = mid;
    }
    const int q = __builtin_amdgcn_readfirstlane(lo);
    const int np = __builtin_amdgcn_readfirstlane(a.probe_n[q]);
    const int* pre_g = a.probe_pre + (long)q * (a.P + 1);
    const int total = __builtin_amdgcn_readfirstlane(pre_g[np]);
    const int len = __builtin_amdgcn_readfirstlane(a.item_len[q]);
    const int i0 = (it - a.item_off[q]) * len;
    const int i1 = min(total, i0 + len);
    __syncthreads();  // the previous item's lists are written out
    TopkShared sh;
    init_topk(sh, smem, T, a.k, a.cap);
    for (int j = tid; j <= np; j += blockDim.x) l_pre[j] = pre_g[j];
    for (int j = tid; j < np; j += blockDim.x) {
      l_pos[j] = a.boff[a.probe_b[(long)q * a.P + j]];
      l_off[j] = a.probe_off[(long)q * a.P + j];
    }
    if (tid == 0) {
      scl[0] = a.lut_scale[q];
      bias[0] = a.bias_total[q];
    }
    const uint8_t* q_src = a.src + (long)q * ((long)n_chunks * kPairStageBytes);
    reinterpret_cast<uint64_t*>(stage)[tid] = reinterpret_cast<const uint64_t*>(q_src)[tid];
    __syncthreads();
    bool built = false;
    const bool wave_lists = a.k <= 64;  // a top-k list per wavefront, merged once per item
    WaveList wl;
    const int step = kScanThreads * V;
    for (int base = i0; base < i1; base += step) {
      long vv[V];
      bool valid[V];
      float voff[V], avp[V];
      const uint4* cp[V];
      float fa[V][T];
      uint32_t acc[V][1], acc2[V][1];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const int i = base + tid + j * kScanThreads;
        valid[j] = i < i1;
        // list entry of local vector i: last e with l_pre[e] <= i
        int e0 = 0, e1 = np;
        while (e1 - e0 > 1) {
          const int mid = (e0 + e1) >> 1;
          if (l_pre[mid] <= i) e0 = mid; else e1 = mid;
        }
        vv[j] = valid[j] ? l_pos[e0] + (i - l_pre[e0]) : 0;
        voff[j] = valid[j] ? l_off[e0] : 0.f;
        avp[j] = a.alpha[vv[j]];  // loaded now, used at the admission
        cp[j] = reinterpret_cast<const uint4*>(a.tiled) + (vv[j] >> 6) * (long)G16 * 64 + (vv[j] & 63);
        fa[j][0] = bias[0];
        acc[j][0] = 0u;
        acc2[j][0] = 0u;
      }
      uint4 ring[NR][V];
#pragma unroll
      for (int u = 0; u < NR; ++u)
#pragma unroll
        for (int j = 0; j < V; ++j) ring[u][j] = cp[j][(long)min(u, G16 - 1) * 64];
      for (int c = 0; c < n_chunks; ++c) {
        const uint8_t* dma = nullptr;
        if (n_chunks > 1 || !built) {
          __syncthreads();
          build_pair_chunk<T>(stage, table);
          __syncthreads();
          if (n_chunks > 1 && (c + 1 < n_chunks || base + step < i1))
            dma = q_src + (long)(c + 1 < n_chunks ? c + 1 : 0) * kPairStageBytes;
          built = true;
        }
        pair_chunk<T, V, NR>(table, stage, dma, scl, cp, c, G16, ring, acc, acc2, fa);
      }
      if (wave_lists) {
        // the score as admit_scores forms it: fa * alpha, then + the offset
#pragma unroll
        for (int j = 0; j < V; ++j)
          wave_list_admit(wl, valid[j] ? fa[j][0] * avp[j] + voff[j] : -INFINITY, (uint32_t)vv[j], valid[j], a.k,
                          tid & 63);
      } else {
        admit_scores<T, V>(sh, fa, vv, valid, a.alpha, 0, 1, a.k, a.cap, &voff, &avp);
      }
    }
    if (wave_lists) {
      wave_lists_write(wl, table, it, 1, 0, a.k, a.part_s, a.part_i, a.part_cnt);
      continue;
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(sh.cnt[0]) > 0) topk_merge(sh, 0, a.k, a.cap);
    const int m = sh.m[0];
    for (int j = tid; j < m; j += blockDim.x) {
      a.part_s[(long)it * a.k + j] = sh.top_s[j];
      a.part_i[(long)it * a.k + j] = sh.top_i[j];
    }
    if (tid == 0) a.part_cnt[it] = m;
  }
}

__device__ void topk_merge_all(const TopkShared& sh, int T, int k, int cap) {
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w < T) {
    const int t = w;
    const int c = __builtin_amdgcn_readfirstlane(min(sh.cnt[t], cap));
    const int m = __builtin_amdgcn_readfirstlane(sh.m[t]);
    if (c > 0 && m + c <= 64) {
      // the usual case (a tile's list starts empty): one bitonic sort across the lanes
      float* ts = sh.top_s + t * k;
      uint32_t* ti = sh.top_i + t * k;
      float s = -INFINITY;
      uint32_t id = 0xffffffffu;
      if (lane < m) {
        s = ts[lane];
        id = ti[lane];
      } else if (lane < m + c) {
        s = sh.cand_s[t * cap + lane - m];
        id = sh.cand_i[t * cap + lane - m];
      }
      wave_sort_desc(s, id, lane);
      const int m2 = min(k, m + c);
      if (lane < m2) {
        ts[lane] = s;
        ti[lane] = id;
      }
      if (lane == m2 - 1 && m2 == k) {
        sh.thr_s[t] = s;
        sh.thr_i[t] = id;
      }
      if (lane == 0) {
        sh.m[t] = m2;
        sh.cnt[t] = 0;
      }
    } else if (c > 0) {
      float* ts = sh.top_s + t * k;
      uint32_t* ti = sh.top_i + t * k;
      float es[4];
      uint32_t ei[4];
      bool have[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = lane + 64 * r;
        have[r] = e < m + c;
        const int ec = min(max(e - m, 0), cap - 1);
        es[r] = e < m ? ts[min(e, k - 1)] : sh.cand_s[t * cap + ec];
        ei[r] = e < m ? ti[min(e, k - 1)] : sh.cand_i[t * cap + ec];
      }
      const int m2 = min(k, m + c);
      for (int r = 0; r < m2; ++r) {
        bool h = false;
        float bs = 0.f;
        uint32_t bi = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (have[j] && (!h || better(es[j], ei[j], bs, bi))) {
            h = true;
            bs = es[j];
            bi = ei[j];
          }
        bool wh = h;
        float ws = bs;
        uint32_t wi = bi;
        wave_best(wh, ws, wi);
        // the ids of a query's entries are distinct: exactly one lane holds the winner
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (have[j] && es[j] == ws && ei[j] == wi) have[j] = false;
        if (lane == 0) {
          ts[r] = ws;
          ti[r] = wi;
        }
      }
      if (lane == 0) {
        sh.m[t] = m2;
        sh.cnt[t] = 0;
        if (m2 == k) {
          sh.thr_s[t] = ts[k - 1];
          sh.thr_i[t] = ti[k - 1];
        }
      }
    }
  }
  __syncthreads();
}

// One wavefront per query: merge its items' lists (at most 512) into k_out
// entries, positions mapped through id_map; counts[q] (when set) = entries
// found, the rest of the row (0, 0).
__global__ __launch_bounds__(64) void merge_items_kernel(const float* __restrict__ part_s,
                                                         const uint32_t* __restrict__ part_i,
                                                         const int* __restrict__ part_cnt,
                                                         const int* __restrict__ item_off, long nq, int k,
                                                         int k_out, const uint64_t* __restrict__ id_map,
                                                         float* __restrict__ out_s, uint64_t* __restrict__ out_i,
                                                         uint64_t* __restrict__ counts) {
  for (long q = blockIdx.x; q < nq; q += gridDim.x) {
    const int f = item_off[q], R = item_off[q + 1] - f;
    merge_query(part_s + (long)f * k, part_i + (long)f * k, part_cnt + f, R, k, k_out, id_map, out_s + q * k_out,
                out_i + q * k_out, threadIdx.x);
    if (counts != nullptr) {
      int c = 0;
      for (int r = threadIdx.x; r < R; r += 64) c += part_cnt[f + r];
      for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
      if (threadIdx.x == 0) counts[q] = (uint64_t)min(c, k_out);
    }
  }
}

// ---------------------------------------------------------------------------
// Bucket-major scan of a large query batch (the multi-bucket plumbing of
// score_prepared_into, src/search/search.cpp:1156-1267, turned inside out).
// The pairs (q, j), j < probe_n[q], are grouped by their bucket
// probe_b[q][j] and cut into tiles of up to kTileQueries queries of one
// bucket. A tile reads its bucket's codes once for all its queries: the
// tile's LUTs sit in LDS interleaved by query (entry (g, e, t) at byte
// g * kTileGroupStride + e * 8 + t), so one 64-bit read fetches a nibble's
// entry for all eight queries (two reads per code byte, no pair table to
// build per tile). The window sums, the float chain and the score
// fa * alpha + offset are those of the query-major scan, and each pair keeps
// its own exact top-k list in slot q * P + j.
// A per-query bound in HBM (the float_order image of the k-th score of any
// full list written so far) seeds each tile's admission bound: a vector
// scoring below it is worse than the k vectors of that list.
// ---------------------------------------------------------------------------
constexpr int kTileChunkBlocks = 16;                      // 256 groups = one flush window per table chunk
constexpr int kTileGroupStride = 264;                     // 256 bytes of entries + 8: spreads the fill over the banks
constexpr int kTileTableBytes = kTileChunkBlocks * 16 * kTileGroupStride;
#ifndef VC_TILE_THREADS
#define VC_TILE_THREADS 1024
#endif
// threads per tile: a bucket of up to this many vectors in one pass (two
// workgroups per CU)
constexpr int kTileThreads = VC_TILE_THREADS;
constexpr int kTileFillTasks = kTileChunkBlocks * 16 * 16 / kTileThreads;  // per thread and chunk
#ifndef VC_TILE_FILL_BATCH
#define VC_TILE_FILL_BATCH 2
#endif
#ifndef VC_TILE_RING
#define VC_TILE_RING 2
#endif
constexpr int kTileFillBatch = VC_TILE_FILL_BATCH;  // fill tasks whose loads are in flight together

// byte E of w shifted right by SH, one SDWA instruction
template <int E, int SH>
__device__ __forceinline__ uint32_t byte_shr(uint32_t w) {
  static_assert(E >= 0 && E < 4 && SH > 0 && SH < 8, "byte and shift");
  uint32_t r;
  if constexpr (E == 0) {
    asm("v_lshrrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
        : "=v"(r) : "v"(w), "i"(SH));
  } else if constexpr (E == 1) {
    asm("v_lshrrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_1"
        : "=v"(r) : "v"(w), "i"(SH));
  } else if constexpr (E == 2) {
    asm("v_lshrrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_2"
        : "=v"(r) : "v"(w), "i"(SH));
  } else {
    asm("v_lshrrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_3"
        : "=v"(r) : "v"(w), "i"(SH));
  }
  return r;
}

// bcnt[b] = 0 (b < L), gbound[q] = the image of -inf (q < nq)
__global__ void tile_init_kernel(int L, int* __restrict__ bcnt, long nq, int* __restrict__ gbound) {
  const long n = max((long)L, nq);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    if (i < L) bcnt[i] = 0;
    if (i < nq) gbound[i] = float_order(-INFINITY);
  }
}

// pairs (q, j), j_lo <= j < min(j_hi, probe_n[q]), per bucket; prank[q * P + j]
// = the pair's rank among its bucket's
__global__ void tile_count_kernel(long nq, int P, int j_lo, int j_hi, const int* __restrict__ probe_b,
                                  const int* __restrict__ probe_n, int* __restrict__ bcnt, int* __restrict__ prank) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nq * P; i += (long)gridDim.x * blockDim.x) {
    const long q = i / P;
    const int j = (int)(i - q * P);
    if (j >= j_lo && j < j_hi && j < probe_n[q]) prank[i] = atomicAdd(&bcnt[probe_b[i]], 1);
  }
}

// Exclusive prefix sums of the pair counts (bstart) and tile counts (tstart)
// of the L buckets; one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void tile_scan_kernel(int L, const int* __restrict__ bcnt, int* __restrict__ bstart,
                                                         int* __restrict__ tstart, int* __restrict__ n_tiles) {
  __shared__ int wsum[2][16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int run_p = 0, run_t = 0;
  for (int b0 = 0; b0 < L; b0 += 1024) {
    const int b = b0 + tid;
    const int c = b < L ? bcnt[b] : 0;
    int vp = c, vt = (c + kTileQueries - 1) / kTileQueries;
    for (int o = 1; o < 64; o <<= 1) {
      const int up = __shfl_up(vp, o), ut = __shfl_up(vt, o);
      if (lane >= o) { vp += up; vt += ut; }
    }
    if (lane == 63) { wsum[0][w] = vp; wsum[1][w] = vt; }
    __syncthreads();
    int bp = run_p, bt = run_t;
    for (int i = 0; i < w; ++i) { bp += wsum[0][i]; bt += wsum[1][i]; }
    if (b < L) {
      bstart[b] = bp + vp - c;
      tstart[b] = bt + vt - (c + kTileQueries - 1) / kTileQueries;
    }
    for (int i = 0; i < 16; ++i) { run_p += wsum[0][i]; run_t += wsum[1][i]; }
    __syncthreads();
  }
  if (tid == 0) {
    bstart[L] = run_p;
    tstart[L] = run_t;
    *n_tiles = run_t;
  }
}

// pair lists grouped by bucket, and each bucket's tiles
__global__ void tile_place_kernel(long nq, int P, int j_lo, int j_hi, const int* __restrict__ probe_b,
                                  const int* __restrict__ probe_n, const int* __restrict__ prank,
                                  const int* __restrict__ bstart, int* __restrict__ pair_q, int* __restrict__ pair_j) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nq * P; i += (long)gridDim.x * blockDim.x) {
    const long q = i / P;
    const int j = (int)(i - q * P);
    if (j >= j_lo && j < j_hi && j < probe_n[q]) {
      const int at = bstart[probe_b[i]] + prank[i];
      pair_q[at] = (int)q;
      pair_j[at] = j;
    }
  }
}

__global__ void tile_list_kernel(int L, const int* __restrict__ bcnt, const int* __restrict__ bstart,
                                 const int* __restrict__ tstart, int* __restrict__ tile_b, int* __restrict__ tile_p0,
                                 int* __restrict__ tile_n) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < L; b += gridDim.x * blockDim.x) {
    const int c = bcnt[b], t0 = tstart[b], p0 = bstart[b];
    for (int i = 0; i * kTileQueries < c; ++i) {
      tile_b[t0 + i] = b;
      tile_p0[t0 + i] = p0 + i * kTileQueries;
      tile_n[t0 + i] = min(kTileQueries, c - i * kTileQueries);
    }
  }
}

// Fill the table with groups [16 * 16 c, 16 * 16 (c + 1)) of the tile's
// eight LUT rows: 4096 tasks of (group, query quad, entry quad), eight per
// thread, each four 32-bit loads (one per query) and a 4 x 4 byte transpose. Groups past G
// (the padding of the last code block) get zero entries.
__device__ __forceinline__ void tile_fill(const BucketTileParams& a, const int* lq, int c, uint8_t* table) {
  const int tid = threadIdx.x;
  const int g0 = c * kTileChunkBlocks * 16;
#pragma unroll
  for (int r0 = 0; r0 < kTileFillTasks; r0 += kTileFillBatch) {
  uint32_t w[kTileFillBatch][4];
#pragma unroll
  for (int r = 0; r < kTileFillBatch; ++r) {
    const int task = tid + (r0 + r) * kTileThreads;
    const int e4 = task & 7, t4 = (task >> 3) & 1, g = g0 + (task >> 4);
    const int gc = min(g, a.G - 1);  // clamped: rows past G are loaded but not stored
#pragma unroll
    for (int i = 0; i < 4; ++i)
      w[r][i] = *reinterpret_cast<const uint32_t*>(a.luts + (long)lq[4 * t4 + i] * a.G * 32 + (long)gc * 32 + e4 * 4);
  }
#pragma unroll
  for (int r = 0; r < kTileFillBatch; ++r) {
    const int task = tid + (r0 + r) * kTileThreads;
    const int e4 = task & 7, t4 = (task >> 3) & 1, gl = task >> 4;
    if (g0 + gl < a.G16 * 16) {
      if (g0 + gl >= a.G) w[r][0] = w[r][1] = w[r][2] = w[r][3] = 0u;
      // w[i] = entries 4 e4 .. 4 e4 + 3 of query 4 t4 + i; out[j] = entry 4 e4 + j of the four queries
      const uint32_t t01l = __builtin_amdgcn_perm(w[r][1], w[r][0], 0x05010400u);
      const uint32_t t01h = __builtin_amdgcn_perm(w[r][1], w[r][0], 0x07030602u);
      const uint32_t t23l = __builtin_amdgcn_perm(w[r][3], w[r][2], 0x05010400u);
      const uint32_t t23h = __builtin_amdgcn_perm(w[r][3], w[r][2], 0x07030602u);
      uint32_t* dst = reinterpret_cast<uint32_t*>(table + gl * kTileGroupStride + e4 * 32 + t4 * 4);
      dst[0] = __builtin_amdgcn_perm(t23l, t01l, 0x05040100u);
      dst[2] = __builtin_amdgcn_perm(t23l, t01l, 0x07060302u);
      dst[4] = __builtin_amdgcn_perm(t23h, t01h, 0x05040100u);
      dst[6] = __builtin_amdgcn_perm(t23h, t01h, 0x07060302u);
    }
  }
  }
}

__device__ __forceinline__ uint2 add64(uint2 x, uint2 y) {
  uint64_t r;
  const uint64_t a = ((uint64_t)x.y << 32) | x.x, b = ((uint64_t)y.y << 32) | y.x;
  asm("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(r) : "v"(a), "v"(b));
  return make_uint2((uint32_t)r, (uint32_t)(r >> 32));
}

__global__ __launch_bounds__(kTileThreads, kTileThreads == 1024 ? VC_PAIR_MIN_WAVES : 4) void bucket_tile_kernel(
    BucketTileParams a) {
  constexpr int T = kTileQueries;
  constexpr int NR = VC_TILE_RING;  // code blocks in flight per thread
  static_assert(T == 8, "eight queries per 64-bit entry");
  __shared__ __align__(16) uint8_t table[kTileTableBytes];
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int G16 = a.G16;
  const int n_chunks = (G16 + kTileChunkBlocks - 1) / kTileChunkBlocks;
  const size_t tk = ((size_t)T * (2 * (size_t)a.k + 2 * (size_t)a.cap) * 4 + 20 * T + 15) & ~size_t(15);
  int* lq = reinterpret_cast<int*>(smem + tk);  // [T] query of each tile slot
  int* lslot = lq + T;                          // [T] its pair's list slot
  float* lscl = reinterpret_cast<float*>(lslot + T);
  float* lbias = lscl + T;
  float* loff = lbias + T;
  const int n_tiles = __builtin_amdgcn_readfirstlane(*a.n_tiles);
  // Tile descriptors are loaded two tiles ahead and the pairs one tile ahead,
  // so no load round trip stands between two tiles. (Slots past a tile's
  // count repeat its last pair.)
  const int G = gridDim.x;
  auto desc = [&](int it, int& db, int& dp0, int& dn) {
    db = 0; dp0 = 0; dn = 1;
    if (it < n_tiles) { db = a.tile_b[it]; dp0 = a.tile_p0[it]; dn = a.tile_n[it]; }
  };
  auto pair = [&](int it, int dp0, int dn, int& pq, int& pj) {
    pq = 0; pj = 0;
    if (it < n_tiles && tid < T) {
      const int pi = dp0 + min(tid, dn - 1);
      pq = a.pair_q[pi];
      pj = a.pair_j[pi];
    }
  };
  int cur_b, cur_p0, cur_n, cur_q, cur_j, nx_b, nx_p0, nx_n;
  // Workgroups go round-robin to the eight XCDs; each round gives XCD x the
  // consecutive tiles [x G / 8, (x + 1) G / 8), so the tiles of one bucket
  // (consecutive) share its L2.
  const int slot0 = G % 8 == 0 ? (int)(blockIdx.x % 8) * (G / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  desc(slot0, cur_b, cur_p0, cur_n);
  pair(slot0, cur_p0, cur_n, cur_q, cur_j);
  desc(slot0 + G, nx_b, nx_p0, nx_n);
  for (int it = slot0; it < n_tiles; it += G) {
    const int b = __builtin_amdgcn_readfirstlane(cur_b);
    const int nt = __builtin_amdgcn_readfirstlane(cur_n);
    const int q_t = cur_q, j_t = cur_j;
    const int nv = __builtin_amdgcn_readfirstlane(a.bsize[b]);
    const long pos0 = a.boff[b];
    int nn_b, nn_p0, nn_n, nx_q, nx_j;
    pair(it + G, nx_p0, nx_n, nx_q, nx_j);
    desc(it + 2 * G, nn_b, nn_p0, nn_n);
    __syncthreads();  // the previous tile's lists are written out
    TopkShared sh;
    init_topk(sh, smem, T, a.k, a.cap);
    if (tid < T) {
      // slots past nt repeat the last query: scanned, never admitted
      const int q = q_t, j = j_t;
      lq[tid] = q;
      lslot[tid] = q * a.P + j;
      lscl[tid] = a.lut_scale[q];
      lbias[tid] = a.bias_total[q];
      loff[tid] = a.probe_off[(long)q * a.P + j];
      if (tid < nt) sh.bound[tid] = a.gbound[q];  // after init_topk's reset by this same thread
    }
    __syncthreads();
    for (int vp0 = 0; vp0 < nv; vp0 += kTileThreads) {
      const int i = vp0 + tid;
      const bool valid = i < nv;
      const bool live = vp0 + (tid & ~63) < nv;  // uniform: the wavefront has a vector
      const long v = pos0 + (valid ? i : 0);
      const uint4* cp = reinterpret_cast<const uint4*>(a.tiled) + (v >> 6) * (long)G16 * 64 + (v & 63);
      const float av = a.alpha[v];
      float fa[1][T];
#pragma unroll
      for (int t = 0; t < T; ++t) fa[0][t] = lbias[t];
      uint32_t accA[2] = {0u, 0u}, accB[2] = {0u, 0u};
      uint4 ring[NR];
#pragma unroll
      for (int u = 0; u < NR; ++u) ring[u] = cp[(long)min(u, G16 - 1) * 64];
      for (int c = 0; c < n_chunks; ++c) {
        __syncthreads();  // the previous chunk's lookups are done
        tile_fill(a, lq, c, table);
        __syncthreads();
        if (!live) continue;
#pragma unroll
        for (int u = 0; u < kTileChunkBlocks; ++u) {
          const int bb = c * kTileChunkBlocks + u;
          if (bb < G16) {  // uniform
            const uint4 cw = ring[u % NR];
            const uint32_t wl[4] = {cw.x & 0x0f0f0f0fu, cw.y & 0x0f0f0f0fu, cw.z & 0x0f0f0f0fu, cw.w & 0x0f0f0f0fu};
            const uint32_t wh[4] = {cw.x & 0xf0f0f0f0u, cw.y & 0xf0f0f0f0u, cw.z & 0xf0f0f0f0u, cw.w & 0xf0f0f0f0u};
            const uint8_t* tb = table + u * 16 * kTileGroupStride;
            // four bytes at a time: their eight reads are issued before the
            // first sum waits on one (the address and add instructions are
            // written out, and the compiler keeps their order)
            static_for<4>([&](auto wc) {
              constexpr int w = decltype(wc)::value;
              uint2 rh[4], rl[4];
              static_for<4>([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                const uint8_t* gp = tb + (4 * w + e) * kTileGroupStride;
                // entries 0-15: the high nibble (first code); 16-31: the low one
                rh[e] = *reinterpret_cast<const uint2*>(gp + byte_shr<e, 1>(wh[w]));
                rl[e] = *reinterpret_cast<const uint2*>(gp + 128 + byte_scaled<e, 3>(wl[w]));
              });
              static_for<4>([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                // bytes of each entry are <= 127: the byte sums have no carries
                const uint2 w2 = add64(rh[e], rl[e]);
                // accA: sum over q of 2^(8q) S_q (mod 2^64); accB: the odd
                // queries' bytes alone as u16 lanes (two v_perm, 0x0c = zero)
                acc64_add<true>(accA, w2);
                acc64_add<true>(accB, make_uint2(__builtin_amdgcn_perm(0u, w2.x, 0x0c030c01u),
                                                 __builtin_amdgcn_perm(0u, w2.y, 0x0c030c01u)));
              });
            });
          }
          ring[u % NR] = cp[(long)min(bb + NR, G16 - 1) * 64];  // clamped, in bounds
        }
        // window flush (a chunk is one 256-group window): every window sum is
        // < 2^16, so accA - 2^8 accB = sum over i of 2^(16i) S_2i exactly
        const uint64_t a64 = ((uint64_t)accA[1] << 32) | accA[0];
        const uint64_t b64 = ((uint64_t)accB[1] << 32) | accB[0];
        const uint64_t e64 = a64 - (b64 << 8);
        const uint32_t evw[2] = {(uint32_t)e64, (uint32_t)(e64 >> 32)};
        const uint32_t odw[2] = {accB[0], accB[1]};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const u16x2 od = __builtin_bit_cast(u16x2, odw[h]);   // queries 4h + 1, 4h + 3
          const u16x2 ev = __builtin_bit_cast(u16x2, evw[h]);   // queries 4h, 4h + 2
          fa[0][4 * h + 0] = __fmaf_rn(lscl[4 * h + 0], (float)ev.x, fa[0][4 * h + 0]);
          fa[0][4 * h + 1] = __fmaf_rn(lscl[4 * h + 1], (float)od.x, fa[0][4 * h + 1]);
          fa[0][4 * h + 2] = __fmaf_rn(lscl[4 * h + 2], (float)ev.y, fa[0][4 * h + 2]);
          fa[0][4 * h + 3] = __fmaf_rn(lscl[4 * h + 3], (float)od.y, fa[0][4 * h + 3]);
        }
        accA[0] = accA[1] = accB[0] = accB[1] = 0u;
      }
      const long vv[1] = {v};
      const bool vl[1] = {valid};
      const float avp[1] = {av};
      admit_scores<T, 1>(sh, fa, vv, vl, a.alpha, 0, nt, a.k, a.cap, nullptr, &avp, loff, vp0 == 0);
    }
    __syncthreads();
    topk_merge_all(sh, T, a.k, a.cap);
    {
      // wavefront t writes query t's list (k <= 64: one entry per lane)
      const int t = tid >> 6, j = tid & 63;
      if (t < nt) {
        const long slot = lslot[t];
        const int m = sh.m[t];
        if (j < m) {
          a.part_s[slot * a.k + j] = sh.top_s[t * a.k + j];
          a.part_i[slot * a.k + j] = sh.top_i[t * a.k + j];
        }
        if (j == 0) {
          a.part_cnt[slot] = m;
          if (m == a.k) atomicMax(&a.gbound[lq[t]], float_order(sh.top_s[t * a.k + a.k - 1]));
        }
      }
    }
    cur_b = nx_b; cur_n = nx_n; cur_q = nx_q; cur_j = nx_j;
    nx_b = nn_b; nx_p0 = nn_p0; nx_n = nn_n;
  }
}

// One wavefront per query: merge its probe_n[q] slot lists (slots q * P ..).
__global__ __launch_bounds__(64) void merge_slots_kernel(const float* __restrict__ part_s,
                                                         const uint32_t* __restrict__ part_i,
                                                         const int* __restrict__ part_cnt,
                                                         const int* __restrict__ probe_n, int P, long nq, int k,
                                                         int k_out, const uint64_t* __restrict__ id_map,
                                                         float* __restrict__ out_s, uint64_t* __restrict__ out_i,
                                                         uint64_t* __restrict__ counts) {
  for (long q = blockIdx.x; q < nq; q += gridDim.x) {
    const long f = q * P;
    const int R = probe_n[q];
    merge_query(part_s + f * k, part_i + f * k, part_cnt + f, R, k, k_out, id_map, out_s + q * k_out,
                out_i + q * k_out, threadIdx.x);
    if (counts != nullptr) {
      int c = 0;
      for (int r = threadIdx.x; r < R; r += 64) c += part_cnt[f + r];
      for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
      if (threadIdx.x == 0) counts[q] = (uint64_t)min(c, k_out);
    }
  }
}

// ---------------------------------------------------------------------------
// score_prepared_into (src/search/search.cpp:1156-1267) for one bucket, the
// reference's AVX2 PERM0 path (src/search/search_avx2.cpp:72-230):
//   1. bucket_scores_kernel: p[i][v] = fa * alpha[v], fa = bias + bias_corr
//      followed by fmaf(scale, window sum, fa) per 256-group window;
//   2. heap_replay_kernel: per listed query, the vectors in order by 32-vector
//      sub-tiles into the query's persistent heap (topk_heap.hpp:19-157).
// ---------------------------------------------------------------------------
constexpr int kWindowGroups = 256;  // kFlushEvery

// grid (ceil(n / 256), queries), 256 threads; one window of LUT rows in LDS.
__global__ __launch_bounds__(256) void bucket_scores_kernel(const uint8_t* __restrict__ tiled,
                                                            const float* __restrict__ alpha, long n, int G,
                                                            int G16, const uint8_t* __restrict__ luts,
                                                            const float* __restrict__ lut_scale,
                                                            const float* __restrict__ bias_total,
                                                            float* __restrict__ p_out) {
  __shared__ __align__(16) uint8_t lut[kWindowGroups * 32];
  const int qi = blockIdx.y;
  const long v = (long)blockIdx.x * 256 + threadIdx.x;
  const long vs = v < n ? v : n - 1;  // lanes past n read a valid row and store nothing
  const uint8_t* ql = luts + (long)qi * G * 32;
  const float scale = lut_scale[qi];
  float fa = bias_total[qi];
  for (int g0 = 0; g0 < G; g0 += kWindowGroups) {
    const int g1 = min(G, g0 + kWindowGroups);
    __syncthreads();
    for (int i = threadIdx.x; i < (g1 - g0) * 8; i += 256)
      reinterpret_cast<uint32_t*>(lut)[i] = reinterpret_cast<const uint32_t*>(ql + (long)g0 * 32)[i];
    __syncthreads();
    uint32_t acc = 0;
    for (int b = g0 / 16; b * 16 < g1; ++b) {
      const uint4 cw = *reinterpret_cast<const uint4*>(tiled + ((vs >> 6) * G16 + b) * 1024 + (vs & 63) * 16);
      const uint32_t words[4] = {cw.x, cw.y, cw.z, cw.w};
      const int ge = min(16, g1 - b * 16);  // padded groups of the last block are skipped
      for (int e = 0; e < ge; ++e) {
        const uint32_t byte = __builtin_amdgcn_ubfe(words[e >> 2], 8 * (e & 3), 8);
        const int gl = b * 16 + e - g0;
        acc += lut[gl * 32 + (byte >> 4)] + lut[gl * 32 + 16 + (byte & 15u)];
      }
    }
    fa = __fmaf_rn(scale, (float)acc, fa);
  }
  if (v < n) p_out[(long)qi * n + v] = fa * alpha[v];
}

// topk_worse (topk_heap.hpp:19-22): lower score is worse; on equal scores the
// larger id
__device__ __forceinline__ bool heap_worse(float sa, uint64_t ia, float sb, uint64_t ib) {
  return sa != sb ? sa < sb : ia > ib;
}

// rescan_min (topk_heap.hpp:26-36) by one wavefront: the first index of the
// worst entry
__device__ __forceinline__ void heap_rescan(const float* hs, const uint64_t* hi, int k, int lane, float& hmin,
                                            long& hmi) {
  int mi = -1;
  float ms = 0.f;
  uint64_t mid = 0;
  for (int j = lane; j < k; j += 64) {
    const float sj = hs[j];
    const uint64_t ij = hi[j];
    if (mi < 0 || heap_worse(sj, ij, ms, mid)) {
      mi = j;
      ms = sj;
      mid = ij;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const int omi = __shfl_xor(mi, o);
    const float os = __shfl_xor(ms, o);
    const uint64_t oid = ((uint64_t)(uint32_t)__shfl_xor((int)(uint32_t)(mid >> 32), o) << 32) |
                         (uint32_t)__shfl_xor((int)(uint32_t)mid, o);
    if (omi >= 0 && (mi < 0 || heap_worse(os, oid, ms, mid) || (!heap_worse(ms, mid, os, oid) && omi < mi))) {
      mi = omi;
      ms = os;
      mid = oid;
    }
  }
  hmin = ms;
  hmi = mi;
}

// topk_sift_down (topk_heap.hpp:38-58), one lane
__device__ void heap_sift_down(float* hs, uint64_t* hi, int k, int i) {
  for (;;) {
    const int l = 2 * i + 1, r = 2 * i + 2;
    int w = i;
    if (l < k && heap_worse(hs[l], hi[l], hs[w], hi[w])) w = l;
    if (r < k && heap_worse(hs[r], hi[r], hs[w], hi[w])) w = r;
    if (w == i) return;
    const float ts = hs[i];
    hs[i] = hs[w];
    hs[w] = ts;
    const uint64_t ti = hi[i];
    hi[i] = hi[w];
    hi[w] = ti;
    i = w;
  }
}

// One wavefront per listed query; its heap (k entries) and state in LDS.
// heap_sz / heap_min / heap_mi are computed identically by every lane.
__global__ __launch_bounds__(64) void heap_replay_kernel(const float* __restrict__ p, long n, int k,
                                                         const uint64_t* __restrict__ id_map,
                                                         const float* __restrict__ offsets,
                                                         float* __restrict__ hs_g, uint64_t* __restrict__ hi_g,
                                                         uint64_t* __restrict__ hsz_g, float* __restrict__ hmin_g,
                                                         uint64_t* __restrict__ hmi_g) {
  extern __shared__ __align__(16) unsigned char smem[];
  float* hs = reinterpret_cast<float*>(smem);
  uint64_t* hi = reinterpret_cast<uint64_t*>(smem + (((size_t)k * 4 + 15) & ~size_t(15)));
  const int qi = blockIdx.x, lane = threadIdx.x;
  for (int j = lane; j < k; j += 64) {
    hs[j] = hs_g[(long)qi * k + j];
    hi[j] = hi_g[(long)qi * k + j];
  }
  long hsz = (long)hsz_g[qi];
  float hmin = hmin_g[qi];
  long hmi = (long)hmi_g[qi];
  const float off = offsets[qi];
  const float* pq = p + (long)qi * n;
  __syncthreads();
  for (long cb = 0; cb < n; cb += 32) {
    const int cnt = (int)min(32L, n - cb);
    // a whole sub-tile starting with a full heap adds the offset only when it
    // is non-zero (search_avx2.cpp:163-180); every other sub-tile adds it
    const bool fast = hsz >= k && cnt == 32;
    float sv = 0.f;
    if (lane < cnt) {
      const float pv = pq[cb + lane];
      sv = fast ? (off != 0.f ? pv + off : pv) : pv + off;
    }
    uint64_t pend = __ballot(lane < cnt && sv > hmin);
    while (pend != 0) {
      const int l = (int)__builtin_ctzll(pend);
      const float s = __shfl(sv, l);
      const uint64_t id = id_map != nullptr ? id_map[cb + l] : (uint64_t)(cb + l);
      // heap_push_or_replace (topk_heap.hpp:113-157); s > hmin holds here
      if (k < 128) {
        const long at = hsz < k ? hsz : hmi;
        if (lane == 0) {
          hs[at] = s;
          hi[at] = id;
        }
        if (hsz < k) ++hsz;
        __syncthreads();
        if (hsz == k) heap_rescan(hs, hi, k, lane, hmin, hmi);
      } else if (hsz < k) {
        if (lane == 0) {
          hs[hsz] = s;
          hi[hsz] = id;
          if (hsz + 1 == k)
            for (int i = k / 2; i-- > 0;) heap_sift_down(hs, hi, k, i);
        }
        ++hsz;
        __syncthreads();
        if (hsz == k) {
          hmin = hs[0];
          hmi = 0;
        }
      } else {
        if (lane == 0) {
          hs[0] = s;
          hi[0] = id;
          heap_sift_down(hs, hi, k, 0);
        }
        __syncthreads();
        hmin = hs[0];
        hmi = 0;
      }
      pend &= pend - 1;
      pend &= __ballot(sv > hmin);  // the threshold only rises
    }
  }
  __syncthreads();
  for (int j = lane; j < k; j += 64) {
    hs_g[(long)qi * k + j] = hs[j];
    hi_g[(long)qi * k + j] = hi[j];
  }
  if (lane == 0) {
    hsz_g[qi] = (uint64_t)hsz;
    hmin_g[qi] = hmin;
    hmi_g[qi] = (uint64_t)hmi;
  }
}

// heaps_to_search_results (src/search/search.cpp:1269-1291): per query the
// heap_sz entries by (score desc, id asc) -- each entry's rank counted against
// all others, equal pairs ordered by position -- then (0, 0) to k.
__global__ __launch_bounds__(256) void heaps_to_results_kernel(long nq, int k, const float* __restrict__ hs_g,
                                                               const uint64_t* __restrict__ hi_g,
                                                               const uint64_t* __restrict__ hsz_g,
                                                               float* __restrict__ out_s,
                                                               uint64_t* __restrict__ out_i) {
  extern __shared__ __align__(16) unsigned char smem[];
  float* hs = reinterpret_cast<float*>(smem);
  uint64_t* hi = reinterpret_cast<uint64_t*>(smem + (((size_t)k * 4 + 15) & ~size_t(15)));
  for (long q = blockIdx.x; q < nq; q += gridDim.x) {
    const int m = (int)min((uint64_t)k, hsz_g[q]);
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      hs[j] = hs_g[q * k + j];
      hi[j] = hi_g[q * k + j];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      const float sj = hs[j];
      const uint64_t ij = hi[j];
      int rank = 0;
      for (int i = 0; i < m; ++i) {
        const float si = hs[i];
        const uint64_t ii = hi[i];
        rank += (si > sj || (si == sj && (ii < ij || (ii == ij && i < j)))) ? 1 : 0;
      }
      out_s[q * k + rank] = sj;
      out_i[q * k + rank] = ij;
    }
    for (int j = m + threadIdx.x; j < k; j += blockDim.x) {
      out_s[q * k + j] = 0.f;
      out_i[q * k + j] = 0u;
    }
  }
}

// float_rescore_topk (src/search/float_rescore.cpp:80-164): one workgroup
// per query row. Each candidate is scored by 16 lanes in the reference's AVX2
// order (:36-57): lanes 0-7 are acc0 and 8-15 acc1 over 16-float steps
// (fused), then acc0 + acc1, the 8-float fused tail on acc0, and the
// horizontal sum ((v0+v4)+(v1+v5)) + ((v2+v6)+(v3+v7)) (:24-33); -inf for an id
// outside the corpus. The row's first n entries are then ordered by
// (score desc, id asc) by rank, entries past n left as they are.
__global__ __launch_bounds__(256) void float_rescore_kernel(long nq, int k, float* __restrict__ scores,
                                                            uint64_t* __restrict__ ids,
                                                            const uint64_t* __restrict__ counts,
                                                            const float* __restrict__ queries,
                                                            const float* __restrict__ corpus, long n_corpus,
                                                            int dim) {
  extern __shared__ __align__(16) unsigned char smem[];
  float* cs = reinterpret_cast<float*>(smem);
  uint64_t* ci = reinterpret_cast<uint64_t*>(smem + (((size_t)k * 4 + 15) & ~size_t(15)));
  const int tid = threadIdx.x, l = tid & 15;
  for (long q = blockIdx.x; q < nq; q += gridDim.x) {
    const int n = counts != nullptr ? (int)min((uint64_t)k, counts[q]) : k;
    const float* qa = queries + q * dim;
    __syncthreads();
    for (int j = tid >> 4; j < n; j += blockDim.x >> 4) {  // the 16 lanes of a group share j
      const uint64_t id = ids[q * k + j];
      float s = -INFINITY;
      if (id < (uint64_t)n_corpus) {
        const float* b = corpus + (long)id * dim;
        float acc = 0.f;
        int i = 0;
        for (; i + 16 <= dim; i += 16) acc = __fmaf_rn(qa[i + l], b[i + l], acc);
        const float a1 = __shfl_down(acc, 8, 16);
        if (l < 8) acc = acc + a1;
        for (; i + 8 <= dim; i += 8)
          if (l < 8) acc = __fmaf_rn(qa[i + l], b[i + l], acc);
        const float a4 = __shfl_down(acc, 4, 16);
        const float lo = acc + a4;  // lanes 0-3: v_l + v_(l+4)
        const float lo1 = __shfl_down(lo, 1, 16), lo2 = __shfl_down(lo, 2, 16), lo3 = __shfl_down(lo, 3, 16);
        s = (lo + lo1) + (lo2 + lo3);  // lane 0
        for (; i < dim; ++i) s = __fmaf_rn(qa[i], b[i], s);
      }
      if (l == 0) {
        cs[j] = s;
        ci[j] = id;
      }
    }
    __syncthreads();
    for (int j = tid; j < n; j += blockDim.x) {
      const float sj = cs[j];
      const uint64_t ij = ci[j];
      int rank = 0;
      for (int i = 0; i < n; ++i) {
        const float si = cs[i];
        const uint64_t ii = ci[i];
        rank += (si > sj || (si == sj && (ii < ij || (ii == ij && i < j)))) ? 1 : 0;
      }
      scores[q * k + rank] = sj;
      ids[q * k + rank] = ij;
    }
  }
}

struct MergeLayout {
  size_t cs, ci, keep, os, oi, total;
};
__host__ __device__ inline size_t align16(size_t v) { return (v + 15) & ~size_t(15); }
__host__ __device__ inline MergeLayout merge_layout(int m, int kd) {
  MergeLayout l;
  l.cs = 0;
  l.ci = align16(l.cs + (size_t)m * 4);
  l.keep = align16(l.ci + (size_t)m * 8);
  l.os = align16(l.keep + (size_t)m * 4);
  l.oi = align16(l.os + (size_t)kd * 4);
  l.total = align16(l.oi + (size_t)kd * 8);
  return l;
}

// merge_search_results (src/search/search.cpp:1551-1600): one workgroup per
// query row; union, a duplicated id keeps its higher score, rank by
// (score desc, id asc), pad with (0, 0).
__global__ __launch_bounds__(256) void merge_results_kernel(long nq, int kd, float* __restrict__ ds,
                                                            uint64_t* __restrict__ di, int ks,
                                                            const float* __restrict__ ss,
                                                            const uint64_t* __restrict__ si) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int m = kd + ks;
  const MergeLayout lay = merge_layout(m, kd);
  float* cs = reinterpret_cast<float*>(smem + lay.cs);
  uint64_t* ci = reinterpret_cast<uint64_t*>(smem + lay.ci);
  int* keep = reinterpret_cast<int*>(smem + lay.keep);
  float* os = reinterpret_cast<float*>(smem + lay.os);
  uint64_t* oi = reinterpret_cast<uint64_t*>(smem + lay.oi);
  for (long q = blockIdx.x; q < nq; q += gridDim.x) {
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
      cs[e] = e < kd ? ds[q * kd + e] : ss[q * ks + (e - kd)];
      ci[e] = e < kd ? di[q * kd + e] : si[q * ks + (e - kd)];
    }
    __syncthreads();
    // an entry is dropped if another entry with the same id has a higher
    // score, or the same score and a lower position (one survivor per id)
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
      int kp = 1;
      for (int f = 0; f < m; ++f)
        if (f != e && ci[f] == ci[e] && (cs[f] > cs[e] || (cs[f] == cs[e] && f < e))) { kp = 0; break; }
      keep[e] = kp;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kd; j += blockDim.x) { os[j] = 0.f; oi[j] = 0u; }
    __syncthreads();
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
      if (!keep[e]) continue;
      int rank = 0;
      for (int f = 0; f < m; ++f)
        if (keep[f] && (cs[f] > cs[e] || (cs[f] == cs[e] && ci[f] < ci[e]))) ++rank;
      if (rank < kd) { os[rank] = cs[e]; oi[rank] = ci[e]; }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kd; j += blockDim.x) { ds[q * kd + j] = os[j]; di[q * kd + j] = oi[j]; }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Format conversions (src/pack/pack.cpp:98-186,403-420)
// ---------------------------------------------------------------------------
// Pseudo-random rows for autotuning inputs: element i = (sum of two
// splitmix64 uniforms in [0, 1)) - 1, a triangular value in (-1, 1).
__global__ __launch_bounds__(256) void fill_value_kernel(float* __restrict__ x, long n, float v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

__global__ __launch_bounds__(256) void fill_rows_kernel(float* __restrict__ x, long total, uint64_t seed) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float u0 = (float)(uint32_t)(z >> 40) * (1.f / 16777216.f);
    const float u1 = (float)(uint32_t)(z & 0xFFFFFF) * (1.f / 16777216.f);
    x[i] = u0 + u1 - 1.f;
  }
}

__device__ __forceinline__ long tiled_offset(long v, int g, int G16) {
  return ((v >> 6) * G16 + (g >> 4)) * 1024 + (v & 63) * 16 + (g & 15);
}

__global__ void tiled_to_flat_kernel(const uint8_t* __restrict__ tiled, long n, int G, int G16,
                                     uint8_t* __restrict__ flat) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n * G; i += (long)gridDim.x * blockDim.x) {
    const long v = i / G;
    const int g = (int)(i - v * G);
    flat[i] = tiled[tiled_offset(v, g, G16)];
  }
}

__global__ void flat_to_tiled_kernel(const uint8_t* __restrict__ flat, long n, int G, int G16,
                                     uint8_t* __restrict__ tiled) {
  const long ntiles = (n + 63) / 64;
  const long total = ntiles * 64 * (long)G16 * 16;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    // i enumerates tiled bytes: tile, chunk, lane, e
    const long tile = i / ((long)G16 * 1024);
    const long rem = i - tile * G16 * 1024;
    const int c16 = (int)(rem / 1024);
    const int lane = (int)((rem / 16) & 63);
    const int e = (int)(rem & 15);
    const long v = tile * 64 + lane;
    const int g = c16 * 16 + e;
    tiled[i] = (v < n && g < G) ? flat[v * G + g] : 0;
  }
}

// The reference CUDA index's code layout (src/gpu/kernels.cuh:18-21,
// src/gpu/kernels.cu:111-133): u32 words [ceil(n/32)][ceil(G/4)][32], word
// (tile, g4, lane) = the byte groups 4 g4 .. 4 g4 + 3 of vector 32 tile +
// lane, little-endian; groups past G and vectors past n are zero.
__global__ void flat_to_gpu32_kernel(const uint8_t* __restrict__ flat, long n, int G, int nbg4,
                                     uint32_t* __restrict__ words) {
  const long total = (n + 31) / 32 * nbg4 * 32;
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < total; w += (long)gridDim.x * blockDim.x) {
    const int lane = (int)(w & 31);
    const long tg = w >> 5;
    const long tile = tg / nbg4;
    const int g4 = (int)(tg - tile * nbg4);
    const long v = tile * 32 + lane;
    uint32_t x = 0u;
    if (v < n) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int g = g4 * 4 + t;
        if (g < G) x |= (uint32_t)flat[v * G + g] << (8 * t);
      }
    }
    words[w] = x;
  }
}

__global__ void gpu32_to_flat_kernel(const uint32_t* __restrict__ words, long n, int G, int nbg4,
                                     uint8_t* __restrict__ flat) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n * G; i += (long)gridDim.x * blockDim.x) {
    const long v = i / G;
    const int g = (int)(i - v * G);
    const uint32_t x = words[((v >> 5) * nbg4 + (g >> 2)) * 32 + (v & 31)];
    flat[i] = (uint8_t)(x >> (8 * (g & 3)));
  }
}

__constant__ unsigned char kPerm0Dev[16] = {0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15};
__constant__ unsigned char kInvPerm0Dev[16] = {0, 2, 4, 6, 8, 10, 12, 14, 1, 3, 5, 7, 9, 11, 13, 15};

// PERM0 192x1: block bk, group g, 32-lane chunk ch, j < 16, va = 192 bk +
// 32 ch + P[j], vb = va + 16: byte j = hi(va) | hi(vb) << 4, byte 16 + j =
// lo(va) | lo(vb) << 4; vectors >= n read as 0.
__global__ void flat_to_perm0_kernel(const uint8_t* __restrict__ flat, long n, int G,
                                     uint8_t* __restrict__ out) {
  const long nb = (n + 191) / 192;
  const long total = nb * G * 96;  // (bk, g, ch, j)
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int j = (int)(i & 15);
    const int ch = (int)((i >> 4) % 6);
    const long bg = (i >> 4) / 6;  // bk * G + g
    const long bk = bg / G;
    const int g = (int)(bg - bk * G);
    const long va = bk * 192 + ch * 32 + kPerm0Dev[j];
    const long vb = va + 16;
    const uint32_t a = va < n ? flat[va * G + g] : 0u;
    const uint32_t b = vb < n ? flat[vb * G + g] : 0u;
    uint8_t* o = out + bg * 192 + ch * 32;
    o[j] = (uint8_t)((a >> 4) | ((b >> 4) << 4));
    o[16 + j] = (uint8_t)((a & 15u) | ((b & 15u) << 4));
  }
}

// sequential 192x1: out[(bk * G + g) * 192 + lane] = flat[bk * 192 + lane][g], 0 past n
__global__ void flat_to_seq_kernel(const uint8_t* __restrict__ flat, long n, int G, uint8_t* __restrict__ out) {
  const long total = (n + 191) / 192 * (long)G * 192;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int lane = (int)(i % 192);
    const long bg = i / 192;
    const long bk = bg / G;
    const int g = (int)(bg - bk * G);
    const long v = bk * 192 + lane;
    out[i] = v < n ? flat[v * G + g] : 0;
  }
}

__global__ void perm0_to_flat_kernel(const uint8_t* __restrict__ blk, long n, int G,
                                     uint8_t* __restrict__ flat) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n * G; i += (long)gridDim.x * blockDim.x) {
    const long v = i / G;
    const int g = (int)(i - v * G);
    const long bk = v / 192;
    const int lane = (int)(v - bk * 192);
    const int ch = lane / 32, l = lane & 31;
    const int j = kInvPerm0Dev[l & 15];
    const uint8_t* tb = blk + (bk * G + g) * 192 + ch * 32;
    const uint32_t hp = tb[j], lp = tb[16 + j];
    const uint32_t hi = l < 16 ? (hp & 15u) : (hp >> 4);
    const uint32_t lo = l < 16 ? (lp & 15u) : (lp >> 4);
    flat[i] = (uint8_t)((hi << 4) | lo);
  }
}

// flat byte groups -> bit-plane rows (the encode() output format).
__global__ void flat_to_packed_kernel(const uint8_t* __restrict__ flat, long n, int D, int bits,
                                      uint8_t* __restrict__ packed) {
  const int G = D / (8 / bits), bpp = D / 8;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n * bpp; i += (long)gridDim.x * blockDim.x) {
    const long v = i / bpp;
    const int c8 = (int)(i - v * bpp);
    unsigned codes[8];
    for (int kk = 0; kk < 8; ++kk) {
      const int j = c8 * 8 + kk;
      if (bits == 2) {
        codes[kk] = (flat[v * G + j / 4] >> (6 - 2 * (j & 3))) & 3u;
      } else {
        const uint32_t byte = flat[v * G + j / 2];
        codes[kk] = (j & 1) ? (byte & 15u) : (byte >> 4);
      }
    }
    for (int p = 0; p < bits; ++p) {
      uint32_t byte = 0;
      for (int kk = 0; kk < 8; ++kk) byte |= ((codes[kk] >> p) & 1u) << (7 - kk);
      packed[v * (long)(bits * bpp) + p * bpp + c8] = (uint8_t)byte;
    }
  }
}

// Bit-planes -> flat bytes (codes_flat_from_packed, src/pack/pack.cpp:98-131,
// 352-356), the inverse of flat_to_packed_kernel: one thread per 8 codes.
__global__ void packed_to_flat_kernel(const uint8_t* __restrict__ packed, long n, int D, int bits,
                                      uint8_t* __restrict__ flat) {
  const int G = D / (8 / bits), bpp = D / 8;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n * bpp; i += (long)gridDim.x * blockDim.x) {
    const long v = i / bpp;
    const int c8 = (int)(i - v * bpp);
    unsigned codes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = 0; p < bits; ++p) {
      const uint32_t byte = packed[v * (long)(bits * bpp) + p * bpp + c8];
      for (int kk = 0; kk < 8; ++kk) codes[kk] |= ((byte >> (7 - kk)) & 1u) << p;
    }
    uint8_t* out = flat + v * G;
    if (bits == 2) {
      out[c8 * 2] = (uint8_t)(codes[0] << 6 | codes[1] << 4 | codes[2] << 2 | codes[3]);
      out[c8 * 2 + 1] = (uint8_t)(codes[4] << 6 | codes[5] << 4 | codes[6] << 2 | codes[7]);
    } else {
      for (int h = 0; h < 4; ++h) out[c8 * 4 + h] = (uint8_t)(codes[2 * h] << 4 | codes[2 * h + 1]);
    }
  }
}

// dst row i = src row idx[i] (rows of row_words 32-bit words).
__global__ void gather_rows_kernel(const uint32_t* __restrict__ src, const uint64_t* __restrict__ idx, long m,
                                   int row_words, uint32_t* __restrict__ dst) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m * row_words; i += (long)gridDim.x * blockDim.x) {
    const long r = i / row_words;
    const int w = (int)(i - r * row_words);
    dst[i] = src[(long)idx[r] * row_words + w];
  }
}

// ---------------------------------------------------------------------------
// Landmark index (reference src/index.cpp:151-265 and the AVX2 path of
// src/search/search.cpp:1304-1546, where sign planes are never built).
// ---------------------------------------------------------------------------

// alpha * y-hat rows (src/index.cpp:88-117), one thread per row, in place:
// level = #{b : x >= b}, alpha = (float) (dot / |c|^2) over sequential double
// sums (the products are exact in double); the squared norm of the scaled row
// sums ROUNDED squares in element order, which is what the reference's
// vectorized fold computes for rows of dim % 8 == 0.
__global__ __launch_bounds__(256) void lm_yhat_kernel(CodeTables t, float* __restrict__ rows, long n,
                                                      float* __restrict__ vnorm2) {
  __shared__ float cen[16];
  if (threadIdx.x == 0)
    for (int i = 0; i < 16; ++i) cen[i] = t.centroids[i];
  __syncthreads();
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int D = t.dim, nb = (1 << t.bits) - 1;
  float* row = rows + r * (long)D;
  auto level = [&](float x) {
    int k = 0;
    for (int b = 0; b < nb; ++b) k += x >= t.boundaries[b] ? 1 : 0;
    return k;
  };
  double dot = 0.0, cc = 0.0;
  for (int i = 0; i < D; i += 8) {  // D % 8 == 0: two 16-byte loads in flight per step
    const float4 a = *reinterpret_cast<const float4*>(row + i);
    const float4 b = *reinterpret_cast<const float4*>(row + i + 4);
    const float xs[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double c = (double)cen[level(xs[u])];
      dot = dot + (double)xs[u] * c;
      cc = cc + c * c;
    }
  }
  const float alpha = cc > 1e-12 ? (float)(dot / cc) : 0.f;
  float vn = 0.f;
  for (int i = 0; i < D; i += 8) {
    const float4 a = *reinterpret_cast<const float4*>(row + i);
    const float4 b = *reinterpret_cast<const float4*>(row + i + 4);
    float xs[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const float vi = alpha * cen[level(xs[u])];
      xs[u] = vi;
      vn = vn + vi * vi;
    }
    *reinterpret_cast<float4*>(row + i) = make_float4(xs[0], xs[1], xs[2], xs[3]);
    *reinterpret_cast<float4*>(row + i + 4) = make_float4(xs[4], xs[5], xs[6], xs[7]);
  }
  vnorm2[r] = vn;
}

// Rows onto landmarks (src/transform/landmarks.cpp:104-121,200-234): out[r][j]
// = <Y_r, L_j> as the AVX2 dot: 8 chains acc[l] = fma(y[8t + l], L[8t + l],
// acc[l]) in t order, then ((a0 + a4) + (a1 + a5)) + ((a2 + a6) + (a3 + a7)).
// Workgroup: 32 rows x 32 landmarks. Lane l = tid & 7 runs chain l of 4 rows
// x 8 landmarks; the 8 chains of an output meet through lane shuffles.
constexpr int kLmK = 64;             // dims per LDS stage
constexpr int kLmPitch = kLmK + 2;   // row pitch: the 4 row blocks of a half-wave hit distinct banks
__global__ __launch_bounds__(256) void lm_project_kernel(const float* __restrict__ Y, long n, int D,
                                                         const float* __restrict__ L, int m,
                                                         float* __restrict__ out, int stride) {
  __shared__ float ys[32 * kLmPitch], ls[32 * kLmPitch];
  const int tid = threadIdx.x, l = tid & 7, oct = tid >> 3, rb = oct & 7, cb = oct >> 3;
  const long r0 = (long)blockIdx.x * 32;
  const int j0 = (int)blockIdx.y * 32;
  float acc[4][8];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[i][c] = 0.f;
  for (int k0 = 0; k0 < D; k0 += kLmK) {
    const int kc = min(kLmK, D - k0);  // a multiple of 8
#pragma unroll
    for (int e = tid; e < 32 * kLmK / 4; e += 256) {
      const int rr = e / (kLmK / 4), c4 = (e % (kLmK / 4)) * 4;
      float4 yv = make_float4(0.f, 0.f, 0.f, 0.f), lv = yv;
      if (c4 < kc) {
        if (r0 + rr < n) yv = *reinterpret_cast<const float4*>(Y + (r0 + rr) * (long)D + k0 + c4);
        if (j0 + rr < m) lv = *reinterpret_cast<const float4*>(L + (long)(j0 + rr) * D + k0 + c4);
      }
      float* yd = ys + rr * kLmPitch + c4;
      float* ld = ls + rr * kLmPitch + c4;
      yd[0] = yv.x; yd[1] = yv.y; yd[2] = yv.z; yd[3] = yv.w;
      ld[0] = lv.x; ld[1] = lv.y; ld[2] = lv.z; ld[3] = lv.w;
    }
    __syncthreads();
    const int steps = kc / 8;
    for (int s8 = 0; s8 < steps; ++s8) {
      float yv[4], lv[8];
#pragma unroll
      for (int i = 0; i < 4; ++i) yv[i] = ys[(rb * 4 + i) * kLmPitch + s8 * 8 + l];
#pragma unroll
      for (int c = 0; c < 8; ++c) lv[c] = ls[(cb * 8 + c) * kLmPitch + s8 * 8 + l];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[i][c] = __fmaf_rn(yv[i], lv[c], acc[i][c]);
    }
    __syncthreads();
  }
  const int j = j0 + cb * 8 + l;  // lane l writes landmark cb * 8 + l of its 4 rows
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float mine = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      float v = acc[i][c];
      v = v + __shfl_xor(v, 4);
      v = v + __shfl_xor(v, 1);
      v = v + __shfl_xor(v, 2);
      mine = c == l ? v : mine;
    }
    const long r = r0 + rb * 4 + i;
    if (r < n && j < stride) out[r * stride + j] = j < m ? mine : 0.f;
  }
}

// Residual norms (src/index.cpp:60-82,200-210), one thread per row:
// |row[0:m)|^2 as the AVX2 norm_sq (8 fused chains over m & ~7 entries,
// summed in order; of the tail, a block of 4 rounded squares added in order
// when at least 4 remain, then fused multiply-adds), rn = sqrtf(max(0,
// vnorm2 - that)).
// ---------------------------------------------------------------------------
// Staged sign-plane search of the landmark index (the reference's branch on an
// AVX-512 VNNI host, src/search/search.cpp:1379-1510).
// ---------------------------------------------------------------------------
__device__ __forceinline__ long tiled_at(long v, int g, int G16) {
  return ((v >> 6) * G16 + (g >> 4)) * 1024 + (v & 63) * 16 + (g & 15);
}

// Sign bytes from the 4-bit landmark codes (both in the tiled layout): byte gs
// of vector v holds the top bit of codes 8 gs .. 8 gs + 7, first code in bit 7
// (the packed codes' top bit-plane, src/pack/pack.cpp:490-507). One thread
// per (vector, sign byte).
__global__ void lm_sign_codes_kernel(const uint8_t* __restrict__ tiled4, long n, int G16_4, int Gs, int G16_s,
                                     uint8_t* __restrict__ tiled_s) {
  const long total = n * (long)Gs;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long v = i / Gs;
    const int gs = (int)(i - v * Gs);
    unsigned sb = 0;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const unsigned b = tiled4[tiled_at(v, 4 * gs + h, G16_4)];
      sb = (sb << 2) | (((b >> 7) & 1u) << 1) | ((b >> 3) & 1u);
    }
    tiled_s[tiled_at(v, gs, G16_s)] = (uint8_t)sb;
  }
}

// build_sign_lut (src/search/search.cpp:862-914) of one query row of S dims
// per workgroup: per byte group and half, the 16 sums of +-q over 4 dims,
// their minimum, the bias (the minima summed in order), scale = largest span
// / 127, then u8 = clamp(round((v - min) / scale ... ), 0, 127).
constexpr int kSignLutThreads = 256;
__global__ __launch_bounds__(kSignLutThreads) void sign_lut_kernel(const float* __restrict__ q_land, int S,
                                                                  uint8_t* __restrict__ luts,
                                                                  float* __restrict__ lut_scale,
                                                                  float* __restrict__ bias_total) {
  extern __shared__ float sl[];  // fv [Gs * 32] | mins [Gs * 2] | span partials [threads]
  const int Gs = S / 8, tid = threadIdx.x;
  const float* q = q_land + (long)blockIdx.x * S;
  float* fv = sl;
  float* mins = fv + Gs * 32;
  float* red = mins + Gs * 2;
  float span = 0.f;
  for (int h = tid; h < Gs * 2; h += blockDim.x) {
    const int d = h * 4;  // (group, half) = (h / 2, h % 2): dims 8 g + 4 half ..
    const float p0 = q[d], p1 = q[d + 1], p2 = q[d + 2], p3 = q[d + 3];
    const float a[4] = {-p0 - p1, p1 - p0, p0 - p1, p0 + p1};
    const float b[4] = {-p2 - p3, p3 - p2, p2 - p3, p2 + p3};
#pragma unroll
    for (int pat = 0; pat < 16; ++pat) fv[h * 16 + pat] = a[pat >> 2] + b[pat & 3];
    const float m01 = fabsf(p0) + fabsf(p1), m23 = fabsf(p2) + fabsf(p3);
    const float mn = -m01 - m23, mx = m01 + m23;
    mins[h] = mn;
    span = fmaxf(span, mx - mn);
  }
  red[tid] = span;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] = fmaxf(red[tid], red[tid + o]);
    __syncthreads();
  }
  float scale = red[0] > 0.f ? div_rn(red[0], 127.f) : 1.f;
  float inv = 1.f;
  if (scale >= FLT_MIN)
    inv = div_rn(1.f, scale);
  else
    scale = 1.f;
  if (tid == 0) {
    float bias = 0.f;
    for (int h = 0; h < Gs * 2; ++h) bias = bias + mins[h];
    lut_scale[blockIdx.x] = scale;
    bias_total[blockIdx.x] = bias;
  }
  uint8_t* out = luts + (long)blockIdx.x * Gs * 32;
  for (int j = tid; j < Gs * 32; j += blockDim.x) {
    float v = roundf((fv[j] - mins[j >> 4]) * inv);
    v = fminf(fmaxf(v, 0.f), 127.f);
    out[j] = (uint8_t)v;
  }
}

// Stage 2: the E shortlisted vectors of one query per workgroup, rescored in
// shortlist order with build_permute_dot's int8 weights and levels
// (src/search/search_vnni.cpp:1218-1360): acc = sum over byte groups of
// lev[lo] * w_lo + lev[hi] * w_hi (exact), f = (float) acc * (cs * qs),
// score = fmaf(f, alpha, q_perp * rn) (the compiled flush, :1194-1217); each
// score goes into a k-heap in shortlist order (topk_heap.hpp:101-157: strict
// admission, rescan below 128, binary heap from 128), then the heap is
// written sorted by (score desc, id asc).
constexpr int kStagedThreads = 256;
struct StagedArgs {
  const float* q_land;
  int S;
  const float* q_perp;
  float cen[16];
  const uint8_t* tiled4;
  int G16;
  const float* alpha;
  const float* rn;
  const uint64_t* cand;  // [nq][E] positions, shortlist order
  int E, k;
  float* out_s;
  uint64_t* out_i;
};

__device__ __forceinline__ bool staged_worse(float sa, uint64_t ia, float sb, uint64_t ib) {
  return sa != sb ? sa < sb : ia > ib;
}

__global__ __launch_bounds__(kStagedThreads) void lm_staged_kernel(StagedArgs a) {
  extern __shared__ __align__(16) unsigned char sm[];
  const int S = a.S, G = S / 2, E = a.E, k = a.k, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long qi = blockIdx.x;
  float* sc = reinterpret_cast<float*>(sm);                     // [E] rescored
  float* hs = sc + ((E + 1) & ~1);                               // [k] heap scores
  uint64_t* hid = reinterpret_cast<uint64_t*>(hs + ((k + 1) & ~1));  // [k] heap ids
  int8_t* wlo = reinterpret_cast<int8_t*>(hid + k);              // [G]
  int8_t* whi = wlo + G;                                         // [G]
  __shared__ float red[kStagedThreads];
  __shared__ int8_t lev[16];
  __shared__ int hsz_sh;
  const float* q = a.q_land + qi * S;
  // query weights: qs = max |q| / 127
  float qm = 0.f;
  for (int d = tid; d < S; d += blockDim.x) qm = fmaxf(qm, fabsf(q[d]));
  red[tid] = qm;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] = fmaxf(red[tid], red[tid + o]);
    __syncthreads();
  }
  float qs = div_rn(red[0], 127.f), inv = 1.f;
  if (qs >= FLT_MIN)
    inv = div_rn(1.f, qs);
  else
    qs = 1.f;
  float cmax = 0.f;
#pragma unroll
  for (int i = 0; i < 16; ++i) cmax = fmaxf(cmax, fabsf(a.cen[i]));
  const float cs = cmax > 0.f ? div_rn(cmax, 127.f) : 1.f;
  if (tid < 16) lev[tid] = (int8_t)fminf(fmaxf(roundf(div_rn(a.cen[tid], cs)), -127.f), 127.f);
  for (int g = tid; g < G; g += blockDim.x) {
    wlo[g] = (int8_t)fminf(fmaxf(roundf(q[2 * g + 1] * inv), -127.f), 127.f);
    whi[g] = (int8_t)fminf(fmaxf(roundf(q[2 * g] * inv), -127.f), 127.f);
  }
  __syncthreads();
  const float pscale = cs * qs, qp = a.q_perp[qi];
  const uint64_t* cand = a.cand + qi * E;
  const int waves = blockDim.x >> 6;
  for (int c = wave; c < E; c += waves) {
    const uint64_t id = cand[c];
    int acc = 0;
    for (int g = lane; g < G; g += 64) {
      const unsigned b = a.tiled4[tiled_at((long)id, g, a.G16)];
      acc += (int)lev[b & 15u] * (int)wlo[g] + (int)lev[b >> 4] * (int)whi[g];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) {
      const float f = (float)acc * pscale;
      sc[c] = __fmaf_rn(f, a.alpha[id], qp * a.rn[id]);
    }
  }
  __syncthreads();
  if (tid == 0) {
    int hsz = 0, hmi = 0;
    float hmin = -INFINITY;
    for (int c = 0; c < E; ++c) {
      const float s = sc[c];
      if (!(s > hmin)) continue;  // strict admission; -inf while the heap fills
      const uint64_t id = cand[c];
      if (k < 128) {
        if (hsz < k) {
          hs[hsz] = s;
          hid[hsz] = id;
          ++hsz;
          if (hsz < k) continue;
        } else {
          hs[hmi] = s;
          hid[hmi] = id;
        }
        int mi = 0;  // rescan for the worst entry
        for (int h = 1; h < k; ++h)
          if (staged_worse(hs[h], hid[h], hs[mi], hid[mi])) mi = h;
        hmin = hs[mi];
        hmi = mi;
      } else {
        int i;
        if (hsz < k) {
          hs[hsz] = s;
          hid[hsz] = id;
          ++hsz;
          if (hsz < k) continue;
          i = k / 2;  // heapify, then the root is the worst
          while (i-- > 0) {
            int j = i;
            for (;;) {
              const int l = 2 * j + 1, r = 2 * j + 2;
              int w = j;
              if (l < k && staged_worse(hs[l], hid[l], hs[w], hid[w])) w = l;
              if (r < k && staged_worse(hs[r], hid[r], hs[w], hid[w])) w = r;
              if (w == j) break;
              const float ts = hs[j];
              hs[j] = hs[w];
              hs[w] = ts;
              const uint64_t ti = hid[j];
              hid[j] = hid[w];
              hid[w] = ti;
              j = w;
            }
          }
        } else {
          hs[0] = s;
          hid[0] = id;
          int j = 0;
          for (;;) {
            const int l = 2 * j + 1, r = 2 * j + 2;
            int w = j;
            if (l < k && staged_worse(hs[l], hid[l], hs[w], hid[w])) w = l;
            if (r < k && staged_worse(hs[r], hid[r], hs[w], hid[w])) w = r;
            if (w == j) break;
            const float ts = hs[j];
            hs[j] = hs[w];
            hs[w] = ts;
            const uint64_t ti = hid[j];
            hid[j] = hid[w];
            hid[w] = ti;
            j = w;
          }
        }
        hmin = hs[0];
        hmi = 0;
      }
    }
    hsz_sh = hsz;
  }
  __syncthreads();
  const int hsz = hsz_sh;
  float* os = a.out_s + qi * k;
  uint64_t* oi = a.out_i + qi * k;
  for (int i = tid; i < k; i += blockDim.x) {
    if (i >= hsz) {
      os[i] = 0.f;
      oi[i] = 0;
      continue;
    }
    int rank = 0;  // entries better than entry i (ids are distinct)
    for (int j = 0; j < hsz; ++j) rank += staged_worse(hs[i], hid[i], hs[j], hid[j]) ? 1 : 0;
    os[rank] = hs[i];
    oi[rank] = hid[i];
  }
}

__global__ __launch_bounds__(256) void lm_resid_kernel(const float* __restrict__ rows, long n, int m,
                                                       int stride, const float* __restrict__ vnorm2,
                                                       float* __restrict__ rn) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float* x = rows + r * (long)stride;
  float c[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int m8 = m & ~7;
  for (int j = 0; j < m8; j += 8) {
    const float4 a = *reinterpret_cast<const float4*>(x + j);
    const float4 b = *reinterpret_cast<const float4*>(x + j + 4);
    c[0] = __fmaf_rn(a.x, a.x, c[0]);
    c[1] = __fmaf_rn(a.y, a.y, c[1]);
    c[2] = __fmaf_rn(a.z, a.z, c[2]);
    c[3] = __fmaf_rn(a.w, a.w, c[3]);
    c[4] = __fmaf_rn(b.x, b.x, c[4]);
    c[5] = __fmaf_rn(b.y, b.y, c[5]);
    c[6] = __fmaf_rn(b.z, b.z, c[6]);
    c[7] = __fmaf_rn(b.w, b.w, c[7]);
  }
  float s = c[0];
#pragma unroll
  for (int u = 1; u < 8; ++u) s = s + c[u];
  int j = m8;
  if (m - j >= 4) {
    for (int u = 0; u < 4; ++u) s = s + x[j + u] * x[j + u];
    j += 4;
  }
  for (; j < m; ++j) s = __fmaf_rn(x[j], x[j], s);
  const float d = vnorm2[r] - s;
  rn[r] = sqrt_rn(0.f < d ? d : 0.f);
}

// Query side of the landmark search (src/transform/landmarks.cpp:271-348,
// src/search/search.cpp:1349-1378), one workgroup per query: int8 levels
// round(x * (127 / amax)) clamped to +-127, exact integer dots with the 4-bit
// basis columns (nibble u = level + 7: dot = sum q u - 7 sum q, four bytes
// per v_dot4), q_land[j] = (q_scale * scale_j) * (float) dot; q_perp =
// (float) sqrt(max(0, |q|^2 - |q_land|^2)), both sums sequential in double.
// codes_t: the basis as u32 words [dim_pad / 8][m] (word w of column j at
// w * m + j), so the lanes of a wavefront, one column each, read consecutive
// words. 1024 threads, one column each for the default 1024 landmarks, with
// the word loop unrolled so that each lane keeps several loads in flight (a
// single query is latency-bound otherwise). raw_R != 0: q_rot holds the raw
// queries; the kernel takes the row norm and rotates them itself (one
// wavefront per rotation block of 64 * raw_R elements, as query preparation
// does), saving a launch.
constexpr int kLmQueryThreads = 1024;
__global__ __launch_bounds__(kLmQueryThreads) void lm_query_kernel(CodeTables t, int raw_R,
                                                                   const float* __restrict__ q_rot, long nq, int D,
                                                       int dim_pad, const uint32_t* __restrict__ codes_t,
                                                       const float* __restrict__ lscales, int m,
                                                       float* __restrict__ q_land, int stride,
                                                       float* __restrict__ q_perp) {
  extern __shared__ __align__(16) unsigned char smem[];
  float* qf = reinterpret_cast<float*>(smem);  // D
  float* ql = qf + D;                          // stride
  int* qe = reinterpret_cast<int*>(ql + stride);  // dims 8w + {0, 2, 4, 6} as bytes
  int* qo = qe + dim_pad / 8;                     // dims 8w + {1, 3, 5, 7}
  float* xs = reinterpret_cast<float*>(qo + dim_pad / 8);  // D: the raw query (raw_R != 0)
  __shared__ float s_norm;
  constexpr int kWaves = kLmQueryThreads / 64;
  __shared__ float s_amax[kWaves];
  __shared__ int s_sum[kWaves];
  __shared__ double s_q2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int W = dim_pad / 8;
  for (long qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    __syncthreads();
    const float* xr = q_rot + qi * (long)D;
    float am = 0.f;
    if (raw_R != 0) {
      stage_to_lds(xs, xr, D);
      __syncthreads();
      if (wave == 0) {
        const float nrm = norm8_lds(xs, D, lane);
        if (lane == 0) s_norm = nrm;
      }
      __syncthreads();
      const float inv = s_norm > kMinNorm ? div_rn(1.f, s_norm) : 0.f;
      if (raw_R == 2) rotate_query_blocks<2>(t, xs, inv, qf, wave, kWaves, lane);
      else if (raw_R == 4) rotate_query_blocks<4>(t, xs, inv, qf, wave, kWaves, lane);
      else if (raw_R == 8) rotate_query_blocks<8>(t, xs, inv, qf, wave, kWaves, lane);
      else rotate_query_blocks<16>(t, xs, inv, qf, wave, kWaves, lane);
      __syncthreads();
      for (int i = tid; i < D; i += kLmQueryThreads) am = fmaxf(am, fabsf(qf[i]));
    } else {
      for (int i = tid; i < D; i += kLmQueryThreads) {
        const float v = xr[i];
        qf[i] = v;
        am = fmaxf(am, fabsf(v));
      }
    }
    for (int o = 32; o > 0; o >>= 1) am = fmaxf(am, __shfl_xor(am, o));
    if (lane == 0) s_amax[wave] = am;
    __syncthreads();
    float amax = s_amax[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) amax = fmaxf(amax, s_amax[w]);
    const bool zero = amax < 1e-30f;
    const float qs = zero ? 0.f : div_rn(amax, 127.f);
    const float inv = zero ? 0.f : div_rn(127.f, amax);
    int sq = 0;
    for (int w = tid; w < W; w += kLmQueryThreads) {
      uint32_t ev = 0u, od = 0u;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int d = 8 * w + u;
        int lv = 0;
        if (!zero && d < D) lv = (int)fmaxf(-127.f, fminf(127.f, roundf(qf[d] * inv)));
        sq += lv;
        const uint32_t b = (uint32_t)lv & 0xffu;
        if (u & 1) od |= b << (8 * (u >> 1));
        else ev |= b << (8 * (u >> 1));
      }
      qe[w] = (int)ev;
      qo[w] = (int)od;
    }
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    if (lane == 0) s_sum[wave] = sq;
    __syncthreads();
    int sumq = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) sumq += s_sum[w];
    if (tid == kLmQueryThreads - 64) {  // |q|^2 in element order while the other lanes take columns
      double a = 0.0;
      for (int d = 0; d < D; ++d) {
        const double v = (double)qf[d];
        a = a + v * v;
      }
      s_q2 = a;
    }
    for (int j = tid; j < stride; j += kLmQueryThreads) {
      float v = 0.f;
      if (j < m) {
        int acc = 0;
#pragma unroll 8
        for (int w = 0; w < W; ++w) {
          const uint32_t cw = codes_t[(long)w * m + j];
          acc = __builtin_amdgcn_sdot4((int)(cw & 0x0F0F0F0Fu), qe[w], acc, false);
          acc = __builtin_amdgcn_sdot4((int)((cw >> 4) & 0x0F0F0F0Fu), qo[w], acc, false);
        }
        v = (qs * lscales[j]) * (float)(acc - 7 * sumq);
      }
      ql[j] = v;
      q_land[qi * (long)stride + j] = v;
    }
    __syncthreads();
    if (tid == 0) {
      double l2 = 0.0;
      for (int j = 0; j < m; ++j) {
        const double v = (double)ql[j];
        l2 = l2 + v * v;
      }
      const double d = s_q2 - l2;
      q_perp[qi] = (float)sqrt(0.0 < d ? d : 0.0);
    }
  }
}

int grid_for(long work, int threads, int cap = 65536) {
  long g = (work + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// Resident workgroups of a persistent launch: workgroups per CU (from the
// occupancy calculator at this block size and LDS) times the CU count.
int resident_grid(const void* fn, int threads, size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return cus * per_cu;
}

// The one-pass encode for its shapes; false when the shape is not one of them.
template <int NX, int CPB, bool TQ>
bool launch_fused_shape(const CodeTables& t, const float* x, long n, long out_offset, uint8_t* tiled, float* alpha,
                        hipStream_t s) {
  using S = FusedShape<NX, CPB>;
  const size_t shared = S::shared_bytes(TQ);
  if (shared + 8 * (size_t)S::SLOT > (size_t)kLdsBudget) return false;
  const int waves = ((size_t)kLdsBudget - shared) / S::SLOT >= 16 ? 16 : 8;
  const size_t lds = shared + (size_t)waves * S::SLOT;
  const void* fn = (const void*)encode_fused_kernel<NX, CPB, TQ>;
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const long octets = (((out_offset + n + 63) >> 6) - (out_offset >> 6)) * 8;
  const long need = (octets + waves - 1) / waves;
  static const int resident = resident_grid(fn, 64 * waves, lds);  // one GPU per process
  const int grid = (int)std::min<long>(need, resident);
  hipLaunchKernelGGL((encode_fused_kernel<NX, CPB, TQ>), dim3(grid), dim3(64 * waves), lds, s, t, x, n, out_offset,
                     tiled, alpha);
  return true;
}

template <int CPB, bool TQ>
bool launch_fused_nx(const CodeTables& t, const float* x, long n, long out_offset, uint8_t* tiled, float* alpha,
                     hipStream_t s) {
  switch (t.dim / 64) {
    case 2: if constexpr (CPB == 2) return launch_fused_shape<2, CPB, TQ>(t, x, n, out_offset, tiled, alpha, s);
            return false;
    case 4: return launch_fused_shape<4, CPB, TQ>(t, x, n, out_offset, tiled, alpha, s);
    case 6: if constexpr (CPB == 2) return launch_fused_shape<6, CPB, TQ>(t, x, n, out_offset, tiled, alpha, s);
            return false;
    case 8: return launch_fused_shape<8, CPB, TQ>(t, x, n, out_offset, tiled, alpha, s);
    case 12: return launch_fused_shape<12, CPB, TQ>(t, x, n, out_offset, tiled, alpha, s);
    case 24: return launch_fused_shape<24, CPB, TQ>(t, x, n, out_offset, tiled, alpha, s);
    default: return false;
  }
}

bool launch_fused(const CodeTables& t, const float* x, long n, long out_offset, uint8_t* tiled, float* alpha,
                  hipStream_t s) {
  if (t.dim % 64 != 0 || t.block != 64 * ((t.dim / 64) & -(t.dim / 64))) return false;
  if ((reinterpret_cast<uintptr_t>(x) & 15) != 0) return false;  // rows are read as 16-byte pieces
  const bool tq = t.shift != nullptr;
  if (t.bits == 2)
    return tq ? launch_fused_nx<4, true>(t, x, n, out_offset, tiled, alpha, s)
              : launch_fused_nx<4, false>(t, x, n, out_offset, tiled, alpha, s);
  return tq ? launch_fused_nx<2, true>(t, x, n, out_offset, tiled, alpha, s)
            : launch_fused_nx<2, false>(t, x, n, out_offset, tiled, alpha, s);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_encode(const CodeTables& t, const float* x, std::size_t n, std::size_t out_offset,
                         std::uint8_t* tiled, std::uint8_t* packed, float* rotated, float* alpha,
                         hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int R = t.block / 64;
  const int NX = t.dim / 64;
  const bool plain = t.prerotated == 0 && t.norm_seq == 0;
  if (plain && tiled != nullptr && alpha != nullptr && packed == nullptr && rotated == nullptr &&
      t.perm_gather != nullptr && t.use_cells && (t.debug_flags & (4 | 64)) == 0 &&
      launch_fused(t, x, (long)n, (long)out_offset, tiled, alpha, s))
    return hipGetLastError();
  if (plain && tiled != nullptr && alpha != nullptr && packed == nullptr && rotated == nullptr && t.dim % 64 == 0 &&
      t.perm_gather != nullptr && (R == 2 || R == 4 || R == 8 || R == 16) && (t.bits != 2 || R >= 4) &&
      NX <= 64 && (t.debug_flags & 4) == 0 &&
      row_shared_bytes(t.dim, t.shift != nullptr) + row_wave_bytes(t.dim, t.groups16, t.block) <= (size_t)kLdsBudget) {
    // two streaming passes: row norms into alpha[], then one wavefront per row
    const size_t shared = row_shared_bytes(t.dim, t.shift != nullptr);
    const size_t per_wave = row_wave_bytes(t.dim, t.groups16, t.block);
    const int waves = (int)std::min<size_t>(kRowMaxWaves, ((size_t)kLdsBudget - std::min(shared, (size_t)kLdsBudget)) / per_wave);
    hipLaunchKernelGGL(norm_rows_kernel, dim3(grid_for((long)(n + 31) / 32, 1, 16384)), dim3(256), 0, s, x, (long)n,
                       t.dim, alpha + out_offset);
    const size_t lds = shared + (size_t)waves * per_wave;
    const long ntile = (long)((out_offset + n + 63) / 64 - out_offset / 64);
    const int grid = grid_for(ntile, 1, 4096);
#define VC_ROW(RR, NN)                                                                                        \
  do {                                                                                                        \
    (void)hipFuncSetAttribute((const void*)encode_row_kernel<RR, NN>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                              (int)lds);                                                                      \
    hipLaunchKernelGGL((encode_row_kernel<RR, NN>), dim3(grid), dim3(64 * waves), lds, s, t, x, (long)n,          \
                       (long)out_offset, tiled, alpha);                                                       \
  } while (0)
    if (NX <= 8) {
      if (R == 2) VC_ROW(2, 8);
      else if (R == 4) VC_ROW(4, 8);
      else VC_ROW(8, 8);
    } else if (NX <= 32) {
      if (R == 2) VC_ROW(2, 32);
      else if (R == 4) VC_ROW(4, 32);
      else if (R == 8) VC_ROW(8, 32);
      else VC_ROW(16, 32);
    } else {
      if (R == 2) VC_ROW(2, 64);
      else if (R == 4) VC_ROW(4, 64);
      else if (R == 8) VC_ROW(8, 64);
      else VC_ROW(16, 64);
    }
#undef VC_ROW
    return hipGetLastError();
  }
  const bool fast_shape = (R == 2 && t.bits != 2) || R == 4 || R == 8 || R == 16;
  if (plain && tiled != nullptr && alpha != nullptr && packed == nullptr && rotated == nullptr && fast_shape &&
      t.perm_sign != nullptr && (t.bits != 2 || R >= 4) && enc_lds_bytes(t.dim, t.groups16) <= (size_t)kLdsBudget) {
    const size_t lds = enc_lds_bytes(t.dim, t.groups16);
    const long tiles = (long)((out_offset + n + kEncodeRows - 1) / kEncodeRows - out_offset / kEncodeRows);
    const int grid = grid_for((tiles + kEncodeWaves - 1) / kEncodeWaves, 1, 8192);
#define VC_ENC(RR)                                                                                          \
  do {                                                                                                      \
    (void)hipFuncSetAttribute((const void*)encode_tile_kernel<RR>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                              (int)lds);                                                                    \
    hipLaunchKernelGGL(encode_tile_kernel<RR>, dim3(grid), dim3(64 * kEncodeWaves), lds, s, t, x, (long)n,      \
                       (long)out_offset, tiled, alpha);                                                     \
  } while (0)
    if (R == 2) VC_ENC(2);
    else if (R == 4) VC_ENC(4);
    else if (R == 8) VC_ENC(8);
    else VC_ENC(16);
#undef VC_ENC
    return hipGetLastError();
  }
  const size_t lds = (size_t)t.dim * 5;
  const int grid = grid_for((long)(n + 7) / 8, 1, 16384);
  hipLaunchKernelGGL(encode_kernel, dim3(grid), dim3(64), lds, s, t, x, (long)n, (long)out_offset, tiled,
                     packed, rotated, alpha);
  return hipGetLastError();
}

hipError_t launch_calibrate(const float* rotated, std::size_t n, int dim, std::size_t lo_idx,
                            std::size_t hi_idx, float qc_lo, float qc_span, std::uint32_t* keys,
                            float* shift, float* scale, float* inv_scale, hipStream_t s) {
  dim3 tg((dim + 31) / 32, (unsigned)((n + 31) / 32));
  hipLaunchKernelGGL(keys_transpose_kernel, tg, dim3(256), 0, s, rotated, (long)n, dim, keys);
  hipLaunchKernelGGL(radix_select_kernel, dim3(dim), dim3(256), 0, s, keys, (long)n, (long)lo_idx,
                     (long)hi_idx, qc_lo, qc_span, shift, scale, inv_scale);
  return hipGetLastError();
}

hipError_t launch_prepare_queries(const CodeTables& t, const float* q, std::size_t nq,
                                  std::uint8_t* luts, float* lut_scale, float* lut_bias,
                                  float* bias_corr, float* bias_total, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  const size_t lds = (size_t)t.dim * 8 + (size_t)t.groups * 8 + (t.shift != nullptr ? (size_t)t.dim * 8 : 0);
  const int grid = grid_for((long)nq, 1, 16384);
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)prepare_queries_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
  hipLaunchKernelGGL(prepare_queries_kernel, dim3(grid), dim3(kPrepThreads), lds, s, t, q, (long)nq, luts,
                     lut_scale, lut_bias, bias_corr, bias_total);
  return hipGetLastError();
}

// top-k state of a workgroup (lists, candidates, counters) + per-query scale
// and bias for the pair kernel
static size_t topk_lds(int T, int k, int cap) { return (size_t)T * (2 * (size_t)k + 2 * (size_t)cap) * 4 + 28 * T + 64; }
static_assert(kCandCap <= kScanThreads && kMaxTopK <= 2 * kScanThreads, "topk_merge: ranks held in registers");

// Vector ranges per query tile: enough workgroups to fill every CU slot when
// the query tiles alone do not; otherwise the split (at most 8) that leaves
// the smallest tail in the last round of workgroups.
static void plan_ranges(ScanPlan& p, std::size_t n, std::size_t nq) {
  const long step = (long)kScanThreads * p.vec_per_thread;
  const long n_qt = (long)((nq + p.tile - 1) / p.tile);
  const long slots = 256L * (p.lds_bytes <= (std::size_t)kLdsBudget / 2 ? 2 : 1);
  const long max_by_n = std::max(1L, (long)((n + step - 1) / step));
  long R;
  if (n_qt < slots) {
    R = (slots + n_qt - 1) / n_qt;
  } else {
    R = 1;
    double best = 1e30;
    for (long c = 1; c <= 8; ++c) {
      const double cost = (double)((n_qt * c + slots - 1) / slots) / (double)c;
      if (cost < best * 0.999) { best = cost; R = c; }
    }
  }
  R = std::min(R, std::min(max_by_n, 512L));
  if (R < 1) R = 1;
  long len = (long)((n + R - 1) / R);
  len = (len + 63) / 64 * 64;
  R = std::max(1L, (long)((n + len - 1) / len));
  p.n_ranges = (int)R;
  p.range_len = (int)len;
}

// Byte-pair plan of tile T at top-k: the largest candidate buffer (at least
// 32) that still lets two workgroups share a CU, else 512 with one workgroup
// per CU (measured: k = 64 at T = 8 scans faster with 32 candidates and two
// workgroups). False when even one workgroup does not fit.
static bool pair_fit(int T, int k, ScanPlan& p) {
  const std::size_t fixed = (std::size_t)kPairTableBytes + kPairStageBytes;
  int cap = kCandCap;
  std::size_t lds = fixed + topk_lds(T, k, cap);
  if (lds > (std::size_t)kLdsBudget / 2) {
    const long room = (long)kLdsBudget / 2 - (long)fixed - (long)topk_lds(T, k, 0);
    const long c2 = room / (8L * T) / 32 * 32;
    if (c2 >= 32) {
      cap = (int)std::min<long>(c2, kCandCap);
      lds = fixed + topk_lds(T, k, cap);
    }
  }
  if (lds > (std::size_t)kLdsBudget) return false;
  p.tile = T;
  p.pair = 1;
  p.cand_cap = cap;
  p.lds_bytes = lds;
  p.lut_chunk16 = 0;
  return true;
}

// plan_ranges' split scaled by `scale` (at least 1 range, at most 512 and
// one step of vectors per range).
static void scale_ranges(ScanPlan& p, std::size_t n, float scale) {
  if (scale == 1.f) return;
  const long step = (long)kScanThreads * p.vec_per_thread;
  const long max_by_n = std::max(1L, (long)((n + step - 1) / step));
  long R = std::lround((double)p.n_ranges * scale);
  R = std::max(1L, std::min(R, std::min(max_by_n, 512L)));
  long len = (long)((n + R - 1) / R);
  len = (len + 63) / 64 * 64;
  p.n_ranges = (int)std::max(1L, (long)((n + len - 1) / len));
  p.range_len = (int)len;
}

ScanPlan plan_scan_with(int groups16, std::size_t n, std::size_t nq, int k, int tile, int vec_per_thread,
                        float range_scale) {
  (void)groups16;
  ScanPlan p{};
  if (tile != 1 && tile != 2 && tile != 4 && tile != 8) return p;
  if (!pair_fit(tile, k, p)) return ScanPlan{};
  p.vec_per_thread = vec_per_thread == 1 ? 1 : 2;
  plan_ranges(p, n, nq);
  scale_ranges(p, n, range_scale);
  return p;
}

ScanPlan plan_scan(int groups16, std::size_t n, std::size_t nq, int k, int forced_tile, bool nibble) {
  TunedScan tuned;
  if (!nibble && forced_tile == 0 && tune_get(tune_key(groups16, n, nq, k), &tuned)) {
    const ScanPlan p = plan_scan_with(groups16, n, nq, k, tuned.tile, tuned.vec_per_thread, tuned.range_scale);
    if (p.tile != 0) return p;
  }
  return plan_scan_rule(groups16, n, nq, k, forced_tile, nibble);
}

ScanPlan plan_scan_rule(int groups16, std::size_t n, std::size_t nq, int k, int forced_tile, bool nibble) {
  ScanPlan p{};
  if (!nibble) {
    // Byte-pair kernel. Measured on 1M x 1536 4-bit: one query per tile is
    // fastest up to two queries, a 4-query tile for three or four, 8 above.
    const int want = nq >= 5 ? 8 : nq >= 3 ? 4 : 1;
    for (int T : {8, 4, 2, 1}) {
      if (T > (forced_tile != 0 ? forced_tile : want)) continue;  // a forced tile is an upper bound
      if (!pair_fit(T, k, p)) continue;
      p.vec_per_thread = T == 1 ? 1 : 2;  // measured; T = 1 keeps a deeper code prefetch instead
      break;
    }
    if (p.tile == 0) return p;  // caller reports the error
    plan_ranges(p, n, nq);
    return p;
  }
  p.cand_cap = kCandCap;
  const int tiles[3] = {8, 4, 1};
  for (int ti = 0; ti < 3; ++ti) {
    const int T = tiles[ti];
    if (forced_tile != 0 && T > forced_tile) continue;
    if (forced_tile == 0) {
      if (T == 8 && nq < 8) continue;
      if (T == 4 && nq < 4) continue;
    }
    const size_t tk = topk_lds(T, k, p.cand_cap);
    if (tk + 16 * 32 * (size_t)T * 16 > (size_t)kLdsBudget) continue;  // not even one window
    const size_t lut_room = (size_t)kLdsBudget - tk - 64;
    int chunk = (int)(lut_room / (16 * 32 * (size_t)T));
    if (chunk >= groups16) {
      chunk = groups16;
    } else {
      // windows (16 chunks of 16 groups) never straddle a reload; balance the
      // chunks so the last one is not a sliver
      const int max_chunk = (chunk / 16) * 16;
      const int n_chunks = (groups16 + max_chunk - 1) / max_chunk;
      chunk = ((groups16 + n_chunks - 1) / n_chunks + 15) / 16 * 16;
    }
    if (chunk < 1) continue;
    p.tile = T;
    p.lut_chunk16 = chunk;
    p.lds_bytes = (((size_t)chunk * 16 * 32 * T + 15) & ~size_t(15)) + tk;
    break;
  }
  if (p.tile == 0) return p;  // caller reports the error
  p.vec_per_thread = p.lut_chunk16 < groups16 ? 2 : 1;  // amortize LUT reloads over more vectors
  plan_ranges(p, n, nq);
  return p;
}

std::size_t scan_lut_bytes(const ScanPlan& p, int groups16, std::size_t nq) {
  const std::size_t n_qt = (nq + p.tile - 1) / p.tile;
  if (p.pair) {
    const std::size_t per_elem = (p.tile == 2 || p.tile == 4) ? 2 : 1;
    return n_qt * (std::size_t)pair_src_blocks(p.tile, groups16) * 16 * 32 * p.tile * per_elem;
  }
  return n_qt * (std::size_t)groups16 * 16 * 32 * p.tile;
}

hipError_t launch_scan(const ScanPlan& p, const std::uint8_t* tiled, const float* alpha, std::size_t n,
                       int groups, int groups16, const std::uint8_t* luts, std::uint8_t* lut_t,
                       const float* lut_scale, const float* bias_total, std::size_t nq, int k, float* part_s,
                       std::uint32_t* part_i, int* part_cnt, hipStream_t s, int phase) {
  const bool sources = phase != 2, scan = phase != 1;
  const long n_qt = (long)((nq + p.tile - 1) / p.tile);
  const dim3 grid((unsigned)(n_qt * p.n_ranges));
  const dim3 block(kScanThreads);
  const long src_elems = n_qt * (long)groups16 * 16 * 32 * p.tile;
  if (p.pair) {
    const long pair_elems = n_qt * pair_src_blocks(p.tile, groups16) * 16 * 32 * p.tile;
#define VC_LAUNCH_PAIR(TT, VV)                                                                            \
  do {                                                                                                    \
    if (sources)                                                                                          \
      hipLaunchKernelGGL(pair_source_kernel<TT>, dim3(grid_for(pair_elems, 256)), dim3(256), 0, s, luts,  \
                         (long)nq, groups, groups16, lut_t);                                              \
    if (scan) {                                                                                           \
      (void)hipFuncSetAttribute((const void*)scan_pair_kernel<TT, VV>,                                     \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)pair_dyn);               \
      hipLaunchKernelGGL((scan_pair_kernel<TT, VV>), grid, block, pair_dyn, s, tiled, alpha, (long)n,       \
                         groups16, lut_t, lut_scale, bias_total, (long)nq, k, p.cand_cap, p.n_ranges,     \
                         (long)p.range_len, part_s, part_i, part_cnt);                                    \
    }                                                                                                     \
  } while (0)
    const size_t pair_dyn = p.lds_bytes - kPairTableBytes - kPairStageBytes;  // table and stage are static LDS
    const bool v2 = p.vec_per_thread == 2;
    switch (p.tile) {
      case 8: if (v2) VC_LAUNCH_PAIR(8, 2); else VC_LAUNCH_PAIR(8, 1); break;
      case 4: if (v2) VC_LAUNCH_PAIR(4, 2); else VC_LAUNCH_PAIR(4, 1); break;
      case 2: if (v2) VC_LAUNCH_PAIR(2, 2); else VC_LAUNCH_PAIR(2, 1); break;
      default: if (v2) VC_LAUNCH_PAIR(1, 2); else VC_LAUNCH_PAIR(1, 1); break;
    }
#undef VC_LAUNCH_PAIR
    return hipGetLastError();
  }
  if (sources)
    hipLaunchKernelGGL(interleave_luts_kernel, dim3(grid_for(src_elems, 256)), dim3(256),
                       0, s, luts, (long)nq, groups, groups16, p.tile, lut_t);
  if (!scan) return hipGetLastError();
#define VC_LAUNCH_SCAN(TT, VV)                                                                          \
  do {                                                                                                  \
    (void)hipFuncSetAttribute((const void*)scan_topk_kernel<TT, VV>,                                     \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds_bytes);            \
    hipLaunchKernelGGL((scan_topk_kernel<TT, VV>), grid, block, p.lds_bytes, s, tiled, alpha, (long)n,    \
                       groups16, p.lut_chunk16, lut_t, lut_scale, bias_total, (long)nq, k, p.cand_cap,  \
                       p.n_ranges, (long)p.range_len, part_s, part_i, part_cnt);                        \
  } while (0)
  if (p.vec_per_thread == 2) {
    if (p.tile == 8) VC_LAUNCH_SCAN(8, 2);
    else if (p.tile == 4) VC_LAUNCH_SCAN(4, 2);
    else VC_LAUNCH_SCAN(1, 2);
  } else {
    if (p.tile == 8) VC_LAUNCH_SCAN(8, 1);
    else if (p.tile == 4) VC_LAUNCH_SCAN(4, 1);
    else VC_LAUNCH_SCAN(1, 1);
  }
#undef VC_LAUNCH_SCAN
  return hipGetLastError();
}

hipError_t launch_merge_ranges(const float* part_s, const std::uint32_t* part_i, const int* part_cnt,
                               std::size_t nq, int n_ranges, int k, int k_out, const std::uint64_t* id_map,
                               float* out_s, std::uint64_t* out_i, hipStream_t s) {
  if (nq == 0 || k_out == 0) return hipSuccess;
  const int grid = grid_for((long)nq, 1, 65536);
  const size_t staged = (size_t)n_ranges * k * 8 + (size_t)n_ranges * 4;
  if (n_ranges > 1 && staged <= (size_t)kMergeStageBytes) {
    hipLaunchKernelGGL(merge_ranges_lds_kernel, dim3(grid), dim3(nq <= 512 ? 1024 : 256), staged, s, part_s, part_i, part_cnt,
                       (long)nq, n_ranges, k, k_out, id_map, out_s, out_i);
  } else {
    hipLaunchKernelGGL(merge_ranges_kernel, dim3(grid), dim3(64), 0, s, part_s, part_i, part_cnt, (long)nq,
                       n_ranges, k, k_out, id_map, out_s, out_i);
  }
  return hipGetLastError();
}

// ---- measured plans ----------------------------------------------------------
namespace {
std::mutex g_tune_mu;
std::map<TuneKey, TunedScan> g_tuned;

int floor_log2(std::size_t v) {
  int r = 0;
  while (v > 1) {
    v >>= 1;
    ++r;
  }
  return r;
}
int ceil_log2(std::size_t v) {
  int r = 0;
  while ((std::size_t(1) << r) < v) ++r;
  return r;
}
}  // namespace

TuneKey tune_key(int groups16, std::size_t n, std::size_t nq, int k) {
  return TuneKey{groups16, floor_log2(std::max<std::size_t>(n, 1)), ceil_log2(std::max<std::size_t>(nq, 1)),
                 ceil_log2((std::size_t)std::max(k, 1))};
}

void tune_put(const TuneKey& key, const TunedScan& v) {
  std::lock_guard lock(g_tune_mu);
  g_tuned[key] = v;
}

bool tune_get(const TuneKey& key, TunedScan* v) {
  std::lock_guard lock(g_tune_mu);
  if (g_tuned.empty()) return false;
  auto it = g_tuned.find(key);
  if (it == g_tuned.end()) return false;
  *v = it->second;
  return true;
}

void tune_clear() {
  std::lock_guard lock(g_tune_mu);
  g_tuned.clear();
}

std::size_t tune_count() {
  std::lock_guard lock(g_tune_mu);
  return g_tuned.size();
}

std::string tune_dump() {
  std::lock_guard lock(g_tune_mu);
  std::string out = "# vectorcache-hip measured scan plans (gfx950)\n";
  char line[256];
  for (const auto& [k, v] : g_tuned) {
    std::snprintf(line, sizeof line, "scan g16=%d nlog=%d nqlog=%d kcls=%d tile=%d vpt=%d rscale=%.4g ms=%.6g\n",
                  k.g16, k.nlog, k.nqlog, k.kcls, v.tile, v.vec_per_thread, (double)v.range_scale, v.ms);
    out += line;
  }
  return out;
}

int tune_parse(const std::string& text) {
  std::vector<std::pair<TuneKey, TunedScan>> got;
  std::size_t pos = 0;
  while (pos < text.size()) {
    std::size_t end = text.find('\n', pos);
    if (end == std::string::npos) end = text.size();
    const std::string line = text.substr(pos, end - pos);
    pos = end + 1;
    if (line.rfind("scan ", 0) != 0) continue;  // comments and other kinds of entries
    TuneKey k{};
    TunedScan v{};
    double rs = 1.0;
    if (std::sscanf(line.c_str(), "scan g16=%d nlog=%d nqlog=%d kcls=%d tile=%d vpt=%d rscale=%lf ms=%lf", &k.g16,
                    &k.nlog, &k.nqlog, &k.kcls, &v.tile, &v.vec_per_thread, &rs, &v.ms) != 8)
      return -1;
    if ((v.tile != 1 && v.tile != 2 && v.tile != 4 && v.tile != 8) || (v.vec_per_thread != 1 && v.vec_per_thread != 2) ||
        !(rs > 0.0 && rs <= 64.0) || k.g16 < 1)
      return -1;
    v.range_scale = (float)rs;
    got.emplace_back(k, v);
  }
  std::lock_guard lock(g_tune_mu);
  for (const auto& [k, v] : got) g_tuned[k] = v;
  return (int)got.size();
}

// ---- many code sets in one launch ------------------------------------------
std::size_t pair_source_bytes(int tile, int groups16, std::size_t nq) {
  const std::size_t n_qt = (nq + tile - 1) / tile;
  const std::size_t per_elem = (tile == 2 || tile == 4) ? 2 : 1;
  return n_qt * (std::size_t)pair_src_blocks(tile, groups16) * 16 * 32 * tile * per_elem;
}

hipError_t launch_pair_source(int tile, const std::uint8_t* luts, std::size_t nq, int groups, int groups16,
                              std::uint8_t* out, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  const long n_qt = (long)((nq + tile - 1) / tile);
  const long elems = n_qt * pair_src_blocks(tile, groups16) * 16 * 32 * tile;
  const dim3 grid(grid_for(elems, 256));
  switch (tile) {
    case 8: hipLaunchKernelGGL(pair_source_kernel<8>, grid, dim3(256), 0, s, luts, (long)nq, groups, groups16, out); break;
    case 4: hipLaunchKernelGGL(pair_source_kernel<4>, grid, dim3(256), 0, s, luts, (long)nq, groups, groups16, out); break;
    case 2: hipLaunchKernelGGL(pair_source_kernel<2>, grid, dim3(256), 0, s, luts, (long)nq, groups, groups16, out); break;
    case 1: hipLaunchKernelGGL(pair_source_kernel<1>, grid, dim3(256), 0, s, luts, (long)nq, groups, groups16, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

std::size_t pair_fixed_lds() { return (std::size_t)kPairTableBytes + kPairStageBytes; }
std::size_t pair_topk_lds(int tile, int k, int cap) { return topk_lds(tile, k, cap); }

hipError_t prepare_multi_scan(int vec_per_thread, std::size_t dyn_lds) {
  const void* f = vec_per_thread == 2 ? (const void*)multi_scan_kernel<2> : (const void*)multi_scan_kernel<1>;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds);
}

hipError_t launch_multi_scan(int vec_per_thread, const MultiColl* colls, const MultiItem* items, long n_items,
                             std::size_t dyn_lds, int grid, hipStream_t s) {
  if (n_items == 0) return hipSuccess;
  const dim3 g((unsigned)std::max(1, grid)), b(kScanThreads);
  if (vec_per_thread == 2)
    hipLaunchKernelGGL(multi_scan_kernel<2>, g, b, dyn_lds, s, colls, items, n_items);
  else
    hipLaunchKernelGGL(multi_scan_kernel<1>, g, b, dyn_lds, s, colls, items, n_items);
  return hipGetLastError();
}

hipError_t launch_fill_value(float* x, std::size_t n, float v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_value_kernel, dim3(grid_for((long)n, 256, 65536)), dim3(256), 0, s, x, (long)n, v);
  return hipGetLastError();
}

hipError_t launch_fill_rows(float* x, std::size_t n, int dim, std::uint64_t seed, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_rows_kernel, dim3(grid_for((long)(n * dim), 256, 65536)), dim3(256), 0, s, x,
                     (long)(n * dim), seed);
  return hipGetLastError();
}

hipError_t launch_merge_results(std::size_t nq, int k_dst, float* dst_s, std::uint64_t* dst_i, int k_src,
                                const float* src_s, const std::uint64_t* src_i, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  const size_t lds = merge_layout(k_dst + k_src, k_dst).total;
  const int grid = grid_for((long)nq, 1, 65536);
  hipLaunchKernelGGL(merge_results_kernel, dim3(grid), dim3(256), lds, s, (long)nq, k_dst, dst_s, dst_i, k_src,
                     src_s, src_i);
  return hipGetLastError();
}

hipError_t launch_bucket_scores(const std::uint8_t* tiled, const float* alpha, std::size_t n, int groups,
                                int groups16, const std::uint8_t* luts, const float* lut_scale,
                                const float* bias_total, std::size_t nq, float* p_out, hipStream_t s) {
  if (n == 0 || nq == 0) return hipSuccess;
  hipLaunchKernelGGL(bucket_scores_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nq), dim3(256), 0, s,
                     tiled, alpha, (long)n, groups, groups16, luts, lut_scale, bias_total, p_out);
  return hipGetLastError();
}

std::size_t heap_lds_bytes(int k) { return (((std::size_t)k * 4 + 15) & ~std::size_t(15)) + (std::size_t)k * 8; }

hipError_t launch_heap_replay(const float* p, std::size_t n, int k, const std::uint64_t* id_map,
                              const float* offsets, std::size_t nq, float* hs, std::uint64_t* hi,
                              std::uint64_t* hsz, float* hmin, std::uint64_t* hmi, hipStream_t s) {
  if (n == 0 || nq == 0 || k == 0) return hipSuccess;
  const size_t lds = heap_lds_bytes(k);
  (void)hipFuncSetAttribute((const void*)heap_replay_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(heap_replay_kernel, dim3((unsigned)nq), dim3(64), lds, s, p, (long)n, k, id_map, offsets, hs,
                     hi, hsz, hmin, hmi);
  return hipGetLastError();
}

// score_prepared's residual post-correction on the AVX2 path (src/search/
// search.cpp:1123-1146), one workgroup per row of k (score, position): each
// score of a position v < n becomes fmaf(rs[row], rn[v], score), the row is
// ranked again by (score desc, position asc), and positions go through id_map
// (when set) as they are written.
__global__ __launch_bounds__(256) void residual_post_kernel(long rows, int k, float* __restrict__ scores,
                                                            uint64_t* __restrict__ ids,
                                                            const float* __restrict__ rn,
                                                            const float* __restrict__ rs, long n,
                                                            const uint64_t* __restrict__ id_map) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* si = reinterpret_cast<uint64_t*>(smem);
  float* ss = reinterpret_cast<float*>(si + k);
  for (long row = blockIdx.x; row < rows; row += gridDim.x) {
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
      const uint64_t v = ids[row * k + j];
      float s = scores[row * k + j];
      if (rn != nullptr && v < (uint64_t)n) s = __fmaf_rn(rs[row], rn[v], s);
      ss[j] = s;
      si[j] = v;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
      const float s = ss[j];
      const uint64_t v = si[j];
      int rank = 0;
      for (int i = 0; i < k; ++i) rank += (ss[i] > s || (ss[i] == s && si[i] < v)) ? 1 : 0;
      scores[row * k + rank] = s;
      ids[row * k + rank] = (id_map != nullptr && v < (uint64_t)n) ? id_map[v] : v;
    }
  }
}

hipError_t launch_residual_post(std::size_t rows, int k, float* scores, std::uint64_t* ids, const float* rn,
                                const float* rs, std::size_t n, const std::uint64_t* id_map, hipStream_t s) {
  if (rows == 0 || k == 0) return hipSuccess;
  const size_t lds = (size_t)k * 12;
  (void)hipFuncSetAttribute((const void*)residual_post_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(residual_post_kernel, dim3((unsigned)std::min<size_t>(rows, 65536)), dim3(256), lds, s,
                     (long)rows, k, scores, ids, rn, rs, (long)n, id_map);
  return hipGetLastError();
}

hipError_t launch_heaps_to_results(std::size_t nq, int k, const float* hs, const std::uint64_t* hi,
                                   const std::uint64_t* hsz, float* out_s, std::uint64_t* out_i, hipStream_t s) {
  if (nq == 0 || k == 0) return hipSuccess;
  const size_t lds = heap_lds_bytes(k);
  (void)hipFuncSetAttribute((const void*)heaps_to_results_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  hipLaunchKernelGGL(heaps_to_results_kernel, dim3((unsigned)std::min<size_t>(nq, 65536)), dim3(256), lds, s,
                     (long)nq, k, hs, hi, hsz, out_s, out_i);
  return hipGetLastError();
}

hipError_t launch_float_rescore(std::size_t nq, int k, float* scores, std::uint64_t* ids,
                                const std::uint64_t* counts, const float* queries, const float* corpus,
                                std::size_t n_corpus, int dim, hipStream_t s) {
  if (nq == 0 || k == 0 || dim == 0) return hipSuccess;
  const size_t lds = heap_lds_bytes(k);
  (void)hipFuncSetAttribute((const void*)float_rescore_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(float_rescore_kernel, dim3((unsigned)std::min<size_t>(nq, 65536)), dim3(256), lds, s, (long)nq,
                     k, scores, ids, counts, queries, corpus, (long)n_corpus, dim);
  return hipGetLastError();
}

hipError_t launch_tiled_to_flat(const std::uint8_t* tiled, std::size_t n, int groups, int groups16,
                                std::uint8_t* flat, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(tiled_to_flat_kernel, dim3(grid_for((long)n * groups, 256)), dim3(256), 0, s, tiled,
                     (long)n, groups, groups16, flat);
  return hipGetLastError();
}

hipError_t launch_flat_to_gpu32(const std::uint8_t* flat, std::size_t n, int groups, std::uint32_t* words,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int nbg4 = (groups + 3) / 4;
  const long total = (long)((n + 31) / 32) * nbg4 * 32;
  hipLaunchKernelGGL(flat_to_gpu32_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, flat, (long)n, groups, nbg4,
                     words);
  return hipGetLastError();
}

hipError_t launch_gpu32_to_flat(const std::uint32_t* words, std::size_t n, int groups, std::uint8_t* flat,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(gpu32_to_flat_kernel, dim3(grid_for((long)n * groups, 256)), dim3(256), 0, s, words, (long)n,
                     groups, (groups + 3) / 4, flat);
  return hipGetLastError();
}

hipError_t launch_flat_to_tiled(const std::uint8_t* flat, std::size_t n, int groups, int groups16,
                                std::uint8_t* tiled, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const long total = (long)((n + 63) / 64) * 64 * groups16 * 16;
  hipLaunchKernelGGL(flat_to_tiled_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, flat, (long)n, groups,
                     groups16, tiled);
  return hipGetLastError();
}

hipError_t launch_flat_to_perm0(const std::uint8_t* flat, std::size_t n, int groups, std::uint8_t* perm0,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  const long total = (long)((n + 191) / 192) * groups * 96;
  hipLaunchKernelGGL(flat_to_perm0_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, flat, (long)n, groups,
                     perm0);
  return hipGetLastError();
}

hipError_t launch_flat_to_seq(const std::uint8_t* flat, std::size_t n, int groups, std::uint8_t* seq,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  const long total = (long)((n + 191) / 192) * groups * 192;
  hipLaunchKernelGGL(flat_to_seq_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, flat, (long)n, groups, seq);
  return hipGetLastError();
}

hipError_t launch_perm0_to_flat(const std::uint8_t* perm0, std::size_t n, int groups, std::uint8_t* flat,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(perm0_to_flat_kernel, dim3(grid_for((long)n * groups, 256)), dim3(256), 0, s, perm0,
                     (long)n, groups, flat);
  return hipGetLastError();
}

hipError_t launch_flat_to_packed(const std::uint8_t* flat, std::size_t n, int dim, int bits,
                                 std::uint8_t* packed, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(flat_to_packed_kernel, dim3(grid_for((long)n * (dim / 8), 256)), dim3(256), 0, s, flat,
                     (long)n, dim, bits, packed);
  return hipGetLastError();
}

hipError_t launch_packed_to_flat(const std::uint8_t* packed, std::size_t n, int dim, int bits, std::uint8_t* flat,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(packed_to_flat_kernel, dim3(grid_for((long)n * (dim / 8), 256)), dim3(256), 0, s, packed,
                     (long)n, dim, bits, flat);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const void* src, const std::uint64_t* idx, std::size_t m, std::size_t row_bytes,
                              void* dst, hipStream_t s) {
  if (m == 0 || row_bytes == 0) return hipSuccess;
  if (row_bytes % 4 != 0) return hipErrorInvalidValue;
  const int words = (int)(row_bytes / 4);
  hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for((long)m * words, 256)), dim3(256), 0, s,
                     static_cast<const uint32_t*>(src), idx, (long)m, words, static_cast<uint32_t*>(dst));
  return hipGetLastError();
}

// ---- landmark index ---------------------------------------------------------
hipError_t launch_lm_yhat(const CodeTables& t, float* rows, std::size_t n, float* vnorm2, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (t.dim % 8 != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lm_yhat_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, rows, (long)n, vnorm2);
  return hipGetLastError();
}

hipError_t launch_lm_project(const float* Y, std::size_t n, int dim, const float* L, int m, float* out,
                             int stride, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (dim % 8 != 0 || stride < m) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lm_project_kernel, dim3((unsigned)((n + 31) / 32), (unsigned)((stride + 31) / 32)), dim3(256),
                     0, s, Y, (long)n, dim, L, m, out, stride);
  return hipGetLastError();
}

hipError_t launch_lm_sign_codes(const std::uint8_t* tiled4, std::size_t n, int groups16_4, int groups_s,
                                int groups16_s, std::uint8_t* tiled_s, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(lm_sign_codes_kernel, dim3(grid_for((long)n * groups_s, 256)), dim3(256), 0, s, tiled4, (long)n,
                     groups16_4, groups_s, groups16_s, tiled_s);
  return hipGetLastError();
}

hipError_t launch_sign_luts(const float* q_land, std::size_t nq, int S, std::uint8_t* luts, float* lut_scale,
                            float* bias_total, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  const int Gs = S / 8;
  const size_t lds = ((size_t)Gs * 34 + kSignLutThreads) * 4;
  if (lds > (size_t)kLdsBudget) return hipErrorInvalidValue;
  (void)hipFuncSetAttribute((const void*)sign_lut_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(sign_lut_kernel, dim3((unsigned)nq), dim3(kSignLutThreads), lds, s, q_land, S, luts, lut_scale,
                     bias_total);
  return hipGetLastError();
}

std::size_t staged_lds_bytes(int S, int E, int k) {
  return ((size_t)((E + 1) & ~1) + (size_t)((k + 1) & ~1)) * 4 + (size_t)k * 8 + (size_t)S;
}

hipError_t launch_lm_staged(const float* q_land, std::size_t nq, int S, const float* q_perp, const float* centroids16,
                            const std::uint8_t* tiled4, int groups16_4, const float* alpha, const float* rn,
                            const std::uint64_t* cand, int E, int k, float* out_s, std::uint64_t* out_i,
                            hipStream_t s) {
  if (nq == 0 || k == 0) return hipSuccess;
  const size_t lds = staged_lds_bytes(S, E, k);
  if (lds > (size_t)kLdsBudget || E < k) return hipErrorInvalidValue;
  StagedArgs a{};
  a.q_land = q_land;
  a.S = S;
  a.q_perp = q_perp;
  for (int i = 0; i < 16; ++i) a.cen[i] = centroids16[i];
  a.tiled4 = tiled4;
  a.G16 = groups16_4;
  a.alpha = alpha;
  a.rn = rn;
  a.cand = cand;
  a.E = E;
  a.k = k;
  a.out_s = out_s;
  a.out_i = out_i;
  (void)hipFuncSetAttribute((const void*)lm_staged_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(lm_staged_kernel, dim3((unsigned)nq), dim3(kStagedThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_lm_resid(const float* rows, std::size_t n, int m, int stride, const float* vnorm2, float* rn,
                           hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (stride % 4 != 0 || stride < m) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lm_resid_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rows, (long)n, m, stride,
                     vnorm2, rn);
  return hipGetLastError();
}

bool lm_query_rotates(const CodeTables& t, int dim, int dim_pad, int stride) {
  const int R = t.block / 64;
  const size_t lds = ((size_t)dim * 2 + (size_t)stride) * 4 + (size_t)(dim_pad / 8) * 8;
  return t.dim == dim && t.perm_gather != nullptr && (R == 2 || R == 4 || R == 8 || R == 16) &&
         lds <= (size_t)kLdsBudget;
}

hipError_t launch_lm_query(const CodeTables* rot, const float* q_rot, std::size_t nq, int dim, int dim_pad,
                           const std::uint32_t* codes_t, const float* lscales, int m, float* q_land, int stride,
                           float* q_perp, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  if (dim_pad % 16 != 0 || dim_pad < dim || stride < m) return hipErrorInvalidValue;
  if (rot != nullptr && !lm_query_rotates(*rot, dim, dim_pad, stride)) return hipErrorInvalidValue;
  const size_t lds = ((size_t)dim * (rot != nullptr ? 2 : 1) + (size_t)stride) * 4 + (size_t)(dim_pad / 8) * 8;
  if (lds > (size_t)kLdsBudget) return hipErrorInvalidValue;
  const CodeTables t = rot != nullptr ? *rot : CodeTables{};
  const int raw_R = rot != nullptr ? rot->block / 64 : 0;
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)lm_query_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(lm_query_kernel, dim3((unsigned)std::min<size_t>(nq, 16384)), dim3(kLmQueryThreads), lds, s, t,
                     raw_R, q_rot,
                     (long)nq, dim, dim_pad, codes_t, lscales, m, q_land, stride, q_perp);
  return hipGetLastError();
}


hipError_t launch_row_inv_norms(const float* x, std::size_t n, int D, float* inv, float* qhat, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(row_inv_norms_kernel, dim3(grid_for((long)(n + 31) / 32, 1, 16384)), dim3(256), 0, s, x, (long)n,
                     D, inv, qhat);
  return hipGetLastError();
}

std::size_t bucket_src_bytes(int groups16, std::size_t nq) {
  return nq * (std::size_t)pair_src_blocks(1, groups16) * 16 * 32;
}

hipError_t launch_bucket_scan(BucketScanParams p, const std::uint8_t* luts, int groups, std::uint8_t* src,
                              hipStream_t s) {
  if (p.nq == 0) return hipSuccess;
  const long elems = p.nq * pair_src_blocks(1, p.G16) * 16 * 32;
  hipLaunchKernelGGL(pair_source_kernel<1>, dim3(grid_for(elems, 256)), dim3(256), 0, s, luts, p.nq, groups, p.G16,
                     src);
  p.src = src;
  // LDS: 72 KiB of table and stage, the top-k state and the probe list; the
  // candidate buffer takes what is left of half a CU (two workgroups per CU)
  const size_t probe = (size_t)4 * ((p.P + 2) & ~1) + 12 * (size_t)p.P;
  const long room = (long)kLdsBudget / 2 - (long)(kPairTableBytes + kPairStageBytes) - (long)topk_lds(1, p.k, 0) -
                    (long)probe - 16;
  int cap = (int)std::min<long>(kCandCap, room / 8 / 32 * 32);
  if (cap < 64) cap = 64;  // one workgroup per CU then
  p.cap = cap;
  const size_t dyn = ((topk_lds(1, p.k, cap) + 15) & ~size_t(15)) + probe;
  if (dyn + kPairTableBytes + kPairStageBytes > (size_t)kLdsBudget) return hipErrorInvalidValue;
  (void)hipFuncSetAttribute((const void*)bucket_scan_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  hipLaunchKernelGGL(bucket_scan_kernel<1>, dim3(kBucketScanGrid), dim3(kScanThreads), dyn, s, p);
  return hipGetLastError();
}

hipError_t launch_merge_items(const float* part_s, const std::uint32_t* part_i, const int* part_cnt,
                              const int* item_off, std::size_t nq, int k, int k_out, const std::uint64_t* id_map,
                              float* out_s, std::uint64_t* out_i, std::uint64_t* counts, hipStream_t s) {
  if (nq == 0 || k_out == 0) return hipSuccess;
  hipLaunchKernelGGL(merge_items_kernel, dim3(grid_for((long)nq, 1, 65536)), dim3(64), 0, s, part_s, part_i,
                     part_cnt, item_off, (long)nq, k, k_out, id_map, out_s, out_i, counts);
  return hipGetLastError();
}

hipError_t launch_bucket_tiles(BucketTileParams p, int k_out, const std::uint64_t* id_map, float* out_s,
                               std::uint64_t* out_i, std::uint64_t* counts, hipStream_t s, hipEvent_t scan_start,
                               hipEvent_t scan_stop) {
  if (p.nq == 0 || k_out == 0) return hipSuccess;
  if (p.k < 1 || p.k > kTileMaxK || p.G > p.G16 * 16 || p.L < 1 || p.P < 1 || p.P > 512)
    return hipErrorInvalidValue;
  const long pairs = p.nq * p.P;
  // LDS: the table, the top-k state of eight queries and the slot data; the
  // candidate buffer takes what is left of half a CU (two workgroups per CU)
  const long room = (long)kLdsBudget / 2 - (long)kTileTableBytes - (long)topk_lds(kTileQueries, p.k, 0) -
                    5 * 4 * kTileQueries - 16;
  p.cap = (int)std::min<long>(std::min<long>(kCandCap, 256 - p.k), room / (8 * kTileQueries)) / 32 * 32;
  if (p.cap < 64) return hipErrorInvalidValue;
  const size_t dyn = ((topk_lds(kTileQueries, p.k, p.cap) + 15) & ~size_t(15)) + 5 * 4 * kTileQueries;
  (void)hipFuncSetAttribute((const void*)bucket_tile_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  hipLaunchKernelGGL(tile_init_kernel, dim3(grid_for(std::max<long>(p.L, p.nq), 256)), dim3(256), 0, s, p.L, p.bcnt,
                     p.nq, p.gbound);
  hipLaunchKernelGGL(tile_count_kernel, dim3(grid_for(pairs, 256)), dim3(256), 0, s, p.nq, p.P, 0, p.P, p.probe_b,
                     p.probe_n, p.bcnt, p.prank);
  hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1024), 0, s, p.L, p.bcnt, p.bstart, p.tstart, p.n_tiles);
  hipLaunchKernelGGL(tile_place_kernel, dim3(grid_for(pairs, 256)), dim3(256), 0, s, p.nq, p.P, 0, p.P, p.probe_b,
                     p.probe_n, p.prank, p.bstart, p.pair_q, p.pair_j);
  hipLaunchKernelGGL(tile_list_kernel, dim3(grid_for(p.L, 256)), dim3(256), 0, s, p.L, p.bcnt, p.bstart, p.tstart,
                     p.tile_b, p.tile_p0, p.tile_n);
  if (scan_start != nullptr) (void)hipEventRecord(scan_start, s);
  hipLaunchKernelGGL(bucket_tile_kernel, dim3(kBucketScanGrid), dim3(kTileThreads), dyn, s, p);
  if (scan_stop != nullptr) (void)hipEventRecord(scan_stop, s);
  hipLaunchKernelGGL(merge_slots_kernel, dim3(grid_for(p.nq, 1, 65536)), dim3(64), 0, s, p.part_s, p.part_i,
                     p.part_cnt, p.probe_n, p.P, p.nq, p.k, k_out, id_map, out_s, out_i, counts);
  return hipGetLastError();
}

}  // namespace vchip
