/* TEST INFRASTRUCTURE — the CPU oracle for the plain flat TurboQuant encode +
 * search path. It is the checker, never the thing measured or shipped: only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * A plain-C restatement of the reference's algorithm. Each function cites the
 * reference file:line it follows (paths relative to the reference root).
 * Built with -ffp-contract=off: every fused multiply-add that the reference's
 * compiled code performs (it builds with -mfma -ffp-contract=fast) is written
 * here as an explicit fma()/fmaf() call, so the semantics are fixed in the
 * source, not by a compiler's contraction choices.
 *
 * Pinning (what checks each part, see tests/test_oracle.py):
 *   - everything is compared bit-for-bit with the reference's own compiled
 *     functions (oracle/_ref, built from the reference sources) on seeded
 *     inputs and against the committed golden vectors in tests/golden/;
 *   - independent of the reference: ChaCha20 block function vs RFC 7539 §2.3.2
 *     (the 8-round variant shares the code path), Hadamard involution, norm
 *     preservation, codebook symmetry/monotonicity, brute-force order
 *     statistics, brute-force LUT sums and top-k by full sort.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants (reference include/vectorcache/constants.hpp:7-28) ------ */
#define OC_BLOCK 192
#define OC_CHUNK 32
#define OC_FLUSH_EVERY 256
#define OC_MAX_DIM 16384
static const float kMinInputNorm = 1e-10f;
static const double kDegenerateInnerEps = 1e-12;

void oc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---- ChaCha8 keystream (reference src/chacha8.cpp:12-85) ----------------
 * IETF layout: 4 constant words, 8 key words (little-endian seed), a 64-bit
 * block counter in words 12..13, zero nonce in 14..15; 8 rounds = 4 double
 * rounds; output = permuted + input; counter increments per block. */
typedef struct {
  uint32_t in[16];
  uint32_t out[16];
  int pos;
} oc_chacha;

static uint32_t rotl(uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }

#define QR(a, b, c, d)   \
  do {                   \
    a += b;              \
    d = rotl(d ^ a, 16); \
    c += d;              \
    b = rotl(b ^ c, 12); \
    a += b;              \
    d = rotl(d ^ a, 8);  \
    c += d;              \
    b = rotl(b ^ c, 7);  \
  } while (0)

/* Exposed for the RFC 7539 known-answer test (rounds = 20) and ChaCha8 use. */
void oc_chacha_block(const uint32_t in[16], uint32_t out[16], int rounds) {
  uint32_t x[16];
  memcpy(x, in, sizeof(x));
  for (int r = 0; r < rounds; r += 2) {
    QR(x[0], x[4], x[8], x[12]);
    QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]);
    QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]);
    QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]);
    QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

static void chacha_init(oc_chacha* g, const uint8_t seed[32]) {
  g->in[0] = 0x61707865u;
  g->in[1] = 0x3320646eu;
  g->in[2] = 0x79622d32u;
  g->in[3] = 0x6b206574u;
  for (int i = 0; i < 8; ++i) {
    const uint8_t* p = seed + 4 * i;
    g->in[4 + i] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                   ((uint32_t)p[3] << 24);
  }
  g->in[12] = g->in[13] = g->in[14] = g->in[15] = 0;
  g->pos = 16;
}

static uint32_t chacha_u32(oc_chacha* g) {
  if (g->pos >= 16) {
    oc_chacha_block(g->in, g->out, 8);
    if (++g->in[12] == 0) ++g->in[13];
    g->pos = 0;
  }
  return g->out[g->pos++];
}

static uint64_t chacha_u64(oc_chacha* g) {
  const uint64_t lo = chacha_u32(g);
  const uint64_t hi = chacha_u32(g);
  return lo | (hi << 32);
}

/* ---- Rotation tables (reference src/transform/rotation.cpp:18-35,139-160) */
/* Fixed 32-byte rotation seed: a format constant of the reference
 * (rotation.cpp:18-20); it must be identical for codes to be compatible. */
static const uint8_t kRotationSeed[32] = {164, 143, 161, 123, 88,  50,  61,  10,
                                          234, 184, 161, 204, 105, 1,   20,  184,
                                          43,  140, 200, 117, 24,  180, 247, 84,
                                          141, 68,  110, 161, 228, 223, 32,  242};

int oc_rotation_tables(size_t dim, uint32_t* perm, float* signs, float* inv_sqrt_block,
                       size_t* block) {
  if (dim == 0 || dim % 8 != 0 || dim > OC_MAX_DIM) return -1;
  oc_chacha g;
  chacha_init(&g, kRotationSeed);
  /* signs are drawn first, one u32 each, low bit set -> -1 */
  for (size_t i = 0; i < dim; ++i) signs[i] = (chacha_u32(&g) & 1u) ? -1.f : 1.f;
  /* then Fisher-Yates from the top, j = u64 mod (i+1) */
  for (size_t i = 0; i < dim; ++i) perm[i] = (uint32_t)i;
  for (size_t i = dim - 1; i >= 1; --i) {
    const size_t j = (size_t)(chacha_u64(&g) % ((uint64_t)i + 1));
    const uint32_t t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
  const size_t b = dim & (~dim + 1); /* largest power of two dividing dim */
  *block = b;
  *inv_sqrt_block = 1.f / sqrtf((float)b);
  return 0;
}

/* ---- norm (reference src/encode/encode.cpp:16-26,290) ------------------
 * 8 interleaved chains c[j % 8] += x_j^2 (fused), pairwise tree, sqrtf. */
float oc_row_norm(const float* row, size_t dim) {
  float c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t j = 0; j < dim; ++j) c[j % 8] = fmaf(row[j], row[j], c[j % 8]);
  const float s = ((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]));
  return sqrtf(s);
}

/* ---- rotation apply (reference src/transform/rotation.cpp:38-121,185-205)
 * y[i] = (x[perm[i]] * inv) * sign[i]; per block of B: butterflies with
 * len = 1, 2, ..., B/2 (a+b, a-b); then y *= 1/sqrt(B). One round. */
static void rotate_with(size_t dim, const uint32_t* perm, const float* signs, size_t block,
                        float isb, const float* src, float inv, float* dst) {
  for (size_t i = 0; i < dim; ++i) dst[i] = (src[perm[i]] * inv) * signs[i];
  for (size_t off = 0; off < dim; off += block) {
    float* y = dst + off;
    for (size_t len = 1; len < block; len <<= 1) {
      for (size_t i = 0; i < block; i += 2 * len) {
        for (size_t j = i; j < i + len; ++j) {
          const float a = y[j];
          const float b = y[j + len];
          y[j] = a + b;
          y[j + len] = a - b;
        }
      }
    }
    for (size_t i = 0; i < block; ++i) y[i] *= isb;
  }
}

typedef struct {
  size_t dim, block;
  float isb;
  uint32_t* perm;
  float* signs;
} oc_rot;

static int rot_make(oc_rot* r, size_t dim) {
  r->dim = dim;
  r->perm = (uint32_t*)malloc(dim * sizeof(uint32_t));
  r->signs = (float*)malloc(dim * sizeof(float));
  if (!r->perm || !r->signs) return -1;
  return oc_rotation_tables(dim, r->perm, r->signs, &r->isb, &r->block);
}

static void rot_free(oc_rot* r) {
  free(r->perm);
  free(r->signs);
}

int oc_rotate_scaled(size_t dim, const float* src, float inv, float* dst) {
  oc_rot r;
  if (rot_make(&r, dim) != 0) return -1;
  rotate_with(dim, r.perm, r.signs, r.block, r.isb, src, inv, dst);
  rot_free(&r);
  return 0;
}

/* rotate_batch_into (reference src/encode/encode.cpp:159-183) */
static void rotate_rows(const oc_rot* r, const float* x, size_t n, float* rotated, float* norms) {
  const size_t dim = r->dim;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    const float* src = x + (size_t)i * dim;
    const float nv = oc_row_norm(src, dim);
    norms[i] = nv;
    const float inv = (nv > kMinInputNorm) ? (1.f / nv) : 0.f;
    rotate_with(dim, r->perm, r->signs, r->block, r->isb, src, inv, rotated + (size_t)i * dim);
  }
}

/* ---- Lloyd-Max Beta((d-1)/2,(d-1)/2) codebook ---------------------------
 * (reference src/quantize/codebook.cpp:12-226). Log-gamma by the Lanczos
 * approximation with the published g = 7, n = 9 coefficients; regularized
 * incomplete beta by the modified-Lentz continued fraction (Numerical
 * Recipes betacf); conditional means by adaptive Simpson; 200 iterations,
 * tolerance 1e-12; f32 centroids, boundaries = f32 midpoints. */
static const double kLanczos[9] = {0.99999999999980993,  676.5203681218851,
                                   -1259.1392167224028,  771.32342877765313,
                                   -176.61502916214059,  12.507343278686905,
                                   -0.13857109526572012, 9.984369654078861e-6,
                                   1.5056327351493116e-7};
static const double kPi = 3.14159265358979323846;

static double lgamma_lanczos(double z) {
  if (z < 0.5) return log(kPi) - log(sin(kPi * z)) - lgamma_lanczos(1.0 - z);
  z -= 1.0;
  double acc = kLanczos[0];
  for (int i = 1; i < 9; ++i) acc += kLanczos[i] / (z + (double)i);
  const double t = z + 7.5;
  return fma(z + 0.5, log(t), 0.5 * log(2.0 * kPi)) - t + log(acc);
}

static double clampmin(double v) { return fabs(v) < 1e-30 ? 1e-30 : v; }

static double beta_contfrac(double a, double b, double x) {
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0;
  double d = 1.0 / clampmin(1.0 - qab * x / qap);
  double h = d;
  for (int m = 1; m <= 200; ++m) {
    const int m2 = 2 * m;
    double aa = (double)m * (b - (double)m) * x / ((qam + m2) * (a + m2));
    d = 1.0 / clampmin(fma(aa, d, 1.0));
    c = clampmin(1.0 + aa / c);
    h *= d * c;
    aa = -(a + (double)m) * (qab + (double)m) * x / ((a + m2) * (qap + m2));
    d = 1.0 / clampmin(fma(aa, d, 1.0));
    c = clampmin(1.0 + aa / c);
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < 3e-14) break;
  }
  return h;
}

static double inc_beta(double a, double b, double x) {
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  const double lbeta = lgamma_lanczos(a) + lgamma_lanczos(b) - lgamma_lanczos(a + b);
  const double front = exp(fma(log(x), a, log(1.0 - x) * b) - lbeta);
  if (x < (a + 1.0) / (a + b + 2.0)) return front / a * beta_contfrac(a, b, x);
  return fma(-(front / b), beta_contfrac(b, a, 1.0 - x), 1.0);
}

double oc_beta_cdf_aa(double a, double t) { return inc_beta(a, a, t); }

typedef struct {
  double a, lbeta;
} beta_pdf_t;

static double beta_pdf(const beta_pdf_t* p, double t) {
  if (t <= 0.0 || t >= 1.0) return 0.0;
  return exp(fma(p->a - 1.0, log(t), (p->a - 1.0) * log(1.0 - t)) - p->lbeta);
}

/* integrand x * pdf((x+1)/2) / 2 on [-1, 1] */
static double cond_mean_f(const beta_pdf_t* p, double x) {
  return x * beta_pdf(p, (x + 1.0) / 2.0) / 2.0;
}

static double simpson_rec(const beta_pdf_t* p, double a, double b, double fa, double fb,
                          double fm, double whole, double tol, int depth) {
  const double mid = (a + b) / 2.0;
  const double lm = (a + mid) / 2.0, rm = (mid + b) / 2.0;
  const double flm = cond_mean_f(p, lm), frm = cond_mean_f(p, rm);
  const double left = (mid - a) / 6.0 * (fma(4.0, flm, fa) + fm);
  const double right = (b - mid) / 6.0 * (fma(4.0, frm, fm) + fb);
  const double both = left + right;
  if (depth == 0 || fabs(both - whole) < 15.0 * tol) return both + (both - whole) / 15.0;
  return simpson_rec(p, a, mid, fa, fm, flm, left, tol / 2.0, depth - 1) +
         simpson_rec(p, mid, b, fm, fb, frm, right, tol / 2.0, depth - 1);
}

static double simpson(const beta_pdf_t* p, double a, double b, double tol, int depth) {
  const double mid = (a + b) / 2.0;
  const double fa = cond_mean_f(p, a), fb = cond_mean_f(p, b), fm = cond_mean_f(p, mid);
  const double whole = (b - a) / 6.0 * (fma(4.0, fm, fa) + fb);
  return simpson_rec(p, a, b, fa, fb, fm, whole, tol, depth);
}

int oc_codebook(size_t bits, size_t dim, float* boundaries, float* centroids) {
  if (bits < 2 || bits > 4 || dim < 2) return -1;
  const double a = ((double)dim - 1.0) / 2.0;
  beta_pdf_t p;
  p.a = a;
  p.lbeta = 2.0 * lgamma_lanczos(a) - lgamma_lanczos(2.0 * a);
  const size_t L = (size_t)1 << bits;
  const double sd = sqrt(2.0 * a / ((2.0 * a + 1.0) * 4.0 * a));
  const double spread = 3.0 * sd;
  double cen[16], nxt[16], edge[17];
  for (size_t i = 0; i < L; ++i)
    cen[i] = -spread + 2.0 * spread * (double)i / (double)(L - 1);
  for (int it = 0; it < 200; ++it) {
    edge[0] = -1.0;
    for (size_t i = 0; i + 1 < L; ++i) edge[i + 1] = (cen[i] + cen[i + 1]) / 2.0;
    edge[L] = 1.0;
    for (size_t i = 0; i < L; ++i) {
      const double lo = edge[i], hi = edge[i + 1];
      const double prob = inc_beta(a, a, (hi + 1.0) / 2.0) - inc_beta(a, a, (lo + 1.0) / 2.0);
      nxt[i] = (prob < 1e-15) ? cen[i] : simpson(&p, lo, hi, 1e-14, 50) / prob;
    }
    double worst = 0.0;
    for (size_t i = 0; i < L; ++i) {
      const double dlt = fabs(cen[i] - nxt[i]);
      if (dlt > worst) worst = dlt;
      cen[i] = nxt[i];
    }
    if (worst < 1e-12) break;
  }
  for (size_t i = 0; i < L; ++i) centroids[i] = (float)cen[i];
  for (size_t i = 0; i + 1 < L; ++i) boundaries[i] = (centroids[i] + centroids[i + 1]) * 0.5f;
  return 0;
}

/* ---- quantize + alpha + bit-planes (reference src/encode/encode.cpp:112-231)
 * code = #{i : y > b_i} (strict); alpha = norm / (float)inner with inner a
 * 4-chain double sum of y * reconstruction, 0 when inner <= 1e-12. Bit-plane
 * p of 8 codes is one byte, code k at bit (7 - k), plane-major per row.
 * With TQ+: compare (y + shift) * scale, reconstruct c * (1/scale) - shift. */
static float quantize_row(const float* y, size_t dim, size_t bits, const float* bnd,
                          const float* cen, const float* shift, const float* scale,
                          const float* inv_scale, int identity, float norm, uint8_t* packed) {
  double ch[4] = {0, 0, 0, 0};
  const size_t nb = ((size_t)1 << bits) - 1;
  const size_t bpp = dim / 8;
  for (size_t c = 0; c < dim / 8; ++c) {
    uint8_t codes[8];
    for (size_t k = 0; k < 8; ++k) {
      const size_t j = c * 8 + k;
      unsigned v = 0;
      if (identity) {
        for (size_t bi = 0; bi < nb; ++bi) v += (y[j] > bnd[bi]) ? 1u : 0u;
        /* the product of two floats is exact in double; fused or not agrees */
        ch[j % 4] = fma((double)y[j], (double)cen[v], ch[j % 4]);
      } else {
        const float cal = (y[j] + shift[j]) * scale[j];
        for (size_t bi = 0; bi < nb; ++bi) v += (cal > bnd[bi]) ? 1u : 0u;
        const double recon = fma((double)cen[v], (double)inv_scale[j], -(double)shift[j]);
        ch[j % 4] = fma((double)y[j], recon, ch[j % 4]);
      }
      codes[k] = (uint8_t)v;
    }
    for (size_t p = 0; p < bits; ++p) {
      unsigned byte = 0;
      for (size_t k = 0; k < 8; ++k) byte |= ((codes[k] >> p) & 1u) << (7 - k);
      packed[p * bpp + c] = (uint8_t)byte;
    }
  }
  const double inner = (ch[0] + ch[1]) + (ch[2] + ch[3]);
  return (inner > kDegenerateInnerEps) ? norm / (float)inner : 0.f;
}

int oc_encode(const float* x, size_t n, size_t dim, size_t bits, const float* shift,
              const float* scale, uint8_t* packed_out, float* alpha_out, float* rotated_out) {
  if (dim == 0 || dim % 8 != 0 || bits < 2 || bits > 4) return -1;
  oc_rot r;
  if (rot_make(&r, dim) != 0) return -1;
  float bnd[15], cen[16];
  oc_codebook(bits, dim, bnd, cen);
  float* rot = rotated_out ? rotated_out : (float*)malloc(n * dim * sizeof(float) + 1);
  float* norms = (float*)malloc(n * sizeof(float) + 1);
  float* inv_s = (float*)malloc(dim * sizeof(float));
  rotate_rows(&r, x, n, rot, norms);
  /* identity also when every shift is 0 and every scale 1 (encode.cpp:204-210) */
  int identity = 1;
  for (size_t d = 0; shift != NULL && scale != NULL && d < dim; ++d) {
    inv_s[d] = 1.f / scale[d];
    if (shift[d] != 0.f || scale[d] != 1.f) identity = 0;
  }
  const size_t bpr = bits * (dim / 8);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    alpha_out[i] = quantize_row(rot + (size_t)i * dim, dim, bits, bnd, cen, shift, scale, inv_s,
                                identity, norms[i], packed_out + (size_t)i * bpr);
  }
  if (!rotated_out) free(rot);
  free(norms);
  free(inv_s);
  rot_free(&r);
  return 0;
}

/* ---- TQ+ calibration (reference src/encode/encode.cpp:142-150,233-286,327-335)
 * Anchors p_hi = I_{(c_outer+1)/2}(a, a), p_lo = 1 - p_hi; per dimension the
 * lo_idx-th and hi_idx-th order statistics of the rotated sample, compared as
 * order-preserving u32 keys of the f32 bits. */
static uint32_t f32_key(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

static float f32_unkey(uint32_t k) {
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float x;
  memcpy(&x, &b, 4);
  return x;
}

static int cmp_u32(const void* pa, const void* pb) {
  const uint32_t a = *(const uint32_t*)pa, b = *(const uint32_t*)pb;
  return (a > b) - (a < b);
}

int oc_fit_calibration(const float* x, size_t n, size_t dim, size_t bits, float* shift_out,
                       float* scale_out) {
  if (n < 2 || dim == 0 || dim % 8 != 0 || bits < 2 || bits > 4) return -1;
  oc_rot r;
  if (rot_make(&r, dim) != 0) return -1;
  float bnd[15], cen[16];
  oc_codebook(bits, dim, bnd, cen);
  float* rot = (float*)malloc(n * dim * sizeof(float));
  float* norms = (float*)malloc(n * sizeof(float));
  rotate_rows(&r, x, n, rot, norms);
  const double a = ((double)dim - 1.0) / 2.0;
  float c_outer = 0.f;
  for (size_t i = 0; i < ((size_t)1 << bits); ++i)
    if (fabsf(cen[i]) > c_outer) c_outer = fabsf(cen[i]);
  const double p_hi = inc_beta(a, a, ((double)c_outer + 1.0) / 2.0);
  const double p_lo = 1.0 - p_hi;
  const float qc_lo = -c_outer, qc_hi = c_outer;
  const float qc_span = qc_hi - qc_lo;
  const size_t lo_idx = (size_t)((double)n * p_lo);
  size_t hi_idx = (size_t)((double)n * p_hi);
  if (hi_idx > n - 1) hi_idx = n - 1;
  if (hi_idx < lo_idx + 1) hi_idx = lo_idx + 1;
#pragma omp parallel
  {
    uint32_t* col = (uint32_t*)malloc(n * sizeof(uint32_t));
#pragma omp for schedule(static)
    for (long long d = 0; d < (long long)dim; ++d) {
      for (size_t i = 0; i < n; ++i) col[i] = f32_key(rot[i * dim + (size_t)d]);
      qsort(col, n, sizeof(uint32_t), cmp_u32);
      const float q_lo = f32_unkey(col[lo_idx]);
      const float q_hi = f32_unkey(col[hi_idx]);
      const float span = q_hi - q_lo;
      shift_out[d] = 0.f;
      scale_out[d] = 1.f;
      if (!(span > 1e-12f) || !(qc_span > 0.f)) continue;
      const float s = qc_span / span;
      scale_out[d] = s;
      shift_out[d] = qc_lo / s - q_lo;
    }
    free(col);
  }
  free(rot);
  free(norms);
  rot_free(&r);
  return 0;
}

/* ---- bit-planes -> row-major byte groups (reference src/pack/pack.cpp:70-131)
 * b=4 and b=3: byte g = code[2g] << 4 | code[2g+1]  (3-bit codes in nibbles);
 * b=2: byte g = c[4g] << 6 | c[4g+1] << 4 | c[4g+2] << 2 | c[4g+3]. */
static size_t groups_of(size_t bits, size_t dim) { return dim / (8 / bits); }

int oc_codes_flat(const uint8_t* packed, size_t n, size_t bits, size_t dim, uint8_t* out) {
  if (dim % 8 != 0 || bits < 2 || bits > 4) return -1;
  const size_t bpp = dim / 8, bpr = bits * bpp, G = groups_of(bits, dim);
  const size_t cpb = 8 / bits;            /* codes per byte: 2, 2, 4 */
  const size_t field = (bits == 3) ? 4 : bits;
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; ++v) {
    const uint8_t* row = packed + (size_t)v * bpr;
    uint8_t* o = out + (size_t)v * G;
    memset(o, 0, G);
    for (size_t j = 0; j < dim; ++j) {
      unsigned code = 0;
      for (size_t p = 0; p < bits; ++p) code |= ((row[p * bpp + j / 8] >> (7 - j % 8)) & 1u) << p;
      const size_t g = j / cpb;
      const size_t shift = (cpb - 1 - j % cpb) * field;
      o[g] = (uint8_t)(o[g] | (code << shift));
    }
  }
  return 0;
}

/* ---- FastScan 192x1 PERM0 layout (reference src/pack/pack.cpp:157-186,
 * include/vectorcache/pack/pack.hpp:14) ------------------------------------
 * For block bk, group g, 32-lane chunk ch, j < 16 with va = 192 bk + 32 ch +
 * P[j], vb = va + 16: byte j = hi(va) | hi(vb) << 4, byte 16 + j = lo(va) |
 * lo(vb) << 4. Vectors >= n read as zero. */
static const unsigned kPerm0[16] = {0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15};

int oc_pack_perm0(const uint8_t* flat, size_t n, size_t G, uint8_t* out) {
  const size_t nb = (n + OC_BLOCK - 1) / OC_BLOCK;
#pragma omp parallel for schedule(static)
  for (long long bk = 0; bk < (long long)nb; ++bk) {
    for (size_t g = 0; g < G; ++g) {
      uint8_t* o = out + ((size_t)bk * G + g) * OC_BLOCK;
      for (size_t ch = 0; ch < OC_BLOCK / OC_CHUNK; ++ch) {
        for (size_t j = 0; j < 16; ++j) {
          const size_t va = (size_t)bk * OC_BLOCK + ch * OC_CHUNK + kPerm0[j];
          const size_t vb = va + 16;
          const unsigned a = va < n ? flat[va * G + g] : 0u;
          const unsigned b = vb < n ? flat[vb * G + g] : 0u;
          o[ch * OC_CHUNK + j] = (uint8_t)((a >> 4) | ((b >> 4) << 4));
          o[ch * OC_CHUNK + 16 + j] = (uint8_t)((a & 15u) | ((b & 15u) << 4));
        }
      }
    }
  }
  return 0;
}

/* inverse of oc_pack_perm0 (reference read_code, src/pack/pack.cpp:189-203,403-420) */
int oc_unpack_perm0(const uint8_t* blocked, size_t n, size_t G, uint8_t* flat_out) {
  static const unsigned kInv[16] = {0, 2, 4, 6, 8, 10, 12, 14, 1, 3, 5, 7, 9, 11, 13, 15};
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; ++v) {
    const size_t bk = (size_t)v / OC_BLOCK, lane = (size_t)v % OC_BLOCK;
    const size_t ch = lane / OC_CHUNK, l = lane % OC_CHUNK;
    const size_t j = kInv[l & 15];
    for (size_t g = 0; g < G; ++g) {
      const uint8_t* t = blocked + (bk * G + g) * OC_BLOCK + ch * OC_CHUNK;
      const unsigned hp = t[j], lp = t[16 + j];
      const unsigned hi = (l < 16) ? (hp & 15u) : (hp >> 4);
      const unsigned lo = (l < 16) ? (lp & 15u) : (lp >> 4);
      flat_out[(size_t)v * G + g] = (uint8_t)((hi << 4) | lo);
    }
  }
  return 0;
}

/* The reference CUDA index's code words (src/gpu/kernels.cuh:18-21,
 * src/gpu/kernels.cu:111-133, zero-filled at allocation :587-595): for vector
 * vi, tile = vi / 32, lane = vi % 32, word (tile * ceil(G/4) + g4) * 32 + lane
 * = bytes[0] | bytes[1] << 8 | bytes[2] << 16 | bytes[3] << 24 of groups
 * 4 g4 .. 4 g4 + 3 (0 past G). */
int oc_flat_to_gpu32(const uint8_t* flat, size_t n, size_t G, uint32_t* words) {
  const size_t nbg4 = (G + 3) / 4, n_tiles = (n + 31) / 32;
  memset(words, 0, n_tiles * nbg4 * 32 * sizeof(uint32_t));
  for (size_t vi = 0; vi < n; ++vi) {
    const size_t tile = vi / 32, lane = vi - tile * 32;
    for (size_t g4 = 0; g4 < nbg4; ++g4) {
      uint32_t bytes[4] = {0, 0, 0, 0};
      for (size_t t = 0; t < 4; ++t) {
        const size_t g = g4 * 4 + t;
        if (g >= G) break;
        bytes[t] = flat[vi * G + g];
      }
      words[(tile * nbg4 + g4) * 32 + lane] = bytes[0] | (bytes[1] << 8) | (bytes[2] << 16) | (bytes[3] << 24);
    }
  }
  return 0;
}

int oc_gpu32_to_flat(const uint32_t* words, size_t n, size_t G, uint8_t* flat_out) {
  const size_t nbg4 = (G + 3) / 4;
  for (size_t vi = 0; vi < n; ++vi)
    for (size_t g = 0; g < G; ++g)
      flat_out[vi * G + g] = (uint8_t)(words[((vi / 32) * nbg4 + g / 4) * 32 + vi % 32] >> (8 * (g % 4)));
  return 0;
}

/* sequential 192x1 layout (reference src/pack/pack.cpp:133-154) */
int oc_pack_seq(const uint8_t* flat, size_t n, size_t G, uint8_t* out) {
  const size_t nb = (n + OC_BLOCK - 1) / OC_BLOCK;
#pragma omp parallel for schedule(static)
  for (long long bk = 0; bk < (long long)nb; ++bk) {
    for (size_t g = 0; g < G; ++g) {
      for (size_t lane = 0; lane < OC_BLOCK; ++lane) {
        const size_t v = (size_t)bk * OC_BLOCK + lane;
        out[((size_t)bk * G + g) * OC_BLOCK + lane] = v < n ? flat[v * G + g] : 0;
      }
    }
  }
  return 0;
}

/* ---- per-query LUT (reference src/search/search.cpp:149-219) -------------
 * For group g the first 16 entries are indexed by the byte's high nibble (the
 * group's first dims), the second 16 by the low nibble. Entry = sum over the
 * nibble's codes of q_d * centroid (sequential from 0.f; for b=2 the second
 * product is fused into the running sum). bias = sequential float sum of
 * (min_hi + min_lo); scale = max span / 127; u8 = clamp(round((v - min) /
 * scale), 0, 127) computed as (v - min) * (1/scale). */
static void build_lut(const float* q, size_t dim, size_t bits, const float* cen, uint8_t* lut,
                      float* scale_out, float* bias_out, float* fv) {
  const size_t cpb = 8 / bits, cpn = cpb / 2, G = dim / cpb;
  const unsigned mask = (1u << bits) - 1u;
  float span = 0.f, bias = 0.f;
  for (size_t g = 0; g < G; ++g) {
    float mn[2], mx[2];
    for (int half = 0; half < 2; ++half) {
      mn[half] = FLT_MAX;
      mx[half] = -FLT_MAX;
      const float* qq = q + g * cpb + (size_t)half * cpn;
      for (unsigned nib = 0; nib < 16; ++nib) {
        float s = 0.f;
        for (size_t c = 0; c < cpn; ++c) {
          const unsigned code = (nib >> ((cpn - 1 - c) * bits)) & mask;
          s = fmaf(qq[c], cen[code], s);
        }
        fv[g * 32 + (size_t)half * 16 + nib] = s;
        if (s < mn[half]) mn[half] = s;
        if (s > mx[half]) mx[half] = s;
      }
    }
    bias += mn[0] + mn[1];
    if (mx[0] - mn[0] > span) span = mx[0] - mn[0];
    if (mx[1] - mn[1] > span) span = mx[1] - mn[1];
    fv[G * 32 + 2 * g] = mn[0];
    fv[G * 32 + 2 * g + 1] = mn[1];
  }
  float scale = (span > 0.f) ? span / 127.f : 1.f;
  float inv = 1.f;
  if (scale >= FLT_MIN)
    inv = 1.f / scale;
  else
    scale = 1.f;
  for (size_t g = 0; g < G; ++g) {
    for (int half = 0; half < 2; ++half) {
      const float mnv = fv[G * 32 + 2 * g + (size_t)half];
      for (unsigned nib = 0; nib < 16; ++nib) {
        float r = roundf((fv[g * 32 + (size_t)half * 16 + nib] - mnv) * inv);
        if (r < 0.f) r = 0.f;
        if (r > 127.f) r = 127.f;
        lut[g * 32 + (size_t)half * 16 + nib] = (uint8_t)r;
      }
    }
  }
  *scale_out = scale;
  *bias_out = bias;
}

/* prepare_queries (reference src/search/search.cpp:762-860, AVX2 backend):
 * norm + rotation as for encode; with TQ+ the LUT is built from q / scale and
 * bias_corr = (float) -sum_d (double) q_d * shift_d. */
int oc_prepare_queries(const float* q, size_t nq, size_t dim, size_t bits, const float* shift,
                       const float* scale, float* q_rot_out, float* bias_corr_out,
                       uint8_t* luts_out, float* lut_scale_out, float* lut_bias_out) {
  if (dim == 0 || dim % 8 != 0 || bits < 2 || bits > 4) return -1;
  oc_rot r;
  if (rot_make(&r, dim) != 0) return -1;
  float bnd[15], cen[16];
  oc_codebook(bits, dim, bnd, cen);
  const size_t G = groups_of(bits, dim);
  const int tq = (shift != NULL && scale != NULL);
#pragma omp parallel
  {
    float* row = (float*)malloc(dim * sizeof(float));
    float* fv = (float*)malloc((G * 34) * sizeof(float));
#pragma omp for schedule(static)
    for (long long i = 0; i < (long long)nq; ++i) {
      const float* src = q + (size_t)i * dim;
      const float nv = oc_row_norm(src, dim);
      const float inv = (nv > kMinInputNorm) ? (1.f / nv) : 0.f;
      rotate_with(dim, r.perm, r.signs, r.block, r.isb, src, inv, row);
      float* dst = q_rot_out + (size_t)i * dim;
      float bc = 0.f;
      if (tq) {
        double acc = 0.0;
        for (size_t d = 0; d < dim; ++d) {
          dst[d] = row[d] / scale[d];
          acc = fma(-(double)row[d], (double)shift[d], acc);
        }
        bc = (float)acc;
      } else {
        memcpy(dst, row, dim * sizeof(float));
      }
      bias_corr_out[i] = bc;
      build_lut(dst, dim, bits, cen, luts_out + (size_t)i * G * 32, &lut_scale_out[i],
                &lut_bias_out[i], fv);
    }
    free(row);
    free(fv);
  }
  rot_free(&r);
  return 0;
}

/* ---- the scan (reference src/search/search_avx2.cpp:95-222) --------------
 * For vector v and 256-group window w the exact integer sum
 *   A_w = sum_g lut[g][hi(byte)] + lut[g][16 + lo(byte)]   (<= 65024)
 * then fa = bias + bias_corr; fa = fmaf(scale, (float) A_w, fa) per window in
 * order; score = fa * alpha_v. */
int oc_window_sums(const uint8_t* lut, const uint8_t* flat, size_t n, size_t G,
                   uint32_t* sums_out) {
  const size_t W = (G + OC_FLUSH_EVERY - 1) / OC_FLUSH_EVERY;
  for (size_t v = 0; v < n; ++v) {
    for (size_t w = 0; w < W; ++w) {
      uint32_t acc = 0;
      const size_t g1 = (w + 1) * OC_FLUSH_EVERY < G ? (w + 1) * OC_FLUSH_EVERY : G;
      for (size_t g = w * OC_FLUSH_EVERY; g < g1; ++g) {
        const unsigned byte = flat[v * G + g];
        acc += lut[g * 32 + (byte >> 4)] + lut[g * 32 + 16 + (byte & 15u)];
      }
      sums_out[v * W + w] = acc;
    }
  }
  return 0;
}

typedef struct {
  float s;
  uint64_t id;
} oc_cand;

/* (score desc, id asc): the reference's merge/sort order
 * (src/search/search.cpp:739-752, include/vectorcache/search/topk_heap.hpp:19-22) */
static int cand_cmp(const void* pa, const void* pb) {
  const oc_cand* a = (const oc_cand*)pa;
  const oc_cand* b = (const oc_cand*)pb;
  if (a->s != b->s) return (a->s > b->s) ? -1 : 1;
  return (a->id < b->id) ? -1 : (a->id > b->id);
}

/* Top-k per query over all n vectors (reference src/search/search.cpp:
 * 600-758,1072-1154: with an id map the ranking breaks ties on the position
 * and maps ids afterwards, remap_ids :324-332). A score is admitted only if it is
 * > -inf (the reference's heaps start at -inf with strict '>' admission, so
 * NaN and -inf never enter). k_eff = min(k, n); rows with fewer admitted
 * candidates are padded with (0.f, 0). -0.f scores are reported as +0.f. */
int oc_scan_topk(const uint8_t* luts, const float* lut_scale, const float* lut_bias_total,
                 size_t nq, const uint8_t* flat, size_t n, size_t G, const float* alphas,
                 const uint64_t* id_map, size_t k, float* scores_out, uint64_t* ids_out,
                 size_t* k_out) {
  const size_t keff = k < n ? k : n;
  *k_out = (n == 0) ? 0 : keff;
  if (n == 0 || nq == 0 || keff == 0) return 0;
  const size_t W = (G + OC_FLUSH_EVERY - 1) / OC_FLUSH_EVERY;
#pragma omp parallel
  {
    oc_cand* c = (oc_cand*)malloc(n * sizeof(oc_cand));
#pragma omp for schedule(dynamic)
    for (long long qi = 0; qi < (long long)nq; ++qi) {
      const uint8_t* lut = luts + (size_t)qi * G * 32;
      size_t cnt = 0;
      for (size_t v = 0; v < n; ++v) {
        float fa = lut_bias_total[qi];
        const uint8_t* row = flat + v * G;
        for (size_t w = 0; w < W; ++w) {
          uint32_t acc = 0;
          const size_t g1 = (w + 1) * OC_FLUSH_EVERY < G ? (w + 1) * OC_FLUSH_EVERY : G;
          for (size_t g = w * OC_FLUSH_EVERY; g < g1; ++g) {
            const unsigned byte = row[g];
            acc += lut[g * 32 + (byte >> 4)] + lut[g * 32 + 16 + (byte & 15u)];
          }
          fa = fmaf(lut_scale[qi], (float)acc, fa);
        }
        const float s = fa * alphas[v] + 0.f;
        if (s > -INFINITY) {
          c[cnt].s = s;
          c[cnt].id = (uint64_t)v; /* ties break on the position, mapped after */
          ++cnt;
        }
      }
      qsort(c, cnt, sizeof(oc_cand), cand_cmp);
      for (size_t j = 0; j < keff; ++j) {
        scores_out[(size_t)qi * keff + j] = j < cnt ? c[j].s : 0.f;
        ids_out[(size_t)qi * keff + j] = j < cnt ? (id_map ? id_map[c[j].id] : c[j].id) : 0;
      }
    }
    free(c);
  }
  return 0;
}

/* search_flat (reference src/search/search.cpp:1293-1302) over PERM0 codes. */
int oc_search_flat(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                   const uint8_t* blocked_perm0, const float* alphas, size_t n,
                   const float* shift, const float* scale, float* scores_out, uint64_t* ids_out,
                   size_t* k_out) {
  if (dim == 0 || dim % 8 != 0 || bits < 2 || bits > 4) return -1;
  const size_t G = groups_of(bits, dim);
  float* q_rot = (float*)malloc(nq * dim * sizeof(float) + 1);
  float* bc = (float*)malloc(nq * sizeof(float) + 1);
  float* ls = (float*)malloc(nq * sizeof(float) + 1);
  float* lb = (float*)malloc(nq * sizeof(float) + 1);
  uint8_t* luts = (uint8_t*)malloc(nq * G * 32 + 1);
  uint8_t* flat = (uint8_t*)malloc(n * G + 1);
  int rc = oc_prepare_queries(q, nq, dim, bits, shift, scale, q_rot, bc, luts, ls, lb);
  if (rc == 0) {
    for (size_t i = 0; i < nq; ++i) lb[i] = lb[i] + bc[i];
    oc_unpack_perm0(blocked_perm0, n, G, flat);
    rc = oc_scan_topk(luts, ls, lb, nq, flat, n, G, alphas, NULL, k, scores_out, ids_out, k_out);
  }
  free(q_rot);
  free(bc);
  free(ls);
  free(lb);
  free(luts);
  free(flat);
  return rc;
}

/* score_prepared on the AVX2 path with every argument (reference
 * src/search/search.cpp:1072-1154): the listed queries (query_indices, or all
 * nq; each < nq, else -2), top k_eff = min(k, n) by local position; then,
 * when both residual_norms (n) and residual_scales (nq) are given, each kept
 * score s of vector vi < n becomes fmaf(residual_scales[query],
 * residual_norms[vi], s) (the reference's `+= rs * resid[vi]` under its
 * contraction flags, :1123-1131) and the row is sorted again by (score desc,
 * position asc) (:1132-1146); ids go through id_map last (remap_ids). */
int oc_score_prepared(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                      const uint8_t* blocked_perm0, const float* alphas, size_t n, const float* shift,
                      const float* scale, const uint64_t* id_map, const size_t* query_indices,
                      size_t n_indices, const float* residual_norms, const float* residual_scales,
                      float* scores_out, uint64_t* ids_out, size_t* k_out) {
  if (dim == 0 || dim % 8 != 0 || bits < 2 || bits > 4) return -1;
  if (query_indices != NULL)
    for (size_t i = 0; i < n_indices; ++i)
      if (query_indices[i] >= nq) return -2;
  const size_t nq_out = query_indices != NULL ? n_indices : nq;
  *k_out = 0;
  if (n == 0 || nq_out == 0) return 0;
  const size_t G = groups_of(bits, dim);
  float* q_rot = (float*)malloc(nq * dim * sizeof(float) + 1);
  float* bc = (float*)malloc(nq * sizeof(float) + 1);
  float* ls = (float*)malloc(nq * sizeof(float) + 1);
  float* lb = (float*)malloc(nq * sizeof(float) + 1);
  uint8_t* luts = (uint8_t*)malloc(nq * G * 32 + 1);
  float* ls_o = (float*)malloc(nq_out * sizeof(float) + 1);
  float* lb_o = (float*)malloc(nq_out * sizeof(float) + 1);
  uint8_t* luts_o = (uint8_t*)malloc(nq_out * G * 32 + 1);
  uint8_t* flat = (uint8_t*)malloc(n * G + 1);
  int rc = oc_prepare_queries(q, nq, dim, bits, shift, scale, q_rot, bc, luts, ls, lb);
  if (rc == 0) {
    for (size_t l = 0; l < nq_out; ++l) {
      const size_t g = query_indices != NULL ? query_indices[l] : l;
      ls_o[l] = ls[g];
      lb_o[l] = lb[g] + bc[g];
      memcpy(luts_o + l * G * 32, luts + g * G * 32, G * 32);
    }
    oc_unpack_perm0(blocked_perm0, n, G, flat);
    rc = oc_scan_topk(luts_o, ls_o, lb_o, nq_out, flat, n, G, alphas, NULL, k, scores_out, ids_out, k_out);
  }
  const size_t ke = *k_out;
  if (rc == 0 && residual_norms != NULL && residual_scales != NULL && ke > 0) {
    oc_cand* row = (oc_cand*)malloc(ke * sizeof(oc_cand));
    for (size_t l = 0; l < nq_out; ++l) {
      const float rs = residual_scales[query_indices != NULL ? query_indices[l] : l];
      for (size_t j = 0; j < ke; ++j) {
        const uint64_t vi = ids_out[l * ke + j];
        float s = scores_out[l * ke + j];
        if (vi < n) s = fmaf(rs, residual_norms[vi], s);
        row[j].s = s;
        row[j].id = vi;
      }
      qsort(row, ke, sizeof(oc_cand), cand_cmp);
      for (size_t j = 0; j < ke; ++j) {
        scores_out[l * ke + j] = row[j].s;
        ids_out[l * ke + j] = row[j].id;
      }
    }
    free(row);
  }
  if (rc == 0 && id_map != NULL)
    for (size_t i = 0; i < nq_out * ke; ++i)
      if (ids_out[i] < n) ids_out[i] = id_map[ids_out[i]];
  free(q_rot);
  free(bc);
  free(ls);
  free(lb);
  free(luts);
  free(ls_o);
  free(lb_o);
  free(luts_o);
  free(flat);
  return rc;
}

/* merge_search_results (reference src/search/search.cpp:1551-1600): union of
 * both rows, duplicates by id keep the higher score, top k_dst by (score desc,
 * id asc), short rows padded with (0.f, 0). */
static int cand_id_cmp(const void* pa, const void* pb) {
  const oc_cand* a = (const oc_cand*)pa;
  const oc_cand* b = (const oc_cand*)pb;
  if (a->id != b->id) return (a->id < b->id) ? -1 : 1;
  if (a->s != b->s) return (a->s > b->s) ? -1 : 1;
  return 0;
}

int oc_merge_results(size_t nq, size_t k_dst, float* dst_scores, uint64_t* dst_ids, size_t k_src,
                     const float* src_scores, const uint64_t* src_ids) {
  if (k_src == 0 || k_dst == 0) return 0;
  oc_cand* c = (oc_cand*)malloc((k_dst + k_src) * sizeof(oc_cand));
  for (size_t qi = 0; qi < nq; ++qi) {
    size_t m = 0;
    for (size_t j = 0; j < k_dst; ++j) {
      c[m].s = dst_scores[qi * k_dst + j];
      c[m++].id = dst_ids[qi * k_dst + j];
    }
    for (size_t j = 0; j < k_src; ++j) {
      c[m].s = src_scores[qi * k_src + j];
      c[m++].id = src_ids[qi * k_src + j];
    }
    qsort(c, m, sizeof(oc_cand), cand_id_cmp);
    size_t u = 0;
    for (size_t j = 0; j < m; ++j)
      if (u == 0 || c[u - 1].id != c[j].id) c[u++] = c[j];
    qsort(c, u, sizeof(oc_cand), cand_cmp);
    for (size_t j = 0; j < k_dst; ++j) {
      dst_scores[qi * k_dst + j] = j < u ? c[j].s : 0.f;
      dst_ids[qi * k_dst + j] = j < u ? c[j].id : 0;
    }
  }
  free(c);
  return 0;
}

/* ---- persistent per-query heaps ------------------------------------------
 * topk_worse / rescan_min / topk_sift_down / topk_heapify /
 * heap_push_or_replace (reference include/vectorcache/search/topk_heap.hpp:
 * 19-66,113-157): the k < 128 "rescan" heap and the k >= 128 binary heap.
 * Admission is strict (score > heap_min); while the heap fills, heap_min
 * keeps the caller's value. The buffered collector (heap_mi == SIZE_MAX,
 * :96-111) keeps its best k/2 through std::nth_element, whose order is the
 * standard library's; it is not restated (OC_ERR_UNSUPPORTED). */
#define OC_ERR_UNSUPPORTED (-3)

static int oc_worse(float sa, uint64_t ia, float sb, uint64_t ib) {
  if (sa != sb) return sa < sb;
  return ia > ib;
}

static void oc_rescan_min(const float* hs, const uint64_t* hi, size_t k, float* hmin, size_t* hmi) {
  size_t mi = 0;
  for (size_t h = 1; h < k; ++h)
    if (oc_worse(hs[h], hi[h], hs[mi], hi[mi])) mi = h;
  *hmin = hs[mi];
  *hmi = mi;
}

static void oc_sift_down(float* hs, uint64_t* hi, size_t k, size_t i) {
  for (;;) {
    const size_t l = 2 * i + 1, r = 2 * i + 2;
    size_t w = i;
    if (l < k && oc_worse(hs[l], hi[l], hs[w], hi[w])) w = l;
    if (r < k && oc_worse(hs[r], hi[r], hs[w], hi[w])) w = r;
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

static void oc_heap_push(float* hs, uint64_t* hi, size_t* hsz, float* hmin, size_t* hmi, size_t k,
                         float s, uint64_t id) {
  if (k < 128) {
    if (*hsz < k) {
      if (!(s > *hmin)) return;
      hs[*hsz] = s;
      hi[*hsz] = id;
      ++*hsz;
      if (*hsz == k) oc_rescan_min(hs, hi, k, hmin, hmi);
    } else if (s > *hmin) {
      hs[*hmi] = s;
      hi[*hmi] = id;
      oc_rescan_min(hs, hi, k, hmin, hmi);
    }
    return;
  }
  if (*hsz < k) {
    if (!(s > *hmin)) return;
    hs[*hsz] = s;
    hi[*hsz] = id;
    ++*hsz;
    if (*hsz == k) {
      for (size_t i = k / 2; i-- > 0;) oc_sift_down(hs, hi, k, i);
      *hmin = hs[0];
      *hmi = 0;
    }
  } else if (s > *hmin) {
    hs[0] = s;
    hi[0] = id;
    oc_sift_down(hs, hi, k, 0);
    *hmin = hs[0];
    *hmi = 0;
  }
}

/* score_prepared_into (reference src/search/search.cpp:1156-1267) on the AVX2
 * PERM0 path (src/search/search_avx2.cpp:72-225): for each listed query, the
 * bucket's vectors in order, by 32-vector sub-tiles (kFlushChunk) with the
 * same window chain as oc_scan_topk; p = fa * alpha. A sub-tile that is whole
 * and starts with a full heap (:163-206) scores p, plus the offset only when
 * the offset is non-zero; every other sub-tile (:208-230) scores p + offset.
 * Ids are id_map[v] (or v). Heaps [nq][k] and their state [nq] are updated in
 * place; query indices must be < nq. */
int oc_score_into(const float* q, size_t nq, size_t dim, size_t k, size_t bits,
                  const uint8_t* blocked_perm0, const float* alphas, size_t n,
                  const uint64_t* id_map, const float* shift, const float* scale,
                  const size_t* qidx, size_t n_qidx, float* heap_s, uint64_t* heap_i,
                  size_t* heap_sz, float* heap_min, size_t* heap_mi, const float* offsets) {
  if (dim == 0 || dim % 8 != 0 || bits < 2 || bits > 4 || n_qidx == 0) return -1;
  for (size_t i = 0; i < n_qidx; ++i)
    if (qidx[i] >= nq) return -2;
  if (n == 0 || k == 0) return 0;
  for (size_t i = 0; i < n_qidx; ++i)
    if (heap_mi[qidx[i]] == (size_t)-1) return OC_ERR_UNSUPPORTED;
  const size_t G = groups_of(bits, dim);
  const size_t W = (G + OC_FLUSH_EVERY - 1) / OC_FLUSH_EVERY;
  float* q_rot = (float*)malloc(nq * dim * sizeof(float) + 1);
  float* bc = (float*)malloc(nq * sizeof(float) + 1);
  float* ls = (float*)malloc(nq * sizeof(float) + 1);
  float* lb = (float*)malloc(nq * sizeof(float) + 1);
  uint8_t* luts = (uint8_t*)malloc(nq * G * 32 + 1);
  uint8_t* flat = (uint8_t*)malloc(n * G + 1);
  int rc = oc_prepare_queries(q, nq, dim, bits, shift, scale, q_rot, bc, luts, ls, lb);
  if (rc == 0) {
    oc_unpack_perm0(blocked_perm0, n, G, flat);
    for (size_t ii = 0; ii < n_qidx; ++ii) {
      const size_t gq = qidx[ii];
      const uint8_t* lut = luts + gq * G * 32;
      const float bias = lb[gq] + bc[gq];
      const float off = offsets ? offsets[gq] : 0.f;
      float* hs = heap_s + gq * k;
      uint64_t* hi = heap_i + gq * k;
      for (size_t cb = 0; cb < n; cb += OC_CHUNK) {
        const size_t end = cb + OC_CHUNK < n ? cb + OC_CHUNK : n;
        const int fast = heap_sz[gq] >= k && end - cb == OC_CHUNK;
        for (size_t v = cb; v < end; ++v) {
          float fa = bias;
          const uint8_t* row = flat + v * G;
          for (size_t w = 0; w < W; ++w) {
            uint32_t acc = 0;
            const size_t g1 = (w + 1) * OC_FLUSH_EVERY < G ? (w + 1) * OC_FLUSH_EVERY : G;
            for (size_t g = w * OC_FLUSH_EVERY; g < g1; ++g) {
              const unsigned byte = row[g];
              acc += lut[g * 32 + (byte >> 4)] + lut[g * 32 + 16 + (byte & 15u)];
            }
            fa = fmaf(ls[gq], (float)acc, fa);
          }
          const float p = fa * alphas[v];
          const uint64_t id = id_map ? id_map[v] : (uint64_t)v;
          if (fast) {
            const float s = off != 0.f ? p + off : p;
            if (s > heap_min[gq]) oc_heap_push(hs, hi, &heap_sz[gq], &heap_min[gq], &heap_mi[gq], k, s, id);
          } else {
            oc_heap_push(hs, hi, &heap_sz[gq], &heap_min[gq], &heap_mi[gq], k, p + off, id);
          }
        }
      }
    }
  }
  free(q_rot);
  free(bc);
  free(ls);
  free(lb);
  free(luts);
  free(flat);
  return rc;
}

/* heaps_to_search_results (reference src/search/search.cpp:1269-1291): each
 * heap's heap_sz entries by (score desc, id asc), the rest of the row (0, 0). */
int oc_heaps_to_results(size_t nq, size_t k, const float* heap_s, const uint64_t* heap_i,
                        const size_t* heap_sz, float* scores_out, uint64_t* ids_out) {
  oc_cand* c = (oc_cand*)malloc((k ? k : 1) * sizeof(oc_cand));
  for (size_t qi = 0; qi < nq; ++qi) {
    const size_t m = heap_sz[qi] < k ? heap_sz[qi] : k;
    for (size_t j = 0; j < m; ++j) {
      c[j].s = heap_s[qi * k + j];
      c[j].id = heap_i[qi * k + j];
    }
    qsort(c, m, sizeof(oc_cand), cand_cmp);
    for (size_t j = 0; j < k; ++j) {
      scores_out[qi * k + j] = j < m ? c[j].s : 0.f;
      ids_out[qi * k + j] = j < m ? c[j].id : 0;
    }
  }
  free(c);
  return 0;
}

/* float_rescore_topk (reference src/search/float_rescore.cpp:80-164): each
 * row's first n = min(k, counts[q]) (or k) candidates get the float inner
 * product of the query with the corpus row (-inf for an id >= n_corpus), in
 * the reference's AVX2 order (:36-57): 8-lane accumulators acc0 / acc1 over
 * 16-float steps (fused), acc0 + acc1, an 8-float fused tail, then the
 * horizontal sum ((v0+v4)+(v1+v5)) + ((v2+v6)+(v3+v7)) (:24-33); then the n
 * entries are ordered by (score desc, id asc). Entries past n are untouched. */
static float oc_dot_avx2_order(const float* a, const float* b, size_t dim) {
  float acc0[8] = {0}, acc1[8] = {0};
  size_t i = 0;
  for (; i + 16 <= dim; i += 16)
    for (int l = 0; l < 8; ++l) {
      acc0[l] = fmaf(a[i + l], b[i + l], acc0[l]);
      acc1[l] = fmaf(a[i + 8 + l], b[i + 8 + l], acc1[l]);
    }
  for (int l = 0; l < 8; ++l) acc0[l] = acc0[l] + acc1[l];
  for (; i + 8 <= dim; i += 8)
    for (int l = 0; l < 8; ++l) acc0[l] = fmaf(a[i + l], b[i + l], acc0[l]);
  float lo[4];
  for (int l = 0; l < 4; ++l) lo[l] = acc0[l] + acc0[l + 4];
  float s = (lo[0] + lo[1]) + (lo[2] + lo[3]);
  for (; i < dim; ++i) s = fmaf(a[i], b[i], s); /* dim % 8 != 0 only */
  return s;
}

int oc_float_rescore(size_t nq, size_t k, float* scores, uint64_t* ids, const size_t* counts,
                     const float* queries, const float* corpus, size_t n_corpus, size_t dim) {
  if (nq == 0 || k == 0 || dim == 0) return 0;
  oc_cand* c = (oc_cand*)malloc(k * sizeof(oc_cand));
  for (size_t q = 0; q < nq; ++q) {
    const size_t n = counts ? (counts[q] < k ? counts[q] : k) : k;
    for (size_t j = 0; j < n; ++j) {
      const uint64_t id = ids[q * k + j];
      c[j].id = id;
      c[j].s = id < n_corpus ? oc_dot_avx2_order(queries + q * dim, corpus + id * dim, dim) : -INFINITY;
    }
    qsort(c, n, sizeof(oc_cand), cand_cmp);
    for (size_t j = 0; j < n; ++j) {
      scores[q * k + j] = c[j].s;
      ids[q * k + j] = c[j].id;
    }
  }
  free(c);
  return 0;
}

/* ==== landmark index (reference src/index.cpp:120-289,
 * src/transform/landmarks.cpp, src/search/search.cpp:1304-1546) ==========
 * On the AVX2 path (TURBOVEC_NO_VNNI=1) planes_wanted() is false
 * (src/pack/pack.cpp:441-449 needs the vector-major kernels), so no sign
 * planes exist and search is: rotate the query, project it onto the 4-bit
 * landmark basis, LUTs from those rows, the PERM0 scan with the residual
 * post-correction. Float semantics follow the reference's object code under
 * its flags (-O3 -mavx2 -mfma -ffp-contract=fast): where the compiler
 * vectorized a reduction it multiplies (rounded) and adds in element order. */

/* mt19937-64 (Matsumoto and Nishimura; the std::mt19937_64 parameters). */
typedef struct {
  uint64_t mt[312];
  int i;
} oc_mt64;

static void mt64_seed(oc_mt64* g, uint64_t s) {
  g->mt[0] = s;
  for (int i = 1; i < 312; ++i) g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
  g->i = 312;
}

static uint64_t mt64_next(oc_mt64* g) {
  if (g->i >= 312) {
    for (int k = 0; k < 312; ++k) {
      const uint64_t y = (g->mt[k] & 0xFFFFFFFF80000000ULL) | (g->mt[(k + 1) % 312] & 0x7FFFFFFFULL);
      g->mt[k] = g->mt[(k + 156) % 312] ^ (y >> 1) ^ ((y & 1ULL) ? 0xB5026F5AA96619E9ULL : 0ULL);
    }
    g->i = 0;
  }
  uint64_t y = g->mt[g->i++];
  y ^= (y >> 29) & 0x5555555555555555ULL;
  y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
  y ^= (y << 37) & 0xFFF7EEE000000000ULL;
  y ^= y >> 43;
  return y;
}

/* libstdc++'s normal_distribution<double> (Marsaglia polar method) over
 * generate_canonical<double, 53>: u = double(x) 2^-64 (below 1); as compiled
 * into the reference, r2 = fma(a, a, b * b); returns b f, keeps a f. */
typedef struct {
  oc_mt64 g;
  double saved;
  int have;
} oc_gauss;

static double gauss_uniform(oc_gauss* s) {
  double u = (double)mt64_next(&s->g) * 0x1p-64;
  if (u >= 1.0) u = nextafter(1.0, 0.0);
  return u;
}

static double gauss_next(oc_gauss* s) {
  if (s->have) {
    s->have = 0;
    return s->saved + 0.0;
  }
  double a, b, r2;
  do {
    a = fma(2.0, gauss_uniform(s), -1.0);
    b = fma(2.0, gauss_uniform(s), -1.0);
    r2 = fma(a, a, b * b);
  } while (r2 > 1.0 || r2 == 0.0);
  const double f = sqrt(-2.0 * log(r2) / r2);
  s->saved = a * f;
  s->have = 1;
  return b * f + 0.0;
}

/* make_ortho_landmarks (landmarks.cpp:134-198): Gaussian dim x m (column
 * major, filled in order), modified Gram-Schmidt in double: dot and norm sums
 * of rounded products in element order, a_i = fma(-dot, a_j, a_i). */
int oc_ortho_landmarks(size_t dim, size_t m, uint64_t seed, float* L) {
  if (dim == 0 || m == 0 || m > dim || dim % 8 != 0) return -1;
  double* A = (double*)malloc(dim * m * sizeof(double));
  if (!A) return -1;
  oc_gauss gs;
  memset(&gs, 0, sizeof(gs));
  mt64_seed(&gs.g, seed);
  for (size_t i = 0; i < dim * m; ++i) A[i] = gauss_next(&gs);
  int rc = 0;
  for (size_t i = 0; i < m && rc == 0; ++i) {
    double* ai = A + i * dim;
    for (size_t j = 0; j < i; ++j) {
      const double* aj = A + j * dim;
      double dot = 0.0;
      for (size_t r = 0; r < dim; ++r) dot = dot + ai[r] * aj[r];
      for (size_t r = 0; r < dim; ++r) ai[r] = fma(-dot, aj[r], ai[r]);
    }
    double nrm = 0.0;
    for (size_t r = 0; r < dim; ++r) nrm = nrm + ai[r] * ai[r];
    nrm = sqrt(nrm);
    if (nrm < 1e-15) rc = -3;
    const double inv = 1.0 / nrm;
    for (size_t r = 0; r < dim; ++r) ai[r] = ai[r] * inv;
  }
  for (size_t i = 0; rc == 0 && i < dim * m; ++i) L[i] = (float)A[i];
  free(A);
  return rc;
}

static size_t pad16(size_t n) { return (n + 15u) & ~(size_t)15u; }

/* quantize_landmarks (landmarks.cpp:236-269): per column amax; scale =
 * amax / 7, level = clamp(round(x * (7 / amax)), -7, 7); byte = (odd + 7) << 4
 * | (even + 7); columns of pad16(dim) / 2 bytes, zero-filled. */
int oc_quantize_landmarks(const float* L, size_t dim, size_t m, uint8_t* codes, float* scales) {
  const size_t bpc = pad16(dim) / 2;
  memset(codes, 0, bpc * m);
  for (size_t j = 0; j < m; ++j) {
    const float* col = L + j * dim;
    float amax = 0.f;
    for (size_t i = 0; i < dim; ++i) amax = fmaxf(amax, fabsf(col[i]));
    scales[j] = 0.f;
    if (amax < 1e-30f) continue;
    scales[j] = amax / 7.f;
    const float inv = 7.f / amax;
    for (size_t i = 0; i < dim; i += 2) {
      const int e = (int)fmaxf(-7.f, fminf(7.f, roundf(col[i] * inv)));
      const int o = i + 1 < dim ? (int)fmaxf(-7.f, fminf(7.f, roundf(col[i + 1] * inv))) : 0;
      codes[j * bpc + i / 2] = (uint8_t)(((o + 7) << 4) | (e + 7));
    }
  }
  return 0;
}

/* dot_f32_avx2 (landmarks.cpp:104-121): 8 fused chains, then
 * ((a0 + a4) + (a1 + a5)) + ((a2 + a6) + (a3 + a7)); dim % 8 == 0. */
static float dot8(const float* a, const float* b, size_t dim) {
  float c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < dim; i += 8)
    for (int l = 0; l < 8; ++l) c[l] = fmaf(a[i + l], b[i + l], c[l]);
  return ((c[0] + c[4]) + (c[1] + c[5])) + ((c[2] + c[6]) + (c[3] + c[7]));
}

/* project_rows_onto_landmarks (landmarks.cpp:200-234) */
int oc_project_rows(const float* Y, size_t n, size_t dim, const float* L, size_t m, float* out,
                    size_t stride) {
  if (dim % 8 != 0 || stride < m) return -1;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    for (size_t j = 0; j < stride; ++j)
      out[(size_t)i * stride + j] = j < m ? dot8(Y + (size_t)i * dim, L + j * dim, dim) : 0.f;
  }
  return 0;
}

/* project_onto_landmarks_q4 (landmarks.cpp:271-348) with the codes/scales of
 * quantize_landmarks: q level = clamp(round(x * (127 / amax)), -127, 127),
 * integer dots over pad16(dim), q_land = (q_scale * scale_j) * (float) dot. */
static void project_q4_rows(const float* q_rot, size_t nq, size_t dim, const uint8_t* codes,
                            const float* scales, size_t m, float* q_land, size_t stride) {
  const size_t dp = pad16(dim), bpc = dp / 2;
#pragma omp parallel
  {
    int8_t* qi8 = (int8_t*)malloc(dp);
#pragma omp for schedule(static)
    for (long long qi = 0; qi < (long long)nq; ++qi) {
      const float* q = q_rot + (size_t)qi * dim;
      float amax = 0.f;
      for (size_t i = 0; i < dim; ++i) amax = fmaxf(amax, fabsf(q[i]));
      memset(qi8, 0, dp);
      float qs = 0.f;
      if (!(amax < 1e-30f)) {
        qs = amax / 127.f;
        const float inv = 127.f / amax;
        for (size_t i = 0; i < dim; ++i) qi8[i] = (int8_t)fmaxf(-127.f, fminf(127.f, roundf(q[i] * inv)));
      }
      float* dst = q_land + (size_t)qi * stride;
      for (size_t j = 0; j < stride; ++j) {
        if (j >= m) {
          dst[j] = 0.f;
          continue;
        }
        int32_t dot = 0;
        for (size_t b = 0; b < bpc; ++b) {
          const uint8_t byte = codes[j * bpc + b];
          dot += (int32_t)qi8[2 * b] * ((int32_t)(byte & 15u) - 7);
          dot += (int32_t)qi8[2 * b + 1] * ((int32_t)(byte >> 4) - 7);
        }
        dst[j] = (qs * scales[j]) * (float)dot;
      }
    }
    free(qi8);
  }
}

int oc_project_q4(const float* q_rot, size_t nq, size_t dim, const float* L, size_t m, float* q_land,
                  size_t stride) {
  if (stride < m) return -1;
  const size_t bpc = pad16(dim) / 2;
  uint8_t* codes = (uint8_t*)malloc(bpc * m);
  float* sc = (float*)malloc(m * sizeof(float));
  oc_quantize_landmarks(L, dim, m, codes, sc);
  project_q4_rows(q_rot, nq, dim, codes, sc, m, q_land, stride);
  free(codes);
  free(sc);
  return 0;
}

/* l2_normalize_into (index.cpp:21-50): 8 fused chains summed in order, the
 * tail in double, norm = (float) sqrt(double); below 1e-10 the row is zero
 * and the norm 0. */
static float lm_unit(const float* src, size_t d, float* dst) {
  float c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t i = 0;
  for (; i + 8 <= d; i += 8)
    for (int l = 0; l < 8; ++l) c[l] = fmaf(src[i + l], src[i + l], c[l]);
  float t = c[0];
  for (int l = 1; l < 8; ++l) t = t + c[l];
  double s = (double)t;
  for (; i < d; ++i) s = fma((double)src[i], (double)src[i], s);
  const float nrm = (float)sqrt(s);
  if (nrm < kMinInputNorm) {
    memset(dst, 0, d * sizeof(float));
    return 0.f;
  }
  const float inv = 1.f / nrm;
  for (size_t k = 0; k < d; ++k) dst[k] = src[k] * inv;
  return nrm;
}

/* encode_scaled_yhat (index.cpp:88-117): unit row, rotation, level = #{b : y
 * >= b}, alpha = (float) (dot / |c|^2) (double sums in order), row = alpha *
 * c, returns the sum of the rounded squares in order (dim % 8 == 0). */
static float lm_yhat(const float* src, size_t dim, const oc_rot* r, const float* bnd, size_t nb,
                     const float* cen, float* unit, float* y) {
  lm_unit(src, dim, unit);
  rotate_with(dim, r->perm, r->signs, r->block, r->isb, unit, 1.f, y);
  double dot = 0.0, cc = 0.0;
  for (size_t i = 0; i < dim; ++i) {
    unsigned k = 0;
    for (size_t b = 0; b < nb; ++b) k += (y[i] >= bnd[b]) ? 1u : 0u;
    const double c = (double)cen[k];
    dot = dot + (double)y[i] * c;
    cc = cc + c * c;
    unit[i] = cen[k];
  }
  const float alpha = (cc > kDegenerateInnerEps) ? (float)(dot / cc) : 0.f;
  float vn = 0.f;
  for (size_t i = 0; i < dim; ++i) {
    const float vi = alpha * unit[i];
    y[i] = vi;
    vn = vn + vi * vi;
  }
  return vn;
}

/* norm_sq_f32 (index.cpp:60-82): 8 fused chains over m & ~7, summed in order;
 * the tail as compiled: 4 rounded squares in order when >= 4 remain, then
 * fused multiply-adds. */
static float lm_norm_sq(const float* x, size_t m) {
  float c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const size_t m8 = m & ~(size_t)7;
  for (size_t i = 0; i < m8; i += 8)
    for (int l = 0; l < 8; ++l) c[l] = fmaf(x[i + l], x[i + l], c[l]);
  float s = c[0];
  for (int l = 1; l < 8; ++l) s = s + c[l];
  size_t j = m8;
  if (m - j >= 4) {
    for (int u = 0; u < 4; ++u) s = s + x[j + u] * x[j + u];
    j += 4;
  }
  for (; j < m; ++j) s = fmaf(x[j], x[j], s);
  return s;
}

/* LandmarkIndex end to end with the reference's state rules: rows added after
 * a prepare() find the earlier rows zeroed (landmark_rows_ was released and is
 * resized), and prepare() encodes every row. Arguments as vcref_lm_search
 * (search_seconds is left 0: the restatement is not timed). */
/* ---- staged sign-plane search (the reference's VNNI-host branch) ----------
 * search_landmark_rerank with sign planes (src/search/search.cpp:1379-1510):
 * on an AVX-512 VNNI host the index keeps the sign plane (the top bit of each
 * 4-bit landmark code; src/pack/pack.cpp:447-449,490-541) once n >= 32768,
 * and a search
 *   1. scans the sign bytes with build_sign_lut's uint8 LUTs (search.cpp:
 *      862-914): per vector A = the integer sum of the LUT entries of its
 *      nibbles, score = fma((float) A, lut_scale, lut_bias) (search_vnni.cpp:
 *      368-389; alpha on sign scores is opt-in and off by default, :1478-1487),
 *      and keeps the E best (E =
 *      planes_exact_len(k, planes_shortlist_len(k, 4)), pack.cpp:451-463) by
 *      (score desc, id asc): the buffered collector with its sample seeds
 *      returns exactly that set, ties included (one query's vectors reach it
 *      in id order, and it drops only what E kept entries beat);
 *   2. rescores those E in that order with build_permute_dot's int8 query
 *      weights and centroid levels (search_vnni.cpp:1318-1360): acc = sum
 *      over byte groups of lev[lo] * w_lo + lev[hi] * w_hi (exact int32),
 *      f = (float) acc * (cs * qs), score = the flush's
 *      f * alpha + q_perp * rn (search_vnni.cpp:1194-1217, one product fused),
 *      pushed into a k-heap in candidate order (topk_heap.hpp:101-157), then
 *      sorted by (score desc, id asc).
 * The fusion sites of steps 1 and 2 are the compiled reference's; the
 * variant flags below exist so that tests can show the other choices fail. */
static int g_staged_variant = 0; /* bit 0: step 1 unfused; bit 1: step 2 fuses the other product */
void oc_set_staged_variant(int v) { g_staged_variant = v; }

size_t oc_planes_exact_len(size_t k) {
  const size_t s_len = 256 > ((k >= 64) ? k * 16 : k * 20) ? 256 : ((k >= 64) ? k * 16 : k * 20);
  const size_t want = 256 > 6 * k ? 256 : 6 * k;
  return s_len < want ? s_len : want;
}

static size_t planes_shortlist_len4(size_t k) {
  const size_t a = (k >= 64) ? k * 16 : k * 20;
  return a > 256 ? a : 256;
}

/* build_sign_lut (search.cpp:862-914) for one query row of S dims */
static void build_sign_lut(const float* q, size_t S, uint8_t* lut, float* scale_out, float* bias_out,
                           float* fv) {
  const size_t nsg = S / 8;
  float* mins = fv + nsg * 32;
  float bias = 0.f, max_span = 0.f;
  for (size_t g = 0; g < nsg; ++g)
    for (size_t half = 0; half < 2; ++half) {
      const size_t d = g * 8 + half * 4;
      const float p0 = q[d], p1 = q[d + 1], p2 = q[d + 2], p3 = q[d + 3];
      const float a[4] = {-p0 - p1, p1 - p0, p0 - p1, p0 + p1};
      const float b[4] = {-p2 - p3, p3 - p2, p2 - p3, p2 + p3};
      for (size_t pat = 0; pat < 16; ++pat) fv[g * 32 + half * 16 + pat] = a[pat >> 2] + b[pat & 3];
      const float m01 = fabsf(p0) + fabsf(p1), m23 = fabsf(p2) + fabsf(p3);
      const float mn = -m01 - m23, mx = m01 + m23;
      mins[g * 2 + half] = mn;
      bias = bias + mn;
      if (mx - mn > max_span) max_span = mx - mn;
    }
  float scale = max_span > 0.f ? max_span / 127.f : 1.f;
  float inv = 1.f;
  if (scale >= FLT_MIN)
    inv = 1.f / scale;
  else
    scale = 1.f;
  for (size_t g = 0; g < nsg; ++g)
    for (size_t half = 0; half < 2; ++half) {
      const float mn = mins[g * 2 + half];
      for (size_t i = 0; i < 16; ++i) {
        const size_t j = g * 32 + half * 16 + i;
        float v = roundf((fv[j] - mn) * inv);
        v = v < 0.f ? 0.f : (v > 127.f ? 127.f : v);
        lut[j] = (uint8_t)v;
      }
    }
  *scale_out = scale;
  *bias_out = bias;
}

typedef struct {
  int8_t lev[16];
  int8_t* w; /* [2 * G]: w[2g] = weight of the low nibble (dim 2g+1), w[2g+1] of the high (dim 2g) */
  float scale;
} oc_pdot;

/* build_permute_dot (search_vnni.cpp:1318-1360) for one query row of S dims */
static void build_permute_dot(const float* q, size_t S, const float* cen, oc_pdot* pd) {
  float cmax = 0.f;
  for (int i = 0; i < 16; ++i) cmax = fabsf(cen[i]) > cmax ? fabsf(cen[i]) : cmax;
  const float cs = cmax > 0.f ? cmax / 127.f : 1.f;
  for (int i = 0; i < 16; ++i) {
    float v = roundf(cen[i] / cs);
    v = v < -127.f ? -127.f : (v > 127.f ? 127.f : v);
    pd->lev[i] = (int8_t)v;
  }
  float qmax = 0.f;
  for (size_t d = 0; d < S; ++d) qmax = fabsf(q[d]) > qmax ? fabsf(q[d]) : qmax;
  float qs = qmax / 127.f, inv = 1.f;
  if (qs >= FLT_MIN)
    inv = 1.f / qs;
  else
    qs = 1.f;
  for (size_t g = 0; g < S / 2; ++g) {
    float lo = roundf(q[2 * g + 1] * inv), hi = roundf(q[2 * g] * inv);
    lo = lo < -127.f ? -127.f : (lo > 127.f ? 127.f : lo);
    hi = hi < -127.f ? -127.f : (hi > 127.f ? 127.f : hi);
    pd->w[2 * g] = (int8_t)lo;
    pd->w[2 * g + 1] = (int8_t)hi;
  }
  pd->scale = cs * qs;
}

/* the staged search of one query (steps 1 and 2 above); rows of ke entries */
static void lm_staged_query(const float* q_land, size_t S, const float* lcen, const uint8_t* flat, size_t n,
                            const float* alpha, const float* rn, float q_perp, size_t ke, float* fv,
                            uint8_t* slut, oc_cand* cand, int8_t* wbuf, float* scores_out, uint64_t* ids_out) {
  const size_t G = S / 2, nsg = S / 8;
  float ls, lb;
  build_sign_lut(q_land, S, slut, &ls, &lb, fv);
  for (size_t v = 0; v < n; ++v) {
    const uint8_t* row = flat + v * G;
    int32_t A = 0;
    for (size_t g = 0; g < nsg; ++g) {
      unsigned sb = 0;
      for (int h = 0; h < 4; ++h) {
        const unsigned b = row[4 * g + h];
        sb = (sb << 2) | ((b >> 7) & 1u) << 1 | ((b >> 3) & 1u);
      }
      A += slut[g * 32 + (sb >> 4)] + slut[g * 32 + 16 + (sb & 15u)];
    }
    cand[v].s = (g_staged_variant & 1) ? ((float)A * ls + lb) : fmaf((float)A, ls, lb);
    cand[v].id = v;
  }
  const size_t E = oc_planes_exact_len(ke) < n ? oc_planes_exact_len(ke) : n;
  qsort(cand, n, sizeof(oc_cand), cand_cmp);
  oc_pdot pd;
  pd.w = wbuf;
  build_permute_dot(q_land, S, lcen, &pd);
  float* hs = (float*)malloc(ke * sizeof(float));
  uint64_t* hi = (uint64_t*)malloc(ke * sizeof(uint64_t));
  size_t hsz = 0, hmi = 0;
  float hmin = -INFINITY;
  for (size_t c = 0; c < E; ++c) {
    const uint64_t id = cand[c].id;
    const uint8_t* row = flat + id * G;
    int32_t acc = 0;
    for (size_t g = 0; g < G; ++g)
      acc += (int32_t)pd.lev[row[g] & 15u] * pd.w[2 * g] + (int32_t)pd.lev[row[g] >> 4] * pd.w[2 * g + 1];
    const float f = (float)acc * pd.scale;
    const float s = (g_staged_variant & 2) ? fmaf(q_perp, rn[id], f * alpha[id]) : fmaf(f, alpha[id], q_perp * rn[id]);
    oc_heap_push(hs, hi, &hsz, &hmin, &hmi, ke, s, id);
  }
  oc_cand* out = (oc_cand*)malloc((hsz + 1) * sizeof(oc_cand));
  for (size_t i = 0; i < hsz; ++i) {
    out[i].s = hs[i];
    out[i].id = hi[i];
  }
  qsort(out, hsz, sizeof(oc_cand), cand_cmp);
  for (size_t j = 0; j < ke; ++j) {
    scores_out[j] = j < hsz ? out[j].s : 0.f;
    ids_out[j] = j < hsz ? out[j].id : 0;
  }
  free(out);
  free(hs);
  free(hi);
}

/* The sign-plane scan alone (step 1 above) over bit-plane codes (n x 4 S/8):
 * top k of every vector's sign score by (score desc, id asc). */
int oc_sign_topk(const float* rows, size_t nq, size_t S, const uint8_t* packed, size_t n, size_t k,
                 float* scores_out, uint64_t* ids_out, size_t* k_out) {
  if (S == 0 || S % 8 != 0) return -1;
  const size_t ke = k < n ? k : n, G = S / 2;
  *k_out = ke;
  uint8_t* flat = (uint8_t*)malloc(n * G + 1);
  oc_codes_flat(packed, n, 4, S, flat);
  float* fv = (float*)malloc((S / 8) * 34 * sizeof(float));
  uint8_t* slut = (uint8_t*)malloc((S / 8) * 32);
  oc_cand* cand = (oc_cand*)malloc((n + 1) * sizeof(oc_cand));
  for (size_t i = 0; i < nq; ++i) {
    float ls, lb;
    build_sign_lut(rows + i * S, S, slut, &ls, &lb, fv);
    for (size_t v = 0; v < n; ++v) {
      const uint8_t* row = flat + v * G;
      int32_t A = 0;
      for (size_t g = 0; g < S / 8; ++g) {
        unsigned sb = 0;
        for (int h = 0; h < 4; ++h) {
          const unsigned b = row[4 * g + h];
          sb = (sb << 2) | ((b >> 7) & 1u) << 1 | ((b >> 3) & 1u);
        }
        A += slut[g * 32 + (sb >> 4)] + slut[g * 32 + 16 + (sb & 15u)];
      }
      cand[v].s = (g_staged_variant & 1) ? ((float)A * ls + lb) : fmaf((float)A, ls, lb);
      cand[v].id = v;
    }
    qsort(cand, n, sizeof(oc_cand), cand_cmp);
    for (size_t j = 0; j < ke; ++j) {
      scores_out[i * ke + j] = cand[j].s;
      ids_out[i * ke + j] = cand[j].id;
    }
  }
  free(flat);
  free(fv);
  free(slut);
  free(cand);
  return 0;
}

static int lm_search_impl(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                          const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                          size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                          uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds, int planes);

int oc_lm_search(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                 const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                 size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                 uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds) {
  return lm_search_impl(x, batch_rows, n_batches, prepare_between, q, nq, dim, scan_bits, m, seed, rerank_bits, k,
                        float_corpus, scores_out, ids_out, k_out, info, search_seconds, 0);
}

/* oc_lm_search as the reference runs on an AVX-512 VNNI host (planes = 1):
 * the staged branch when scan_bits == 4, at least 32768 vectors are prepared
 * and the shortlist is shorter than the index (search.cpp:1379-1384,
 * pack.cpp:447-449). Below that the reference's VNNI host scans with its
 * permute-dot kernels, which this restatement does not cover (-2). */
int oc_lm_search_planes(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                        const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                        size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                        uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds) {
  return lm_search_impl(x, batch_rows, n_batches, prepare_between, q, nq, dim, scan_bits, m, seed, rerank_bits, k,
                        float_corpus, scores_out, ids_out, k_out, info, search_seconds, 1);
}

static int lm_search_impl(const float* x, const size_t* batch_rows, size_t n_batches, int prepare_between,
                          const float* q, size_t nq, size_t dim, size_t scan_bits, size_t m, uint64_t seed,
                          size_t rerank_bits, size_t k, const float* float_corpus, float* scores_out,
                          uint64_t* ids_out, size_t* k_out, size_t* info, double* search_seconds, int planes) {
  if (search_seconds != NULL) *search_seconds = 0.0;
  if (rerank_bits == 0) rerank_bits = scan_bits;
  if (dim == 0 || dim % 8 != 0 || scan_bits < 2 || scan_bits > 4 || rerank_bits < 2 || rerank_bits > 4 ||
      m == 0 || m > dim)
    return -1;
  const size_t S = (m + 7u) & ~(size_t)7u;
  const size_t G = groups_of(scan_bits, S);
  float* L = (float*)malloc(dim * m * sizeof(float));
  if (oc_ortho_landmarks(dim, m, seed, L) != 0) {
    free(L);
    return -1;
  }
  const size_t bpc = pad16(dim) / 2;
  uint8_t* lq = (uint8_t*)malloc(bpc * m);
  float* lsc = (float*)malloc(m * sizeof(float));
  oc_quantize_landmarks(L, dim, m, lq, lsc);
  float bnd[15], cen[16], lbnd[15], lcen[16];
  oc_codebook(rerank_bits, dim, bnd, cen);
  oc_codebook(scan_bits, S, lbnd, lcen);
  oc_rot r;
  rot_make(&r, dim);
  size_t n_total = 0;
  for (size_t b = 0; b < n_batches; ++b) n_total += batch_rows[b];
  float* rows = (float*)calloc(n_total * S + 1, sizeof(float));
  float* rn = (float*)calloc(n_total + 1, sizeof(float));
  float* alpha = (float*)calloc(n_total + 1, sizeof(float));
  uint8_t* flat = (uint8_t*)calloc(n_total * G + 1, 1);
  size_t n = 0, n_enc = 0;
  int rows_live = 1;
  for (size_t b = 0; b < n_batches; ++b) {
    if (!rows_live) {
      memset(rows, 0, n_total * S * sizeof(float));
      rows_live = 1;
    }
    const size_t nb = batch_rows[b];
#pragma omp parallel
    {
      float* unit = (float*)malloc(dim * sizeof(float));
      float* y = (float*)malloc(dim * sizeof(float));
#pragma omp for schedule(static)
      for (long long ii = 0; ii < (long long)nb; ++ii) {
        const size_t i = n + (size_t)ii;
        const float vn = lm_yhat(x + i * dim, dim, &r, bnd, ((size_t)1 << rerank_bits) - 1, cen, unit, y);
        float* row = rows + i * S;
        for (size_t j = 0; j < S; ++j) row[j] = j < m ? dot8(y, L + j * dim, dim) : 0.f;
        const float d = vn - lm_norm_sq(row, m);
        rn[i] = sqrtf(0.f < d ? d : 0.f);
      }
      free(unit);
      free(y);
    }
    n += nb;
    if (prepare_between || b + 1 == n_batches) {
      /* prepare(): every row, then the rows are released */
      const size_t bpr = scan_bits * (S / 8);
      uint8_t* packed = (uint8_t*)malloc(n * bpr + 1);
#pragma omp parallel
      {
        float* unit = (float*)malloc(S * sizeof(float));
#pragma omp for schedule(static)
        for (long long ii = 0; ii < (long long)n; ++ii) {
          const float nrm = lm_unit(rows + (size_t)ii * S, S, unit);
          alpha[ii] = quantize_row(unit, S, scan_bits, lbnd, lcen, NULL, NULL, NULL, 1, nrm,
                                   packed + (size_t)ii * bpr);
        }
        free(unit);
      }
      oc_codes_flat(packed, n, scan_bits, S, flat);
      free(packed);
      n_enc = n;
      rows_live = 0;
    }
  }
  int rc = 0;
  const size_t ke = k < n ? k : n;
  *k_out = ke;
  if (nq > 0 && ke > 0) {
    float* q_rot = (float*)malloc(nq * dim * sizeof(float));
    float* q_land = (float*)malloc(nq * S * sizeof(float));
    float* q_perp = (float*)malloc(nq * sizeof(float));
    uint8_t* luts = (uint8_t*)malloc(nq * G * 32);
    float* ls = (float*)malloc(nq * sizeof(float));
    float* lb = (float*)malloc(nq * sizeof(float));
    for (size_t i = 0; i < nq; ++i) {
      const float nv = oc_row_norm(q + i * dim, dim);
      const float inv = (nv > kMinInputNorm) ? (1.f / nv) : 0.f;
      rotate_with(dim, r.perm, r.signs, r.block, r.isb, q + i * dim, inv, q_rot + i * dim);
    }
    project_q4_rows(q_rot, nq, dim, lq, lsc, m, q_land, S);
    float* fv = (float*)malloc(G * 34 * sizeof(float));
    for (size_t i = 0; i < nq; ++i) {
      double q2 = 0.0, l2 = 0.0;
      for (size_t d = 0; d < dim; ++d) q2 = q2 + (double)q_rot[i * dim + d] * (double)q_rot[i * dim + d];
      for (size_t j = 0; j < m; ++j) l2 = l2 + (double)q_land[i * S + j] * (double)q_land[i * S + j];
      const double d2 = q2 - l2;
      q_perp[i] = (float)sqrt(0.0 < d2 ? d2 : 0.0);
      float bias;
      build_lut(q_land + i * S, S, scan_bits, lcen, luts + i * G * 32, &ls[i], &bias, fv);
      lb[i] = bias + 0.f;
    }
    /* S % 32 != 0: the reference's sign scan reads past its 32-wide sign
     * blocks with the 192-wide kernel (search_vnni.cpp:1466-1469, 803-843), an
     * undefined result; the exact scan below stands in for it. */
    const int sign_ok = S % 32 == 0;
    const int staged = planes && sign_ok && scan_bits == 4 && n_enc >= 32768 && planes_shortlist_len4(ke) < n_enc;
    if (planes && sign_ok && !staged) rc = -2;
    if (!sign_ok) planes = 0;
    if (staged) {
      uint8_t* slut = (uint8_t*)malloc((S / 8) * 32);
      oc_cand* cand = (oc_cand*)malloc(n_enc * sizeof(oc_cand));
      int8_t* wbuf = (int8_t*)malloc(S);
      for (size_t i = 0; i < nq; ++i)
        lm_staged_query(q_land + i * S, S, lcen, flat, n_enc, alpha, rn, q_perp[i], ke, fv, slut, cand, wbuf,
                        scores_out + i * ke, ids_out + i * ke);
      free(slut);
      free(cand);
      free(wbuf);
    }
    free(fv);
    size_t kk = 0;
    if (!planes) rc = oc_scan_topk(luts, ls, lb, nq, flat, n_enc, G, alpha, NULL, ke, scores_out, ids_out, &kk);
    /* residual post-correction and re-rank (search.cpp:1123-1146) */
    oc_cand* rowc = (oc_cand*)malloc(ke * sizeof(oc_cand));
    for (size_t i = 0; rc == 0 && !planes && i < nq; ++i) {
      for (size_t j = 0; j < ke; ++j) {
        const uint64_t vi = ids_out[i * ke + j];
        float s = scores_out[i * ke + j];
        if (vi < n_enc) s = fmaf(q_perp[i], rn[vi], s);
        rowc[j].s = s;
        rowc[j].id = vi;
      }
      qsort(rowc, ke, sizeof(oc_cand), cand_cmp);
      for (size_t j = 0; j < ke; ++j) {
        scores_out[i * ke + j] = rowc[j].s;
        ids_out[i * ke + j] = rowc[j].id;
      }
    }
    free(rowc);
    if (rc == 0 && float_corpus != NULL)
      rc = oc_float_rescore(nq, ke, scores_out, ids_out, NULL, q, float_corpus, n, dim);
    free(q_rot);
    free(q_land);
    free(q_perp);
    free(luts);
    free(ls);
    free(lb);
  }
  if (info != NULL) {
    info[0] = S;
    info[1] = (n_enc + 191) / 192 * 192 * G;
    info[2] = info[1] + 4 * n + 12 * n_enc;
  }
  rot_free(&r);
  free(L);
  free(lq);
  free(lsc);
  free(rows);
  free(rn);
  free(alpha);
  free(flat);
  return rc;
}

/* ---- bucketed index ------------------------------------------------------
 * The reference describes its bucketed index in prose only (README.md:56,
 * 114-120): rows are L2-normalized, each bucket keeps a frozen encode
 * centroid c, rows are encoded as residuals x_hat - c, and a bucket's
 * vectors score <q_hat, c> + alpha <q_hat, r_hat> (score_prepared_into with
 * score_offsets, src/search/search.cpp:1156-1267). x_hat = x * inv, inv =
 * 1 / row_norm(x) (0 below 1e-10), each product and difference rounded. */
int oc_bucket_residuals(const float* x, size_t dim, const uint64_t* pos_rows, const uint32_t* pos_bucket,
                        size_t n, const float* cent, float* out) {
  for (size_t p = 0; p < n; ++p) {
    const float* xr = x + (size_t)pos_rows[p] * dim;
    const float nv = oc_row_norm(xr, dim);
    const float inv = (nv > kMinInputNorm) ? (1.f / nv) : 0.f;
    for (size_t d = 0; d < dim; ++d) {
      const float v = xr[d] * inv;
      out[p * dim + d] = cent != NULL ? v - cent[(size_t)pos_bucket[p] * dim + d] : v;
    }
  }
  return 0;
}

/* Score offset of query q and probed bucket b: (float) of the sequential
 * double sum over d of q_hat[d] * c_b[d] (each product exact in double). */
int oc_probe_offsets(const float* q, size_t nq, size_t dim, const float* cent, const int32_t* probe_b,
                     size_t P, const int32_t* probe_n, float* off) {
  for (size_t i = 0; i < nq; ++i) {
    const float* qr = q + i * dim;
    const float nv = oc_row_norm(qr, dim);
    const float inv = (nv > kMinInputNorm) ? (1.f / nv) : 0.f;
    for (size_t j = 0; j < (size_t)probe_n[i] && j < P; ++j) {
      const float* c = cent + (size_t)probe_b[i * P + j] * dim;
      double acc = 0.0;
      for (size_t d = 0; d < dim; ++d) acc = fma((double)(qr[d] * inv), (double)c[d], acc);
      off[i * P + j] = (float)acc;
    }
  }
  return 0;
}
