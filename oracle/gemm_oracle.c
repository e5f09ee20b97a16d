/*
 * gemm_oracle.c -- CPU oracle for the SGEMM/DGEMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's
 * cpu_baseline leg as the checker; never linked into DGEMM.so / SGEMM.so.
 * Built with -ffp-contract=off so the error-free transformations below are
 * not fused by the compiler.  Every function names the reference lines it
 * restates; the emulated forms have no reference counterpart and restate the
 * algorithm of this library's emulated kernels.
 */
#include "gemm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* Counter-based generator: SplitMix64.  The stream key is the SplitMix64   */
/* finaliser of (seed, id); element idx is the idx-th SplitMix64 output of  */
/* that key, so any entry can be drawn without drawing the others.          */
/* ---------------------------------------------------------------------- */
#define SM64_GAMMA 0x9E3779B97F4A7C15ULL

static uint64_t sm64_fin(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t oracle_counter_u64(uint64_t seed, uint64_t id, uint64_t idx) {
  uint64_t key = sm64_fin(seed * SM64_GAMMA + (id + 1) * 0xD1B54A32D192ED03ULL);
  return sm64_fin(key + (idx + 1) * SM64_GAMMA);
}

static double u01_f64(uint64_t x) { return (double)(x >> 11) * 0x1p-53; }

double oracle_uniform_f64(uint64_t seed, uint64_t id, uint64_t idx) {
  return u01_f64(oracle_counter_u64(seed, id, idx)) * 2.0 - 1.0;
}

float oracle_uniform_f32(uint64_t seed, uint64_t id, uint64_t idx) {
  float u = (float)(oracle_counter_u64(seed, id, idx) >> 40) * 0x1p-24f;
  return u * 2.0f - 1.0f;
}

static double gauss(uint64_t seed, uint64_t id, uint64_t i1, uint64_t i2) {
  double u1 = 1.0 - u01_f64(oracle_counter_u64(seed, id, i1)); /* (0,1] */
  double u2 = u01_f64(oracle_counter_u64(seed, id, i2));
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

double oracle_emu_f64(uint64_t seed, uint64_t id, uint64_t idx, double phi) {
  double u = u01_f64(oracle_counter_u64(seed, id, 3 * idx));
  double g = gauss(seed, id, 3 * idx + 1, 3 * idx + 2);
  return (u - 0.5) * exp(phi * g);
}

float oracle_lognormal_f32(uint64_t seed, uint64_t id, uint64_t idx, double sigma) {
  return (float)exp(sigma * gauss(seed, id, 2 * idx, 2 * idx + 1));
}

static double gen_f64(uint64_t seed, uint64_t id, uint64_t idx, int dist, double param) {
  if (dist == 1) return oracle_emu_f64(seed, id, idx, param);
  if (dist == 2) return (double)oracle_lognormal_f32(seed, id, idx, param);
  return oracle_uniform_f64(seed, id, idx);
}

static float gen_f32(uint64_t seed, uint64_t id, uint64_t idx, int dist, double param) {
  if (dist == 1) return (float)oracle_emu_f64(seed, id, idx, param);
  if (dist == 2) return oracle_lognormal_f32(seed, id, idx, param);
  return oracle_uniform_f32(seed, id, idx);
}

void oracle_fill_f64(double *p, int rows, int cols, int ld, uint64_t seed,
                     uint64_t id, int dist, double param) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < cols; j++) {
    double *col = p + (int64_t)j * ld;
    for (int i = 0; i < rows; i++)
      col[i] = gen_f64(seed, id, (uint64_t)i + (uint64_t)j * rows, dist, param);
    for (int i = rows; i < ld; i++) col[i] = NAN;
  }
}

void oracle_fill_f32(float *p, int rows, int cols, int ld, uint64_t seed,
                     uint64_t id, int dist, double param) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < cols; j++) {
    float *col = p + (int64_t)j * ld;
    for (int i = 0; i < rows; i++)
      col[i] = gen_f32(seed, id, (uint64_t)i + (uint64_t)j * rows, dist, param);
    for (int i = rows; i < ld; i++) col[i] = NAN;
  }
}

/* ---------------------------------------------------------------------- */
/* Error-free transformations (Knuth TwoSum, FMA TwoProd) and Dot2          */
/* (Ogita, Rump, Oishi 2005): the result is as accurate as if computed in   */
/* twice the working precision and then rounded.                            */
/* ---------------------------------------------------------------------- */
typedef struct { double hi, lo; } dd_t;

static inline void dot2_step(dd_t *s, double a, double b) {
  double p = a * b;
  double pe = fma(a, b, -p);
  double t = s->hi + p;
  double bb = t - s->hi;
  double se = (s->hi - (t - bb)) + (p - bb);
  s->hi = t;
  s->lo += se + pe;
}

static inline double dd_round(dd_t s) { return s.hi + s.lo; }

/* op(A)[i,p] and op(B)[p,j] for the reference's transpose convention:
 * only the first character matters, 'N'/'n' is no-transpose, anything else
 * transposes (src/gemm_driver.c:17-18). */
static inline int is_notrans(char c) { return c == 'N' || c == 'n'; }

/* ---------------------------------------------------------------------- */
/* Reference semantics with Dot2 accumulation.                              */
/* Restates GEMM_DRIVER, src/gemm_driver.c:118-197:                         */
/*   :133  if beta != 1, C[0:m,0:n] *= beta (C_MULT_BETA :106-116)          */
/*   :134  if alpha != 0, C += A * (alpha*B) with alpha folded into the      */
/*         packed B block (src/gemm_copy.c:65 outb=in*ALPHA)                 */
/* ---------------------------------------------------------------------- */
void oracle_dgemm(const char *transa, const char *transb, const int *pm,
                  const int *pn, const int *pk, const double *palpha,
                  const double *a, const int *plda, const double *b,
                  const int *pldb, const double *pbeta, double *c,
                  const int *pldc) {
  const int m = *pm, n = *pn, k = *pk, lda = *plda, ldb = *pldb, ldc = *pldc;
  const double alpha = *palpha, beta = *pbeta;
  const int na = is_notrans(*transa), nb = is_notrans(*transb);
  if (m <= 0 || n <= 0) return;
#pragma omp parallel
  {
    double *bcol = (double *)malloc(sizeof(double) * (k > 0 ? k : 1));
    dd_t *acc = (dd_t *)malloc(sizeof(dd_t) * m);
#pragma omp for schedule(dynamic, 1)
    for (int j = 0; j < n; j++) {
      double *cj = c + (int64_t)j * ldc;
      for (int i = 0; i < m; i++) {
        double c1 = (beta != 1.0) ? beta * cj[i] : cj[i];
        acc[i].hi = c1;
        acc[i].lo = 0.0;
      }
      if (alpha != 0.0) {
        for (int p = 0; p < k; p++) {
          double bv = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
          bcol[p] = bv * alpha;
        }
        if (na) {
          for (int p = 0; p < k; p++) {
            const double *ap = a + (int64_t)p * lda;
            double bp = bcol[p];
            for (int i = 0; i < m; i++) dot2_step(&acc[i], ap[i], bp);
          }
        } else {
          for (int i = 0; i < m; i++) {
            const double *ai = a + (int64_t)i * lda;
            for (int p = 0; p < k; p++) dot2_step(&acc[i], ai[p], bcol[p]);
          }
        }
      }
      for (int i = 0; i < m; i++) cj[i] = dd_round(acc[i]);
    }
    free(bcol);
    free(acc);
  }
}

void oracle_sgemm(const char *transa, const char *transb, const int *pm,
                  const int *pn, const int *pk, const float *palpha,
                  const float *a, const int *plda, const float *b,
                  const int *pldb, const float *pbeta, float *c,
                  const int *pldc) {
  const int m = *pm, n = *pn, k = *pk, lda = *plda, ldb = *pldb, ldc = *pldc;
  const float alpha = *palpha, beta = *pbeta;
  const int na = is_notrans(*transa), nb = is_notrans(*transb);
  if (m <= 0 || n <= 0) return;
#pragma omp parallel
  {
    float *bcol = (float *)malloc(sizeof(float) * (k > 0 ? k : 1));
    dd_t *acc = (dd_t *)malloc(sizeof(dd_t) * m);
#pragma omp for schedule(dynamic, 1)
    for (int j = 0; j < n; j++) {
      float *cj = c + (int64_t)j * ldc;
      for (int i = 0; i < m; i++) {
        float c1 = (beta != 1.0f) ? beta * cj[i] : cj[i];
        acc[i].hi = (double)c1;
        acc[i].lo = 0.0;
      }
      if (alpha != 0.0f) {
        for (int p = 0; p < k; p++) {
          float bv = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
          bcol[p] = bv * alpha; /* one FP32 rounding, as the packed B block */
        }
        if (na) {
          for (int p = 0; p < k; p++) {
            const float *ap = a + (int64_t)p * lda;
            double bp = (double)bcol[p];
            for (int i = 0; i < m; i++) dot2_step(&acc[i], (double)ap[i], bp);
          }
        } else {
          for (int i = 0; i < m; i++) {
            const float *ai = a + (int64_t)i * lda;
            for (int p = 0; p < k; p++)
              dot2_step(&acc[i], (double)ai[p], (double)bcol[p]);
          }
        }
      }
      for (int i = 0; i < m; i++) cj[i] = (float)dd_round(acc[i]);
    }
    free(bcol);
    free(acc);
  }
}

void oracle_dgemm_longdouble(char transa, char transb, int m, int n, int k,
                             double alpha, const double *a, int lda,
                             const double *b, int ldb, double beta,
                             double *c, int ldc) {
  const int na = is_notrans(transa), nb = is_notrans(transb);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      long double s = 0.0L;
      for (int p = 0; p < k; p++) {
        long double av = na ? a[i + (int64_t)p * lda] : a[p + (int64_t)i * lda];
        long double bv = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
        s += av * bv;
      }
      long double cv = c[i + (int64_t)j * ldc];
      long double r = (beta != 1.0 ? (long double)beta * cv : cv);
      if (alpha != 0.0) r += (long double)alpha * s;
      c[i + (int64_t)j * ldc] = (double)r;
    }
}

/* ---------------------------------------------------------------------- */
/* Sampled entries for generated inputs (A id 0, B id 1, C id 2).  Stored   */
/* shapes: A is m x k (transa N) or k x m (T); B is k x n (N) or n x k (T); */
/* the counter index of a stored element (r, s) is r + s*rows.              */
/* ---------------------------------------------------------------------- */
void oracle_sampled_f64(char transa, char transb, int m, int n, int k,
                        double alpha, double beta, uint64_t seed, int dist,
                        double param, int count, const int *is, const int *js,
                        double *out, double *absprod) {
  const int na = is_notrans(transa), nb = is_notrans(transb);
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < count; t++) {
    const uint64_t i = (uint64_t)is[t], j = (uint64_t)js[t];
    double cv = gen_f64(seed, 2, i + j * (uint64_t)m, dist, param);
    dd_t s = {(beta != 1.0) ? beta * cv : cv, 0.0};
    double ab = 0.0;
    if (alpha != 0.0) {
      for (uint64_t p = 0; p < (uint64_t)k; p++) {
        double av = gen_f64(seed, 0, na ? i + p * (uint64_t)m : p + i * (uint64_t)k, dist, param);
        double bv = gen_f64(seed, 1, nb ? p + j * (uint64_t)k : j + p * (uint64_t)n, dist, param);
        dot2_step(&s, av, bv * alpha);
        ab += fabs(av * bv);
      }
    }
    out[t] = dd_round(s);
    if (absprod) absprod[t] = ab;
  }
}

void oracle_sampled_f32(char transa, char transb, int m, int n, int k,
                        float alpha, float beta, uint64_t seed, int dist,
                        double param, int count, const int *is, const int *js,
                        double *out, double *absprod) {
  const int na = is_notrans(transa), nb = is_notrans(transb);
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < count; t++) {
    const uint64_t i = (uint64_t)is[t], j = (uint64_t)js[t];
    float cv = gen_f32(seed, 2, i + j * (uint64_t)m, dist, param);
    float c1 = (beta != 1.0f) ? beta * cv : cv;
    dd_t s = {(double)c1, 0.0};
    double ab = 0.0;
    if (alpha != 0.0f) {
      for (uint64_t p = 0; p < (uint64_t)k; p++) {
        float av = gen_f32(seed, 0, na ? i + p * (uint64_t)m : p + i * (uint64_t)k, dist, param);
        float bv = gen_f32(seed, 1, nb ? p + j * (uint64_t)k : j + p * (uint64_t)n, dist, param);
        dot2_step(&s, (double)av, (double)(bv * alpha));
        ab += fabs((double)av * (double)bv);
      }
    }
    out[t] = dd_round(s);
    if (absprod) absprod[t] = ab;
  }
}

/* The same Dot2 entries for explicit operands: rows (nr x k) are whole rows
 * of op(A), cols (nc x k) whole columns of op(B), entry t pairs rows[ri[t]]
 * with cols[cj[t]] and starts from c1[t] (C already multiplied by beta);
 * alpha is folded into the B entries as in oracle_sampled_f64. */
void oracle_dot2_pairs(const double *rows, const double *cols, int k, int count,
                       const int *ri, const int *cj, double alpha, const double *c1,
                       double *out, double *absprod) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < count; t++) {
    const double *x = rows + (size_t)ri[t] * k, *y = cols + (size_t)cj[t] * k;
    dd_t s = {c1[t], 0.0};
    double ab = 0.0;
    if (alpha != 0.0)
      for (int p = 0; p < k; p++) {
        dot2_step(&s, x[p], y[p] * alpha);
        ab += fabs(x[p] * y[p]);
      }
    out[t] = dd_round(s);
    if (absprod) absprod[t] = ab;
  }
}

/* ---------------------------------------------------------------------- */
/* Ozaki-I INT8 slicing.  No reference counterpart: this restates the       */
/* emulated DGEMM kernel of this library.                                   */
/*   op(A)[i,:] = 2^ea_i * x_i,  op(B)[:,j] = 2^fb_j * y_j,  |x|,|y| < 1     */
/*   x = sum_{t=1..s} d_t 2^-7t (+ dropped residual), digits truncated       */
/*   S_w = sum_{t+u=w} sum_p d_t(i,p) e_u(p,j)  for w = 2..s+1 (exact int32) */
/*   acc = 0; for w = s+1 .. 2: acc = acc*2^-7 + S_w   (FP64, one rounding   */
/*   per step; acc*2^-7 is exact)                                           */
/*   C = fma(alpha, ldexp(acc, ea_i + fb_j - 14), c1),                       */
/*   c1 = beta != 1 ? beta*C : C (same rule as the native path).             */
/* K is processed in chunks of oracle_ozaki_kchunk(s) so that every S_w     */
/* fits int32; chunk q > 0 uses c1 = C (the previous chunk's result).       */
/* ---------------------------------------------------------------------- */
int oracle_ozaki_exponent(const double *x, int count, int stride) {
  double amax = 0.0;
  for (int t = 0; t < count; t++) {
    double v = fabs(x[(int64_t)t * stride]);
    if (!(v <= DBL_MAX)) return OZAKI_NONFINITE; /* Inf or NaN */
    if (v > amax) amax = v;
  }
  if (amax == 0.0) return 0;
  int e;
  frexp(amax, &e); /* amax = f*2^e, f in [0.5,1) */
  return e;
}

void oracle_ozaki_digits(double x, int e, int slices, int8_t *d) {
  double r = (e == OZAKI_NONFINITE) ? 0.0 : ldexp(x, -e);
  for (int t = 0; t < slices; t++) {
    double y = r * 128.0;
    double q = trunc(y);
    d[t] = (int8_t)q;
    r = y - q;
  }
}

int oracle_ozaki_kchunk(int slices) {
  int64_t lim = (int64_t)2147483647 / ((int64_t)slices * 127 * 127);
  return (int)(lim - lim % 128);
}

void oracle_ozaki_dgemm(char transa, char transb, int m, int n, int k,
                        double alpha, const double *a, int lda,
                        const double *b, int ldb, double beta, double *c,
                        int ldc, int s) {
  const int na = is_notrans(transa), nb = is_notrans(transb);
  if (m <= 0 || n <= 0) return;
  if (alpha == 0.0 || k == 0) {
    if (beta != 1.0)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) c[i + (int64_t)j * ldc] *= beta;
    return;
  }
  int *ea = (int *)malloc(sizeof(int) * m);
  int *fb = (int *)malloc(sizeof(int) * n);
  /* digits laid out [t][i][p] and [u][j][p] */
  int8_t *da = (int8_t *)malloc((size_t)s * m * k);
  int8_t *db = (int8_t *)malloc((size_t)s * n * k);
  for (int i = 0; i < m; i++)
    ea[i] = na ? oracle_ozaki_exponent(a + i, k, lda)
               : oracle_ozaki_exponent(a + (int64_t)i * lda, k, 1);
  for (int j = 0; j < n; j++)
    fb[j] = nb ? oracle_ozaki_exponent(b + (int64_t)j * ldb, k, 1)
               : oracle_ozaki_exponent(b + j, k, ldb);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m; i++) {
    int8_t dig[16];
    for (int p = 0; p < k; p++) {
      double v = na ? a[i + (int64_t)p * lda] : a[p + (int64_t)i * lda];
      oracle_ozaki_digits(v, ea[i], s, dig);
      for (int t = 0; t < s; t++) da[((size_t)t * m + i) * k + p] = dig[t];
    }
  }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++) {
    int8_t dig[16];
    for (int p = 0; p < k; p++) {
      double v = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
      oracle_ozaki_digits(v, fb[j], s, dig);
      for (int u = 0; u < s; u++) db[((size_t)u * n + j) * k + p] = dig[u];
    }
  }
  const int kc = oracle_ozaki_kchunk(s);
  for (int k0 = 0; k0 < k; k0 += kc) {
    const int k1 = (k0 + kc < k) ? k0 + kc : k;
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < n; j++) {
      for (int i = 0; i < m; i++) {
        double acc = 0.0;
        for (int w = s + 1; w >= 2; w--) {
          int32_t sw = 0;
          for (int t = 1; t <= s; t++) {
            int u = w - t;
            if (u < 1 || u > s) continue;
            const int8_t *pa = da + ((size_t)(t - 1) * m + i) * k;
            const int8_t *pb = db + ((size_t)(u - 1) * n + j) * k;
            int32_t part = 0;
            for (int p = k0; p < k1; p++) part += (int32_t)pa[p] * (int32_t)pb[p];
            sw += part;
          }
          acc = fma(acc, 0x1p-7, (double)sw);
        }
        double *cij = c + i + (int64_t)j * ldc;
        double c1 = (k0 == 0 && beta != 1.0) ? beta * *cij : *cij;
        double scaled;
        if (ea[i] == OZAKI_NONFINITE || fb[j] == OZAKI_NONFINITE)
          scaled = NAN;
        else
          scaled = ldexp(acc, ea[i] + fb[j] - 14);
        *cij = fma(alpha, scaled, c1);
      }
    }
  }
  free(ea);
  free(fb);
  free(da);
  free(db);
}

/* ---------------------------------------------------------------------- */
/* split-BF16x3.  No reference counterpart: restates the emulated SGEMM     */
/* kernel.  x = hi + lo + r with hi = bf16_rne(x), lo = bf16_rne(x - hi)    */
/* (x - hi is exact in FP32); the product keeps hi*hi + hi*lo + lo*hi.      */
/* ---------------------------------------------------------------------- */
uint16_t oracle_bf16_rne(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) /* Inf/NaN: keep, quiet NaN */
    return (uint16_t)((u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

float oracle_bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float x;
  memcpy(&x, &u, 4);
  return x;
}

void oracle_bf16_split(float x, uint16_t *hi, uint16_t *lo) {
  uint16_t h = oracle_bf16_rne(x);
  float r = x - oracle_bf16_to_f32(h);
  *hi = h;
  *lo = oracle_bf16_rne(r);
}

void oracle_bf16x3_sgemm(char transa, char transb, int m, int n, int k,
                         float alpha, const float *a, int lda, const float *b,
                         int ldb, float beta, float *c, int ldc) {
  const int na = is_notrans(transa), nb = is_notrans(transb);
  if (m <= 0 || n <= 0) return;
#pragma omp parallel
  {
    float *bh = (float *)malloc(sizeof(float) * (k > 0 ? k : 1));
    float *bl = (float *)malloc(sizeof(float) * (k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 1)
    for (int j = 0; j < n; j++) {
      for (int p = 0; p < k; p++) {
        float bv = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
        uint16_t h, l;
        oracle_bf16_split(bv, &h, &l);
        bh[p] = oracle_bf16_to_f32(h);
        bl[p] = oracle_bf16_to_f32(l);
      }
      for (int i = 0; i < m; i++) {
        float *cij = c + i + (int64_t)j * ldc;
        float c1 = (beta != 1.0f) ? beta * *cij : *cij;
        if (alpha == 0.0f) {
          *cij = c1;
          continue;
        }
        dd_t s = {0.0, 0.0};
        for (int p = 0; p < k; p++) {
          float av = na ? a[i + (int64_t)p * lda] : a[p + (int64_t)i * lda];
          uint16_t h, l;
          oracle_bf16_split(av, &h, &l);
          double ah = oracle_bf16_to_f32(h), al = oracle_bf16_to_f32(l);
          dot2_step(&s, ah, bh[p]);
          dot2_step(&s, ah, bl[p]);
          dot2_step(&s, al, bh[p]);
        }
        *cij = (float)((double)alpha * dd_round(s) + (double)c1);
      }
    }
    free(bh);
    free(bl);
  }
}

/* ---------------------------------------------------------------------- */
/* Complex GEMM (zgemm_/cgemm_ and the 3m forms of the library).  The      */
/* reference keeps its complex kernels outside its build                   */
/* (OpenBLAS-like_implementation/zgemm_kernel_4x2_haswell.c,                */
/* cgemm_kernel_8x2_haswell.c); this restates the BLAS operation with the   */
/* library's conventions: (re, im) pairs, 'N'/'n' no transpose, 'C'/'c'     */
/* conjugate transpose, other characters transpose, and dgemm_'s beta rule  */
/* (beta != 1 scales C first, also beta = 0).  Re and Im of op(A) op(B) are */
/* Dot2-accumulated (2k products each) and rounded once; then              */
/* C = beta C + alpha P in double.                                           */
/* ---------------------------------------------------------------------- */
static void complex_gemm(char transa, char transb, int m, int n, int k, const double alpha[2],
                         const void *a, int lda, const void *b, int ldb, const double beta[2],
                         void *c, int ldc, int single) {
  const int na = is_notrans(transa), nb = is_notrans(transb);
  const double sa = (transa == 'C' || transa == 'c') ? -1.0 : 1.0;
  const double sb = (transb == 'C' || transb == 'c') ? -1.0 : 1.0;
  if (m <= 0 || n <= 0) return;
  const int product = (alpha[0] != 0.0 || alpha[1] != 0.0) && k > 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < m; i++) {
      double pr = 0.0, pi = 0.0;
      if (product) {
        dd_t re = {0.0, 0.0}, im = {0.0, 0.0};
        for (int p = 0; p < k; p++) {
          const int64_t ia = na ? i + (int64_t)p * lda : p + (int64_t)i * lda;
          const int64_t ib = nb ? p + (int64_t)j * ldb : j + (int64_t)p * ldb;
          double ar, ai, br, bi;
          if (single) {
            ar = ((const float *)a)[2 * ia];
            ai = ((const float *)a)[2 * ia + 1];
            br = ((const float *)b)[2 * ib];
            bi = ((const float *)b)[2 * ib + 1];
          } else {
            ar = ((const double *)a)[2 * ia];
            ai = ((const double *)a)[2 * ia + 1];
            br = ((const double *)b)[2 * ib];
            bi = ((const double *)b)[2 * ib + 1];
          }
          ai *= sa;
          bi *= sb;
          dot2_step(&re, ar, br);
          dot2_step(&re, -ai, bi);
          dot2_step(&im, ar, bi);
          dot2_step(&im, ai, br);
        }
        pr = dd_round(re);
        pi = dd_round(im);
      }
      const int64_t ic = i + (int64_t)j * ldc;
      double cr, ci;
      if (single) {
        cr = ((float *)c)[2 * ic];
        ci = ((float *)c)[2 * ic + 1];
      } else {
        cr = ((double *)c)[2 * ic];
        ci = ((double *)c)[2 * ic + 1];
      }
      if (beta[0] != 1.0 || beta[1] != 0.0) {
        const double nr = beta[0] * cr - beta[1] * ci, ni = beta[0] * ci + beta[1] * cr;
        cr = nr;
        ci = ni;
      }
      if (product) {
        cr += alpha[0] * pr - alpha[1] * pi;
        ci += alpha[0] * pi + alpha[1] * pr;
      }
      if (single) {
        ((float *)c)[2 * ic] = (float)cr;
        ((float *)c)[2 * ic + 1] = (float)ci;
      } else {
        ((double *)c)[2 * ic] = cr;
        ((double *)c)[2 * ic + 1] = ci;
      }
    }
  }
}

void oracle_zgemm(char transa, char transb, int m, int n, int k, const double alpha[2],
                  const double *a, int lda, const double *b, int ldb, const double beta[2],
                  double *c, int ldc) {
  complex_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, 0);
}

void oracle_cgemm(char transa, char transb, int m, int n, int k, const float alpha[2],
                  const float *a, int lda, const float *b, int ldb, const float beta[2],
                  float *c, int ldc) {
  const double al[2] = {alpha[0], alpha[1]}, be[2] = {beta[0], beta[1]};
  complex_gemm(transa, transb, m, n, k, al, a, lda, b, ldb, be, c, ldc, 1);
}

/* ---------------------------------------------------------------------- */
/* Triangular multiply and solve (dtrmm_/dtrsm_, strmm_/strsm_ of the      */
/* library; the reference's own kernels for them live outside its build:   */
/* OpenBLAS-like_implementation/trmm_kernel_4x8_simplified.c,               */
/* TRSM_KERNEL_{LN,LT,RN,RT}_simplified.c).  BLAS operations with the       */
/* library's character rules (side 'L'/'l' left else right, uplo 'U'/'u'   */
/* upper else lower, transa as dgemm_, diag 'U'/'u' unit); alpha = 0 sets B */
/* to zero.  Sums are Dot2-accumulated; the solve is substitution.          */
/* ---------------------------------------------------------------------- */
typedef struct {
  const void *a;
  int64_t lda;
  int upper, trans, unit, single;
} tri_t;

/* op(A)(r, c) within op(A)'s triangle (0 outside it; 1 on a unit diagonal) */
static double tri_at(const tri_t *t, int r, int c) {
  const int sr = t->trans ? c : r, sc = t->trans ? r : c; /* stored indices */
  if (sr == sc && t->unit) return 1.0;
  if (t->upper ? sr > sc : sr < sc) return 0.0;
  const int64_t i = sr + (int64_t)sc * t->lda;
  return t->single ? (double)((const float *)t->a)[i] : ((const double *)t->a)[i];
}

static double bget(const void *b, int single, int64_t i) {
  return single ? (double)((const float *)b)[i] : ((const double *)b)[i];
}
static void bset(void *b, int single, int64_t i, double v) {
  if (single) ((float *)b)[i] = (float)v;
  else ((double *)b)[i] = v;
}

static void tri_op(int solve, char side, char uplo, char transa, char diag, int m, int n,
                   double alpha, const void *a, int lda, void *b, int ldb, int single) {
  if (m <= 0 || n <= 0) return;
  const int left = side == 'L' || side == 'l';
  const tri_t t = {a, lda, uplo == 'U' || uplo == 'u', !is_notrans(transa),
                   diag == 'U' || diag == 'u', single};
  if (alpha == 0.0) {
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) bset(b, single, i + (int64_t)j * ldb, 0.0);
    return;
  }
  const int k = left ? m : n;
  /* op(A) lower: (lower stored, no transpose) or (upper, transposed) */
  const int lower = t.upper ? t.trans : !t.trans;
  if (left) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < n; j++) {
      double *col = (double *)malloc(sizeof(double) * k);
      double *out = (double *)malloc(sizeof(double) * k);
      for (int i = 0; i < k; i++) col[i] = bget(b, single, i + (int64_t)j * ldb);
      if (!solve) {
        for (int i = 0; i < k; i++) {
          dd_t s = {0.0, 0.0};
          for (int q = 0; q < k; q++) {
            const double v = tri_at(&t, i, q);
            if (v != 0.0) dot2_step(&s, v, col[q]);
          }
          out[i] = alpha * dd_round(s);
        }
      } else {
        for (int step = 0; step < k; step++) {
          const int i = lower ? step : k - 1 - step;
          dd_t s = {alpha * col[i], 0.0};
          for (int q = 0; q < k; q++) {
            if (lower ? q >= i : q <= i) continue;
            const double v = tri_at(&t, i, q);
            if (v != 0.0) dot2_step(&s, -v, out[q]);
          }
          out[i] = t.unit ? dd_round(s) : dd_round(s) / tri_at(&t, i, i);
        }
      }
      for (int i = 0; i < k; i++) bset(b, single, i + (int64_t)j * ldb, out[i]);
      free(col);
      free(out);
    }
  } else {
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < m; i++) {
      double *row = (double *)malloc(sizeof(double) * k);
      double *out = (double *)malloc(sizeof(double) * k);
      for (int j = 0; j < k; j++) row[j] = bget(b, single, i + (int64_t)j * ldb);
      if (!solve) {
        for (int j = 0; j < k; j++) {
          dd_t s = {0.0, 0.0};
          for (int q = 0; q < k; q++) {
            const double v = tri_at(&t, q, j);
            if (v != 0.0) dot2_step(&s, row[q], v);
          }
          out[j] = alpha * dd_round(s);
        }
      } else {
        /* x op(A) = alpha b: column j of op(A) has its entries at rows
         * q <= j (upper) or q >= j (lower) */
        for (int step = 0; step < k; step++) {
          const int j = lower ? k - 1 - step : step;
          dd_t s = {alpha * row[j], 0.0};
          for (int q = 0; q < k; q++) {
            if (lower ? q <= j : q >= j) continue;
            const double v = tri_at(&t, q, j);
            if (v != 0.0) dot2_step(&s, -out[q], v);
          }
          out[j] = t.unit ? dd_round(s) : dd_round(s) / tri_at(&t, j, j);
        }
      }
      for (int j = 0; j < k; j++) bset(b, single, i + (int64_t)j * ldb, out[j]);
      free(row);
      free(out);
    }
  }
}

void oracle_trmm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                 const double *a, int lda, double *b, int ldb) {
  tri_op(0, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, 0);
}
void oracle_trsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                 const double *a, int lda, double *b, int ldb) {
  tri_op(1, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, 0);
}
void oracle_trmm_f32(char side, char uplo, char transa, char diag, int m, int n, float alpha,
                     const float *a, int lda, float *b, int ldb) {
  tri_op(0, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, 1);
}
void oracle_trsm_f32(char side, char uplo, char transa, char diag, int m, int n, float alpha,
                     const float *a, int lda, float *b, int ldb) {
  tri_op(1, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, 1);
}

/* ---------------------------------------------------------------------- */
/* Ozaki-II (CRT).  No reference counterpart: restates the emulated DGEMM  */
/* of the library's modular form.  Per row i of op(A) the exponent e_i      */
/* (oracle_ozaki_exponent), the bit count P_i (oracle_ozaki2_vector_bits)   */
/* and the integers a_ip = rint(a * 2^(P_i - e_i)); likewise f_j, P_j and   */
/* b_pj per column j.  The GPU forms the exact integer X_ij = sum_p a_ip    */
/* b_pj from residues modulo pairwise-coprime moduli (INT8 residue products */
/* on the matrix cores, then the Chinese remainder theorem); the bit counts */
/* keep |X| < M/2 by the Cauchy-Schwarz inequality.  This restatement       */
/* computes X directly with 256-bit integer arithmetic instead (no CRT), so */
/* it checks the GPU's reconstruction independently.  The result is RNE(X)  */
/* to 53 bits, scaled by 2^(e_i - P_i + f_j - P_j) with one ldexp.          */
/* ---------------------------------------------------------------------- */
typedef __int128 i128;
typedef unsigned __int128 u128;

int oracle_ozaki2_modulus(int j) {
  /* the j-th of the pairwise-coprime integers taken greedily downward from
   * 256 */
  static int list[OZAKI2_MAX_MODULI];
  static int count;
  if (!count) {
    for (int c = 256; c > 1 && count < OZAKI2_MAX_MODULI; c--) {
      int ok = 1;
      for (int t = 0; t < count && ok; t++) {
        int x = c, y = list[t];
        while (y) {
          int r = x % y;
          x = y;
          y = r;
        }
        ok = (x == 1);
      }
      if (ok) list[count++] = c;
    }
  }
  return (j >= 0 && j < count) ? list[j] : -1;
}

/* The bit-count budget of the first nmod moduli: T = 2^19 Mr with
 * Mr = M - floor(M / 2^10) - 1, M the product of the moduli, rounded down to
 * its top 64 bits, T >= t64 2^e (e = bitlen(T) - 64).  Computed with 32-bit
 * limbs, one bit at a time. */
static void ozaki2_budget(int nmod, uint64_t *t64, int *e) {
  uint64_t limb[8] = {1};
  for (int j = 0; j < nmod; j++) {
    uint64_t carry = 0;
    for (int t = 0; t < 8; t++) {
      const uint64_t v = limb[t] * (uint64_t)oracle_ozaki2_modulus(j) + carry;
      limb[t] = v & 0xffffffffu;
      carry = v >> 32;
    }
  }
  /* Mr = M - (M >> 10) - 1, limb by limb with borrows */
  uint64_t q[8];
  for (int t = 0; t < 8; t++) q[t] = ((limb[t] >> 10) | (t < 7 ? limb[t + 1] << 22 : 0)) & 0xffffffffu;
  int64_t borrow = 1;
  for (int t = 0; t < 8; t++) {
    int64_t v = (int64_t)limb[t] - (int64_t)q[t] - borrow;
    borrow = v < 0;
    limb[t] = (uint64_t)(v + (borrow ? ((int64_t)1 << 32) : 0));
  }
  int bl = 0;
  for (int bit = 255; bit >= 0; bit--)
    if ((limb[bit / 32] >> (bit % 32)) & 1) {
      bl = bit + 1;
      break;
    }
  uint64_t t = 0;
  for (int bit = bl - 1; bit >= bl - 64; bit--) t = (t << 1) | ((limb[bit / 32] >> (bit % 32)) & 1);
  *t64 = t;
  *e = bl - 64 + 19;
}

/* P for a vector whose bound sum is S (see oracle_ozaki2_vector_bits): the
 * largest P <= OZAKI2_MAX_BITS with S 4^P <= t64 2^e.  Then |a_p| <=
 * t_p 2^(P-10) (P >= 10) gives ||a||^2 <= 2^(2P-20) S <= Mr / 2 and, for two
 * vectors, |sum_p a_p b_p| <= ||a|| ||b|| <= Mr / 2 < M / 2. */
static int ozaki2_bits_for_sum(int nmod, uint64_t S) {
  uint64_t t64;
  int e;
  ozaki2_budget(nmod, &t64, &e);
  for (int P = OZAKI2_MAX_BITS; P > 0; P--) {
    const int d = 2 * P - e;
    if (d <= 0) return P; /* S < 2^52 <= t64 2^(e - 2P) */
    if (d < 64 && ((u128)S << d) <= (u128)t64) return P;
  }
  return 0;
}

int oracle_ozaki2_bits(int nmod, int k) {
  if (nmod < OZAKI2_MIN_MODULI || nmod > OZAKI2_MAX_MODULI || k < 1) return -1;
  /* every t_p = 1024 */
  return ozaki2_bits_for_sum(nmod, (uint64_t)k << 20);
}

int oracle_ozaki2_vector_bits(const double *x, int count, int stride, int nmod) {
  if (nmod < OZAKI2_MIN_MODULI || nmod > OZAKI2_MAX_MODULI) return -1;
  const int e = oracle_ozaki_exponent(x, count, stride);
  uint64_t S = 0;
  if (e != OZAKI_NONFINITE)
    for (int p = 0; p < count; p++) {
      /* t = ceil(|x| 2^(10-e)) <= 1024: the scaling and ceil are exact */
      const uint64_t t = (uint64_t)ceil(ldexp(fabs(x[(int64_t)p * stride]), 10 - e));
      S += t * t;
    }
  return ozaki2_bits_for_sum(nmod, S);
}

static i128 ozaki2_scaled(double x, int e, int P) {
  if (x == 0.0) return 0;
  uint64_t bits;
  memcpy(&bits, &x, 8);
  const int neg = (int)(bits >> 63);
  const int be = (int)((bits >> 52) & 0x7ff);
  uint64_t mant = bits & 0xfffffffffffffull;
  int q;
  if (be == 0) {
    q = -1074; /* subnormal: x = mant * 2^-1074 */
  } else {
    mant |= 1ull << 52;
    q = be - 1075;
  }
  const int sh = q + P - e; /* x * 2^(P-e) = mant * 2^sh */
  u128 mag;
  if (sh >= 0) {
    mag = (u128)mant << sh;
  } else {
    const int s = -sh;
    if (s > 63) {
      mag = 0; /* mant < 2^53 <= 2^(s-1): rounds to 0 */
    } else {
      uint64_t qv = mant >> s, rem = mant & ((1ull << s) - 1), half = 1ull << (s - 1);
      if (rem > half || (rem == half && (qv & 1))) qv++;
      mag = qv;
    }
  }
  return neg ? -(i128)mag : (i128)mag;
}

void oracle_ozaki2_scaled(double x, int e, int P, uint64_t *lo, int64_t *hi) {
  const i128 v = ozaki2_scaled(x, e, P);
  *lo = (uint64_t)(u128)v;
  *hi = (int64_t)(v >> 64);
}

/* 256-bit two's complement accumulator, little-endian 64-bit limbs */
static void acc_add_u128_shifted(uint64_t acc[4], u128 v, int word, int negate) {
  uint64_t add[4] = {0, 0, 0, 0};
  if (word < 4) add[word] = (uint64_t)v;
  if (word + 1 < 4) add[word + 1] = (uint64_t)(v >> 64);
  if (negate) {
    /* acc -= add */
    uint64_t borrow = 0;
    for (int t = 0; t < 4; t++) {
      const u128 d = (u128)acc[t] - add[t] - borrow;
      acc[t] = (uint64_t)d;
      borrow = (uint64_t)(d >> 64) ? 1 : 0;
    }
  } else {
    uint64_t carry = 0;
    for (int t = 0; t < 4; t++) {
      const u128 s = (u128)acc[t] + add[t] + carry;
      acc[t] = (uint64_t)s;
      carry = (uint64_t)(s >> 64);
    }
  }
}

static void acc_add_product(uint64_t acc[4], i128 a, i128 b) {
  if (a == 0 || b == 0) return;
  const int neg = (a < 0) != (b < 0);
  const u128 ua = a < 0 ? (u128)(-a) : (u128)a, ub = b < 0 ? (u128)(-b) : (u128)b;
  const uint64_t a0 = (uint64_t)ua, a1 = (uint64_t)(ua >> 64);
  const uint64_t b0 = (uint64_t)ub, b1 = (uint64_t)(ub >> 64);
  acc_add_u128_shifted(acc, (u128)a0 * b0, 0, neg);
  acc_add_u128_shifted(acc, (u128)a0 * b1, 1, neg);
  acc_add_u128_shifted(acc, (u128)a1 * b0, 1, neg);
  acc_add_u128_shifted(acc, (u128)a1 * b1, 2, neg);
}

double oracle_ozaki2_round(const uint64_t limbs[4], int scale) {
  uint64_t v[4] = {limbs[0], limbs[1], limbs[2], limbs[3]};
  const int neg = (int)(v[3] >> 63);
  if (neg) { /* magnitude = -v */
    uint64_t carry = 1;
    for (int t = 0; t < 4; t++) {
      const u128 s = (u128)(~v[t]) + carry;
      v[t] = (uint64_t)s;
      carry = (uint64_t)(s >> 64);
    }
  }
  int top = -1;
  for (int t = 3; t >= 0; t--)
    if (v[t]) {
      top = 64 * t + 63 - __builtin_clzll(v[t]);
      break;
    }
  if (top < 0) return 0.0;
  uint64_t mant;
  int shift = top - 52 > 0 ? top - 52 : 0;
  if (shift == 0) {
    mant = v[0];
  } else {
    /* mant = v >> shift (53 bits), rem = low shift bits compared with half */
    const int w = shift / 64, b = shift % 64;
    mant = v[w] >> b;
    if (b && w + 1 < 4) mant |= v[w + 1] << (64 - b);
    mant &= (1ull << 53) - 1;
    const int hw = (shift - 1) / 64, hb = (shift - 1) % 64; /* the half bit */
    const int half_bit = (int)((v[hw] >> hb) & 1);
    int sticky = 0;
    for (int t = 0; t < hw; t++) sticky |= v[t] != 0;
    if (hb) sticky |= (v[hw] & ((1ull << hb) - 1)) != 0;
    if (half_bit && (sticky || (mant & 1))) {
      mant++;
      if (mant == (1ull << 53)) {
        mant >>= 1;
        shift++;
      }
    }
  }
  const double r = ldexp((double)mant, shift + scale);
  return neg ? -r : r;
}

void oracle_ozaki2_dgemm(char transa, char transb, int m, int n, int k,
                         double alpha, const double *a, int lda,
                         const double *b, int ldb, double beta, double *c,
                         int ldc, int nmod) {
  const int na = is_notrans(transa), nb = is_notrans(transb);
  if (m <= 0 || n <= 0) return;
  if (alpha == 0.0 || k == 0) {
    if (beta != 1.0)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) c[i + (int64_t)j * ldc] *= beta;
    return;
  }
  const int kc = k < OZAKI2_KCHUNK ? k : OZAKI2_KCHUNK;
  int *ea = (int *)malloc(sizeof(int) * m), *pa = (int *)malloc(sizeof(int) * m);
  int *fb = (int *)malloc(sizeof(int) * n), *pb = (int *)malloc(sizeof(int) * n);
  i128 *ia = (i128 *)malloc(sizeof(i128) * (size_t)m * k); /* [i][p] */
  i128 *ib = (i128 *)malloc(sizeof(i128) * (size_t)n * k); /* [j][p] */
  for (int i = 0; i < m; i++) {
    const double *row = na ? a + i : a + (int64_t)i * lda;
    const int inc = na ? lda : 1;
    ea[i] = oracle_ozaki_exponent(row, k, inc);
    pa[i] = oracle_ozaki2_vector_bits(row, k, inc, nmod);
  }
  for (int j = 0; j < n; j++) {
    const double *col = nb ? b + (int64_t)j * ldb : b + j;
    const int inc = nb ? 1 : ldb;
    fb[j] = oracle_ozaki_exponent(col, k, inc);
    pb[j] = oracle_ozaki2_vector_bits(col, k, inc, nmod);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m; i++)
    for (int p = 0; p < k; p++) {
      const double v = na ? a[i + (int64_t)p * lda] : a[p + (int64_t)i * lda];
      ia[(size_t)i * k + p] = ea[i] == OZAKI_NONFINITE ? 0 : ozaki2_scaled(v, ea[i], pa[i]);
    }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int p = 0; p < k; p++) {
      const double v = nb ? b[p + (int64_t)j * ldb] : b[j + (int64_t)p * ldb];
      ib[(size_t)j * k + p] = fb[j] == OZAKI_NONFINITE ? 0 : ozaki2_scaled(v, fb[j], pb[j]);
    }
  for (int k0 = 0; k0 < k; k0 += kc) {
    const int k1 = (k0 + kc < k) ? k0 + kc : k;
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < n; j++) {
      for (int i = 0; i < m; i++) {
        uint64_t acc[4] = {0, 0, 0, 0};
        for (int p = k0; p < k1; p++)
          acc_add_product(acc, ia[(size_t)i * k + p], ib[(size_t)j * k + p]);
        double *cij = c + i + (int64_t)j * ldc;
        const double c1 = (k0 == 0 && beta != 1.0) ? beta * *cij : *cij;
        double scaled;
        if (ea[i] == OZAKI_NONFINITE || fb[j] == OZAKI_NONFINITE)
          scaled = NAN;
        else
          scaled = oracle_ozaki2_round(acc, ea[i] - pa[i] + fb[j] - pb[j]);
        *cij = fma(alpha, scaled, c1);
      }
    }
  }
  free(ea);
  free(pa);
  free(fb);
  free(pb);
  free(ia);
  free(ib);
}

/* ---------------------------------------------------------------------- */
/* Sampled entries of the two emulated forms (one K chunk), for products   */
/* too large to restate whole.  Each entry repeats, for its one row and    */
/* column, exactly the steps of oracle_ozaki_dgemm / oracle_ozaki2_dgemm.  */
/* ---------------------------------------------------------------------- */
int oracle_ozaki_pairs(const double *rows, const double *cols, int k, int count,
                       const int *ri, const int *cj, double alpha, const double *c1,
                       int slices, double *out) {
  if (slices < 1 || slices > 16 || k < 1 || k > oracle_ozaki_kchunk(slices)) return -1;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
  for (int t = 0; t < count; t++) {
    const double *x = rows + (size_t)ri[t] * k, *y = cols + (size_t)cj[t] * k;
    const int ex = oracle_ozaki_exponent(x, k, 1), ey = oracle_ozaki_exponent(y, k, 1);
    int8_t *dx = (int8_t *)malloc((size_t)slices * k), *dy = (int8_t *)malloc((size_t)slices * k);
    if (!dx || !dy) {
      bad = 1;
      free(dx);
      free(dy);
      continue;
    }
    int8_t dig[16];
    for (int p = 0; p < k; p++) {
      oracle_ozaki_digits(x[p], ex, slices, dig);
      for (int u = 0; u < slices; u++) dx[(size_t)u * k + p] = dig[u];
      oracle_ozaki_digits(y[p], ey, slices, dig);
      for (int u = 0; u < slices; u++) dy[(size_t)u * k + p] = dig[u];
    }
    double acc = 0.0;
    for (int w = slices + 1; w >= 2; w--) {
      int32_t sw = 0;
      for (int ta = 1; ta <= slices; ta++) {
        const int ub = w - ta;
        if (ub < 1 || ub > slices) continue;
        int32_t part = 0;
        for (int p = 0; p < k; p++)
          part += (int32_t)dx[(size_t)(ta - 1) * k + p] * (int32_t)dy[(size_t)(ub - 1) * k + p];
        sw += part;
      }
      acc = fma(acc, 0x1p-7, (double)sw);
    }
    const double scaled = (ex == OZAKI_NONFINITE || ey == OZAKI_NONFINITE)
                              ? NAN : ldexp(acc, ex + ey - 14);
    out[t] = fma(alpha, scaled, c1[t]);
    free(dx);
    free(dy);
  }
  return bad ? -1 : 0;
}

int oracle_ozaki2_pairs(const double *rows, const double *cols, int k, int count,
                        const int *ri, const int *cj, double alpha, const double *c1,
                        int nmod, double *out) {
  if (nmod < OZAKI2_MIN_MODULI || nmod > OZAKI2_MAX_MODULI || k < 1 || k > OZAKI2_KCHUNK)
    return -1;
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < count; t++) {
    const double *x = rows + (size_t)ri[t] * k, *y = cols + (size_t)cj[t] * k;
    const int ex = oracle_ozaki_exponent(x, k, 1), ey = oracle_ozaki_exponent(y, k, 1);
    const int px = oracle_ozaki2_vector_bits(x, k, 1, nmod);
    const int py = oracle_ozaki2_vector_bits(y, k, 1, nmod);
    double scaled;
    if (ex == OZAKI_NONFINITE || ey == OZAKI_NONFINITE) {
      scaled = NAN;
    } else {
      uint64_t acc[4] = {0, 0, 0, 0};
      for (int p = 0; p < k; p++)
        acc_add_product(acc, ozaki2_scaled(x[p], ex, px), ozaki2_scaled(y[p], ey, py));
      scaled = oracle_ozaki2_round(acc, ex - px + ey - py);
    }
    out[t] = fma(alpha, scaled, c1[t]);
  }
  return 0;
}
