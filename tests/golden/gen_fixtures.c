/*
 * gen_fixtures.c -- writes the golden vectors of tests/golden/ from the
 * reference's own sgemm_/dgemm_ (wjc404/GEMM_AVX2 src/gemm_driver.c:118),
 * built from its sources by oracle/Makefile into oracle/_ref/.
 *
 * Usage (see make_fixtures.sh):  gen_fixtures <DGEMM.so> <SGEMM.so> <outdir>
 *
 * Inputs come from the counter-based generator of the oracle (A id 0, B id 1,
 * C id 2; element (r,s) of a stored rows x cols matrix has counter index
 * r + s*rows; padding rows hold NaN).  Outputs:
 *   manifest.json      one record per case (shape, flags, scalars, seed, kind,
 *                      offsets, Frobenius norm of the reference C, whether
 *                      the reference touched the ld padding)
 *   values.f64         little-endian float64 values (FP32 results widened)
 *   index.i32          little-endian int32 (i, j) pairs for sampled cases
 * Only data travels with the repository; the reference itself does not.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/gemm_oracle.h"

typedef void (*dgemm_fn)(const char *, const char *, const int *, const int *,
                         const int *, const double *, const double *,
                         const int *, const double *, const int *,
                         const double *, double *, const int *);
typedef void (*sgemm_fn)(const char *, const char *, const int *, const int *,
                         const int *, const float *, const float *,
                         const int *, const float *, const int *,
                         const float *, float *, const int *);

static dgemm_fn ref_dgemm;
static sgemm_fn ref_sgemm;
static FILE *fvals, *fidx, *fman;
static long nvals, nidx;
static int ncases;

static void pick_samples(int m, int n, int count, uint64_t salt, int *is, int *js) {
  int t = 0;
  int corner_i[4] = {0, m - 1, 0, m - 1}, corner_j[4] = {0, n - 1, n - 1, 0};
  for (; t < 4 && t < count; t++) {
    is[t] = corner_i[t];
    js[t] = corner_j[t];
  }
  for (; t < count; t++) {
    uint64_t x = oracle_counter_u64(salt, 99, (uint64_t)t);
    is[t] = (int)((x & 0xffffffffu) % (uint64_t)m);
    js[t] = (int)((x >> 32) % (uint64_t)n);
  }
}

/* dtype 0 = FP64, 1 = FP32.  kind: full (count = 0) or sampled (count > 0). */
static void run_case(const char *group, int dtype, char ta, char tb, int m,
                     int n, int k, int pad, double alpha, double beta,
                     uint64_t seed, int dist, double param, int count) {
  const int na = (ta == 'N' || ta == 'n'), nb = (tb == 'N' || tb == 'n');
  const int ra = na ? m : k, ca = na ? k : m, rb = nb ? k : n, cb = nb ? n : k;
  /* pad >= 0: ld = rows + pad; pad < 0: every ld is -pad */
  const int lda = pad < 0 ? -pad : (ra > 0 ? ra : 1) + pad;
  const int ldb = pad < 0 ? -pad : (rb > 0 ? rb : 1) + pad;
  const int ldc = pad < 0 ? -pad : (m > 0 ? m : 1) + pad;
  const size_t es = dtype == 0 ? 8 : 4;
  void *A = malloc(es * (size_t)lda * (ca > 0 ? ca : 1));
  void *B = malloc(es * (size_t)ldb * (cb > 0 ? cb : 1));
  void *C = malloc(es * (size_t)ldc * (n > 0 ? n : 1));
  if (dtype == 0) {
    oracle_fill_f64(A, ra, ca, lda, seed, 0, dist, param);
    oracle_fill_f64(B, rb, cb, ldb, seed, 1, dist, param);
    oracle_fill_f64(C, m, n, ldc, seed, 2, dist, param);
    ref_dgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
  } else {
    float al = (float)alpha, be = (float)beta;
    oracle_fill_f32(A, ra, ca, lda, seed, 0, dist, param);
    oracle_fill_f32(B, rb, cb, ldb, seed, 1, dist, param);
    oracle_fill_f32(C, m, n, ldc, seed, 2, dist, param);
    ref_sgemm(&ta, &tb, &m, &n, &k, &al, A, &lda, B, &ldb, &be, C, &ldc);
  }
  double fro = 0.0;
  int pad_touched = 0;
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < ldc; i++) {
      double v = dtype == 0 ? ((double *)C)[i + (size_t)j * ldc]
                            : (double)((float *)C)[i + (size_t)j * ldc];
      if (i < m) fro += v * v;
      else if (!isnan(v)) pad_touched = 1;
    }
  }
  fro = sqrt(fro);
  long voff = nvals, ioff = nidx;
  int stored;
  if (count == 0) {
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) {
        double v = dtype == 0 ? ((double *)C)[i + (size_t)j * ldc]
                              : (double)((float *)C)[i + (size_t)j * ldc];
        fwrite(&v, 8, 1, fvals);
      }
    stored = m * n;
    nvals += stored;
  } else {
    int *is = malloc(sizeof(int) * count), *js = malloc(sizeof(int) * count);
    pick_samples(m, n, count, seed * 131 + (uint64_t)(m ^ (n << 8) ^ (k << 16)), is, js);
    for (int t = 0; t < count; t++) {
      double v = dtype == 0 ? ((double *)C)[is[t] + (size_t)js[t] * ldc]
                            : (double)((float *)C)[is[t] + (size_t)js[t] * ldc];
      int32_t ij[2] = {is[t], js[t]};
      fwrite(&v, 8, 1, fvals);
      fwrite(ij, 4, 2, fidx);
    }
    stored = count;
    nvals += count;
    nidx += 2L * count;
    free(is);
    free(js);
  }
  fprintf(fman,
          "%s  {\"group\": \"%s\", \"dtype\": \"%s\", \"transa\": \"%c\", "
          "\"transb\": \"%c\", \"m\": %d, \"n\": %d, \"k\": %d, \"lda\": %d, "
          "\"ldb\": %d, \"ldc\": %d, \"alpha\": %.17g, \"beta\": %.17g, "
          "\"seed\": %llu, \"dist\": %d, \"param\": %.17g, \"kind\": \"%s\", "
          "\"count\": %d, \"voff\": %ld, \"ioff\": %ld, \"fro\": %.17g, "
          "\"pad_touched\": %d}",
          ncases ? ",\n" : "", group, dtype == 0 ? "f64" : "f32", ta, tb, m, n,
          k, lda, ldb, ldc, alpha, beta, (unsigned long long)seed, dist, param,
          count == 0 ? "full" : "sampled", stored, voff, ioff, fro,
          pad_touched);
  ncases++;
  fflush(fman);
  fprintf(stderr, "case %d %s %s %c%c %dx%dx%d done\n", ncases, group,
          dtype == 0 ? "f64" : "f32", ta, tb, m, n, k);
  free(A);
  free(B);
  free(C);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s DGEMM.so SGEMM.so outdir [big]\n", argv[0]);
    return 2;
  }
  void *hd = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
  void *hs = dlopen(argv[2], RTLD_NOW | RTLD_LOCAL);
  if (!hd || !hs) {
    fprintf(stderr, "dlopen failed: %s\n", dlerror());
    return 1;
  }
  ref_dgemm = (dgemm_fn)dlsym(hd, "dgemm_");
  ref_sgemm = (sgemm_fn)dlsym(hs, "sgemm_");
  const int big = argc > 4 && atoi(argv[4]) != 0;
  char path[4096];
  snprintf(path, sizeof path, "%s/values.f64", argv[3]);
  fvals = fopen(path, "wb");
  snprintf(path, sizeof path, "%s/index.i32", argv[3]);
  fidx = fopen(path, "wb");
  snprintf(path, sizeof path, "%s/manifest.json", argv[3]);
  fman = fopen(path, "w");
  if (!fvals || !fidx || !fman) return 1;
  fprintf(fman, "[\n");

  static const char tr[4][2] = {{'N', 'N'}, {'N', 'T'}, {'T', 'N'}, {'T', 'T'}};
  static const double ab[4][2] = {{1.0, 0.0}, {1.5, -0.5}, {0.0, 2.0}, {1.0, 1.0}};
  /* F1: small and odd shapes, full C, every trans pair and (alpha, beta). */
  static const int small[][3] = {{1, 1, 1},   {2, 3, 4},   {7, 5, 3},
                                 {8, 8, 8},   {9, 7, 13},  {12, 12, 12},
                                 {16, 16, 16}, {17, 15, 33}, {31, 33, 32},
                                 {1, 40, 17}, {33, 1, 64},  {29, 31, 1},
                                 {5, 6, 0}};
  for (size_t s = 0; s < sizeof small / sizeof small[0]; s++)
    for (int t = 0; t < 4; t++)
      for (int q = 0; q < 4; q++)
        for (int dt = 0; dt < 2; dt++)
          run_case("small", dt, tr[t][0], tr[t][1], small[s][0], small[s][1],
                   small[s][2], 3, ab[q][0], ab[q][1], 10 + s, 0, 0.0, 0);
  /* F1b: block-edge shapes around the reference's Mr/Nb/Kb (8/96/288 FP64,
   * 16/192/336 FP32), sampled entries. */
  static const int medium[][3] = {{63, 65, 64},   {95, 97, 96},
                                  {97, 96, 289},  {193, 191, 337},
                                  {257, 193, 577}, {130, 260, 1000}};
  for (size_t s = 0; s < sizeof medium / sizeof medium[0]; s++)
    for (int t = 0; t < 4; t++)
      for (int dt = 0; dt < 2; dt++)
        run_case("medium", dt, tr[t][0], tr[t][1], medium[s][0], medium[s][1],
                 medium[s][2], 5, 1.5, -0.5, 40 + s, 0, 0.0, 512);
  /* F2: DGEMM NN 1024^3, ld 1024, alpha 1, beta 0, seed 0. */
  run_case("nn1024", 0, 'N', 'N', 1024, 1024, 1024, 0, 1.0, 0.0, 0, 0, 0.0, 4096);
  if (big) {
    /* F3: 8192^3 ld 8200, alpha 1.5, beta -0.5, every trans, S and D, plus
     * the ragged 8191 x 8189 x 8187 shape (ld 8200 in both). */
    for (int dt = 0; dt < 2; dt++) {
      for (int t = 0; t < 4; t++)
        run_case("sq8192", dt, tr[t][0], tr[t][1], 8192, 8192, 8192, -8200, 1.5,
                 -0.5, 1, 0, 0.0, 2048);
      run_case("ragged8192", dt, 'N', 'N', 8191, 8189, 8187, -8200, 1.5, -0.5, 1,
               0, 0.0, 2048);
    }
    /* F4: DGEMM 8192^3 NN alpha 1 beta 0, emulation inputs phi in {0,1,4}. */
    static const double phis[3] = {0.0, 1.0, 4.0};
    for (int p = 0; p < 3; p++)
      run_case("emu8192", 0, 'N', 'N', 8192, 8192, 8192, 0, 1.0, 0.0, 3, 1,
               phis[p], 2048);
  }
  fprintf(fman, "\n]\n");
  fclose(fman);
  fclose(fvals);
  fclose(fidx);
  fprintf(stderr, "%d cases, %ld values, %ld index words\n", ncases, nvals, nidx);
  return 0;
}
