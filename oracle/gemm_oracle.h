/*
 * gemm_oracle.h -- CPU oracle for the SGEMM/DGEMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (DGEMM.so /
 * SGEMM.so) links, loads or calls this code.  Only tests/, the smoke check and
 * the cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * What it restates:
 *   - oracle_dgemm / oracle_sgemm: the semantics of the reference's Fortran-ABI
 *     entry point GEMM_DRIVER (wjc404/GEMM_AVX2 src/gemm_driver.c:118-197):
 *       * only the first character of transa/transb is read, 'N'/'n' is
 *         no-transpose and any other character is transpose (:17-18);
 *       * beta != 1 scales C[0:m,0:n] in place first, including beta == 0,
 *         so NaN/Inf already in C propagate (:106-116, :133);
 *       * alpha != 0 gates the product (:134); alpha is folded into B with one
 *         rounding before the product (src/gemm_copy.c:65, :88, :117, :136);
 *       * rows m..ldc-1 of C are never written (masked stores,
 *         src/kernel/gemm_kernel_unroll6.S:145-155,192-229).
 *     The products are accumulated with Ogita-Rump-Oishi Dot2 (double-double),
 *     so the oracle is *more* accurate than the reference; parity is by
 *     tolerance, not bits (the reference sums in a cyclic blocked order,
 *     src/kernel/gemm_kernel_unroll6.S:70-125).
 *   - the counter-based input generator (SplitMix64 on (seed, id, index)).
 *   - the emulated forms this build adds (no reference counterpart):
 *     Ozaki-I INT8 slicing of FP64, the Ozaki-II CRT form of FP64 and the
 *     split-BF16x3 form of FP32.  These restate the algorithm the GPU kernels
 *     implement, bit for bit where the GPU arithmetic is exact (integer slice
 *     and residue products, FP64 recombination, exact CRT reconstruction).
 *
 * Pinning (see tests/test_oracle.py): golden outputs of the reference built
 * from source (tests/golden/), the closed-form 12x12x12 known-answer test of
 * OpenBLAS-like_implementation/dgemm_kernel_4x4_haswell.c:472-492, brute force
 * in long double on tiny shapes, and invariants (ones, alpha=0, k=0,
 * transpose identities, linearity).
 */
#ifndef GEMM_ORACLE_H
#define GEMM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based generator ------------------------------------------ */
uint64_t oracle_counter_u64(uint64_t seed, uint64_t id, uint64_t idx);
double oracle_uniform_f64(uint64_t seed, uint64_t id, uint64_t idx);   /* U[-1,1) */
float oracle_uniform_f32(uint64_t seed, uint64_t id, uint64_t idx);    /* U[-1,1) */
/* (u-0.5)*exp(phi*g), u~U[0,1), g~N(0,1) by Box-Muller */
double oracle_emu_f64(uint64_t seed, uint64_t id, uint64_t idx, double phi);
/* exp(sigma*g), g~N(0,1) */
float oracle_lognormal_f32(uint64_t seed, uint64_t id, uint64_t idx, double sigma);

/* Fill a rows x cols column-major matrix with leading dimension ld.  Element
 * (i,j) takes counter index i + j*rows; rows..ld-1 of every column are NaN.
 * dist: 0 uniform[-1,1), 1 emu(phi), 2 lognormal(sigma=param). */
void oracle_fill_f64(double *p, int rows, int cols, int ld, uint64_t seed,
                     uint64_t id, int dist, double param);
void oracle_fill_f32(float *p, int rows, int cols, int ld, uint64_t seed,
                     uint64_t id, int dist, double param);

/* ---- reference semantics, Dot2 accumulation --------------------------- */
void oracle_dgemm(const char *transa, const char *transb, const int *m,
                  const int *n, const int *k, const double *alpha,
                  const double *a, const int *lda, const double *b,
                  const int *ldb, const double *beta, double *c,
                  const int *ldc);
void oracle_sgemm(const char *transa, const char *transb, const int *m,
                  const int *n, const int *k, const float *alpha,
                  const float *a, const int *lda, const float *b,
                  const int *ldb, const float *beta, float *c, const int *ldc);

/* Brute force in long double, no alpha pre-rounding, for tiny shapes. */
void oracle_dgemm_longdouble(char transa, char transb, int m, int n, int k,
                             double alpha, const double *a, int lda,
                             const double *b, int ldb, double beta,
                             double *c, int ldc);

/* Sampled entries of C for inputs drawn from the generator (A id 0, B id 1,
 * C id 2, all uniform[-1,1) unless dist/param say otherwise), without
 * materialising the matrices.  out[t] = expected C[is[t], js[t]] after the
 * call, computed with Dot2.  Also returns sum|a||b| per entry in absprod
 * (may be NULL) for componentwise error bounds. */
void oracle_sampled_f64(char transa, char transb, int m, int n, int k,
                        double alpha, double beta, uint64_t seed, int dist,
                        double param, int count, const int *is, const int *js,
                        double *out, double *absprod);
void oracle_sampled_f32(char transa, char transb, int m, int n, int k,
                        float alpha, float beta, uint64_t seed, int dist,
                        double param, int count, const int *is, const int *js,
                        double *out, double *absprod);

/* The same Dot2 entries for explicit operands (whole rows of op(A) and
 * columns of op(B), k contiguous entries each; entry t pairs rows[ri[t]]
 * with cols[cj[t]] and starts from c1[t] = beta-scaled C), for inputs that
 * are not drawn on the host. */
void oracle_dot2_pairs(const double *rows, const double *cols, int k, int count,
                       const int *ri, const int *cj, double alpha, const double *c1,
                       double *out, double *absprod);

/* ---- Ozaki-I INT8 slicing (emulated DGEMM restatement) ---------------- */
/* Exponent of a row/column: the e with max|x| < 2^e, max|x| >= 2^(e-1); 0 for
 * an all-zero vector; OZAKI_NONFINITE when any entry is Inf/NaN. */
#define OZAKI_NONFINITE 0x7fffffff
int oracle_ozaki_exponent(const double *x, int count, int stride);
/* Signed 7-bit digits of x*2^-e (|.| < 1), truncated toward zero:
 * x*2^-e = sum_t d[t]*2^(-7(t+1)) + r, |d[t]| <= 127. */
void oracle_ozaki_digits(double x, int e, int slices, int8_t *d);
/* Largest K chunk (a multiple of 128) whose weight sums stay exact in int32
 * for this slice count. */
int oracle_ozaki_kchunk(int slices);
/* The emulated product, same semantics as oracle_dgemm. */
void oracle_ozaki_dgemm(char transa, char transb, int m, int n, int k,
                        double alpha, const double *a, int lda,
                        const double *b, int ldb, double beta, double *c,
                        int ldc, int slices);

/* ---- complex GEMM (zgemm_/cgemm_ restatement) -------------------------- */
/* C = beta C (beta != 1, also 0) + alpha op(A) op(B), (re, im) pairs, sizes
 * and leading dimensions in complex elements; 'N'/'n' no transpose, 'C'/'c'
 * conjugate transpose, other characters transpose.  Re and Im of the product
 * are Dot2-accumulated and rounded once. */
void oracle_zgemm(char transa, char transb, int m, int n, int k, const double alpha[2],
                  const double *a, int lda, const double *b, int ldb, const double beta[2],
                  double *c, int ldc);
void oracle_cgemm(char transa, char transb, int m, int n, int k, const float alpha[2],
                  const float *a, int lda, const float *b, int ldb, const float beta[2],
                  float *c, int ldc);

/* ---- triangular multiply / solve (dtrmm_/dtrsm_ restatement) ---------- */
/* TRMM: B := alpha op(A) B (side 'L'/'l') or alpha B op(A) (else); TRSM:
 * op(A) X = alpha B or X op(A) = alpha B, X over B.  uplo 'U'/'u' upper
 * stored triangle else lower (the other triangle is never read), transa as
 * for dgemm, diag 'U'/'u' unit (diagonal not read); alpha = 0 sets B to
 * zero.  Dot2-accumulated sums; substitution for the solve.  The _f32 forms
 * read and write float and compute the same way. */
void oracle_trmm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                 const double *a, int lda, double *b, int ldb);
void oracle_trsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                 const double *a, int lda, double *b, int ldb);
void oracle_trmm_f32(char side, char uplo, char transa, char diag, int m, int n, float alpha,
                     const float *a, int lda, float *b, int ldb);
void oracle_trsm_f32(char side, char uplo, char transa, char diag, int m, int n, float alpha,
                     const float *a, int lda, float *b, int ldb);

/* ---- Ozaki-II, CRT over INT8 residues (emulated DGEMM restatement) ---- */
#define OZAKI2_MIN_MODULI 8
#define OZAKI2_MAX_MODULI 22
#define OZAKI2_KCHUNK 65536 /* K chunk: kc * 128^2 < 2^31 */
#define OZAKI2_MAX_BITS 77   /* cap on P: scaled integers below 2^78 */
/* The j-th modulus: pairwise-coprime integers taken greedily downward from
 * 256 (256, 255, 253, 251, 247, ...); -1 past the end. */
int oracle_ozaki2_modulus(int j);
/* Bits P of the scaled integers of a vector of `count` entries (stride
 * `stride`) for nmod moduli: with e its exponent (oracle_ozaki_exponent),
 * t_p = ceil(|x_p| 2^(10-e)) <= 1024 and S = sum_p t_p^2 (exact), the largest
 * P with 2P + bitlen(S - 1) <= floor(log2(M)) + 19, M the product of the
 * moduli, capped at OZAKI2_MAX_BITS.  Any two vectors' integers then have
 * |dot| <= 2^(floor(log2 M) - 1) < M/2 (Cauchy-Schwarz); S and P do not
 * depend on the summation order.  -1 for nmod outside
 * [OZAKI2_MIN_MODULI, OZAKI2_MAX_MODULI]. */
int oracle_ozaki2_vector_bits(const double *x, int count, int stride, int nmod);
/* The smallest vector bit count over all vectors of k entries (every
 * t_p = 1024), -1 as above or for k < 1. */
int oracle_ozaki2_bits(int nmod, int k);
/* rint(x * 2^(P-e)) (ties to even) as a signed 128-bit integer (lo, hi). */
void oracle_ozaki2_scaled(double x, int e, int P, uint64_t *lo, int64_t *hi);
/* A signed 256-bit integer (little-endian two's complement 64-bit limbs)
 * rounded to 53 bits (ties to even) and scaled by 2^scale with one ldexp. */
double oracle_ozaki2_round(const uint64_t limbs[4], int scale);
/* The emulated product, same semantics as oracle_ozaki_dgemm, computed with
 * exact 256-bit integer dot products (no residues). */
void oracle_ozaki2_dgemm(char transa, char transb, int m, int n, int k,
                         double alpha, const double *a, int lda,
                         const double *b, int ldb, double beta, double *c,
                         int ldc, int nmod);

/* ---- sampled entries of the emulated forms ---------------------------- */
/* alpha * (op(A) op(B))[i,j] + c1 at `count` entries of a product too large
 * to restate whole.  rows holds nr full rows of op(A) and cols nc full
 * columns of op(B), each k contiguous entries (whole vectors: the exponents
 * and bit counts depend on every entry); entry t pairs rows[ri[t]] with
 * cols[cj[t]] and adds c1[t] (C already multiplied by beta).  The arithmetic
 * is that of oracle_ozaki_dgemm / oracle_ozaki2_dgemm for one K chunk, so
 * every out[t] equals the whole-matrix restatement's entry bit for bit.
 * Returns 0, or -1 when k exceeds one K chunk or the count is out of range. */
int oracle_ozaki_pairs(const double *rows, const double *cols, int k, int count,
                       const int *ri, const int *cj, double alpha, const double *c1,
                       int slices, double *out);
int oracle_ozaki2_pairs(const double *rows, const double *cols, int k, int count,
                        const int *ri, const int *cj, double alpha, const double *c1,
                        int nmod, double *out);

/* ---- split-BF16x3 (emulated SGEMM restatement) ------------------------ */
uint16_t oracle_bf16_rne(float x);
float oracle_bf16_to_f32(uint16_t h);
void oracle_bf16_split(float x, uint16_t *hi, uint16_t *lo);
/* C = alpha*(Ah*Bh + Ah*Bl + Al*Bh) + beta*C with exact (Dot2) accumulation
 * of the three bf16 products; same beta/alpha semantics as oracle_sgemm
 * except that alpha is applied after the product, as the GPU form does. */
void oracle_bf16x3_sgemm(char transa, char transb, int m, int n, int k,
                         float alpha, const float *a, int lda, const float *b,
                         int ldb, float beta, float *c, int ldc);

#ifdef __cplusplus
}
#endif
#endif
