/*
 * gemm_mi355x.h -- public C ABI of the MI355X (gfx950) SGEMM/DGEMM library.
 *
 * Shipped as two shared objects, like the reference (wjc404/GEMM_AVX2
 * Makefile:12-18, README.md:10):
 *   DGEMM.so  exports dgemm_  and the dgemm_x_* extensions
 *   SGEMM.so  exports sgemm_  and the sgemm_x_* extensions
 * Link against or LD_PRELOAD either one.  All matrices are column-major.
 *
 * The two Fortran-ABI symbols are drop-ins for the reference's; everything
 * named *_x_* is an extension with no reference counterpart.
 */
#ifndef GEMM_MI355X_H
#define GEMM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Drop-in entry points.
 *
 * Replace GEMM_DRIVER = dgemm_/sgemm_ of wjc404/GEMM_AVX2
 * src/gemm_driver.c:118 (built twice, -DDOUBLE selecting dgemm_, :8-14).
 * Same Fortran calling convention (every argument by pointer, LP64 int, no
 * hidden string lengths) and the same semantics:
 *   - only *transa / *transb are read; 'N'/'n' = no transpose, any other
 *     character = transpose (src/gemm_driver.c:17-18);
 *   - if *beta != 1, C[0:m,0:n] is scaled by beta first, also for beta == 0,
 *     so NaN/Inf in C propagate (:106-116, :133);
 *   - if *alpha != 0, C += alpha * op(A) * op(B) (:134);
 *   - rows m..ldc-1 of C are never written;
 *   - C is complete when the call returns (synchronous).
 * A, B and C may be host memory (staged to the GPU and back) or device
 * memory of the current HIP device (used in place), independently.
 * Additive behaviour: invalid arguments (negative sizes, ld too small, NULL
 * operands that would be read) and GPU errors leave C untouched and are
 * reported through *_x_last_error() and, once per process, on stderr
 * (silenced by GEMMX_QUIET=1).  The reference does not check arguments.
 * ------------------------------------------------------------------------ */
void dgemm_(const char *transa, const char *transb, const int *m, const int *n,
            const int *k, const double *alpha, const double *a, const int *lda,
            const double *b, const int *ldb, const double *beta, double *c,
            const int *ldc);
void sgemm_(const char *transa, const char *transb, const int *m, const int *n,
            const int *k, const float *alpha, const float *a, const int *lda,
            const float *b, const int *ldb, const float *beta, float *c,
            const int *ldc);

/* Complex GEMM, the level-3 complex siblings that the reference keeps outside
 * its build as OpenBLAS kernels (OpenBLAS-like_implementation/
 * zgemm_kernel_4x2_haswell.c, cgemm_kernel_8x2_haswell.c; three-product
 * forms zgemm3m_kernel_4x4_haswell.c, cgemm3m_kernel_8x4_haswell.c), with
 * the BLAS argument order: C = alpha op(A) op(B) + beta C, complex elements
 * and scalars stored as (re, im) pairs, sizes and leading dimensions in
 * complex elements.  'N'/'n' = no transpose, 'C'/'c' = conjugate transpose,
 * any other character = transpose; the rest of dgemm_'s semantics as above
 * (beta != 1 scales C first, also beta = 0; alpha = 0 skips the product;
 * rows m..ldc-1 untouched; synchronous; host or device operands).
 * zgemm_/cgemm_ use four real MFMA products (Re = ArBr - AiBi,
 * Im = ArBi + AiBr); zgemm3m_/cgemm3m_ three ((Ar + Ai)(Br + Bi) - ArBr -
 * AiBi for Im), faster, with a larger error when real and imaginary parts
 * cancel. */
void zgemm_(const char *transa, const char *transb, const int *m, const int *n,
            const int *k, const double *alpha, const double *a, const int *lda,
            const double *b, const int *ldb, const double *beta, double *c,
            const int *ldc);
void zgemm3m_(const char *transa, const char *transb, const int *m, const int *n,
              const int *k, const double *alpha, const double *a, const int *lda,
              const double *b, const int *ldb, const double *beta, double *c,
              const int *ldc);
void cgemm_(const char *transa, const char *transb, const int *m, const int *n,
            const int *k, const float *alpha, const float *a, const int *lda,
            const float *b, const int *ldb, const float *beta, float *c,
            const int *ldc);
void cgemm3m_(const char *transa, const char *transb, const int *m, const int *n,
              const int *k, const float *alpha, const float *a, const int *lda,
              const float *b, const int *ldb, const float *beta, float *c,
              const int *ldc);

/* Triangular multiply (TRMM: B := alpha op(A) B or alpha B op(A)) and solve
 * (TRSM: op(A) X = alpha B or X op(A) = alpha B, X overwriting B), the
 * triangular level-3 siblings the reference keeps outside its build as
 * OpenBLAS kernels (OpenBLAS-like_implementation/trmm_kernel_4x8_simplified.c,
 * strsm_kernel_8x4_haswell_{LN,LT,RN,RT}.c), with the BLAS argument order.
 * side 'L'/'l' = op(A) on the left (A is m x m), anything else = right (A is
 * n x n); uplo 'U'/'u' = A's upper triangle is stored, else the lower; the
 * other triangle is never used; transa as for dgemm_; diag 'U'/'u' = unit
 * diagonal (not read); alpha = 0 sets B to zero without reading A or B (as
 * BLAS specifies).  Synchronous; host or device operands; errors as for
 * dgemm_.  TRMM runs one MFMA GEMM on a dense copy of the triangle; TRSM
 * inverts 64 x 64 diagonal blocks and updates the rest with MFMA GEMMs. */
void dtrmm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const double *alpha, const double *a, const int *lda,
            double *b, const int *ldb);
void dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const double *alpha, const double *a, const int *lda,
            double *b, const int *ldb);
void strmm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const float *alpha, const float *a, const int *lda,
            float *b, const int *ldb);
void strsm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const int *m, const int *n, const float *alpha, const float *a, const int *lda,
            float *b, const int *ldb);

/* ILP64 forms of the same entry points (no counterpart in the reference,
 * which is LP64 only: src/gemm_driver.c:118 takes int): every integer
 * argument is a 64-bit integer, as for Fortran callers built with 64-bit
 * default integers against an ILP64 BLAS ("_64_" symbol suffix).  Same
 * semantics as dgemm_/sgemm_; leading dimensions are used at full width,
 * and m, n, k above 2^31 - 1 are rejected (GEMMX_ERR_ARG, C untouched). */
void dgemm_64_(const char *transa, const char *transb, const int64_t *m,
               const int64_t *n, const int64_t *k, const double *alpha,
               const double *a, const int64_t *lda, const double *b,
               const int64_t *ldb, const double *beta, double *c,
               const int64_t *ldc);
void sgemm_64_(const char *transa, const char *transb, const int64_t *m,
               const int64_t *n, const int64_t *k, const float *alpha,
               const float *a, const int64_t *lda, const float *b,
               const int64_t *ldb, const float *beta, float *c,
               const int64_t *ldc);

/* ------------------------------------------------------------------------
 * Extensions (no reference counterpart).
 * ------------------------------------------------------------------------ */

/* Error codes returned by the extension entry points. */
#define GEMMX_OK 0
#define GEMMX_ERR_ARG 1      /* invalid argument */
#define GEMMX_ERR_NODEVICE 2 /* no HIP device */
#define GEMMX_ERR_ALLOC 3    /* device or pinned allocation failed */
#define GEMMX_ERR_LAUNCH 4   /* kernel launch or HIP runtime error */
#define GEMMX_ERR_UNSUPPORTED 5

/* Arithmetic used by dgemm_ / sgemm_ (process-wide, default native).
 * Also set by the environment at first use:
 *   GEMMX_DGEMM_MODE = native | ozaki[:slices] (slices 3..8, default 8)
 *                    | ozaki2[:moduli]         (moduli 8..22, default 20)
 *                    | auto  (the tuning table's `emu` entry of the shape's
 *                             bucket, written by the autotuner; native when
 *                             there is none)
 *   GEMMX_SGEMM_MODE = native | bf16x3
 *                    | auto  (the tuning table's `emu32` entry of the
 *                             shape's bucket; native when there is none)
 * For GEMMX_MODE_OZAKI2 the second argument of dgemm_x_set_mode is the
 * modulus count. */
#define GEMMX_MODE_NATIVE 0
#define GEMMX_MODE_OZAKI 1  /* FP64 by INT8 slices on INT8 MFMA */
#define GEMMX_MODE_BF16X3 2 /* FP32 by split-BF16x3 on BF16 MFMA */
#define GEMMX_MODE_OZAKI2 3 /* FP64 by INT8 residues (CRT) on INT8 MFMA */
#define GEMMX_MODE_AUTO 4   /* per-shape choice from the tuning table */
int dgemm_x_set_mode(int mode, int slices);
int sgemm_x_set_mode(int mode);

/* Tile configuration of the native kernels, in place of the reference's
 * compile-time tuning headers (dgemm_tune.h / sgemm_tune.h, clamped in
 * src/gemm_set_parameters.h:10-61): cfg in 0..GEMMX_TILE_CONFIGS-1 forces one
 * configuration for every call, -1 returns to the tuning table.
 * *_x_load_tuning reads a table written by the autotuner (path; NULL clears
 * it) and returns the number of entries read, or -1 (previous table kept)
 * when the file is unreadable, its format version, arch or ROCm release
 * does not match, or it names no ROCm release; the table is shared by both
 * element types.  GEMMX_TUNE_FILE (else gemm_mi355x_tune.txt beside the
 * library), GEMMX_NATIVE_CFG and GEMMX_NATIVE_SPLIT are read at first use.
 * *_x_tune_lookup returns the configuration a call with these arguments
 * would use (no device needed), or -1 for negative sizes. */
#define GEMMX_TILE_CONFIGS 10
int dgemm_x_set_tile_config(int cfg);
int sgemm_x_set_tile_config(int cfg);
/* Split-K of the native kernels, in place of the reference's N-split of
 * every call over all its threads (src/gemm_driver.c:128-151): a product
 * with few C tiles divides K into `split` ranges computed by separate
 * workgroups into partial tiles, which a second kernel sums in range order
 * (the result does not depend on scheduling) and combines with alpha and
 * beta*C.  split >= 1 (1..64) forces a factor for every later call of the
 * element type, -1 returns to the tuning table's factor or, without one,
 * a heuristic (about four workgroups per CU when the tiles give fewer than
 * two, K ranges of at least 256, partial tiles within 64 MiB).  Only tile
 * configurations 0, 1, 5, 7, 8 and 9 have split-K kernels; a strided batch is never
 * split.  *_x_tune_lookup_split returns the number of K ranges a call with
 * these arguments runs (1 = one pass) and, in *bytes (may be NULL), the
 * bytes of partial tiles: the Fortran-ABI entries take them from the
 * library's device scratch, the device-pointer entries from a stream-ordered
 * pool of the current device (hipMallocFromPoolAsync on the call's stream,
 * released to the pool after the reduction; when that memory is not
 * available the call runs in one pass).  No device needed for the lookup. */
int dgemm_x_set_split_config(int split);
int sgemm_x_set_split_config(int split);
int dgemm_x_tune_lookup_split(char transa, char transb, int m, int n, int k, size_t *bytes);
int sgemm_x_tune_lookup_split(char transa, char transb, int m, int n, int k, size_t *bytes);
/* "major.minor" of the ROCm release the library was built with: the value
 * of the `rocm` directive a tuning table must carry to be used. */
const char *dgemm_x_build_rocm(void);
const char *sgemm_x_build_rocm(void);
/* Product kernel of the emulated forms (Ozaki-I, Ozaki-II, BF16x3), for every
 * later call: 1 = direct-to-LDS loads with 128-byte K stages (the default),
 * 0 = register-staged; -1 returns to GEMMX_SLICE_CFG (read at first use) or
 * the default.  Both give identical integer products.  No counterpart in
 * the reference (its blocking is fixed at compile time, dgemm_tune.h). */
int dgemm_x_set_product_config(int cfg);
int sgemm_x_set_product_config(int cfg);
int dgemm_x_load_tuning(const char *path);
int sgemm_x_load_tuning(const char *path);
int dgemm_x_tune_lookup(char transa, char transb, int m, int n, int k);
/* The arithmetic dgemm_ uses in GEMMX_MODE_AUTO for these arguments
 * (GEMMX_MODE_NATIVE, _OZAKI or _OZAKI2; count = slices or moduli, 0 for
 * native); no device needed. */
int dgemm_x_tune_lookup_emulated(char transa, char transb, int m, int n, int k,
                                 int *count);
int sgemm_x_tune_lookup(char transa, char transb, int m, int n, int k);
/* The arithmetic sgemm_ uses in GEMMX_MODE_AUTO for these arguments
 * (GEMMX_MODE_NATIVE or GEMMX_MODE_BF16X3); no device needed. */
int sgemm_x_tune_lookup_emulated(char transa, char transb, int m, int n, int k);
/* The tile configuration a strided batch of `batch` problems with these
 * arguments uses (*_x_strided_batched_dev): a forced configuration, else
 * the table's `strided` entry of the shape's bucket (measured on batches
 * that fill the GPU many times over) when the batch holds at least 256
 * tiles of 128x128, else what *_x_tune_lookup returns.  No device needed;
 * -1 for negative arguments.  No counterpart in the reference (it has no
 * batched entry point). */
int dgemm_x_tune_lookup_strided(char transa, char transb, int m, int n, int k, int batch);
int sgemm_x_tune_lookup_strided(char transa, char transb, int m, int n, int k, int batch);

/* Last error of the calling thread: returns its code and, when msg is not
 * NULL, a static description.  Cleared by the next successful call. */
int dgemm_x_last_error(const char **msg);
int sgemm_x_last_error(const char **msg);

/* Device-resident form: A, B, C are device pointers on the current device;
 * the work is enqueued on `stream` (a hipStream_t, NULL = legacy default
 * stream) and the call returns without waiting.  Same semantics as dgemm_.
 * Returns a GEMMX_* code. */
int dgemm_x_dev(char transa, char transb, int m, int n, int k, double alpha,
                const double *a, int lda, const double *b, int ldb, double beta,
                double *c, int ldc, void *stream);
int sgemm_x_dev(char transa, char transb, int m, int n, int k, float alpha,
                const float *a, int lda, const float *b, int ldb, float beta,
                float *c, int ldc, void *stream);

/* Allocation-free form of dgemm_x_dev / sgemm_x_dev.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels, split-K choice and results as the _x_dev form, but the
 * split-K partial tiles go to a static device array of the library
 * (64 MiB per GPU, held until the process ends) instead of the stream-ordered
 * pool; the call makes no device allocation or free.  A call whose partial
 * tiles would exceed 64 MiB (only a tuning-table or forced split can ask for
 * that) returns GEMMX_ERR_UNSUPPORTED and leaves C untouched.  All such calls
 * on one GPU share the array: calls that overlap on a GPU (on different
 * streams or from several host threads) race. */
int dgemm_x_dev_noalloc(char transa, char transb, int m, int n, int k, double alpha,
                        const double *a, int lda, const double *b, int ldb, double beta,
                        double *c, int ldc, void *stream);
int sgemm_x_dev_noalloc(char transa, char transb, int m, int n, int k, float alpha,
                        const float *a, int lda, const float *b, int ldb, float beta,
                        float *c, int ldc, void *stream);

/* Strided batch on device pointers (no counterpart in the reference; the
 * BLAS-style batched interface): `batch` problems of one shape, problem i
 * reading A + i*stride_a and B + i*stride_b and updating C + i*stride_c
 * (strides in elements, >= 0; 0 shares one A or B across the batch;
 * stride_c >= ldc*n so that no two C overlap), each with the semantics of
 * dgemm_x_dev, native FP64/FP32 MFMA arithmetic, all in the launches of one
 * problem (a grid row per problem).  Enqueued on `stream`; returns a
 * GEMMX_* code. */
int dgemm_x_strided_batched_dev(char transa, char transb, int m, int n, int k,
                                double alpha, const double *a, int lda, int64_t stride_a,
                                const double *b, int ldb, int64_t stride_b, double beta,
                                double *c, int ldc, int64_t stride_c, int batch,
                                void *stream);
int sgemm_x_strided_batched_dev(char transa, char transb, int m, int n, int k,
                                float alpha, const float *a, int lda, int64_t stride_a,
                                const float *b, int ldb, int64_t stride_b, float beta,
                                float *c, int ldc, int64_t stride_c, int batch,
                                void *stream);

/* Emulated DGEMM on device pointers: Ozaki-I with `slices` signed 7-bit INT8
 * slices per operand (3..8), exact INT32 accumulation on INT8 MFMA, FP64 only
 * for the row/column exponent scaling and the recombination.  `work` is
 * device scratch of at least dgemm_x_ozaki_workspace() bytes, or NULL to use
 * the library's own (then the call synchronises `stream`). */
size_t dgemm_x_ozaki_workspace(int m, int n, int k, int slices);
int dgemm_x_ozaki_dev(char transa, char transb, int m, int n, int k,
                      double alpha, const double *a, int lda, const double *b,
                      int ldb, double beta, double *c, int ldc, int slices,
                      void *work, size_t work_bytes, void *stream);

/* Emulated DGEMM on device pointers, Ozaki-II: every row of op(A) and
 * column of op(B) is scaled by a power of two to integers rint(x 2^(P-e))
 * (e: the vector's exponent, max |x| < 2^e); their exact integer product is
 * formed from residues modulo `moduli` (8..22) pairwise-coprime integers
 * <= 256 (one INT8 GEMM per modulus on the matrix cores, exact in INT32) by
 * the Chinese remainder theorem, then rounded once to FP64 and scaled back.
 * Each vector's bit count P comes from a bound on its 2-norm
 * (S = sum ceil(|x| 2^(10-e))^2, the largest P <= 77 with
 * 2P + bitlen(S - 1) <= floor(log2 M) + 19, M the product of the moduli), so
 * that every integer dot product stays inside the CRT range by
 * Cauchy-Schwarz: vectors whose mass sits in few entries keep up to
 * log2(sqrt(k)) more bits.  dgemm_x_ozaki2_bits gives the smallest P over
 * vectors of k entries (all at the top of their binade: 24 bits at 8
 * moduli, 55 at 16, 63 at 18 for k = 8192; -1 when unsupported); entries
 * smaller than 2^(e-P) are lost.  `work` as for dgemm_x_ozaki_dev, of at
 * least dgemm_x_ozaki2_workspace() bytes. */
size_t dgemm_x_ozaki2_workspace(int m, int n, int k, int moduli);
int dgemm_x_ozaki2_bits(int moduli, int k);
int dgemm_x_ozaki2_dev(char transa, char transb, int m, int n, int k,
                       double alpha, const double *a, int lda, const double *b,
                       int ldb, double beta, double *c, int ldc, int moduli,
                       void *work, size_t work_bytes, void *stream);

/* Phase timing of the emulated DGEMMs (Ozaki-I and Ozaki-II), for the
 * calling thread: after dgemm_x_set_phase_timing(1), each emulated call
 * records HIP events on its stream; dgemm_x_phase_times waits for the last
 * call's events and gives, in ms, the split, the slice/residue products
 * (with more than one K chunk: products and recombination together) and the
 * recombination.  Off by default; costs four event records per call. */
int dgemm_x_set_phase_timing(int on);
int dgemm_x_phase_times(float *ms);

/* Emulated SGEMM on device pointers: split-BF16x3 (a = hi + lo in BF16,
 * C ~ hi*hi + hi*lo + lo*hi) on BF16 MFMA with FP32 accumulation.  `work` is
 * device scratch of at least sgemm_x_bf16x3_workspace() bytes, or NULL to
 * use the library's own (then the call synchronises `stream`). */
size_t sgemm_x_bf16x3_workspace(int m, int n, int k);
int sgemm_x_bf16x3_dev(char transa, char transb, int m, int n, int k,
                       float alpha, const float *a, int lda, const float *b,
                       int ldb, float beta, float *c, int ldc, void *work,
                       size_t work_bytes, void *stream);

/* Grouped batch on device pointers.  No counterpart in the reference, which
 * computes one GEMM per call (wjc404/GEMM_AVX2 src/gemm_driver.c:118); each
 * problem i has the semantics of dgemm_/sgemm_ with transa[i], transb[i],
 * m[i] ... ldc[i] (alpha/beta by value in arrays, a/b/c device pointers).
 *
 * _create checks every problem, orders them heaviest first, uploads the
 * descriptors and captures one persistent grouped launch into a hipGraph on
 * the current device; the argument arrays may be freed afterwards.  _run
 * replays it on `stream` (NULL = null stream; the current device must be the
 * plan's) and returns without synchronising; replays of one plan must not
 * overlap.  Problems with m == 0, n == 0, or with no product and beta == 1
 * are dropped; C blocks of different problems must not overlap.  _flops
 * gives 2*m*n*k summed over the problems that have a product.  _destroy
 * releases the plan (after its last replay has finished). */
int dgemm_x_batch_create(void **plan, int count, const char *transa,
                         const char *transb, const int *m, const int *n,
                         const int *k, const double *alpha,
                         const double *const *a, const int *lda,
                         const double *const *b, const int *ldb,
                         const double *beta, double *const *c, const int *ldc);
int dgemm_x_batch_run(void *plan, void *stream);
double dgemm_x_batch_flops(const void *plan);
int dgemm_x_batch_destroy(void *plan);
int sgemm_x_batch_create(void **plan, int count, const char *transa,
                         const char *transb, const int *m, const int *n,
                         const int *k, const float *alpha,
                         const float *const *a, const int *lda,
                         const float *const *b, const int *ldb,
                         const float *beta, float *const *c, const int *ldc);
int sgemm_x_batch_run(void *plan, void *stream);
double sgemm_x_batch_flops(const void *plan);
int sgemm_x_batch_destroy(void *plan);
/* Tile configuration of the batch kernel for plans created afterwards
 * (0 64x64x16, 1 64x64x32, 2 128x64x16, 3 32x32x32; -1 = the environment's
 * GEMMX_BATCH_CFG, else the tuning table's `batch` entry written by the
 * autotuner, else the default).  *_x_batch_config returns the configuration
 * a plan created now would use.  No reference counterpart (the reference's
 * blocking is fixed per build, dgemm_tune.h / sgemm_tune.h). */
#define GEMMX_BATCH_CONFIGS 4
int dgemm_x_batch_set_config(int cfg);
int sgemm_x_batch_set_config(int cfg);
int dgemm_x_batch_config(void);
int sgemm_x_batch_config(void);

#ifdef __cplusplus
}
#endif
#endif
