/*
 * gemmx_internal.h -- internal C ABI between the host C layer (gemmx_api.c)
 * and the gfx950 HIP kernels (*.hip).  Not installed; the public surface is
 * include/gemm_mi355x.h.
 *
 * All matrices are column-major with the Fortran BLAS conventions of the
 * reference entry point (wjc404/GEMM_AVX2 src/gemm_driver.c:118-127):
 * op(A) is m x k, op(B) is k x n, C is m x n; lda/ldb/ldc are in elements.
 */
#ifndef GEMMX_INTERNAL_H
#define GEMMX_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__ 1
#endif
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One GEMM on device-resident operands.  Pointers are device addresses. */
typedef struct gemmx_problem {
  int ta, tb;      /* 1 = transpose (any first char other than 'N'/'n') */
  int m, n, k;
  int64_t lda, ldb, ldc;
  double alpha, beta; /* FP32 problems hold the float values widened */
  const void *a;
  const void *b;
  void *c;
  /* Strided batch (native arithmetic only): `batch` problems of this shape,
   * problem i at a + i*sa, b + i*sb, c + i*sc (elements); batch 0 or 1 is
   * a single problem. */
  int batch;
  int64_t sa, sb, sc;
} gemmx_problem;

/* A complex GEMM (interleaved re, im pairs; sizes and leading dimensions in
 * complex elements): ta/tb transpose, conja/conjb conjugate (transa 'C'). */
typedef struct gemmx_cproblem {
  int ta, tb, conja, conjb;
  int m, n, k;
  int64_t lda, ldb, ldc;
  double alpha[2], beta[2];
  const void *a;
  const void *b;
  void *c;
} gemmx_cproblem;
/* Device workspace of gemmx_complex (real planes of the operands and of the
 * products; three = the three-product form), and the complex GEMM itself
 * (gemm_complex.hip; es 8 = double complex, 4 = float complex). */
size_t gemmx_complex_workspace(int es, int ta, int tb, int m, int n, int k, int three);
hipError_t gemmx_complex(int es, const gemmx_cproblem *p, int three, void *work, hipStream_t s);

/* A triangular multiply or solve (dtrmm_/dtrsm_ and the s forms): side
 * left (op(A) B / op(A) X = alpha B) or right, A's stored triangle upper or
 * lower, op(A) = A or A^T, unit diagonal or not. */
typedef struct gemmx_tproblem {
  int left, upper, trans, unit;
  int m, n;
  int64_t lda, ldb;
  double alpha;
  const void *a;
  void *b;
} gemmx_tproblem;
/* Device workspace of gemmx_tri, and the operation itself (gemm_tri.hip;
 * solve 1 = TRSM, 0 = TRMM; es 8 or 4). */
size_t gemmx_tri_workspace(int es, int solve, int left, int m, int n);
hipError_t gemmx_tri(int es, int solve, const gemmx_tproblem *p, void *work, hipStream_t s);

/* Kernel configurations of the native MFMA tile kernels.  The tuning table
 * (gemmx_tune.c) picks one per (dtype, trans, shape bucket). */
enum {
  GEMMX_NCFG_128x128 = 0, /* 128x128x16, 4 waves of 64x64, >= 2 waves/SIMD */
  GEMMX_NCFG_64x64 = 1,   /* 64x64x16, 4 waves of 32x32 */
  GEMMX_NCFG_256x128 = 2, /* 256x128x16, 8 waves of 64x64, >= 2 waves/SIMD */
  /* 3: 128x128x32 1 wave/SIMD; 4: 128x128x16 1 wave/SIMD; 5: 128x128 with
   * 8 waves of 32x64, K slices of 16 (FP64) or 32 (FP32); 6: 256x128x16 1 wave/SIMD; 7: 64x64 with 8 waves of
   * 32x16, K slices of 16 (FP64) or 32 (FP32); 8: 64x64 with two K groups of
   * 8 waves of 32x16, K slices of 32 (FP64) or 64 (FP32); 9: 64x64x32, 4 waves
   * of 32x32 (see gemm_native_impl.h) */
  GEMMX_NCFG_COUNT = 10
};

/* Host-side error state and device lookup shared by the host sources
 * (gemmx_api.c).  gemmx_report records the per-thread last error and
 * returns `code`. */
int gemmx_report(int code, const char *fmt, ...);
/* Nonzero when GEMMX_FAIL_ALLOC asks the library's device allocations to
 * fail (failure-path testing). */
int gemmx_alloc_failure_injected(void);
void gemmx_clear_error(void);
int gemmx_last_code(void);
int gemmx_current_device(int *dev);
/* One-time library set-up (modes from the environment, tuning table). */
void gemmx_init(void);

/* Staging of host operands through pinned buffers (gemmx_stage.c): cols
 * columns of row_bytes bytes, leading dimensions in bytes.  The copies start
 * after event `after`, or, when it is NULL, after the work queued on
 * `compute` at the call; h2d makes `compute` wait for its copies, d2h returns
 * when the host block is complete.  (Blocks below 4 MB go through one runtime
 * copy on `compute` instead.)  The caller holds the device lock.
 * GEMMX_STAGE_CHUNK_MB (default 32) sets the chunk, GEMMX_COPY_THREADS
 * (default 8) the host copy threads. */
int gemmx_stage_h2d(int dev, void *dst, size_t dld, const void *src, size_t sld,
                    size_t row_bytes, size_t cols, hipStream_t compute, hipEvent_t after);
int gemmx_stage_d2h(int dev, void *dst, size_t dld, const void *src, size_t sld,
                    size_t row_bytes, size_t cols, hipStream_t compute, hipEvent_t after);
size_t gemmx_stage_chunk_bytes(void);

/* Native tile kernels (gemm_native.hip).  alpha != 0 and k > 0 here; the
 * host layer handles the alpha == 0 / k == 0 / empty cases.  split > 1 (a
 * single problem, configurations with split-K kernels) divides K over that
 * many grid rows writing partial tiles to `work` (gemmx_native_workspace
 * bytes), then reduces them into C; split 1 or work NULL is one pass. */
hipError_t gemmx_native_f64(const gemmx_problem *p, int cfg, int split, void *work,
                            hipStream_t s);
hipError_t gemmx_native_f32(const gemmx_problem *p, int cfg, int split, void *work,
                            hipStream_t s);
/* BM x BN tile of a configuration; returns 1 when it has split-K kernels,
 * 0 when not, -1 for an unknown configuration. */
int gemmx_native_tile(int cfg, int *bm, int *bn);
/* Splits actually run for a requested split (K ranges of whole BK slices of
 * the element size's kernels), and the bytes of partial tiles they need (0
 * for one pass). */
int gemmx_native_splits(int es, int cfg, int k, int split);
size_t gemmx_native_workspace(int es, int cfg, int m, int n, int k, int split);
/* Static device scratch of the allocation-free entry points (a file-scope
 * array in gemm_native.hip, GEMMX_NOALLOC_SCRATCH bytes per device): its
 * address on device `dev`, which must be current; taken on the first call
 * per device. */
#define GEMMX_NOALLOC_SCRATCH (64u << 20)
#define GEMMX_NOALLOC_MAX_DEV 64
hipError_t gemmx_noalloc_scratch(int dev, void **ptr);
/* C[0:m,0:n] *= beta (beta != 1), the reference's C_MULT_BETA
 * (src/gemm_driver.c:106-116), for `batch` matrices sc elements apart. */
hipError_t gemmx_scale_f64(double *c, int64_t ldc, int m, int n, double beta, int batch,
                           int64_t sc, hipStream_t s);
hipError_t gemmx_scale_f32(float *c, int64_t ldc, int m, int n, float beta, int batch,
                           int64_t sc, hipStream_t s);

/* Grouped batch (gemm_batch.hip): one problem of a persistent grouped
 * launch.  Problems are stored heaviest first; tiles of problem p are
 * tile_begin .. tile_begin + tiles_m*tiles_n - 1 of the launch. */
enum {
  GEMMX_BF_TA = 1,      /* transpose A */
  GEMMX_BF_TB = 2,      /* transpose B */
  GEMMX_BF_VEC_A = 4,   /* A loadable with 16-byte loads */
  GEMMX_BF_VEC_B = 8,
  GEMMX_BF_SCALE = 16   /* alpha == 0 or k == 0: C = beta*C only */
};
typedef struct gemmx_batch_desc {
  const void *a;
  const void *b;
  void *c;
  int64_t lda, ldb, ldc;
  double alpha, beta;
  int m, n, k, flags;
  int tiles_m, tiles_n, tile_begin, pad;
} gemmx_batch_desc;

/* Batch kernel configurations (gemm_batch.hip): 0 64x64x16, 1 64x64x32,
 * 2 128x64x16, 3 32x32x32.  gemmx_batch_tile gives the tile of `cfg` (-1 for
 * an unknown one); gemmx_batch_grid the number of its workgroups that fit on
 * the current device at once (the persistent grid, 0 on failure). */
#define GEMMX_BATCH_CONFIGS 4
int gemmx_batch_tile(int cfg, int *bm, int *bn);
int gemmx_batch_grid(int es, int cfg);
/* Tile queues of a grouped launch, one per XCD of MI355X: queue x is tiles
 * [qbeg[x], qbeg[x+1]) of d_tile_prob.  d_sched holds the queues' ticket
 * counters, GEMMX_BATCH_CTR_STRIDE words apart (GEMMX_BATCH_QUEUES *
 * GEMMX_BATCH_CTR_STRIDE words).  The launch zeroes the counters, then runs
 * the persistent grouped kernel. */
#define GEMMX_BATCH_QUEUES 8
#define GEMMX_BATCH_CTR_STRIDE 64
typedef struct gemmx_batch_queues {
  int qbeg[GEMMX_BATCH_QUEUES + 1];
} gemmx_batch_queues;
hipError_t gemmx_batch_launch(int es, int cfg, const gemmx_batch_desc *d_desc,
                              const int *d_tile_prob, unsigned *d_sched,
                              const gemmx_batch_queues *q, int grid, hipStream_t s);

/* Emulated DGEMM, Ozaki-I INT8 slices (gemm_slices.hip).  `work` is device
 * scratch of gemmx_ozaki_workspace() bytes. */
size_t gemmx_ozaki_workspace(int m, int n, int k, int slices);
int gemmx_ozaki_kchunk(int slices);
hipError_t gemmx_ozaki_f64(const gemmx_problem *p, int slices, void *work,
                           hipStream_t s, hipEvent_t *ev);

/* Emulated DGEMM, Ozaki-II: residues modulo `moduli` (8..22) pairwise-coprime
 * integers <= 256 (gemm_emulated.hip).  `work` is device scratch of
 * gemmx_ozaki2_workspace() bytes; gemmx_ozaki2_bits gives the integer bits P
 * used for K = k (-1 for an unsupported modulus count). */
size_t gemmx_ozaki2_workspace(int m, int n, int k, int moduli);
int gemmx_ozaki2_bits(int moduli, int k);
hipError_t gemmx_ozaki2_f64(const gemmx_problem *p, int moduli, void *work,
                            hipStream_t s, hipEvent_t *ev);
/* ev (NULL, or four events of the current device): the emulated DGEMMs
 * record ev[0] before the split, ev[1] after it, ev[2] after the products
 * (with more than one K chunk, after the last chunk's recombination) and
 * ev[3] at the end, on stream s. */

/* Emulated SGEMM, split-BF16x3 (gemm_slices.hip).  `work` is device scratch
 * of gemmx_bf16x3_workspace() bytes. */
size_t gemmx_bf16x3_workspace(int m, int n, int k);
hipError_t gemmx_bf16x3_f32(const gemmx_problem *p, void *work, hipStream_t s);

/* Product kernel of the emulated forms (gemm_products.hip): 1 direct-to-LDS
 * (default), 0 register-staged, -1 back to GEMMX_SLICE_CFG / the default.
 * Returns 0, or -1 for another value. */
int gemmx_set_product_config(int cfg);

#ifdef __cplusplus
}
#endif
#endif
