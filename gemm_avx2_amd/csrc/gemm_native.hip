/*
 * gemm_native.hip -- native DGEMM / SGEMM tile kernels for gfx950.
 *
 * Replaces the reference's compute layer: the AVX2 micro-kernels
 * gemmblkregccc / gemmblktailccc (wjc404/GEMM_AVX2 src/kernel/gemm_kernel.S:
 * 117-238) and the edge kernels gemmblkirreg{k,n,}ccc
 * (src/edge_kernel/gemm_edge_kernel.c:39-117), and the dispatch between them
 * (GEMM_CALC_COLUMN[_EDGE], src/gemm_driver.c:80-105).  One kernel covers
 * every case: edges in m, n and k are predicated tiles, not separate kernels.
 *
 * Per workgroup: one BM x BN tile of C.  K is walked in BK slices; each slice
 * of op(A) and op(B) is copied global->registers->LDS (double-buffered, the
 * next slice's loads in flight during the current slice's MFMAs) and consumed
 * by 16x16x4 MFMAs (FP64: v_mfma_f64_16x16x4_f64, FP32:
 * v_mfma_f32_16x16x4_f32).  The epilogue applies C = alpha*acc + beta*C with
 * one fma per element (alpha moves from the reference's packed B block,
 * src/gemm_copy.c:65, into the epilogue).
 *
 * Transpose handling: the LDS layout follows the global layout, so the
 * global->LDS copy is always a straight copy of contiguous runs, and the
 * transpose is absorbed by how the MFMA operands are read from LDS:
 *   op(A), transa N: LDS [k][m]     transa T: LDS [m][k]
 *   op(B), transb N: LDS [n][k]     transb T: LDS [k][n]
 * In MFMA step s of a BK slice, lane group g = lane/16 supplies k = g*BK/4+s,
 * identically for A and B, so the four k-groups of each MFMA together cover
 * the slice; where LDS is K-contiguous, consecutive steps are read with one
 * 16-byte LDS read.
 *
 * A strided batch of problems of one shape runs in the same launches: grid
 * row blockIdx.y takes problem (batch base + blockIdx.y), its operands offset
 * by the per-problem strides.
 *
 * Every workgroup runs a branch-free loop over the full K slices (the K
 * remainder, if any, is one peeled predicated slice); an edge tile's loads
 * past M or N are clamped back into the operand and its stores masked.
 * Operands that allow 16-byte loads are copied in 16-byte chunks, others one
 * element per load (the LDS layout is the same).  In the
 * branch-free loop the issue order is set explicitly: the next slice's global
 * loads first, each MFMA step's fragment reads spread over the previous
 * step's MFMAs, and the next slice's LDS writes spread over the last step.
 */
#include <type_traits>

#include "gemmx_device.h"
#include "gemmx_internal.h"
#include "gemmx_tile.h"

using namespace gemmx;

namespace {

/* tiles_m x tiles_n are the interior tiles, all inside C.  The first `lead`
 * workgroups take the edge tiles -- the partial tile-row tm = tiles_m (when
 * M % BM != 0, tiles_n + (N % BN != 0) tiles) then the partial tile-column
 * tn = tiles_n for tm < tiles_m -- on the branch-free path with loads clamped
 * into the operands; the rest take the interior tiles.  NARROW: operands
 * that do not allow 16-byte loads, loaded one element per load.  Dispatched first, the edge tiles run beside the interior
 * ones instead of in a trailing wave of their own.  Interior tile ids start
 * at `lead`, which shifts every workgroup's XCD (blockIdx.x % 8) by the same
 * amount, so the workgroups of one XCD still share tile rows and columns.
 * The grid holds exactly one workgroup per tile: a padding workgroup would
 * give one XCD an extra tile and, when the tiles fill whole waves, a wave of
 * its own. */
template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WM, int WN, int MINW,
          int KVMAX, bool NARROW>
__global__ void __launch_bounds__(64 * WM * WN, MINW)
gemm_native_kernel(const T *__restrict__ A, const T *__restrict__ B,
                   T *__restrict__ C, int M, int N, int K, int64_t lda,
                   int64_t ldb, int64_t ldc, T alpha, T beta, int tiles_m,
                   int tiles_n, int vecA, int vecB, int lead, int64_t sa, int64_t sb,
                   int64_t sc) {
  __shared__ __align__(16) T lds[tile_lds_elems<T, TA, TB, BM, BN, BK>()];
  A += blockIdx.y * sa;
  B += blockIdx.y * sb;
  C += blockIdx.y * sc;
  int tm, tn;
  if ((int)blockIdx.x < lead) {
    const int row_tiles = (M % BM) ? tiles_n + ((N % BN) ? 1 : 0) : 0;
    const int col_tiles = (N % BN) ? tiles_m : 0;
    const int e = blockIdx.x;
    if (e >= row_tiles + col_tiles) return; /* lead == row_tiles + col_tiles */
    if (e < row_tiles) {
      tm = tiles_m;
      tn = e;
    } else {
      tm = e - row_tiles;
      tn = tiles_n;
    }
    gemm_tile<T, TA, TB, BM, BN, BK, WM, WN, KVMAX, true, true, NARROW>(
        A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, tm * BM, tn * BN, vecA, vecB, lds);
    return;
  }
  tile_coords(blockIdx.x - lead, gridDim.x - lead, tiles_m, tiles_n, &tm, &tn);
  gemm_tile<T, TA, TB, BM, BN, BK, WM, WN, KVMAX, true, false, NARROW>(
      A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, tm * BM, tn * BN, vecA, vecB, lds);
}

template <typename T>
__global__ void __launch_bounds__(256)
scale_kernel(T *__restrict__ C, int64_t ldc, int M, int N, T beta, int64_t sc) {
  C += blockIdx.y * sc;
  const int64_t total = (int64_t)M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / M, i = t - j * M;
    C[i + j * ldc] *= beta;
  }
}

/* One launch covers C: its leading workgroups the edge tiles (clamped
 * loads), the rest the interior tiles; 16-byte loads where both operands
 * allow them, else one element per load.  A strided batch goes in grid rows
 * of at most GEMMX_GRID_Y problems per launch. */
constexpr int GEMMX_GRID_Y = 65535;

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WM, int WN, int MINW, int KVMAX>
hipError_t launch_trans(const gemmx_problem *p, hipStream_t s) {
  constexpr int VE = Vec16<T>::N;
  const int batch = p->batch > 1 ? p->batch : 1;
  const int64_t sa = batch > 1 ? p->sa : 0, sb = batch > 1 ? p->sb : 0;
  const int64_t sc = batch > 1 ? p->sc : 0;
  const int vecA = ((uintptr_t)p->a % 16 == 0) && (p->lda % VE == 0) && (sa % VE == 0);
  const int vecB = ((uintptr_t)p->b % 16 == 0) && (p->ldb % VE == 0) && (sb % VE == 0);
  const T alpha = (T)p->alpha, beta = (T)p->beta;
  const dim3 block(64 * WM * WN);
  const int tiles_m = (p->m + BM - 1) / BM, tiles_n = (p->n + BN - 1) / BN;
  const int full_m = p->m / BM, full_n = p->n / BN;
  for (int b0 = 0; b0 < batch; b0 += GEMMX_GRID_Y) {
    const unsigned nb = (unsigned)(batch - b0 < GEMMX_GRID_Y ? batch - b0 : GEMMX_GRID_Y);
    const T *A = (const T *)p->a + b0 * sa, *B = (const T *)p->b + b0 * sb;
    T *C = (T *)p->c + b0 * sc;
    const int lead = tiles_m * tiles_n - full_m * full_n; /* edge tiles */
    const dim3 grid(lead + full_m * full_n, nb);
    if (vecA && vecB)
      hipLaunchKernelGGL((gemm_native_kernel<T, TA, TB, BM, BN, BK, WM, WN, MINW, KVMAX, false>),
                         grid, block, 0, s, A, B, C, p->m, p->n, p->k, p->lda, p->ldb, p->ldc,
                         alpha, beta, full_m, full_n, vecA, vecB, lead, sa, sb, sc);
    else
      hipLaunchKernelGGL((gemm_native_kernel<T, TA, TB, BM, BN, BK, WM, WN, MINW, KVMAX, true>),
                         grid, block, 0, s, A, B, C, p->m, p->n, p->k, p->lda, p->ldb, p->ldc,
                         alpha, beta, full_m, full_n, vecA, vecB, lead, sa, sb, sc);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T, int BM, int BN, int BK, int WM, int WN, int MINW, int KVMAX>
hipError_t launch_cfg(const gemmx_problem *p, hipStream_t s) {
  if (!p->ta && !p->tb) return launch_trans<T, false, false, BM, BN, BK, WM, WN, MINW, KVMAX>(p, s);
  if (!p->ta && p->tb) return launch_trans<T, false, true, BM, BN, BK, WM, WN, MINW, KVMAX>(p, s);
  if (p->ta && !p->tb) return launch_trans<T, true, false, BM, BN, BK, WM, WN, MINW, KVMAX>(p, s);
  return launch_trans<T, true, true, BM, BN, BK, WM, WN, MINW, KVMAX>(p, s);
}

template <typename T>
hipError_t launch_native(const gemmx_problem *p, int cfg, hipStream_t s) {
  /* BM, BN, BK, waves (M x N), min waves per SIMD, K steps per LDS read */
  switch (cfg) {
  case 1: return launch_cfg<T, 64, 64, 16, 2, 2, 2, 1>(p, s);
  case 2: return launch_cfg<T, 256, 128, 16, 4, 2, 2, 1>(p, s);
  /* FP32 128x128 with K-vector reads fits two workgroups per CU's register
   * file; FP64 (twice the accumulator registers) takes one */
  case 3: return launch_cfg<T, 128, 128, 32, 2, 2, std::is_same<T, float>::value ? 2 : 1, 4>(p, s);
  case 4: return launch_cfg<T, 128, 128, 16, 2, 2, 1, 4>(p, s);
  case 5: return launch_cfg<T, 128, 128, 16, 4, 2, 2, 1>(p, s);
  case 6: return launch_cfg<T, 256, 128, 16, 4, 2, 1, 4>(p, s);
  case 0:
  default: return launch_cfg<T, 128, 128, 16, 2, 2, 2, 1>(p, s);
  }
}

template <typename T>
hipError_t launch_scale(T *c, int64_t ldc, int m, int n, T beta, int batch, int64_t sc,
                        hipStream_t s) {
  const int64_t total = (int64_t)m * n;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) return hipSuccess;
  if (batch < 1) batch = 1;
  for (int b0 = 0; b0 < batch; b0 += GEMMX_GRID_Y) {
    const unsigned nb = (unsigned)(batch - b0 < GEMMX_GRID_Y ? batch - b0 : GEMMX_GRID_Y);
    hipLaunchKernelGGL((scale_kernel<T>), dim3((unsigned)blocks, nb), dim3(256), 0, s,
                       c + b0 * sc, ldc, m, n, beta, batch > 1 ? sc : (int64_t)0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

} // namespace

extern "C" hipError_t gemmx_native_f64(const gemmx_problem *p, int cfg, hipStream_t s) {
  return launch_native<double>(p, cfg, s);
}
extern "C" hipError_t gemmx_native_f32(const gemmx_problem *p, int cfg, hipStream_t s) {
  return launch_native<float>(p, cfg, s);
}
extern "C" hipError_t gemmx_scale_f64(double *c, int64_t ldc, int m, int n,
                                      double beta, int batch, int64_t sc, hipStream_t s) {
  return launch_scale<double>(c, ldc, m, n, beta, batch, sc, s);
}
extern "C" hipError_t gemmx_scale_f32(float *c, int64_t ldc, int m, int n,
                                      float beta, int batch, int64_t sc, hipStream_t s) {
  return launch_scale<float>(c, ldc, m, n, beta, batch, sc, s);
}
