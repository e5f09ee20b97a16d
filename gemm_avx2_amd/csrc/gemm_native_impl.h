/*
 * gemm_native_impl.h -- the native DGEMM / SGEMM tile kernels for gfx950 and
 * their launchers, instantiated per element type and transpose pair by the
 * gemm_native_{f64,f32}_{nn,nt,tn,tt}.hip translation units (one kernel set
 * each, so the eight compile in parallel); gemm_native.hip dispatches.
 * See gemm_native.hip for the design.
 */
#ifndef GEMM_NATIVE_IMPL_H
#define GEMM_NATIVE_IMPL_H

#include <type_traits>

#include "gemmx_device.h"
#include "gemmx_internal.h"
#include "gemmx_tile.h"

namespace gemmx {
/* internal linkage: every translation unit has its own kernels and stubs */
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
          int KVMAX, bool NARROW, bool SPLIT, int KG>
__global__ void __launch_bounds__(64 * WM * WN * KG, MINW)
gemm_native_kernel(const T *__restrict__ A, const T *__restrict__ B,
                   T *__restrict__ C, int M, int N, int K, int64_t lda,
                   int64_t ldb, int64_t ldc, T alpha, T beta, int tiles_m,
                   int tiles_n, int vecA, int vecB, int lead, int64_t sa, int64_t sb,
                   int64_t sc, int kc) {
  __shared__ __align__(16) T lds[tile_lds_need<T, TA, TB, BM, BN, BK, WM, WN, KG>()];
  /* C through LDS in the epilogue (store_c_tile_lds) for FP32 and for the
   * 64x64 FP64 tiles; the 128-wide FP64 tiles, whose epilogue is a small
   * part of a tile, keep the direct one (measured: 2-4% slower at 2048^3 and
   * 4096^3 with the LDS form) */
  constexpr bool EPI = std::is_same<T, float>::value || BM == 64;
  const int tiles_m_all = tiles_m + ((M % BM) ? 1 : 0);
  const int tiles_all = tiles_m_all * (tiles_n + ((N % BN) ? 1 : 0));
  /* bx: the tile (edge tiles first), by: the grid row (K range or problem),
   * nx: the tiles of a row */
  int bx = blockIdx.x, by = blockIdx.y, nx = gridDim.x;
  if (SPLIT) {
    /* one-dimensional grid of K ranges x tiles, renumbered so that
     * consecutive numbers run on one XCD (workgroups are dispatched
     * round-robin over the 8 XCDs); number n takes tile n % tiles_all of
     * K range n / tiles_all, so the tiles of one K range share that range's
     * operand panels in one XCD's L2.  The partial tiles go to C (the
     * workspace), range by range, tile by tile. */
    constexpr int XCDS = 8;
    const int total = gridDim.x, xcd = blockIdx.x % XCDS, local = blockIdx.x / XCDS;
    const int per = total / XCDS, rem = total % XCDS;
    const int n = xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
    by = n / tiles_all;
    bx = n - by * tiles_all;
    nx = tiles_all;
    const int k0 = by * kc;
    K = K - k0 < kc ? K - k0 : kc;
    A += TA ? (int64_t)k0 : (int64_t)k0 * lda;
    B += TB ? (int64_t)k0 * ldb : (int64_t)k0;
  } else {
    A += blockIdx.y * sa;
    B += blockIdx.y * sb;
    C += blockIdx.y * sc;
  }
  auto part = [&](int tm, int tn) {
    return C + ((int64_t)by * tiles_all + tm + (int64_t)tn * tiles_m_all) * (BM * BN);
  };
  int tm, tn;
  if (bx < lead) {
    const int row_tiles = (M % BM) ? tiles_n + ((N % BN) ? 1 : 0) : 0;
    const int col_tiles = (N % BN) ? tiles_m : 0;
    const int e = bx;
    if (e >= row_tiles + col_tiles) return; /* lead == row_tiles + col_tiles */
    if (e < row_tiles) {
      tm = tiles_m;
      tn = e;
    } else {
      tm = e - row_tiles;
      tn = tiles_n;
    }
    gemm_tile<T, TA, TB, BM, BN, BK, WM, WN, KVMAX, true, true, NARROW, SPLIT,
              (BM == 64 ? 2 : 1), KG, EPI>(
        A, B, SPLIT ? part(tm, tn) : C, M, N, K, lda, ldb, ldc, alpha, beta, tm * BM, tn * BN,
        vecA, vecB, lds);
    return;
  }
  tile_coords(bx - lead, nx - lead, tiles_m, tiles_n, &tm, &tn);
  gemm_tile<T, TA, TB, BM, BN, BK, WM, WN, KVMAX, true, false, NARROW, SPLIT,
              (BM == 64 ? 2 : 1), KG, EPI>(
      A, B, SPLIT ? part(tm, tn) : C, M, N, K, lda, ldb, ldc, alpha, beta, tm * BM, tn * BN,
      vecA, vecB, lds);
}

/* Split-K reduction: C = alpha * (P_0 + P_1 + ... + P_{S-1}) + beta * C for
 * the tiles' valid elements, the partial tiles summed in split order (so the
 * result does not depend on which split finished first); beta*C is formed
 * as in the one-pass epilogue, so beta = 0 still multiplies C (NaN and Inf
 * already in C propagate).  Thread t of the grid takes element t % (BM*BN)
 * of tile t / (BM*BN) in the partials' fragment order: loads of consecutive
 * threads are consecutive. */
template <typename T, int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(256)
splitk_reduce_kernel(const T *__restrict__ P, T *__restrict__ C, int M, int N, int64_t ldc,
                     T alpha, T beta, int tiles_m_all, int tiles_all, int splits) {
  constexpr int TE = BM * BN;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)tiles_all * TE) return;
  const int tile = (int)(t / TE), e = (int)(t % TE);
  int row, col;
  splitk_partial_coords<T, BM, BN, WM, WN>(e, &row, &col);
  row += (tile % tiles_m_all) * BM;
  col += (tile / tiles_m_all) * BN;
  if (row >= M || col >= N) return;
  T sum = P[t];
  for (int q = 1; q < splits; ++q) sum += P[(int64_t)q * tiles_all * TE + t];
  T *cp = C + row + (int64_t)col * ldc;
  *cp = fma_t(alpha, sum, beta * *cp);
}

/* One launch covers C: its leading workgroups the edge tiles (clamped
 * loads), the rest the interior tiles; 16-byte loads where both operands
 * allow them, else one element per load.  A strided batch goes in grid rows
 * of at most GEMMX_GRID_Y problems per launch.  With split > 1 (one problem)
 * grid row q computes the K range [q kc, (q+1) kc) of every tile into the
 * workspace and a second launch reduces the partials into C. */
constexpr int GEMMX_GRID_Y = 65535;

/* K depth per split: ceil(K / split) rounded up to whole BK slices. */
inline int split_depth(int k, int split, int bk) {
  const int kc = (k + split - 1) / split;
  return (kc + bk - 1) / bk * bk;
}

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WM, int WN, int MINW, int KVMAX,
          bool SPLITOK, int KG = 1>
hipError_t launch_trans(const gemmx_problem *p, int split, void *work, hipStream_t s) {
  constexpr int VE = Vec16<T>::N;
  const int batch = p->batch > 1 ? p->batch : 1;
  const int64_t sa = batch > 1 ? p->sa : 0, sb = batch > 1 ? p->sb : 0;
  const int64_t sc = batch > 1 ? p->sc : 0;
  const int vecA = ((uintptr_t)p->a % 16 == 0) && (p->lda % VE == 0) && (sa % VE == 0);
  const int vecB = ((uintptr_t)p->b % 16 == 0) && (p->ldb % VE == 0) && (sb % VE == 0);
  const T alpha = (T)p->alpha, beta = (T)p->beta;
  const dim3 block(64 * WM * WN * KG);
  const int tiles_m = (int)(((int64_t)p->m + BM - 1) / BM);
  const int tiles_n = (int)(((int64_t)p->n + BN - 1) / BN);
  const int full_m = p->m / BM, full_n = p->n / BN;
  const int lead = tiles_m * tiles_n - full_m * full_n; /* edge tiles */
  if constexpr (SPLITOK) if (split > 1 && batch == 1 && work) {
    const int kc = split_depth(p->k, split, BK);
    const unsigned ns = (unsigned)((p->k + kc - 1) / kc);
    T *P = (T *)work;
    const dim3 grid((unsigned)((lead + full_m * full_n) * ns));
    if (vecA && vecB)
      hipLaunchKernelGGL((gemm_native_kernel<T, TA, TB, BM, BN, BK, WM, WN, MINW, KVMAX, false, true, KG>),
                         grid, block, 0, s, (const T *)p->a, (const T *)p->b, P, p->m, p->n,
                         p->k, p->lda, p->ldb, p->ldc, alpha, beta, full_m, full_n, vecA, vecB,
                         lead, (int64_t)0, (int64_t)0, (int64_t)0, kc);
    else
      hipLaunchKernelGGL((gemm_native_kernel<T, TA, TB, BM, BN, BK, WM, WN, MINW, KVMAX, true, true, KG>),
                         grid, block, 0, s, (const T *)p->a, (const T *)p->b, P, p->m, p->n,
                         p->k, p->lda, p->ldb, p->ldc, alpha, beta, full_m, full_n, vecA, vecB,
                         lead, (int64_t)0, (int64_t)0, (int64_t)0, kc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t elems = (int64_t)tiles_m * tiles_n * BM * BN;
    hipLaunchKernelGGL((splitk_reduce_kernel<T, BM, BN, WM, WN>),
                       dim3((unsigned)((elems + 255) / 256)), dim3(256), 0, s, (const T *)P,
                       (T *)p->c, p->m, p->n, p->ldc, alpha, beta, tiles_m, tiles_m * tiles_n,
                       (int)ns);
    return hipGetLastError();
  }
  for (int b0 = 0; b0 < batch; b0 += GEMMX_GRID_Y) {
    const unsigned nb = (unsigned)(batch - b0 < GEMMX_GRID_Y ? batch - b0 : GEMMX_GRID_Y);
    const T *A = (const T *)p->a + b0 * sa, *B = (const T *)p->b + b0 * sb;
    T *C = (T *)p->c + b0 * sc;
    const dim3 grid(lead + full_m * full_n, nb);
    if (vecA && vecB)
      hipLaunchKernelGGL((gemm_native_kernel<T, TA, TB, BM, BN, BK, WM, WN, MINW, KVMAX, false, false, KG>),
                         grid, block, 0, s, A, B, C, p->m, p->n, p->k, p->lda, p->ldb, p->ldc,
                         alpha, beta, full_m, full_n, vecA, vecB, lead, sa, sb, sc, 0);
    else
      hipLaunchKernelGGL((gemm_native_kernel<T, TA, TB, BM, BN, BK, WM, WN, MINW, KVMAX, true, false, KG>),
                         grid, block, 0, s, A, B, C, p->m, p->n, p->k, p->lda, p->ldb, p->ldc,
                         alpha, beta, full_m, full_n, vecA, vecB, lead, sa, sb, sc, 0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

/* Tile of each configuration: BM, BN, the K slice depth of the FP64 and of
 * the FP32 kernels, and whether it has split-K kernels.  The launchers take
 * their slice depth from here, so the host's split arithmetic
 * (gemmx_native_splits) and the kernels' agree. */
struct NativeCfg {
  int bm, bn, bk64, bk32, split_ok;
};
constexpr NativeCfg kNativeCfg[GEMMX_NCFG_COUNT] = {
    {128, 128, 16, 16, 1}, {64, 64, 16, 16, 1}, {256, 128, 16, 16, 0}, {128, 128, 32, 32, 0},
    {128, 128, 16, 16, 0}, {128, 128, 16, 32, 1}, {256, 128, 16, 16, 0}, {64, 64, 16, 32, 1},
    {64, 64, 32, 64, 1}, {64, 64, 32, 32, 1}};
template <typename T>
constexpr int cfg_bk(int cfg) {
  return std::is_same<T, double>::value ? kNativeCfg[cfg].bk64 : kNativeCfg[cfg].bk32;
}

/* One transpose pair: the configuration switch. */
template <typename T, bool TA, bool TB>
hipError_t launch_native_trans(const gemmx_problem *p, int cfg, int split, void *work,
                               hipStream_t s) {
  if (cfg < 0 || cfg >= GEMMX_NCFG_COUNT) cfg = 0;
  if (!kNativeCfg[cfg].split_ok) split = 1;
  /* BM, BN, BK, waves (M x N), min waves per SIMD, K steps per LDS read */
  switch (cfg) {
  case 1: return launch_trans<T, TA, TB, 64, 64, cfg_bk<T>(1), 2, 2, 2, 1, true>(p, split, work, s);
  case 2: return launch_trans<T, TA, TB, 256, 128, cfg_bk<T>(2), 4, 2, 2, 1, false>(p, 1, NULL, s);
  /* FP32 128x128 with K-vector reads fits two workgroups per CU's register
   * file; FP64 (twice the accumulator registers) takes one */
  case 3: return launch_trans<T, TA, TB, 128, 128, cfg_bk<T>(3), 2, 2, std::is_same<T, float>::value ? 2 : 1, 4, false>(p, 1, NULL, s);
  case 4: return launch_trans<T, TA, TB, 128, 128, cfg_bk<T>(4), 2, 2, 1, 4, false>(p, 1, NULL, s);
  case 5: return launch_trans<T, TA, TB, 128, 128, cfg_bk<T>(5), 4, 2, 2, 1, true>(p, split, work, s);
  case 6: return launch_trans<T, TA, TB, 256, 128, cfg_bk<T>(6), 4, 2, 1, 4, false>(p, 1, NULL, s);
  /* 64x64 with two waves per SIMD (8 waves of 32x16) for products with
   * about one tile per CU; FP32 takes 32-deep K slices so that each of the
   * 512 threads copies a whole 16-byte chunk of each operand slice */
  case 7: return launch_trans<T, TA, TB, 64, 64, cfg_bk<T>(7), 2, 4, 2, 1, true>(p, split, work, s);
  /* the same 64x64 tile with four waves per SIMD: two K groups of 8 waves
   * of 32x16 share each (twice as deep) LDS slice, one multiplying its
   * first half and one its second, and add their accumulators at the end */
  case 8: return launch_trans<T, TA, TB, 64, 64, cfg_bk<T>(8), 2, 4, 4, 1, true, 2>(p, split, work, s);
  /* configuration 1 with 32-deep slices: half the barriers per flop (FP32
   * 1024^3: 5-10% faster than 16-deep slices; ragged K: up to 10% slower) */
  case 9: return launch_trans<T, TA, TB, 64, 64, cfg_bk<T>(9), 2, 2, 2, 1, true>(p, split, work, s);
  case 0:
  default: return launch_trans<T, TA, TB, 128, 128, cfg_bk<T>(0), 2, 2, 2, 1, true>(p, split, work, s);
  }
}

} // namespace
} // namespace gemmx

/* The entry of one translation unit: element type T, transposes TA, TB. */
#define GEMMX_NATIVE_INSTANCE(T, TA, TB, NAME)                                                  \
  extern "C" hipError_t NAME(const gemmx_problem *p, int cfg, int split, void *work,           \
                             hipStream_t s) {                                                  \
    return gemmx::launch_native_trans<T, TA, TB>(p, cfg, split, work, s);                      \
  }

#endif
