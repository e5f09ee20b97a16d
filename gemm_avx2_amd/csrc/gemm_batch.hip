/*
 * gemm_batch.hip -- grouped batch: many independent GEMMs of uneven shapes
 * and mixed transposes in one persistent launch.  No reference counterpart
 * (the reference computes one GEMM per dgemm_ call, wjc404/GEMM_AVX2
 * src/gemm_driver.c:118); every problem keeps dgemm_'s semantics, including
 * the beta != 1 rule and the alpha == 0 gate.
 *
 * The grid is as large as the device holds at once; each workgroup takes the
 * next tile with an atomic ticket until the tickets run out, so every wave
 * reaches the exit.  Tiles are numbered heaviest problem first (the host
 * sorts), which with dynamic tickets approximates longest-processing-time
 * scheduling.  A tile is computed by the same gemm_tile as the single-GEMM
 * kernels, instantiated for the four transpose pairs and chosen per tile.
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"
#include "gemmx_tile.h"

using namespace gemmx;

namespace {

constexpr int BT_BM = 64, BT_BN = 64, BT_BK = 16, BT_WM = 2, BT_WN = 2;
constexpr int BT_THREADS = 64 * BT_WM * BT_WN;

__global__ void batch_reset_kernel(unsigned *counter) { *counter = 0u; }

/* Operand pointers read from the descriptors are generic; marking them as
 * global lets every load through them be a global (not flat) load. */
template <typename T>
__device__ __forceinline__ T *as_global(T *p) {
  return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}

template <typename T>
__global__ void __launch_bounds__(BT_THREADS, 2)
gemm_batch_kernel(const gemmx_batch_desc *__restrict__ desc,
                  const int *__restrict__ tile_prob, int total,
                  unsigned *__restrict__ counter) {
  __shared__ __align__(16) T lds[tile_lds_max<T, BT_BM, BT_BN, BT_BK>()];
  __shared__ int ticket;
  for (;;) {
    if (threadIdx.x == 0) ticket = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(ticket);
    __syncthreads(); /* everyone has read the ticket before it is replaced */
    if (t >= total) break;
    const int p = __builtin_amdgcn_readfirstlane(tile_prob[t]);
    const gemmx_batch_desc &q = desc[p];
    const int local = t - q.tile_begin;
    const int tm = local % q.tiles_m, tn = local / q.tiles_m;
    const int m0 = tm * BT_BM, n0 = tn * BT_BN;
    const T alpha = (T)q.alpha, beta = (T)q.beta;
    const T *A = as_global((const T *)q.a);
    const T *B = as_global((const T *)q.b);
    T *C = as_global((T *)q.c);
    const int f = q.flags;
    if (f & GEMMX_BF_SCALE) {
      /* alpha == 0 or k == 0: C = beta*C on this tile (beta != 1 here) */
      for (int e = threadIdx.x; e < BT_BM * BT_BN; e += BT_THREADS) {
        const int i = m0 + e % BT_BM, j = n0 + e / BT_BM;
        if (i < q.m && j < q.n) C[i + j * q.ldc] *= beta;
      }
    } else {
      const int va = (f & GEMMX_BF_VEC_A) != 0, vb = (f & GEMMX_BF_VEC_B) != 0;
#define GEMMX_TILE(TA, TB)                                                              \
  gemm_tile<T, TA, TB, BT_BM, BT_BN, BT_BK, BT_WM, BT_WN, 1, false>(                    \
      A, B, C, q.m, q.n, q.k, q.lda, q.ldb, q.ldc, alpha, beta, m0, n0, va, vb, lds)
      switch (f & (GEMMX_BF_TA | GEMMX_BF_TB)) {
      case 0: GEMMX_TILE(false, false); break;
      case GEMMX_BF_TB: GEMMX_TILE(false, true); break;
      case GEMMX_BF_TA: GEMMX_TILE(true, false); break;
      default: GEMMX_TILE(true, true); break;
      }
#undef GEMMX_TILE
    }
    __syncthreads(); /* LDS of this tile is free before the next tile's stores */
  }
}

template <typename T>
int batch_grid() {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gemm_batch_kernel<T>, BT_THREADS, 0) !=
      hipSuccess)
    return 0;
  return cus * (per > 0 ? per : 1);
}

} // namespace

extern "C" int gemmx_batch_tile(void) { return BT_BM; }

extern "C" int gemmx_batch_grid(int es) {
  return es == 8 ? batch_grid<double>() : batch_grid<float>();
}

extern "C" hipError_t gemmx_batch_launch(int es, const gemmx_batch_desc *d_desc,
                                         const int *d_tile_prob, int total_tiles,
                                         unsigned *d_counter, int grid, hipStream_t s) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, s, d_counter);
  if (es == 8)
    hipLaunchKernelGGL(gemm_batch_kernel<double>, dim3(grid), dim3(BT_THREADS), 0, s, d_desc,
                       d_tile_prob, total_tiles, d_counter);
  else
    hipLaunchKernelGGL(gemm_batch_kernel<float>, dim3(grid), dim3(BT_THREADS), 0, s, d_desc,
                       d_tile_prob, total_tiles, d_counter);
  return hipGetLastError();
}
