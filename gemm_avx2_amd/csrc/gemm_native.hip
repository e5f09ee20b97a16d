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
 * Split-K (a single problem with few tiles; the reference instead divides
 * every call's N range over all its threads, src/gemm_driver.c:128-151):
 * grid row q computes the K range [q kc, (q+1) kc) of every tile (kc whole
 * BK slices) and stores the raw accumulators, in the registers' fragment
 * order, to a workspace of partial tiles; splitk_reduce_kernel then sums the
 * ranges in order q = 0, 1, ... (so the result does not depend on which
 * workgroup finished first) and writes C = alpha*sum + beta*C.
 *
 * The kernels and launchers are in gemm_native_impl.h, instantiated by one
 * translation unit per element type and transpose pair
 * (gemm_native_{f64,f32}_{nn,nt,tn,tt}.hip); this file dispatches.
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
#include <pthread.h>

#include "gemm_native_impl.h"

using namespace gemmx;

extern "C" hipError_t gemmx_native_f64_nn(const gemmx_problem *p, int cfg, int split, void *work,
                                      hipStream_t s);
extern "C" hipError_t gemmx_native_f64_nt(const gemmx_problem *p, int cfg, int split, void *work,
                                      hipStream_t s);
extern "C" hipError_t gemmx_native_f64_tn(const gemmx_problem *p, int cfg, int split, void *work,
                                      hipStream_t s);
extern "C" hipError_t gemmx_native_f64_tt(const gemmx_problem *p, int cfg, int split, void *work,
                                      hipStream_t s);
extern "C" hipError_t gemmx_native_f32_nn(const gemmx_problem *p, int cfg, int split, void *work,
                                      hipStream_t s);
extern "C" hipError_t gemmx_native_f32_nt(const gemmx_problem *p, int cfg, int split, void *work,
                                      hipStream_t s);
extern "C" hipError_t gemmx_native_f32_tn(const gemmx_problem *p, int cfg, int split, void *work,
                                      hipStream_t s);
extern "C" hipError_t gemmx_native_f32_tt(const gemmx_problem *p, int cfg, int split, void *work,
                                      hipStream_t s);

namespace {

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

/* Device memory of the allocation-free entry points (dgemm_x_dev_noalloc,
 * sgemm_x_dev_noalloc): the split-K partial tiles of their calls.  One array
 * per GPU the library runs on, held until the process ends; every such call
 * on a GPU uses the same array, so calls that overlap on one GPU (on
 * different streams, or from several host threads) race.  Nothing in it has
 * to start at zero: the partials are written before they are read. */
static __device__ __align__(256) unsigned char scratch[GEMMX_NOALLOC_SCRATCH];

/* The address of `scratch` on each device: the array's name is a host-side
 * stand-in, so its device address is taken once per device, with that device
 * current, and only that address is handed to kernels. */
static void *g_scratch_ptr[GEMMX_NOALLOC_MAX_DEV];
static pthread_mutex_t g_scratch_lock = PTHREAD_MUTEX_INITIALIZER;

extern "C" hipError_t gemmx_noalloc_scratch(int dev, void **ptr) {
  if (dev < 0 || dev >= GEMMX_NOALLOC_MAX_DEV) return hipErrorInvalidDevice;
  pthread_mutex_lock(&g_scratch_lock);
  hipError_t e = hipSuccess;
  if (!g_scratch_ptr[dev]) {
    void *p = NULL;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) g_scratch_ptr[dev] = p;
    else if (e == hipSuccess) e = hipErrorInvalidSymbol;
  }
  *ptr = g_scratch_ptr[dev];
  pthread_mutex_unlock(&g_scratch_lock);
  return e;
}

extern "C" hipError_t gemmx_native_f64(const gemmx_problem *p, int cfg, int split, void *work,
                                       hipStream_t s) {
  if (!p->ta) return p->tb ? gemmx_native_f64_nt(p, cfg, split, work, s)
                           : gemmx_native_f64_nn(p, cfg, split, work, s);
  return p->tb ? gemmx_native_f64_tt(p, cfg, split, work, s)
               : gemmx_native_f64_tn(p, cfg, split, work, s);
}
extern "C" hipError_t gemmx_native_f32(const gemmx_problem *p, int cfg, int split, void *work,
                                       hipStream_t s) {
  if (!p->ta) return p->tb ? gemmx_native_f32_nt(p, cfg, split, work, s)
                           : gemmx_native_f32_nn(p, cfg, split, work, s);
  return p->tb ? gemmx_native_f32_tt(p, cfg, split, work, s)
               : gemmx_native_f32_tn(p, cfg, split, work, s);
}
extern "C" int gemmx_native_tile(int cfg, int *bm, int *bn) {
  if (cfg < 0 || cfg >= GEMMX_NCFG_COUNT) return -1;
  *bm = kNativeCfg[cfg].bm;
  *bn = kNativeCfg[cfg].bn;
  return kNativeCfg[cfg].split_ok;
}
extern "C" int gemmx_native_splits(int es, int cfg, int k, int split) {
  if (cfg < 0 || cfg >= GEMMX_NCFG_COUNT || !kNativeCfg[cfg].split_ok || split <= 1 || k < 1)
    return 1;
  const int kc = split_depth(k, split, es == 8 ? kNativeCfg[cfg].bk64 : kNativeCfg[cfg].bk32);
  return (k + kc - 1) / kc;
}
extern "C" size_t gemmx_native_workspace(int es, int cfg, int m, int n, int k, int split) {
  const int ns = gemmx_native_splits(es, cfg, k, split);
  if (ns <= 1 || m <= 0 || n <= 0) return 0;
  const NativeCfg &c = kNativeCfg[cfg];
  return (size_t)ns * (((size_t)m + c.bm - 1) / c.bm) * (((size_t)n + c.bn - 1) / c.bn) *
         (size_t)c.bm * (size_t)c.bn * (size_t)es;
}
extern "C" hipError_t gemmx_scale_f64(double *c, int64_t ldc, int m, int n,
                                      double beta, int batch, int64_t sc, hipStream_t s) {
  return launch_scale<double>(c, ldc, m, n, beta, batch, sc, s);
}
extern "C" hipError_t gemmx_scale_f32(float *c, int64_t ldc, int m, int n,
                                      float beta, int batch, int64_t sc, hipStream_t s) {
  return launch_scale<float>(c, ldc, m, n, beta, batch, sc, s);
}
