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
 * the slice.
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"

using namespace gemmx;

namespace {

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
gemm_native_kernel(const T *__restrict__ A, const T *__restrict__ B,
                   T *__restrict__ C, int M, int N, int K, int64_t lda,
                   int64_t ldb, int64_t ldc, T alpha, T beta, int tiles_m,
                   int tiles_n, int vecA, int vecB) {
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  constexpr int KS = BK / 4;
  constexpr int VE = Vec16<T>::N;
  constexpr int A_NL = TA ? BM : BK, A_LEN = TA ? BK : BM, A_LS = A_LEN + VE;
  constexpr int B_NL = TB ? BK : BN, B_LEN = TB ? BN : BK, B_LS = B_LEN + VE;
  static_assert(WTM % 16 == 0 && WTN % 16 == 0 && BK % 4 == 0, "tile shape");

  __shared__ __align__(16) T sA[2][A_NL * A_LS];
  __shared__ __align__(16) T sB[2][B_NL * B_LS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, kg = lane >> 4;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, &tm, &tn);
  const int m0 = tm * BM, n0 = tn * BN;

  TileCopy<T, A_NL, A_LEN, NT> ca;
  TileCopy<T, B_NL, B_LEN, NT> cb;
  auto load_slice = [&](int k0) {
    if (!TA) ca.load(A + m0 + k0 * lda, lda, K - k0, M - m0, vecA, tid);
    else     ca.load(A + k0 + m0 * lda, lda, M - m0, K - k0, vecA, tid);
    if (!TB) cb.load(B + k0 + n0 * ldb, ldb, N - n0, K - k0, vecB, tid);
    else     cb.load(B + n0 + k0 * ldb, ldb, K - k0, N - n0, vecB, tid);
  };

  typedef typename Mfma16<T>::acc_t acc_t;
  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = T(0);

  const int nk = (K + BK - 1) / BK;
  load_slice(0);
  ca.template store<A_LS>(sA[0], tid);
  cb.template store<B_LS>(sB[0], tid);
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_slice((kt + 1) * BK);
    const T *as = sA[cur];
    const T *bs = sB[cur];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int kx = kg * KS + s;
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int row = wm * WTM + i * 16 + li;
        af[i] = TA ? as[row * A_LS + kx] : as[kx * A_LS + row];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = wn * WTN + j * 16 + li;
        bf[j] = TB ? bs[kx * B_LS + col] : bs[col * B_LS + kx];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = Mfma16<T>::mma(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < nk) {
      ca.template store<A_LS>(sA[cur ^ 1], tid);
      cb.template store<B_LS>(sB[cur ^ 1], tid);
    }
    __syncthreads();
  }

  /* Epilogue: rows m..ldc-1 and columns >= n are never touched, matching the
   * reference's masked stores (src/kernel/gemm_kernel_unroll6.S:145-155). */
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn * WTN + j * 16 + li;
      const int row0 = m0 + wm * WTM + i * 16;
      if (col < N) {
        T *cp = C + row0 + col * ldc;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int dr = Mfma16<T>::row(kg, r);
          if (row0 + dr < M) cp[dr] = fma_t(alpha, acc[i][j][r], beta * cp[dr]);
        }
      }
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
scale_kernel(T *__restrict__ C, int64_t ldc, int M, int N, T beta) {
  const int64_t total = (int64_t)M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / M, i = t - j * M;
    C[i + j * ldc] *= beta;
  }
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
hipError_t launch_cfg(const gemmx_problem *p, hipStream_t s) {
  const int tiles_m = (p->m + BM - 1) / BM, tiles_n = (p->n + BN - 1) / BN;
  const dim3 grid(tiles_m * tiles_n), block(64 * WM * WN);
  constexpr int VE = Vec16<T>::N;
  const int vecA = ((uintptr_t)p->a % 16 == 0) && (p->lda % VE == 0);
  const int vecB = ((uintptr_t)p->b % 16 == 0) && (p->ldb % VE == 0);
  const T *A = (const T *)p->a, *B = (const T *)p->b;
  T *C = (T *)p->c;
  const T alpha = (T)p->alpha, beta = (T)p->beta;
#define GEMMX_LAUNCH(TA, TB)                                                  \
  hipLaunchKernelGGL((gemm_native_kernel<T, TA, TB, BM, BN, BK, WM, WN>), grid, \
                     block, 0, s, A, B, C, p->m, p->n, p->k, p->lda, p->ldb,    \
                     p->ldc, alpha, beta, tiles_m, tiles_n, vecA, vecB)
  if (!p->ta && !p->tb) GEMMX_LAUNCH(false, false);
  else if (!p->ta && p->tb) GEMMX_LAUNCH(false, true);
  else if (p->ta && !p->tb) GEMMX_LAUNCH(true, false);
  else GEMMX_LAUNCH(true, true);
#undef GEMMX_LAUNCH
  return hipGetLastError();
}

template <typename T>
hipError_t launch_native(const gemmx_problem *p, int cfg, hipStream_t s) {
  switch (cfg) {
  case GEMMX_NCFG_64x64: return launch_cfg<T, 64, 64, 16, 2, 2>(p, s);
  case GEMMX_NCFG_256x128: return launch_cfg<T, 256, 128, 16, 4, 2>(p, s);
  case GEMMX_NCFG_128x128:
  default: return launch_cfg<T, 128, 128, 16, 2, 2>(p, s);
  }
}

template <typename T>
hipError_t launch_scale(T *c, int64_t ldc, int m, int n, T beta, hipStream_t s) {
  const int64_t total = (int64_t)m * n;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) return hipSuccess;
  hipLaunchKernelGGL((scale_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s,
                     c, ldc, m, n, beta);
  return hipGetLastError();
}

} // namespace

extern "C" hipError_t gemmx_native_f64(const gemmx_problem *p, int cfg, hipStream_t s) {
  return launch_native<double>(p, cfg, s);
}
extern "C" hipError_t gemmx_native_f32(const gemmx_problem *p, int cfg, hipStream_t s) {
  return launch_native<float>(p, cfg, s);
}
extern "C" hipError_t gemmx_scale_f64(double *c, int64_t ldc, int m, int n,
                                      double beta, hipStream_t s) {
  return launch_scale<double>(c, ldc, m, n, beta, s);
}
extern "C" hipError_t gemmx_scale_f32(float *c, int64_t ldc, int m, int n,
                                      float beta, hipStream_t s) {
  return launch_scale<float>(c, ldc, m, n, beta, s);
}
