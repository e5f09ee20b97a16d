/*
 * gemm_bf16x3.hip -- emulated SGEMM: split-BF16x3 on BF16 MFMA.
 *
 * No reference counterpart (the reference's sgemm_ is native FP32,
 * wjc404/GEMM_AVX2 src/gemm_driver.c:118 built without -DDOUBLE); the
 * boundary semantics are the same as the native path:
 * C = alpha*op(A)*op(B) + (beta != 1 ? beta*C : C).
 *
 * Every FP32 operand x is split while it is staged into LDS:
 *   hi = bf16_rne(x),  lo = bf16_rne(x - hi)    (x - hi is exact in FP32)
 * and the product keeps hi*hi + hi*lo + lo*hi (lo*lo, ~2^-16 relative, is
 * dropped), each a v_mfma_f32_16x16x32_bf16 accumulating in FP32.  The
 * transpose is folded into the staging: each thread gathers 8 consecutive K
 * values of one row of op(A) (or column of op(B)) and writes them as one
 * 16-byte LDS line segment, so the MFMA operands are always K-contiguous.
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"

using namespace gemmx;

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

/* Round-to-nearest-even FP32 -> BF16 bits; Inf stays Inf, NaN stays NaN. */
__device__ __forceinline__ unsigned bf16_rne_bits(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

/* Gathers 8 consecutive K values of NV vectors (rows of op(A) or columns of
 * op(B)) per thread and splits them.  Element p of vector v is at
 * g[v*vld + p*eld] (one of vld/eld is 1).  Units are (vector, 8-wide K
 * group); consecutive threads take consecutive vectors when eld != 1 so that
 * each scalar load coalesces across the wave. */
template <int NV, int BK, int NT>
struct SplitStage {
  static constexpr int KG = BK / 8;
  static constexpr int UNITS = NV * KG;
  static constexpr int PER = UNITS / NT;
  static_assert(UNITS % NT == 0, "units must divide evenly over the threads");
  i4_t hi[PER], lo[PER];

  __device__ __forceinline__ void unit(int u, bool v_fast, int *v, int *g) const {
    if (v_fast) { *v = u % NV; *g = u / NV; }
    else        { *v = u / KG; *g = u % KG; }
  }

  __device__ __forceinline__ void load(const float *__restrict__ base, int64_t vld,
                                       int64_t eld, int nv_valid, int k_valid,
                                       bool vec_ok, int tid) {
    const bool v_fast = eld != 1;
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      int v, g;
      unit(tid + c * NT, v_fast, &v, &g);
      float x[8];
      if (!v_fast && vec_ok && v < nv_valid && g * 8 + 8 <= k_valid) {
        const f4_t *p = reinterpret_cast<const f4_t *>(base + v * vld + g * 8);
        const f4_t x0 = p[0], x1 = p[1];
#pragma unroll
        for (int e = 0; e < 4; ++e) { x[e] = x0[e]; x[4 + e] = x1[e]; }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int kk = g * 8 + e;
          x[e] = (v < nv_valid && kk < k_valid) ? base[v * vld + kk * eld] : 0.0f;
        }
      }
      unsigned h[8], l[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        h[e] = bf16_rne_bits(x[e]);
        l[e] = bf16_rne_bits(x[e] - __uint_as_float(h[e] << 16));
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        hi[c][q] = (int)(h[2 * q] | (h[2 * q + 1] << 16));
        lo[c][q] = (int)(l[2 * q] | (l[2 * q + 1] << 16));
      }
    }
  }

  /* LDS lines hold LS bf16 per vector; stores are 16-byte aligned. */
  template <int LS>
  __device__ __forceinline__ void store(unsigned short *sh, unsigned short *sl,
                                        bool v_fast, int tid) const {
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      int v, g;
      unit(tid + c * NT, v_fast, &v, &g);
      *reinterpret_cast<i4_t *>(sh + v * LS + g * 8) = hi[c];
      *reinterpret_cast<i4_t *>(sl + v * LS + g * 8) = lo[c];
    }
  }
};

template <bool TA, bool TB, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
bf16x3_kernel(const float *__restrict__ A, const float *__restrict__ B,
              float *__restrict__ C, int M, int N, int K, int64_t lda,
              int64_t ldb, int64_t ldc, float alpha, float beta, int tiles_m,
              int tiles_n, int vecA, int vecB) {
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  constexpr int LS = BK + 8; /* bf16 per LDS line: +16 bytes against conflicts */
  constexpr int KSTEPS = BK / 32;
  __shared__ __align__(16) unsigned short sAh[2][BM * LS], sAl[2][BM * LS];
  __shared__ __align__(16) unsigned short sBh[2][BN * LS], sBl[2][BN * LS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, kg = lane >> 4;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, &tm, &tn);
  const int m0 = tm * BM, n0 = tn * BN;

  /* op(A)[i,p]: transa N at A[i + p*lda], T at A[p + i*lda].
   * op(B)[p,j]: transb N at B[p + j*ldb], T at B[j + p*ldb]. */
  const int64_t a_vld = TA ? lda : 1, a_eld = TA ? 1 : lda;
  const int64_t b_vld = TB ? 1 : ldb, b_eld = TB ? ldb : 1;
  SplitStage<BM, BK, NT> sa;
  SplitStage<BN, BK, NT> sb;
  auto load_slice = [&](int k0) {
    sa.load(A + m0 * a_vld + k0 * a_eld, a_vld, a_eld, M - m0, K - k0, vecA, tid);
    sb.load(B + n0 * b_vld + k0 * b_eld, b_vld, b_eld, N - n0, K - k0, vecB, tid);
  };

  f4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  const int nk = (K + BK - 1) / BK;
  load_slice(0);
  sa.template store<LS>(sAh[0], sAl[0], a_eld != 1, tid);
  sb.template store<LS>(sBh[0], sBl[0], b_eld != 1, tid);
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_slice((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) {
      const int koff = ks * 32 + kg * 8;
      bf16x8_t ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int o = (wm * WTM + i * 16 + li) * LS + koff;
        ah[i] = *reinterpret_cast<const bf16x8_t *>(sAh[cur] + o);
        al[i] = *reinterpret_cast<const bf16x8_t *>(sAl[cur] + o);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int o = (wn * WTN + j * 16 + li) * LS + koff;
        bh[j] = *reinterpret_cast<const bf16x8_t *>(sBh[cur] + o);
        bl[j] = *reinterpret_cast<const bf16x8_t *>(sBl[cur] + o);
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nk) {
      sa.template store<LS>(sAh[cur ^ 1], sAl[cur ^ 1], a_eld != 1, tid);
      sb.template store<LS>(sBh[cur ^ 1], sBl[cur ^ 1], b_eld != 1, tid);
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn * WTN + j * 16 + li;
      const int row = m0 + wm * WTM + i * 16 + kg * 4;
      if (col < N) {
        float *cp = C + row + col * ldc;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row + r < M) cp[r] = __builtin_fmaf(alpha, acc[i][j][r], beta * cp[r]);
      }
    }
}

} // namespace

extern "C" hipError_t gemmx_bf16x3_f32(const gemmx_problem *p, hipStream_t s) {
  constexpr int BM = 128, BN = 128, BK = 32, WM = 2, WN = 2;
  const int tiles_m = (p->m + BM - 1) / BM, tiles_n = (p->n + BN - 1) / BN;
  const dim3 grid(tiles_m * tiles_n), block(64 * WM * WN);
  const int vecA = ((uintptr_t)p->a % 16 == 0) && (p->lda % 4 == 0);
  const int vecB = ((uintptr_t)p->b % 16 == 0) && (p->ldb % 4 == 0);
  const float *A = (const float *)p->a, *B = (const float *)p->b;
  float *C = (float *)p->c;
  const float alpha = (float)p->alpha, beta = (float)p->beta;
#define GEMMX_LAUNCH(TA, TB)                                                   \
  hipLaunchKernelGGL((bf16x3_kernel<TA, TB, BM, BN, BK, WM, WN>), grid, block, \
                     0, s, A, B, C, p->m, p->n, p->k, p->lda, p->ldb, p->ldc,  \
                     alpha, beta, tiles_m, tiles_n, vecA, vecB)
  if (!p->ta && !p->tb) GEMMX_LAUNCH(false, false);
  else if (!p->ta && p->tb) GEMMX_LAUNCH(false, true);
  else if (p->ta && !p->tb) GEMMX_LAUNCH(true, false);
  else GEMMX_LAUNCH(true, true);
#undef GEMMX_LAUNCH
  return hipGetLastError();
}
