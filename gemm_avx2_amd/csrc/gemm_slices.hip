/*
 * gemm_slices.hip -- the emulated forms: products of low-precision slices on
 * the INT8 and BF16 matrix cores.  No reference counterpart (the reference's
 * sgemm_/dgemm_, wjc404/GEMM_AVX2 src/gemm_driver.c:118, are native); the
 * boundary semantics are the native ones: C = alpha*op(A)op(B) + beta*C with
 * the beta != 1 rule.
 *
 * Both forms run in three steps:
 *   1. split: every row of op(A) and column of op(B) is cut into slices that
 *      are stored K-contiguous ([slice][row][kpad], K zero-padded to a
 *      multiple of SL_BK), folding any transpose into this pass;
 *   2. products: one tile kernel multiplies slice pairs (t, u) on MFMA;
 *   3. epilogue: alpha/beta (and, for Ozaki, the FP64 recombination).
 *
 * Ozaki-I (DGEMM): s signed 7-bit digits per FP64 entry of x*2^-e (row or
 * column exponent e, truncation toward zero), INT8 slices, weight
 * w = t + u in 2..s+1 kept.  The product kernel sums the pairs of one weight
 * exactly in INT32 (K chunked so that s*k*127^2 < 2^31) and stores S_w; the
 * recombination kernel evaluates acc = acc*2^-7 + S_w for w = s+1..2 in FP64
 * and writes C = fma(alpha, ldexp(acc, ea_i + fb_j - 14), beta*C).  Every
 * step is exact or the same FP64 operation as the oracle's restatement, so
 * the result is bit-identical to it.
 *
 * split-BF16x3 (SGEMM): x = hi + lo with hi = bf16_rne(x), lo =
 * bf16_rne(x - hi); pairs (lo,hi), (hi,lo), (hi,hi) accumulate in one FP32
 * MFMA accumulator; the epilogue writes C = fma(alpha, acc, beta*C).
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"

using namespace gemmx;

namespace {

constexpr int SL_BK = 128;          /* K bytes per LDS stage (2 MFMA k-steps) */
constexpr int OZ_NONFINITE = 0x7fffffff;
constexpr int MAXPAIRS = 36;

inline size_t rup(size_t v, size_t a) { return (v + a - 1) / a * a; }

/* ------------------------------------------------------------------------ */
/* Split kernels                                                            */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ int exponent_of_absbits(unsigned long long b) {
  if (b >= 0x7ff0000000000000ull) return OZ_NONFINITE;
  if (b == 0) return 0;
  int e;
  (void)frexp(__longlong_as_double((long long)b), &e);
  return e;
}

/* Exponent per vector with elements contiguous (x[v*vld + p]): a wave per
 * vector. */
__global__ void __launch_bounds__(256)
exp_contig_kernel(const double *__restrict__ x, int64_t vld, int nvec, int len,
                  int *__restrict__ out) {
  const int v = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (v >= nvec) return;
  const double *p = x + v * vld;
  unsigned long long mx = 0;
  for (int i = lane; i < len; i += 64) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(p[i]) & 0x7fffffffffffffffull;
    mx = b > mx ? b : mx;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  if (lane == 0) out[v] = exponent_of_absbits(mx);
}

/* Exponent per vector with vectors adjacent (x[v + p*eld]): a thread per
 * vector, the wave's reads of one p coalesce. */
__global__ void __launch_bounds__(256)
exp_strided_kernel(const double *__restrict__ x, int64_t eld, int nvec, int len,
                   int *__restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvec) return;
  unsigned long long mx = 0;
  for (int i = 0; i < len; ++i) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x[v + i * eld]) & 0x7fffffffffffffffull;
    mx = b > mx ? b : mx;
  }
  out[v] = exponent_of_absbits(mx);
}

/* 16 consecutive K entries of one vector per thread -> 16 bytes per slice.
 * v_fast: consecutive threads take consecutive vectors (strided input). */
template <int S>
__global__ void __launch_bounds__(256)
ozaki_split_kernel(const double *__restrict__ x, int64_t vld, int64_t eld,
                   int nvec, int len, int64_t kpad, const int *__restrict__ ex,
                   int v_fast, int8_t *__restrict__ out) {
  const int64_t nblk = kpad / 16;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (tid >= nblk * nvec) return;
  int64_t v, blk;
  if (v_fast) { v = tid % nvec; blk = tid / nvec; }
  else        { v = tid / nblk; blk = tid % nblk; }
  const int e = ex[v];
  i4_t dig[S];
#pragma unroll
  for (int t = 0; t < S; ++t) dig[t] = i4_t{0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int64_t p = blk * 16 + q;
    const double val = (p < len) ? x[v * vld + p * eld] : 0.0;
    double r = (e == OZ_NONFINITE) ? 0.0 : ldexp(val, -e);
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const double y = r * 128.0;
      const double d = trunc(y);
      r = y - d;
      const unsigned byte = (unsigned)((int)d) & 0xffu;
      dig[t][q >> 2] |= (int)(byte << (8 * (q & 3)));
    }
  }
#pragma unroll
  for (int t = 0; t < S; ++t)
    *reinterpret_cast<i4_t *>(out + ((int64_t)t * nvec + v) * kpad + blk * 16) = dig[t];
}

__device__ __forceinline__ unsigned bf16_rne_bits(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

/* 8 consecutive K entries of one vector per thread -> 16 bytes of hi and of
 * lo.  out is [2][nvec][kpad] bf16 (hi slice first). */
__global__ void __launch_bounds__(256)
bf16_split_kernel(const float *__restrict__ x, int64_t vld, int64_t eld,
                  int nvec, int len, int64_t kpad, int v_fast,
                  unsigned short *__restrict__ out) {
  const int64_t nblk = kpad / 8;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (tid >= nblk * nvec) return;
  int64_t v, blk;
  if (v_fast) { v = tid % nvec; blk = tid / nvec; }
  else        { v = tid / nblk; blk = tid % nblk; }
  i4_t hi, lo;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned h2[2], l2[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int64_t p = blk * 8 + 2 * q + e;
      const float val = (p < len) ? x[v * vld + p * eld] : 0.0f;
      h2[e] = bf16_rne_bits(val);
      l2[e] = bf16_rne_bits(val - __uint_as_float(h2[e] << 16));
    }
    hi[q] = (int)(h2[0] | (h2[1] << 16));
    lo[q] = (int)(l2[0] | (l2[1] << 16));
  }
  *reinterpret_cast<i4_t *>(out + v * kpad + blk * 8) = hi;
  *reinterpret_cast<i4_t *>(out + ((int64_t)nvec + v) * kpad + blk * 8) = lo;
}

/* ------------------------------------------------------------------------ */
/* Slice-pair product kernel                                                */
/* ------------------------------------------------------------------------ */
struct Pairs {
  int count;
  signed char t[MAXPAIRS], u[MAXPAIRS], last[MAXPAIRS], w[MAXPAIRS];
};

/* MFMA with 64 K-bytes per lane group step: INT8 16x16x64 -> INT32, BF16
 * 16x16x32 -> FP32.  Operand: 16 bytes per lane, lane l covers row l%16 and
 * K bytes [16*(l/16), 16*(l/16)+16) of the 64; A and B use the same K bytes
 * per lane, so any fixed hardware order of those bytes cancels out.
 * D: lane l, element r holds row 4*(l/16)+r, column l%16. */
template <bool BF16> struct SliceMfma;
template <> struct SliceMfma<false> {
  typedef i4_t acc_t;
  static __device__ __forceinline__ acc_t mma(i4_t a, i4_t b, acc_t c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
  }
};
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
template <> struct SliceMfma<true> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t mma(i4_t a, i4_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                   __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
  }
};

/* One BM x BN tile of C per workgroup, walking (pair, K stage).
 * Ozaki (BF16 = false): when a pair closes its weight the INT32 sums are
 * stored to S[w-2] (column-major m x n, ld m) and reset.
 * BF16x3 (BF16 = true): all pairs accumulate; the epilogue writes
 * C = fma(alpha, acc, beta*C). */
template <bool BF16, int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
slice_products_kernel(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs,
                      int M, int N, int64_t kpad_bytes, int kt_begin, int kt_end,
                      Pairs pairs, int tiles_m, int tiles_n,
                      int *__restrict__ S, float *__restrict__ C, int64_t ldc,
                      float alpha, float beta) {
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  constexpr int LEN = SL_BK / 4; /* ints per LDS line */
  constexpr int LS = LEN + 4;    /* +16 bytes against bank conflicts */
  constexpr int KSTEPS = SL_BK / 64;
  typedef SliceMfma<BF16> MM;
  typedef typename MM::acc_t acc_t;
  extern __shared__ __align__(16) int smem[];
  /* LDS: A stage 0 | A stage 1 | B stage 0 | B stage 1 */
  auto sA = [&](int st) { return smem + st * (BM * LS); };
  auto sB = [&](int st) { return smem + 2 * BM * LS + st * (BN * LS); };

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, kg = lane >> 4;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, &tm, &tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int64_t ld_i = kpad_bytes / 4;

  TileCopy<int, BM, LEN, NT> ca;
  TileCopy<int, BN, LEN, NT> cb;
  const int nkt = kt_end - kt_begin;
  const int iters = pairs.count * nkt;
  auto load_iter = [&](int q) {
    const int pi = q / nkt, kt = kt_begin + q % nkt;
    const int t = pairs.t[pi], u = pairs.u[pi];
    const int *ga = reinterpret_cast<const int *>(As + ((int64_t)t * M + m0) * kpad_bytes) + kt * LEN;
    const int *gb = reinterpret_cast<const int *>(Bs + ((int64_t)u * N + n0) * kpad_bytes) + kt * LEN;
    ca.load(ga, ld_i, M - m0, LEN, true, tid);
    cb.load(gb, ld_i, N - n0, LEN, true, tid);
  };

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  load_iter(0);
  ca.template store<LS>(sA(0), tid);
  cb.template store<LS>(sB(0), tid);
  __syncthreads();

  for (int q = 0; q < iters; ++q) {
    const int cur = q & 1;
    if (q + 1 < iters) load_iter(q + 1);
    const int *as = sA(cur);
    const int *bs = sB(cur);
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) {
      const int koff = ks * 16 + kg * 4;
      i4_t af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
        af[i] = *reinterpret_cast<const i4_t *>(as + (wm * WTM + i * 16 + li) * LS + koff);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bf[j] = *reinterpret_cast<const i4_t *>(bs + (wn * WTN + j * 16 + li) * LS + koff);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MM::mma(af[i], bf[j], acc[i][j]);
    }
    if (!BF16 && (q % nkt) == nkt - 1 && pairs.last[q / nkt]) {
      int *Sw = S + (int64_t)(pairs.w[q / nkt] - 2) * M * N;
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int col = n0 + wn * WTN + j * 16 + li;
          const int row = m0 + wm * WTM + i * 16 + kg * 4;
          if (col < N) {
            int *sp = Sw + row + (int64_t)col * M;
            if (row + 3 < M && (M & 3) == 0) {
              *reinterpret_cast<i4_t *>(sp) = *reinterpret_cast<i4_t *>(&acc[i][j]);
            } else {
#pragma unroll
              for (int r = 0; r < 4; ++r)
                if (row + r < M) sp[r] = (int)acc[i][j][r];
            }
          }
          acc[i][j] = acc_t{0, 0, 0, 0};
        }
    }
    if (q + 1 < iters) {
      ca.template store<LS>(sA(cur ^ 1), tid);
      cb.template store<LS>(sB(cur ^ 1), tid);
    }
    __syncthreads();
  }

  if (BF16) {
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
            if (row + r < M) cp[r] = __builtin_fmaf(alpha, (float)acc[i][j][r], beta * cp[r]);
        }
      }
  }
}

/* FP64 recombination of the weight sums and the alpha/beta epilogue. */
__global__ void __launch_bounds__(256)
ozaki_recombine_kernel(const int *__restrict__ S, int s, const int *__restrict__ ea,
                       const int *__restrict__ fb, double *__restrict__ C,
                       int64_t ldc, int M, int N, double alpha, double beta) {
  const int64_t total = (int64_t)M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / M, i = t - j * M;
    double acc = 0.0;
    for (int w = s + 1; w >= 2; --w)
      acc = __builtin_fma(acc, 0x1p-7, (double)S[(int64_t)(w - 2) * total + t]);
    const int e = ea[i], f = fb[j];
    const double scaled = (e == OZ_NONFINITE || f == OZ_NONFINITE) ? __builtin_nan("")
                                                                  : ldexp(acc, e + f - 14);
    double *cp = C + i + j * ldc;
    *cp = __builtin_fma(alpha, scaled, beta * *cp);
  }
}

/* ------------------------------------------------------------------------ */
/* Host side                                                                */
/* ------------------------------------------------------------------------ */
constexpr int P_BM = 256, P_BN = 256, P_WM = 2, P_WN = 4;

size_t products_lds() {
  return (size_t)2 * (P_BM + P_BN) * (SL_BK / 4 + 4) * 4;
}

template <bool BF16>
hipError_t launch_products(const int8_t *as, const int8_t *bs, int m, int n,
                           int64_t kpad, int kt0, int kt1, const Pairs &pairs,
                           int *S, float *C, int64_t ldc, float alpha,
                           float beta, hipStream_t st) {
  auto kern = slice_products_kernel<BF16, P_BM, P_BN, P_WM, P_WN>;
  const size_t lds = products_lds();
  hipError_t e = hipFuncSetAttribute((const void *)kern,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int tiles_m = (m + P_BM - 1) / P_BM, tiles_n = (n + P_BN - 1) / P_BN;
  hipLaunchKernelGGL(kern, dim3(tiles_m * tiles_n), dim3(64 * P_WM * P_WN), lds, st,
                     as, bs, m, n, kpad, kt0, kt1, pairs, tiles_m, tiles_n, S, C, ldc,
                     alpha, beta);
  return hipGetLastError();
}

struct OzLayout {
  size_t ea, fb, as, bs, sw, total;
  int64_t kpad;
};
OzLayout oz_layout(int m, int n, int k, int s) {
  OzLayout L;
  L.kpad = (int64_t)rup((size_t)k, SL_BK);
  L.ea = 0;
  L.fb = rup((size_t)m * 4, 256);
  L.as = L.fb + rup((size_t)n * 4, 256);
  L.bs = L.as + rup((size_t)s * m * L.kpad, 256);
  L.sw = L.bs + rup((size_t)s * n * L.kpad, 256);
  L.total = L.sw + rup((size_t)s * m * n * 4, 256);
  return L;
}

Pairs ozaki_pairs(int s) {
  Pairs P;
  P.count = 0;
  for (int w = s + 1; w >= 2; --w) {
    const int tlo = w - s > 1 ? w - s : 1, thi = w - 1 < s ? w - 1 : s;
    for (int t = tlo; t <= thi; ++t) {
      P.t[P.count] = (signed char)(t - 1); /* slice index, 0-based */
      P.u[P.count] = (signed char)(w - t - 1);
      P.last[P.count] = (signed char)(t == thi);
      P.w[P.count] = (signed char)w;
      P.count++;
    }
  }
  return P;
}

template <int S>
hipError_t launch_ozaki_split(const double *x, int64_t vld, int64_t eld, int nvec,
                              int len, int64_t kpad, const int *ex, int v_fast,
                              int8_t *out, hipStream_t st) {
  const int64_t threads = (kpad / 16) * nvec;
  hipLaunchKernelGGL((ozaki_split_kernel<S>), dim3((unsigned)((threads + 255) / 256)),
                     dim3(256), 0, st, x, vld, eld, nvec, len, kpad, ex, v_fast, out);
  return hipGetLastError();
}

hipError_t ozaki_split(int s, const double *x, int64_t vld, int64_t eld, int nvec,
                       int len, int64_t kpad, const int *ex, int v_fast, int8_t *out,
                       hipStream_t st) {
  switch (s) {
  case 3: return launch_ozaki_split<3>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  case 4: return launch_ozaki_split<4>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  case 5: return launch_ozaki_split<5>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  case 6: return launch_ozaki_split<6>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  case 7: return launch_ozaki_split<7>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  default: return launch_ozaki_split<8>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  }
}

/* Exponents and slices of op(A) rows (vectors = rows) or op(B) columns.
 * contiguous: vector v at x + v*ld (elements adjacent); else at x + v with
 * element stride ld. */
hipError_t ozaki_prepare(const double *x, int64_t ld, bool contiguous, int nvec,
                         int k, int64_t kpad, int s, int *ex, int8_t *slices,
                         hipStream_t st) {
  if (contiguous)
    hipLaunchKernelGGL(exp_contig_kernel, dim3((nvec + 3) / 4), dim3(256), 0, st, x,
                       ld, nvec, k, ex);
  else
    hipLaunchKernelGGL(exp_strided_kernel, dim3((nvec + 255) / 256), dim3(256), 0,
                       st, x, ld, nvec, k, ex);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return contiguous ? ozaki_split(s, x, ld, 1, nvec, k, kpad, ex, 0, slices, st)
                    : ozaki_split(s, x, 1, ld, nvec, k, kpad, ex, 1, slices, st);
}

hipError_t bf16_prepare(const float *x, int64_t ld, bool contiguous, int nvec,
                        int k, int64_t kpad, unsigned short *out, hipStream_t st) {
  const int64_t threads = (kpad / 8) * nvec;
  const dim3 grid((unsigned)((threads + 255) / 256));
  if (contiguous)
    hipLaunchKernelGGL(bf16_split_kernel, grid, dim3(256), 0, st, x, ld, (int64_t)1,
                       nvec, k, kpad, 0, out);
  else
    hipLaunchKernelGGL(bf16_split_kernel, grid, dim3(256), 0, st, x, (int64_t)1, ld,
                       nvec, k, kpad, 1, out);
  return hipGetLastError();
}

} // namespace

extern "C" int gemmx_ozaki_kchunk(int slices) {
  const int64_t lim = (int64_t)2147483647 / ((int64_t)slices * 127 * 127);
  return (int)(lim - lim % SL_BK);
}

extern "C" size_t gemmx_ozaki_workspace(int m, int n, int k, int slices) {
  return oz_layout(m, n, k, slices).total;
}

extern "C" hipError_t gemmx_ozaki_f64(const gemmx_problem *p, int s, void *work,
                                      hipStream_t st) {
  if (s < 3 || s > 8 || !work) return hipErrorInvalidValue;
  const OzLayout L = oz_layout(p->m, p->n, p->k, s);
  char *w = (char *)work;
  int *ea = (int *)(w + L.ea), *fb = (int *)(w + L.fb);
  int8_t *as = (int8_t *)(w + L.as), *bs = (int8_t *)(w + L.bs);
  int *S = (int *)(w + L.sw);
  /* op(A) rows: transa N -> row i at a + i, stride lda; T -> contiguous.
   * op(B) columns: transb N -> contiguous at b + j*ldb; T -> stride ldb. */
  hipError_t e = ozaki_prepare((const double *)p->a, p->lda, p->ta != 0, p->m, p->k,
                               L.kpad, s, ea, as, st);
  if (e != hipSuccess) return e;
  e = ozaki_prepare((const double *)p->b, p->ldb, p->tb == 0, p->n, p->k, L.kpad, s,
                    fb, bs, st);
  if (e != hipSuccess) return e;
  const Pairs pairs = ozaki_pairs(s);
  const int kc_tiles = gemmx_ozaki_kchunk(s) / SL_BK;
  const int nkt = (int)(L.kpad / SL_BK);
  const int64_t total = (int64_t)p->m * p->n;
  int64_t rblocks = (total + 255) / 256;
  if (rblocks > 65536) rblocks = 65536;
  for (int kt0 = 0; kt0 < nkt; kt0 += kc_tiles) {
    const int kt1 = kt0 + kc_tiles < nkt ? kt0 + kc_tiles : nkt;
    e = launch_products<false>(as, bs, p->m, p->n, L.kpad, kt0, kt1, pairs, S, NULL, 0,
                               0.f, 0.f, st);
    if (e != hipSuccess) return e;
    /* chunks after the first add onto C: beta = 1 leaves C as it is */
    hipLaunchKernelGGL(ozaki_recombine_kernel, dim3((unsigned)rblocks), dim3(256), 0, st,
                       S, s, ea, fb, (double *)p->c, p->ldc, p->m, p->n, p->alpha,
                       kt0 == 0 ? p->beta : 1.0);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

extern "C" size_t gemmx_bf16x3_workspace(int m, int n, int k) {
  const size_t kpad = rup((size_t)k, SL_BK / 2); /* bf16 elements */
  return rup((size_t)2 * m * kpad * 2, 256) + rup((size_t)2 * n * kpad * 2, 256);
}

extern "C" hipError_t gemmx_bf16x3_f32(const gemmx_problem *p, void *work, hipStream_t st) {
  if (!work) return hipErrorInvalidValue;
  const int64_t kpad = (int64_t)rup((size_t)p->k, SL_BK / 2); /* bf16 elements */
  unsigned short *as = (unsigned short *)work;
  unsigned short *bs = (unsigned short *)((char *)work + rup((size_t)2 * p->m * kpad * 2, 256));
  hipError_t e = bf16_prepare((const float *)p->a, p->lda, p->ta != 0, p->m, p->k, kpad,
                              as, st);
  if (e != hipSuccess) return e;
  e = bf16_prepare((const float *)p->b, p->ldb, p->tb == 0, p->n, p->k, kpad, bs, st);
  if (e != hipSuccess) return e;
  Pairs pairs;
  pairs.count = 3;
  /* slice 0 = hi, 1 = lo: (lo,hi), (hi,lo), (hi,hi) */
  const signed char tt[3] = {1, 0, 0}, uu[3] = {0, 1, 0};
  for (int i = 0; i < 3; ++i) {
    pairs.t[i] = tt[i];
    pairs.u[i] = uu[i];
    pairs.last[i] = 0;
    pairs.w[i] = 0;
  }
  const int64_t kpad_bytes = kpad * 2;
  return launch_products<true>((const int8_t *)as, (const int8_t *)bs, p->m, p->n,
                               kpad_bytes, 0, (int)(kpad_bytes / SL_BK), pairs, NULL,
                               (float *)p->c, p->ldc, (float)p->alpha, (float)p->beta, st);
}
