/*
 * gemm_ozaki.hip -- emulated DGEMM: Ozaki-I INT8 slicing on INT8 MFMA.
 *
 * No reference counterpart (the reference only has the native FP64 path,
 * wjc404/GEMM_AVX2 src/gemm_driver.c:118-197); the semantics at the boundary
 * are the same: C = alpha*op(A)*op(B) + (beta != 1 ? beta*C : C).
 *
 * Algorithm (s = slices, 3..8):
 *   1. Per row i of op(A) the exponent ea_i with max|a_i.| < 2^ea_i, per
 *      column j of op(B) the exponent fb_j (FP64 only here).
 *   2. x = a*2^-ea (|x| < 1) is cut into s signed 7-bit digits by truncation:
 *      x = sum_t d_t 2^-7t + r, |d_t| <= 127, all steps exact in FP64.
 *      Slices are stored K-contiguous: As[t][i][p], Bs[u][j][p], K padded with
 *      zeros to a multiple of 128.
 *   3. For each weight w = t+u (s+1 down to 2), the INT8 MFMA
 *      v_mfma_i32_16x16x64_i8 accumulates S_w = sum_{t+u=w} sum_p d_t e_u
 *      exactly in INT32 (K is split so that s*k*127^2 < 2^31).
 *   4. acc = acc*2^-7 + S_w in FP64 (acc*2^-7 is exact), then
 *      C = fma(alpha, ldexp(acc, ea_i + fb_j - 14), beta*C).
 * The result is deterministic and bit-identical to the CPU restatement in the
 * oracle (integer sums are exact; every FP64 step is the same operation).
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"

using namespace gemmx;

namespace {

constexpr int OZ_NONFINITE = 0x7fffffff;
constexpr int OZ_BK = 128; /* K bytes per LDS stage; K is padded to this */
constexpr int OZ_MAXPAIRS = 36;

inline size_t rup(size_t v, size_t a) { return (v + a - 1) / a * a; }
inline int64_t kpad_of(int k) { return (int64_t)rup((size_t)k, OZ_BK); }

struct OzLayout {
  size_t ea, fb, as, bs, total;
};
inline OzLayout oz_layout(int m, int n, int k, int s) {
  OzLayout L;
  const size_t kp = (size_t)kpad_of(k);
  L.ea = 0;
  L.fb = rup((size_t)m * 4, 256);
  L.as = L.fb + rup((size_t)n * 4, 256);
  L.bs = L.as + rup((size_t)s * m * kp, 256);
  L.total = L.bs + rup((size_t)s * n * kp, 256);
  return L;
}

__device__ __forceinline__ int exponent_of_absbits(unsigned long long b) {
  if (b >= 0x7ff0000000000000ull) return OZ_NONFINITE;
  if (b == 0) return 0;
  int e;
  (void)frexp(__longlong_as_double((long long)b), &e);
  return e;
}

/* Exponent per vector, vectors contiguous along the element axis
 * (element p of vector v at x[v*vld + p]): one wave per vector. */
__global__ void __launch_bounds__(256)
exp_contig_kernel(const double *__restrict__ x, int64_t vld, int nvec, int len,
                  int *__restrict__ out) {
  const int v = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (v >= nvec) return;
  const double *p = x + v * vld;
  unsigned long long mx = 0;
  for (int i = lane; i < len; i += 64) {
    unsigned long long b = (unsigned long long)__double_as_longlong(p[i]) & 0x7fffffffffffffffull;
    mx = b > mx ? b : mx;
  }
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  if (lane == 0) out[v] = exponent_of_absbits(mx);
}

/* Exponent per vector, vectors adjacent in memory (element p of vector v at
 * x[v + p*eld]): one thread per vector, a wave reads 64 adjacent values. */
__global__ void __launch_bounds__(256)
exp_strided_kernel(const double *__restrict__ x, int64_t eld, int nvec, int len,
                   int *__restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvec) return;
  unsigned long long mx = 0;
  for (int i = 0; i < len; ++i) {
    unsigned long long b = (unsigned long long)__double_as_longlong(x[v + i * eld]) & 0x7fffffffffffffffull;
    mx = b > mx ? b : mx;
  }
  out[v] = exponent_of_absbits(mx);
}

/* Slices of 16 consecutive K entries of one vector per thread.  Element p of
 * vector v at x[v*vld + p*eld].  v_fast: consecutive threads take
 * consecutive vectors (eld != 1, reads coalesce across vectors); otherwise
 * consecutive K blocks of one vector. */
template <int S>
__global__ void __launch_bounds__(256)
split_kernel(const double *__restrict__ x, int64_t vld, int64_t eld, int nvec,
             int len, int64_t kpad, const int *__restrict__ ex, int v_fast,
             int8_t *__restrict__ out) {
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

struct OzPairs {
  int count;
  signed char t[OZ_MAXPAIRS], u[OZ_MAXPAIRS], last[OZ_MAXPAIRS];
};

/* INT8 slice products and FP64 recombination for one BM x BN tile of C.
 * Iterations run over (slice pair, K stage); a pair flagged `last` closes its
 * weight: the INT32 sum is folded into the FP64 accumulator and reset. */
template <int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
ozaki_gemm_kernel(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs,
                  const int *__restrict__ ea, const int *__restrict__ fb,
                  double *__restrict__ C, int64_t ldc, int M, int N,
                  int64_t kpad, int kt_begin, int kt_end, double alpha,
                  double beta, int tiles_m, int tiles_n, OzPairs pairs) {
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  constexpr int LEN = OZ_BK / 4;   /* ints per LDS line */
  constexpr int LS = LEN + 4;      /* +16 bytes against bank conflicts */
  constexpr int KSTEPS = OZ_BK / 64;
  __shared__ __align__(16) int sA[2][BM * LS];
  __shared__ __align__(16) int sB[2][BN * LS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, kg = lane >> 4;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, &tm, &tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int64_t ld_i = kpad / 4; /* line stride in ints */

  TileCopy<int, BM, LEN, NT> ca;
  TileCopy<int, BN, LEN, NT> cb;
  const int nkt = kt_end - kt_begin;
  const int iters = pairs.count * nkt;
  auto load_iter = [&](int q) {
    const int pi = q / nkt, kt = kt_begin + q % nkt;
    const int t = pairs.t[pi], u = pairs.u[pi];
    const int *ga = reinterpret_cast<const int *>(As + ((int64_t)(t - 1) * M + m0) * kpad) + kt * LEN;
    const int *gb = reinterpret_cast<const int *>(Bs + ((int64_t)(u - 1) * N + n0) * kpad) + kt * LEN;
    ca.load(ga, ld_i, M - m0, LEN, true, tid);
    cb.load(gb, ld_i, N - n0, LEN, true, tid);
  };

  i4_t acc[FM][FN];
  double accd[FM][FN][4];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc[i][j] = i4_t{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) accd[i][j][r] = 0.0;
    }

  load_iter(0);
  ca.template store<LS>(sA[0], tid);
  cb.template store<LS>(sB[0], tid);
  __syncthreads();

  for (int q = 0; q < iters; ++q) {
    const int cur = q & 1;
    if (q + 1 < iters) load_iter(q + 1);
    const int *as = sA[cur];
    const int *bs = sB[cur];
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) {
      const int koff = ks * 16 + kg * 4; /* ints: 64 bytes per step, 16 per lane group */
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
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if ((q % nkt) == nkt - 1 && pairs.last[q / nkt]) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            accd[i][j][r] = __builtin_fma(accd[i][j][r], 0x1p-7, (double)acc[i][j][r]);
          acc[i][j] = i4_t{0, 0, 0, 0};
        }
    }
    if (q + 1 < iters) {
      ca.template store<LS>(sA[cur ^ 1], tid);
      cb.template store<LS>(sB[cur ^ 1], tid);
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
        const int f = fb[col];
        double *cp = C + row + col * ldc;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (row + r < M) {
            const int e = ea[row + r];
            const double scaled = (e == OZ_NONFINITE || f == OZ_NONFINITE)
                                      ? __builtin_nan("")
                                      : ldexp(accd[i][j][r], e + f - 14);
            cp[r] = __builtin_fma(alpha, scaled, beta * cp[r]);
          }
        }
      }
    }
}

OzPairs make_pairs(int s) {
  OzPairs P;
  P.count = 0;
  for (int w = s + 1; w >= 2; --w) {
    const int tlo = w - s > 1 ? w - s : 1, thi = w - 1 < s ? w - 1 : s;
    for (int t = tlo; t <= thi; ++t) {
      P.t[P.count] = (signed char)t;
      P.u[P.count] = (signed char)(w - t);
      P.last[P.count] = (signed char)(t == thi);
      P.count++;
    }
  }
  return P;
}

template <int S>
hipError_t launch_split(const double *x, int64_t vld, int64_t eld, int nvec,
                        int len, int64_t kpad, const int *ex, int v_fast,
                        int8_t *out, hipStream_t st) {
  const int64_t threads = (kpad / 16) * nvec;
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  hipLaunchKernelGGL((split_kernel<S>), dim3(blocks), dim3(256), 0, st, x, vld,
                     eld, nvec, len, kpad, ex, v_fast, out);
  return hipGetLastError();
}

hipError_t split_any(int s, const double *x, int64_t vld, int64_t eld, int nvec,
                     int len, int64_t kpad, const int *ex, int v_fast,
                     int8_t *out, hipStream_t st) {
  switch (s) {
  case 3: return launch_split<3>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  case 4: return launch_split<4>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  case 5: return launch_split<5>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  case 6: return launch_split<6>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  case 7: return launch_split<7>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  default: return launch_split<8>(x, vld, eld, nvec, len, kpad, ex, v_fast, out, st);
  }
}

/* Exponents and slices of one operand.  For op(A) the vectors are rows
 * (transa N: row i at a + i, stride lda; T: contiguous at a + i*lda); for
 * op(B) the vectors are columns (transb N: contiguous at b + j*ldb; T: at
 * b + j, stride ldb). */
hipError_t prepare_operand(const double *x, int64_t ld, bool contiguous,
                           int nvec, int k, int64_t kpad, int s, int *ex,
                           int8_t *slices, hipStream_t st) {
  if (contiguous)
    hipLaunchKernelGGL(exp_contig_kernel, dim3((nvec + 3) / 4), dim3(256), 0, st,
                       x, ld, nvec, k, ex);
  else
    hipLaunchKernelGGL(exp_strided_kernel, dim3((nvec + 255) / 256), dim3(256),
                       0, st, x, ld, nvec, k, ex);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return contiguous ? split_any(s, x, ld, 1, nvec, k, kpad, ex, 0, slices, st)
                    : split_any(s, x, 1, ld, nvec, k, kpad, ex, 1, slices, st);
}

} // namespace

extern "C" int gemmx_ozaki_kchunk(int slices) {
  const int64_t lim = (int64_t)2147483647 / ((int64_t)slices * 127 * 127);
  return (int)(lim - lim % OZ_BK);
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
  const int64_t kpad = kpad_of(p->k);
  hipError_t e = prepare_operand((const double *)p->a, p->lda, p->ta != 0, p->m,
                                 p->k, kpad, s, ea, as, st);
  if (e != hipSuccess) return e;
  e = prepare_operand((const double *)p->b, p->ldb, p->tb == 0, p->n, p->k, kpad,
                      s, fb, bs, st);
  if (e != hipSuccess) return e;

  constexpr int BM = 128, BN = 128, WM = 2, WN = 4;
  const int tiles_m = (p->m + BM - 1) / BM, tiles_n = (p->n + BN - 1) / BN;
  const OzPairs pairs = make_pairs(s);
  const int kc_tiles = gemmx_ozaki_kchunk(s) / OZ_BK;
  const int nkt = (int)(kpad / OZ_BK);
  for (int kt0 = 0; kt0 < nkt; kt0 += kc_tiles) {
    const int kt1 = kt0 + kc_tiles < nkt ? kt0 + kc_tiles : nkt;
    /* Chunks after the first add onto C: beta = 1 leaves C as it is. */
    const double beta = kt0 == 0 ? p->beta : 1.0;
    hipLaunchKernelGGL((ozaki_gemm_kernel<BM, BN, WM, WN>),
                       dim3(tiles_m * tiles_n), dim3(64 * WM * WN), 0, st, as,
                       bs, ea, fb, (double *)p->c, p->ldc, p->m, p->n, kpad,
                       kt0, kt1, p->alpha, beta, tiles_m, tiles_n, pairs);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
