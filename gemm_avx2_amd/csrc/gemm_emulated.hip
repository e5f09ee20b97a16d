/*
 * gemm_emulated.hip -- host orchestration of the emulated forms and the
 * Ozaki-I recombination.  No reference counterpart: the reference has only
 * native sgemm_/dgemm_ (wjc404/GEMM_AVX2 src/gemm_driver.c:118); the boundary
 * semantics are the native ones, C = alpha*op(A)op(B) + beta*C with the
 * beta != 1 rule.
 *
 * Ozaki-I (DGEMM): split (gemm_split.hip) -> one INT8 GEMM per weight
 * w = t + u in 2..s+1 of the slice pairs, exact in INT32 (gemm_products.hip;
 * K chunked so that s*k*127^2 < 2^31) -> recombination here:
 * acc = acc*2^-7 + S_w for w = s+1..2 in FP64 (acc*2^-7 is exact), then
 * C = fma(alpha, ldexp(acc, ea_i + fb_j - 14), beta*C).  Every step is exact
 * or the same FP64 operation as the oracle's restatement, so results are
 * bit-identical to it.
 *
 * split-BF16x3 (SGEMM): split -> pairs (lo,hi), (hi,lo), (hi,hi) in one FP32
 * accumulator -> C = fma(alpha, acc, beta*C) in the products epilogue.
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"
#include "gemmx_slices.h"

using namespace gemmx;

namespace {

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

struct OzLayout {
  size_t ea, fb, mb, as, bs, sw, total;
  int64_t kpad;
};
OzLayout oz_layout(int m, int n, int k, int s) {
  OzLayout L;
  L.kpad = (int64_t)rup((size_t)k, SL_BK);
  L.ea = 0;
  L.fb = rup((size_t)m * 4, 256);
  L.mb = L.fb + rup((size_t)n * 4, 256);
  L.as = L.mb + rup((size_t)(m > n ? m : n) * 8, 256);
  L.bs = L.as + rup((size_t)s * m * L.kpad, 256);
  L.sw = L.bs + rup((size_t)s * n * L.kpad, 256);
  L.total = L.sw + rup((size_t)s * m * n * 4, 256);
  return L;
}

/* One group per weight w = t + u (s+1 down to 2, the heaviest first), its
 * sums going to slot w - 2 of S; slice indices are 0-based. */
SliceGroups ozaki_groups(int s) {
  SliceGroups G;
  G.ngroups = 0;
  for (int w = s + 1; w >= 2; --w) {
    const int tlo = w - s > 1 ? w - s : 1, thi = w - 1 < s ? w - 1 : s;
    int t[MAX_GROUP_PAIRS], u[MAX_GROUP_PAIRS], c = 0;
    for (int ts = tlo; ts <= thi; ++ts, ++c) {
      t[c] = ts - 1;
      u[c] = w - ts - 1;
    }
    G.add(w - 2, c, t, u);
  }
  return G;
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
  unsigned long long *mb = (unsigned long long *)(w + L.mb);
  hipError_t e = ozaki_prepare((const double *)p->a, p->lda, p->ta != 0, p->m, p->k,
                               L.kpad, s, ea, mb, as, st);
  if (e != hipSuccess) return e;
  e = ozaki_prepare((const double *)p->b, p->ldb, p->tb == 0, p->n, p->k, L.kpad, s,
                    fb, mb, bs, st);
  if (e != hipSuccess) return e;
  const SliceGroups groups = ozaki_groups(s);
  const int kc = gemmx_ozaki_kchunk(s); /* bytes, a multiple of SL_BK */
  const int kp = (int)L.kpad;
  const int64_t total = (int64_t)p->m * p->n;
  int64_t rblocks = (total + 255) / 256;
  if (rblocks > 65536) rblocks = 65536;
  for (int kb0 = 0; kb0 < kp; kb0 += kc) {
    const int kb1 = kb0 + kc < kp ? kb0 + kc : kp;
    e = slice_products(false, as, bs, p->m, p->n, L.kpad, kb0, kb1, groups, S, NULL, 0,
                       0.f, 0.f, st);
    if (e != hipSuccess) return e;
    /* chunks after the first add onto C: beta = 1 leaves C as it is */
    hipLaunchKernelGGL(ozaki_recombine_kernel, dim3((unsigned)rblocks), dim3(256), 0, st,
                       S, s, ea, fb, (double *)p->c, p->ldc, p->m, p->n, p->alpha,
                       kb0 == 0 ? p->beta : 1.0);
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
  /* slice 0 = hi, 1 = lo; one group: (lo,hi), (hi,lo), (hi,hi) */
  SliceGroups G;
  G.ngroups = 0;
  const int tt[3] = {1, 0, 0}, uu[3] = {0, 1, 0};
  G.add(0, 3, tt, uu);
  const int64_t kpad_bytes = kpad * 2;
  return slice_products(true, (const int8_t *)as, (const int8_t *)bs, p->m, p->n, kpad_bytes,
                        0, (int)kpad_bytes, G, NULL, (float *)p->c, p->ldc, (float)p->alpha,
                        (float)p->beta, st);
}
