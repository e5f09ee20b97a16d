/*
 * gemm_complex.hip -- complex GEMM (zgemm_/cgemm_ and the three-product
 * forms zgemm3m_/cgemm3m_) on the native real MFMA tile kernels.
 *
 * The level-3 complex siblings of the reference's real GEMM live outside its
 * build, as OpenBLAS kernels (wjc404/GEMM_AVX2
 * OpenBLAS-like_implementation/zgemm_kernel_4x2_haswell.c,
 * cgemm_kernel_8x2_haswell.c, zgemm3m_kernel_4x4_haswell.c,
 * cgemm3m_kernel_8x4_haswell.c).  Here the product goes to the matrix cores
 * through real GEMMs on split operands:
 *
 *   split:    interleaved (re, im) pairs of the stored A and B -> dense real
 *             and imaginary planes (leading dimension = stored rows); the
 *             conjugation of transa/transb 'C' negates the imaginary plane,
 *             so the real GEMMs only see 'N' or 'T';
 *   4 real products (zgemm_/cgemm_):
 *             Pr = Ar Br - Ai Bi,  Pi = Ar Bi + Ai Br
 *             (the second product of each accumulates onto the first through
 *             the real kernel's beta = 1 epilogue);
 *   3 real products (zgemm3m_/cgemm3m_):
 *             T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi),
 *             Pr = T1 - T2, Pi = T3 - T1 - T2 (fewer products, larger
 *             rounding error on operands with cancelling parts);
 *   combine:  C = beta C (first, when beta != 1, also for beta = 0, as the
 *             real entry points do) + alpha P, complex alpha and beta.
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"
#include "gemmx_tune.h"

namespace {

/* Planes of a stored rows x cols complex matrix (element (i, j) at
 * x[2 (i + j ld)], its imaginary part next): re, im and, for the
 * three-product form, re + im, each rows x cols with leading dimension rows. */
template <typename T>
__global__ void __launch_bounds__(256)
cplx_split_kernel(const T *__restrict__ x, int64_t ld, int rows, int cols, T *__restrict__ re,
                  T *__restrict__ im, T *__restrict__ sum, int conj) {
  const int64_t total = (int64_t)rows * cols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / rows, i = t - j * rows;
    const T *p = x + 2 * (i + j * ld);
    const T r = p[0], m = conj ? -p[1] : p[1];
    re[t] = r;
    im[t] = m;
    if (sum) sum[t] = r + m;
  }
}

/* C = beta C (beta != 1) + alpha P with P = (pr, pi), or, for the
 * three-product form (t3 != NULL), P = (pr - pi, t3 - pr - pi) from
 * pr = T1, pi = T2. */
template <typename T>
__global__ void __launch_bounds__(256)
cplx_combine_kernel(const T *__restrict__ pr, const T *__restrict__ pi,
                    const T *__restrict__ t3, T *__restrict__ c, int64_t ldc, int m, int n,
                    T ar, T ai, T br, T bi, int has_p) {
  const int64_t total = (int64_t)m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / m, i = t - j * m;
    T *cp = c + 2 * (i + j * ldc);
    T cr = cp[0], ci = cp[1];
    if (br != (T)1 || bi != (T)0) {
      const T nr = br * cr - bi * ci, ni = br * ci + bi * cr;
      cr = nr;
      ci = ni;
    }
    if (has_p) {
      T qr, qi;
      if (t3) {
        qr = pr[t] - pi[t];
        qi = t3[t] - pr[t] - pi[t];
      } else {
        qr = pr[t];
        qi = pi[t];
      }
      cr += ar * qr - ai * qi;
      ci += ar * qi + ai * qr;
    }
    cp[0] = cr;
    cp[1] = ci;
  }
}

size_t rup(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct CplxLayout {
  size_t a[3], b[3], p[3], total;
};

CplxLayout cplx_layout(int es, int ta, int tb, int m, int n, int k, int three) {
  CplxLayout L;
  const size_t ra = ta ? k : m, ca = ta ? m : k, rb = tb ? n : k, cb = tb ? k : n;
  const size_t sa = rup(ra * ca * es, 256), sb = rup(rb * cb * es, 256), sp = rup((size_t)m * n * es, 256);
  const int planes = three ? 3 : 2;
  size_t off = 0;
  for (int q = 0; q < 3; ++q) L.a[q] = q < planes ? (off += sa) - sa : 0;
  for (int q = 0; q < 3; ++q) L.b[q] = q < planes ? (off += sb) - sb : 0;
  for (int q = 0; q < 3; ++q) L.p[q] = q < planes ? (off += sp) - sp : 0;
  L.total = off;
  return L;
}

unsigned grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 8192) b = 8192;
  return (unsigned)(b < 1 ? 1 : b);
}

template <typename T>
hipError_t real_gemm(int ta, int tb, int m, int n, int k, double alpha, const T *a, int64_t lda,
                     const T *b, int64_t ldb, double beta, T *c, hipStream_t s) {
  gemmx_problem q = {};
  q.ta = ta;
  q.tb = tb;
  q.m = m;
  q.n = n;
  q.k = k;
  q.lda = lda;
  q.ldb = ldb;
  q.ldc = m;
  q.alpha = alpha;
  q.beta = beta;
  q.a = a;
  q.b = b;
  q.c = c;
  constexpr int es = sizeof(T);
  const int cfg = gemmx_tune_pick(es, ta, tb, m, n, k);
  return es == 8 ? gemmx_native_f64(&q, cfg, 1, NULL, s)
                 : gemmx_native_f32(&q, cfg, 1, NULL, s);
}

template <typename T>
hipError_t complex_gemm(const gemmx_cproblem *p, int three, void *work, hipStream_t s) {
  const int m = p->m, n = p->n, k = p->k;
  const T ar = (T)p->alpha[0], ai = (T)p->alpha[1], br = (T)p->beta[0], bi = (T)p->beta[1];
  const bool product = (ar != (T)0 || ai != (T)0) && k > 0;
  if (!product) {
    if (br == (T)1 && bi == (T)0) return hipSuccess;
    hipLaunchKernelGGL((cplx_combine_kernel<T>), dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                       (const T *)nullptr, (const T *)nullptr, (const T *)nullptr, (T *)p->c,
                       p->ldc, m, n, ar, ai, br, bi, 0);
    return hipGetLastError();
  }
  const int ta = p->ta != 0, tb = p->tb != 0;
  const CplxLayout L = cplx_layout(sizeof(T), ta, tb, m, n, k, three);
  char *w = (char *)work;
  T *A[3], *B[3], *P[3];
  for (int q = 0; q < 3; ++q) {
    A[q] = (T *)(w + L.a[q]);
    B[q] = (T *)(w + L.b[q]);
    P[q] = (T *)(w + L.p[q]);
  }
  const int ra = ta ? k : m, ca = ta ? m : k, rb = tb ? n : k, cb = tb ? k : n;
  hipLaunchKernelGGL((cplx_split_kernel<T>), dim3(grid_for((int64_t)ra * ca)), dim3(256), 0, s,
                     (const T *)p->a, p->lda, ra, ca, A[0], A[1], three ? A[2] : (T *)nullptr,
                     p->conja);
  hipLaunchKernelGGL((cplx_split_kernel<T>), dim3(grid_for((int64_t)rb * cb)), dim3(256), 0, s,
                     (const T *)p->b, p->ldb, rb, cb, B[0], B[1], three ? B[2] : (T *)nullptr,
                     p->conjb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  /* The first product into each plane has beta = 0, which (as for dgemm_)
   * still multiplies the plane: clear it so nothing of the scratch survives
   * (0 * NaN). */
  for (int q = 0; q < (three ? 3 : 2); ++q)
    if ((e = hipMemsetAsync(P[q], 0, (size_t)m * n * sizeof(T), s)) != hipSuccess) return e;
  if (three) {
    if ((e = real_gemm<T>(ta, tb, m, n, k, 1.0, A[0], ra, B[0], rb, 0.0, P[0], s)) != hipSuccess ||
        (e = real_gemm<T>(ta, tb, m, n, k, 1.0, A[1], ra, B[1], rb, 0.0, P[1], s)) != hipSuccess ||
        (e = real_gemm<T>(ta, tb, m, n, k, 1.0, A[2], ra, B[2], rb, 0.0, P[2], s)) != hipSuccess)
      return e;
  } else {
    if ((e = real_gemm<T>(ta, tb, m, n, k, 1.0, A[0], ra, B[0], rb, 0.0, P[0], s)) != hipSuccess ||
        (e = real_gemm<T>(ta, tb, m, n, k, -1.0, A[1], ra, B[1], rb, 1.0, P[0], s)) != hipSuccess ||
        (e = real_gemm<T>(ta, tb, m, n, k, 1.0, A[0], ra, B[1], rb, 0.0, P[1], s)) != hipSuccess ||
        (e = real_gemm<T>(ta, tb, m, n, k, 1.0, A[1], ra, B[0], rb, 1.0, P[1], s)) != hipSuccess)
      return e;
  }
  hipLaunchKernelGGL((cplx_combine_kernel<T>), dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                     (const T *)P[0], (const T *)P[1], three ? (const T *)P[2] : (const T *)nullptr,
                     (T *)p->c, p->ldc, m, n, ar, ai, br, bi, 1);
  return hipGetLastError();
}

} // namespace

extern "C" size_t gemmx_complex_workspace(int es, int ta, int tb, int m, int n, int k,
                                          int three) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  return cplx_layout(es, ta != 0, tb != 0, m, n, k, three).total;
}

extern "C" hipError_t gemmx_complex(int es, const gemmx_cproblem *p, int three, void *work,
                                    hipStream_t s) {
  if (p->m <= 0 || p->n <= 0) return hipSuccess;
  return es == 8 ? complex_gemm<double>(p, three, work, s) : complex_gemm<float>(p, three, work, s);
}
