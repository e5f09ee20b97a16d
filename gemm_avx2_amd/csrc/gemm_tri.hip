/*
 * gemm_tri.hip -- triangular multiply and solve (dtrmm_/strmm_,
 * dtrsm_/strsm_) on the native real MFMA GEMM.
 *
 * The reference keeps these level-3 siblings outside its build as OpenBLAS
 * kernels (wjc404/GEMM_AVX2 OpenBLAS-like_implementation/
 * trmm_kernel_4x8_simplified.c, strsm_kernel_8x4_haswell_{LN,LT,RN,RT}.c,
 * TRSM_KERNEL_{LN,LT,RN,RT}_simplified.c).  Here:
 *
 *   TRMM  B := alpha op(A) B (side L) or alpha B op(A) (side R), in place,
 *         recursive on halves (trmm_rec): the half no other half feeds is
 *         updated first, then the other receives the off-diagonal product
 *         (one deep GEMM), then is updated; leaves of up to TRMM_LEAF are
 *         one GEMM on a dense copy of the diagonal block (unit diagonal
 *         written, zeros outside the stored triangle) through a workspace.
 *   TRSM  op(A) X = alpha B (side L) or X op(A) = alpha B (side R), X over
 *         B.  The TRI_NB x TRI_NB diagonal blocks of op(A) are inverted by
 *         substitution (one thread per column of the inverse, the block in
 *         LDS); the solve recurses on halves (trsm_rec): a block is
 *         Y_i = inv(op(A)_ii) B_i (a GEMM into the workspace Y), and between
 *         the two halves of a range one GEMM with beta = 1 subtracts the
 *         solved half from the other in place on B; finally B = alpha Y.
 *         The inverted blocks are only as accurate as their conditioning
 *         allows (the usual trade of GPU TRSMs).
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"
#include "gemmx_tune.h"

namespace {

constexpr int TRI_NB = 64;

/* Dense k x k copy (leading dimension k) of the stored triangle of A:
 * upper (row <= col) or lower, unit diagonal when `unit`, zeros elsewhere. */
template <typename T>
__global__ void __launch_bounds__(256)
tri_dense_kernel(const T *__restrict__ A, int64_t lda, int k, int upper, int unit,
                 T *__restrict__ D) {
  const int64_t total = (int64_t)k * k;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / k, r = t - c * k;
    T v = (T)0;
    if (r == c) v = unit ? (T)1 : A[r + c * lda];
    else if (upper ? r < c : r > c) v = A[r + c * lda];
    D[t] = v;
  }
}

/* Inverses of the diagonal blocks of op(A) (k x k, blocks of TRI_NB, the
 * last possibly smaller): block b's inverse, dense TRI_NB x TRI_NB with
 * leading dimension TRI_NB, at Dinv + b TRI_NB^2.  op(A)(r, c) is A(r, c), or
 * A(c, r) when `trans`; `lower` tells which triangle op(A) has.  One thread
 * per column of the inverse: forward substitution on e_c for a lower block,
 * backward for an upper one. */
template <typename T>
__global__ void __launch_bounds__(TRI_NB)
tri_inv_kernel(const T *__restrict__ A, int64_t lda, int k, int trans, int lower, int unit,
               T *__restrict__ Dinv) {
  __shared__ T blk[TRI_NB][TRI_NB + 1]; /* blk[c][r] = op(A)(i0 + r, i0 + c) */
  __shared__ T xs[TRI_NB][TRI_NB + 1];  /* xs[c][r]: column c of the inverse */
  const int b = blockIdx.x, i0 = b * TRI_NB;
  const int kb = k - i0 < TRI_NB ? k - i0 : TRI_NB;
  const int t = threadIdx.x;
  for (int c = 0; c < TRI_NB; ++c) {
    T v = (T)0;
    if (t < kb && c < kb) {
      const int64_t gr = i0 + t, gc = i0 + c;
      v = trans ? A[gc + gr * lda] : A[gr + gc * lda];
    }
    blk[c][t] = v;
    xs[c][t] = (T)0;
  }
  __syncthreads();
  const int c = t;
  if (c < kb) {
    T *x = xs[c];
    if (lower) {
      for (int r = c; r < kb; ++r) {
        T s = r == c ? (T)1 : (T)0;
        for (int q = c; q < r; ++q) s -= blk[q][r] * x[q];
        x[r] = unit ? s : s / blk[r][r];
      }
    } else {
      for (int r = c; r >= 0; --r) {
        T s = r == c ? (T)1 : (T)0;
        for (int q = r + 1; q <= c; ++q) s -= blk[q][r] * x[q];
        x[r] = unit ? s : s / blk[r][r];
      }
    }
  }
  __syncthreads();
  T *out = Dinv + (int64_t)b * TRI_NB * TRI_NB;
  for (int cc = 0; cc < TRI_NB; ++cc) out[t + (int64_t)cc * TRI_NB] = xs[cc][t];
}

/* B = alpha Y (m x n), or B = 0 when `zero`. */
template <typename T>
__global__ void __launch_bounds__(256)
scale_copy_kernel(const T *__restrict__ Y, int64_t ldy, T *__restrict__ B, int64_t ldb, int m,
                  int n, T alpha, int zero) {
  const int64_t total = (int64_t)m * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / m, i = t - j * m;
    B[i + j * ldb] = zero ? (T)0 : alpha * Y[i + j * ldy];
  }
}

unsigned grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 8192) b = 8192;
  return (unsigned)(b < 1 ? 1 : b);
}

size_t rup(size_t v, size_t a) { return (v + a - 1) / a * a; }

/* C (ldc) = alpha op(A) op(B) + beta C on the native kernels. */
template <typename T>
hipError_t gemm(int ta, int tb, int m, int n, int k, double alpha, const T *a, int64_t lda,
                const T *b, int64_t ldb, double beta, T *c, int64_t ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  gemmx_problem q = {};
  q.ta = ta;
  q.tb = tb;
  q.m = m;
  q.n = n;
  q.k = k;
  q.lda = lda;
  q.ldb = ldb;
  q.ldc = ldc;
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

constexpr int TRMM_LEAF = 256;

/* op(A)(r0.., c0..) as a GEMM operand: stored A at (r0, c0), or at (c0, r0)
 * read transposed. */
template <typename T>
const T *op_at(const gemmx_tproblem *p, int r0, int c0) {
  const T *A = (const T *)p->a;
  return p->trans ? A + c0 + (int64_t)r0 * p->lda : A + r0 + (int64_t)c0 * p->lda;
}

/* B's rows (side L) or columns (side R) i0 .. i0 + kk := alpha op(A)_range
 * times them.  A range of at most TRMM_LEAF is one GEMM on a dense copy D of
 * its diagonal block of A (zeros outside the stored triangle) into W, copied
 * back.  A larger one splits at a multiple of TRI_NB near its middle: the
 * half that no other half feeds is updated first (recursively), then the
 * other half receives its off-diagonal product (one GEMM of depth = that
 * half, beta = 1, reading the not-yet-updated half), then that half is
 * updated.  W and D are workspace for the largest leaf. */
template <typename T>
hipError_t trmm_rec(const gemmx_tproblem *p, int lower, int i0, int kk, T *D, T *W,
                    hipStream_t s) {
  const int m = p->m, n = p->n;
  T *B = (T *)p->b;
  if (kk <= TRMM_LEAF) {
    const T *Ablk = (const T *)p->a + i0 + (int64_t)i0 * p->lda;
    hipLaunchKernelGGL((tri_dense_kernel<T>), dim3(grid_for((int64_t)kk * kk)), dim3(256), 0, s,
                       Ablk, p->lda, kk, p->upper, p->unit, D);
    const int64_t wsz = p->left ? (int64_t)kk * n : (int64_t)m * kk;
    hipError_t e = hipMemsetAsync(W, 0, (size_t)wsz * sizeof(T), s); /* beta = 0 reads W */
    if (e != hipSuccess) return e;
    if (p->left) {
      e = gemm<T>(p->trans, 0, kk, n, kk, p->alpha, D, kk, B + i0, p->ldb, 0.0, W, kk, s);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((scale_copy_kernel<T>), dim3(grid_for(wsz)), dim3(256), 0, s,
                         (const T *)W, (int64_t)kk, B + i0, p->ldb, kk, n, (T)1, 0);
    } else {
      e = gemm<T>(0, p->trans, m, kk, kk, p->alpha, B + (int64_t)i0 * p->ldb, p->ldb, D, kk,
                  0.0, W, m, s);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((scale_copy_kernel<T>), dim3(grid_for(wsz)), dim3(256), 0, s,
                         (const T *)W, (int64_t)m, B + (int64_t)i0 * p->ldb, p->ldb, m, kk, (T)1,
                         0);
    }
    return hipGetLastError();
  }
  const int h = ((kk / 2 + TRI_NB - 1) / TRI_NB) * TRI_NB; /* 0 < h < kk */
  /* Left, op(A) lower: new B2 = L21 B1 + L22 B2 needs the old B1, so B2 goes
   * first; upper: new B1 = U11 B1 + U12 B2 needs the old B2, so B1 first.
   * Right: upper -> new B2 = B1 U12 + B2 U22, B2 first; lower -> B1 first. */
  const bool second_first = p->left ? lower : !lower;
  const int a0 = i0, a1 = i0 + h;           /* the two halves' starts */
  const int ka = h, kb2 = kk - h;
  const int u0 = second_first ? a1 : a0, ku = second_first ? kb2 : ka; /* updated first */
  const int v0 = second_first ? a0 : a1, kv = second_first ? ka : kb2; /* the feeding half */
  hipError_t e = trmm_rec<T>(p, lower, u0, ku, D, W, s);
  if (e != hipSuccess) return e;
  e = p->left ? gemm<T>(p->trans, 0, ku, n, kv, p->alpha, op_at<T>(p, u0, v0), p->lda, B + v0,
                        p->ldb, 1.0, B + u0, p->ldb, s)
              : gemm<T>(0, p->trans, m, ku, kv, p->alpha, B + (int64_t)v0 * p->ldb, p->ldb,
                        op_at<T>(p, v0, u0), p->lda, 1.0, B + (int64_t)u0 * p->ldb, p->ldb, s);
  if (e != hipSuccess) return e;
  return trmm_rec<T>(p, lower, v0, kv, D, W, s);
}

template <typename T>
hipError_t trmm(const gemmx_tproblem *p, void *work, hipStream_t s) {
  const int m = p->m, n = p->n, k = p->left ? m : n;
  T *B = (T *)p->b;
  if (p->alpha == 0.0) {
    hipLaunchKernelGGL((scale_copy_kernel<T>), dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                       (const T *)nullptr, (int64_t)0, B, p->ldb, m, n, (T)0, 1);
    return hipGetLastError();
  }
  const int leaf = k < TRMM_LEAF ? k : TRMM_LEAF;
  T *D = (T *)work;
  T *W = (T *)((char *)work + rup((size_t)leaf * leaf * sizeof(T), 256));
  const int lower = p->upper ? p->trans : !p->trans;
  return trmm_rec<T>(p, lower, 0, k, D, W, s);
}

/* Solves the rows (side L) or columns (side R) i0 .. i0 + kk of op(A) Y = B
 * (Y op(A) = B) into Y, B holding the right-hand sides less the parts already
 * solved: one TRI_NB block is Y_i = inv(op(A)_ii) B_i (or B_i inv(op(A)_ii));
 * a larger range splits at a multiple of TRI_NB near its middle, solves the
 * half op(A)'s triangle puts first, subtracts its contribution from the
 * other half with one GEMM of depth = the solved half, and solves that
 * half.  Most flops thus land in GEMMs of large depth. */
template <typename T>
hipError_t trsm_rec(const gemmx_tproblem *p, const T *Dinv, T *Y, int lower, int i0, int kk,
                    hipStream_t s) {
  const int m = p->m, n = p->n;
  T *B = (T *)p->b;
  const T *A = (const T *)p->a;
  if (kk <= TRI_NB) {
    const T *Di = Dinv + (int64_t)(i0 / TRI_NB) * TRI_NB * TRI_NB;
    return p->left
               ? gemm<T>(0, 0, kk, n, kk, 1.0, Di, TRI_NB, B + i0, p->ldb, 0.0, Y + i0, m, s)
               : gemm<T>(0, 0, m, kk, kk, 1.0, B + (int64_t)i0 * p->ldb, p->ldb, Di, TRI_NB, 0.0,
                         Y + (int64_t)i0 * m, m, s);
  }
  const int h = ((kk / 2 + TRI_NB - 1) / TRI_NB) * TRI_NB; /* 0 < h < kk */
  /* left: rows in order when op(A) is lower; right: columns in order when
   * op(A) is upper */
  const bool forward = p->left ? lower : !lower;
  const int s1 = forward ? i0 : i0 + h, k1 = forward ? h : kk - h; /* solved first */
  const int s2 = forward ? i0 + h : i0, k2 = kk - k1;              /* then */
  hipError_t e = trsm_rec<T>(p, Dinv, Y, lower, s1, k1, s);
  if (e != hipSuccess) return e;
  /* op(A)(r0.., c0..) as a GEMM operand: stored A at (r0, c0), or at (c0, r0)
   * read transposed */
  auto opA = [&](int r0, int c0) {
    return p->trans ? A + c0 + (int64_t)r0 * p->lda : A + r0 + (int64_t)c0 * p->lda;
  };
  e = p->left ? gemm<T>(p->trans, 0, k2, n, k1, -1.0, opA(s2, s1), p->lda, Y + s1, m, 1.0,
                        B + s2, p->ldb, s)
              : gemm<T>(0, p->trans, m, k2, k1, -1.0, Y + (int64_t)s1 * m, m, opA(s1, s2),
                        p->lda, 1.0, B + (int64_t)s2 * p->ldb, p->ldb, s);
  if (e != hipSuccess) return e;
  return trsm_rec<T>(p, Dinv, Y, lower, s2, k2, s);
}

template <typename T>
hipError_t trsm(const gemmx_tproblem *p, void *work, hipStream_t s) {
  const int m = p->m, n = p->n, k = p->left ? m : n;
  T *B = (T *)p->b;
  const T *A = (const T *)p->a;
  if (p->alpha == 0.0) {
    hipLaunchKernelGGL((scale_copy_kernel<T>), dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                       (const T *)nullptr, (int64_t)0, B, p->ldb, m, n, (T)0, 1);
    return hipGetLastError();
  }
  const int nblk = (k + TRI_NB - 1) / TRI_NB;
  T *Dinv = (T *)work;
  T *Y = (T *)((char *)work + rup((size_t)nblk * TRI_NB * TRI_NB * sizeof(T), 256));
  /* op(A) is lower when (lower stored, no transpose) or (upper, transposed) */
  const int lower = p->upper ? p->trans : !p->trans;
  hipLaunchKernelGGL((tri_inv_kernel<T>), dim3(nblk), dim3(TRI_NB), 0, s, A, p->lda, k, p->trans,
                     lower, p->unit, Dinv);
  hipError_t e = hipMemsetAsync(Y, 0, (size_t)m * n * sizeof(T), s); /* beta = 0 reads Y */
  if (e != hipSuccess) return e;
  e = trsm_rec<T>(p, Dinv, Y, lower, 0, k, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((scale_copy_kernel<T>), dim3(grid_for((int64_t)m * n)), dim3(256), 0, s,
                     (const T *)Y, (int64_t)m, B, p->ldb, m, n, (T)p->alpha, 0);
  return hipGetLastError();
}

} // namespace

extern "C" size_t gemmx_tri_workspace(int es, int solve, int left, int m, int n) {
  if (m <= 0 || n <= 0) return 0;
  const size_t k = left ? m : n;
  if (solve) /* diagonal-block inverses, Y */
    return rup(((k + TRI_NB - 1) / TRI_NB) * TRI_NB * TRI_NB * es, 256) +
           rup((size_t)m * n * es, 256);
  /* the largest leaf's dense triangle and product */
  const size_t leaf = k < (size_t)TRMM_LEAF ? k : (size_t)TRMM_LEAF;
  return rup(leaf * leaf * es, 256) + rup(leaf * (left ? (size_t)n : (size_t)m) * es, 256);
}

extern "C" hipError_t gemmx_tri(int es, int solve, const gemmx_tproblem *p, void *work,
                                hipStream_t s) {
  if (p->m <= 0 || p->n <= 0) return hipSuccess;
  if (solve) return es == 8 ? trsm<double>(p, work, s) : trsm<float>(p, work, s);
  return es == 8 ? trmm<double>(p, work, s) : trmm<float>(p, work, s);
}
