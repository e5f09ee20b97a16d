/*
 * gemmx_slices.h -- shared pieces of the emulated forms (split, products,
 * recombination), which live in gemm_split.hip, gemm_products.hip and
 * gemm_emulated.hip.  Internal to the library.
 *
 * Slices are stored K-contiguous: slice t of a matrix with `nvec` vectors
 * (rows of op(A) or columns of op(B)) is nvec lines of kpad bytes, line v at
 * base + (t*nvec + v)*kpad.  kpad is K padded with zeros to a multiple of
 * SL_BK bytes.
 */
#ifndef GEMMX_SLICES_H
#define GEMMX_SLICES_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gemmx {

constexpr int SL_BK = 128;               /* K padding granule in bytes */
constexpr int OZ_NONFINITE = 0x7fffffff; /* exponent of a row/column with Inf/NaN */
constexpr int MAX_PAIRS = 36;
constexpr int MAX_GROUPS = 24;
constexpr int MAX_GROUP_PAIRS = 8; /* pairs in one group (an Ozaki weight: <= slices) */

inline size_t rup(size_t v, size_t a) { return (v + a - 1) / a * a; }

/* Slice-pair work of the products kernel.  Group g multiplies gcnt[g] pairs
 * of slices (t_c, u_c), c < gcnt[g], into one accumulator and writes it to
 * output slot gout[g]; t_c = (tpack[g] >> 4c) & 15, u_c likewise from
 * upack[g].  All fields are 32-bit so that the kernel reads them with scalar
 * loads and unpacks the pair of each iteration with shifts. */
struct SliceGroups {
  int ngroups;
  int gcnt[MAX_GROUPS], gout[MAX_GROUPS];
  unsigned tpack[MAX_GROUPS], upack[MAX_GROUPS];

  void add(int out, int count, const int *t, const int *u) {
    unsigned tp = 0, up = 0;
    for (int c = 0; c < count; ++c) {
      tp |= (unsigned)t[c] << (4 * c);
      up |= (unsigned)u[c] << (4 * c);
    }
    gcnt[ngroups] = count;
    gout[ngroups] = out;
    tpack[ngroups] = tp;
    upack[ngroups] = up;
    ngroups++;
  }
};

/* Exponents (per row of op(A) / column of op(B)) and s INT8 slices of an FP64
 * operand.  contiguous: vector v at x + v*ld, entries adjacent; otherwise
 * vector v at x + v with entry stride ld.  mbits: nvec words of scratch. */
hipError_t ozaki_prepare(const double *x, int64_t ld, bool contiguous, int nvec, int k,
                         int64_t kpad, int s, int *ex, unsigned long long *mbits,
                         int8_t *slices, hipStream_t st);

/* hi and lo BF16 slices ([2][nvec][kpad] bf16, kpad in elements) of an FP32
 * operand, same vector convention. */
hipError_t bf16_prepare(const float *x, int64_t ld, bool contiguous, int nvec, int k,
                        int64_t kpad, unsigned short *out, hipStream_t st);

/* Ozaki-II: residues of rint(x * 2^(P - e)) modulo OZ2_MODULI[0..nmod-1]
 * (pairwise coprime, taken greedily downward from 256), each stored as a
 * signed byte in [-m/2, m/2), one plane per modulus in the slice layout.
 * Each vector has its own P: with t_p = ceil(|x_p| 2^(10-e)) <= 1024 and
 * S = sum_p t_p^2 (exact integer, any order), the largest P (capped at
 * OZ2_MAX_BITS) with S 4^P <= T, T = 2^19 Mr, Mr = M - floor(M / 2^10) - 1
 * and M the product of the moduli (T rounded down to its top 64 bits,
 * Oz2Bound).  |rint(x 2^(P-e))| <= t 2^(P-10) (P >= 10) bounds the vector's
 * squared 2-norm by 2^(2P-20) S <= Mr / 2, so any dot product of two scaled
 * vectors stays within Mr / 2 < (1 - 2^-10) M / 2 (Cauchy-Schwarz): the
 * reconstruction's X / M stays 2^-11 or more from a half-integer.  Compared
 * with one P from k max-size terms this keeps up to log2(sqrt(k)) more bits
 * of vectors whose mass sits in few entries. */
constexpr int OZ2_MAX_MODULI = 22;
constexpr int OZ2_MIN_MODULI = 8;
constexpr int OZ2_KCHUNK = 65536; /* K bytes per chunk: kc * 128^2 < 2^31 */
constexpr int OZ2_MAX_BITS = 77;  /* cap on P: scaled integers below 2^78 */
constexpr int OZ2_MODULI[OZ2_MAX_MODULI] = {256, 255, 253, 251, 247, 241, 239, 233,
                                            229, 227, 223, 217, 211, 199, 197, 193,
                                            191, 181, 179, 173, 167, 163};

/* The bit-count budget T >= t64 2^e of the first nmod moduli (above). */
struct Oz2Bound {
  unsigned long long t64;
  int e;
};

/* Host: Oz2Bound of the first nmod moduli (8 <= nmod <= OZ2_MAX_MODULI). */
inline Oz2Bound oz2_bound(int nmod) {
  constexpr int W = 8; /* 32-bit words: M < 2^176 */
  unsigned m[W] = {1u};
  for (int j = 0; j < nmod; ++j) {
    unsigned long long carry = 0;
    for (int l = 0; l < W; ++l) {
      const unsigned long long v = (unsigned long long)m[l] * (unsigned)OZ2_MODULI[j] + carry;
      m[l] = (unsigned)v;
      carry = v >> 32;
    }
  }
  /* Mr = M - (M >> 10) - 1 */
  unsigned sh[W];
  for (int l = 0; l < W; ++l) sh[l] = (m[l] >> 10) | (l + 1 < W ? m[l + 1] << 22 : 0u);
  long long borrow = 1; /* the trailing - 1 */
  for (int l = 0; l < W; ++l) {
    long long v = (long long)m[l] - (long long)sh[l] - borrow;
    borrow = v < 0;
    m[l] = (unsigned)(v + (borrow ? (1ll << 32) : 0));
  }
  int top = W - 1;
  while (top > 0 && m[top] == 0u) --top;
  const int bl = 32 * top + 32 - __builtin_clz(m[top]); /* bitlen(Mr) >= 64 */
  /* t64 = floor(Mr / 2^(bl - 64)), e = bl - 64 + 19 */
  const int s = bl - 64;
  unsigned long long t = 0;
  for (int b = 63; b >= 0; --b) {
    const int bit = s + b, w = bit >> 5;
    t = (t << 1) | ((m[w] >> (bit & 31)) & 1u);
  }
  return Oz2Bound{t, s + 19};
}

/* Per vector the shift D = P - e (OZ_NONFINITE for a vector holding Inf or
 * NaN) into ex, and nmod residue planes of an FP64 operand; bd = oz2_bound(nmod).
 * mbits and sums: nvec scratch words each. */
hipError_t ozaki2_prepare(const double *x, int64_t ld, bool contiguous, int nvec, int k,
                          int64_t kpad, int nmod, Oz2Bound bd, int *ex, unsigned long long *mbits,
                          unsigned long long *sums, int8_t *res, hipStream_t st);
/* The vector bit count P for bound sum S (host and device): the largest
 * P <= OZ2_MAX_BITS with S 4^P <= t64 2^e.  A fitting P has
 * 2P + bitlen(S) - 1 < 64 + e, and the start value below fits or its
 * predecessor does, so the loop runs at most twice. */
__host__ __device__ inline int ozaki2_bits_for_sum(Oz2Bound b, unsigned long long S) {
  int bl = 0; /* bitlen(S) */
  for (unsigned long long v = S; v; v >>= 1) ++bl;
  int P = (64 + b.e - bl) / 2;
  if (P > OZ2_MAX_BITS) P = OZ2_MAX_BITS;
  for (; P > 0; --P) {
    const int d = 2 * P - b.e; /* S 4^P <= t64 2^e  <=>  S 2^d <= t64 */
    if (d <= 0 ? S <= b.t64 : (d < 64 && S <= (b.t64 >> d))) break;
  }
  return P;
}

/* Kinds of slice products. */
enum SliceKind {
  SP_I8_SUM = 0, /* INT8 slices, group g's INT32 sums to S + gout[g]*m*n */
  SP_BF16 = 1,   /* BF16 slices, one group, C = fma(alpha, acc, beta*C) */
  SP_I8_MOD = 2  /* INT8 residues, group g multiplies plane g by plane g and
                    stores the sum modulo gout[g] as a byte to
                    (uint8_t *)S + g*m*n */
};

/* Products of slice pairs over the K byte range [kb0, kb1) (multiples of
 * SL_BK); outputs column-major with leading dimension m (S) or ldc (C). */
hipError_t slice_products(int kind, const int8_t *as, const int8_t *bs, int m, int n,
                          int64_t kpad_bytes, int kb0, int kb1, const SliceGroups &g,
                          void *S, float *C, int64_t ldc, float alpha, float beta,
                          hipStream_t st);

} // namespace gemmx
#endif
