/*
 * gemm_split.hip -- the split pass of the emulated forms: every row of op(A)
 * and column of op(B) is cut into low-precision slices stored K-contiguous
 * (see gemmx_slices.h), with any transpose folded into this pass.
 *
 * Ozaki-I (FP64 -> INT8): per vector the exponent e with max|x| < 2^e, then s
 * signed 7-bit digits of x*2^-e by truncation toward zero,
 * x*2^-e = sum_t d_t 2^-7t + r, |d_t| <= 127; every step is exact in FP64.
 * Ozaki-II (FP64 -> INT8 residues): per vector the same exponent e and a bit
 * count P from the vector's 2-norm bound, then the residues of
 * rint(x * 2^(P-e)) modulo each modulus (exact integer work).
 * BF16x3 (FP32 -> 2 x BF16): hi = bf16_rne(x), lo = bf16_rne(x - hi), the
 * difference being exact in FP32.
 * No reference counterpart (the reference packs A and B for its AVX2
 * kernels, wjc404/GEMM_AVX2 src/gemm_copy.c:4-143, without changing their
 * precision).
 */
#include "gemmx_device.h"
#include "gemmx_slices.h"

namespace gemmx {
namespace {

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

/* Max |x| bits per vector with vectors adjacent (x[v + p*eld]).  Block
 * (64 vectors) x (256 K entries): each thread scans 64 entries of one vector
 * (a wave reads 64 adjacent vectors per entry), the block reduces over its 4
 * K quarters and folds the result into maxbits[v] with an atomic max.
 * maxbits must start at zero. */
__global__ void __launch_bounds__(256)
absmax_strided_kernel(const double *__restrict__ x, int64_t eld, int nvec, int len,
                      unsigned long long *__restrict__ maxbits) {
  __shared__ unsigned long long part[4][64];
  const int vl = threadIdx.x & 63, kq = threadIdx.x >> 6;
  const int v = blockIdx.x * 64 + vl;
  const int p0 = blockIdx.y * 256 + kq * 64;
  unsigned long long mx = 0;
  if (v < nvec) {
    const int p1 = p0 + 64 < len ? p0 + 64 : len;
    for (int p = p0; p < p1; ++p) {
      const unsigned long long b = (unsigned long long)__double_as_longlong(x[v + p * eld]) & 0x7fffffffffffffffull;
      mx = b > mx ? b : mx;
    }
  }
  part[kq][vl] = mx;
  __syncthreads();
  if (kq == 0 && v < nvec) {
#pragma unroll
    for (int q = 1; q < 4; ++q) mx = part[q][vl] > mx ? part[q][vl] : mx;
    atomicMax(maxbits + v, mx);
  }
}

__global__ void __launch_bounds__(256)
zero_u64_kernel(unsigned long long *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void __launch_bounds__(256)
exp_from_bits_kernel(const unsigned long long *__restrict__ maxbits, int nvec,
                     int *__restrict__ out) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nvec) out[v] = exponent_of_absbits(maxbits[v]);
}


__device__ __forceinline__ int oz2_shift(Oz2Bound bd, int e, unsigned long long S) {
  return e == OZ_NONFINITE ? OZ_NONFINITE : ozaki2_bits_for_sum(bd, S) - e;
}

/* Ozaki-II shift D = P - e per vector with elements contiguous: a block of
 * 256 threads per vector.  Vectors of up to 256 * OZ2_SHIFT_REGS entries are
 * read once, into registers (every load issued before the first is used),
 * and both the exponent and the bound sum come from there; longer vectors
 * are read twice.  The bound terms are summed in FP64, exactly (integers
 * below 2^53), then in 64-bit integers. */
constexpr int OZ2_SHIFT_REGS = 32;

/* Ozaki-II bound term of one entry: ceil(|x| 2^(10-e)) squared, exact in
 * FP64 (the scaling by a power of two and the ceil are exact, t <= 1024). */
__device__ __forceinline__ double oz2_term_d(double x, int e) {
  const double t = ceil(ldexp(fabs(x), 10 - e));
  return t * t;
}

__device__ __forceinline__ unsigned long long max_u64(unsigned long long a,
                                                      unsigned long long b) {
  return a > b ? a : b;
}

__global__ void __launch_bounds__(256)
oz2_shift_contig_kernel(const double *__restrict__ x, int64_t vld, int nvec, int len, Oz2Bound bd,
                        int *__restrict__ out) {
  constexpr int R = OZ2_SHIFT_REGS;
  __shared__ unsigned long long red[4];
  __shared__ int eshared;
  const int v = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double *p = x + v * vld;
  const bool in_regs = len <= 256 * R;
  double r[R];
  unsigned long long mx = 0;
  if (in_regs) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int j = tid + 256 * i;
      r[i] = j < len ? p[j] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
      mx = max_u64(mx, (unsigned long long)__double_as_longlong(r[i]) & 0x7fffffffffffffffull);
  } else {
    for (int j = tid; j < len; j += 256)
      mx = max_u64(mx, (unsigned long long)__double_as_longlong(p[j]) & 0x7fffffffffffffffull);
  }
  for (int off = 32; off > 0; off >>= 1) mx = max_u64(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) red[wave] = mx;
  __syncthreads();
  if (tid == 0)
    eshared = exponent_of_absbits(max_u64(max_u64(red[0], red[1]), max_u64(red[2], red[3])));
  __syncthreads();
  const int e = eshared;
  double sd = 0.0;
  if (e != OZ_NONFINITE) {
    if (in_regs) {
#pragma unroll
      for (int i = 0; i < R; ++i) sd += oz2_term_d(r[i], e);
    } else {
      for (int j = tid; j < len; j += 256) sd += oz2_term_d(p[j], e);
    }
  }
  unsigned long long S = (unsigned long long)sd;
  for (int off = 32; off > 0; off >>= 1) S += __shfl_xor(S, off, 64);
  __syncthreads(); /* red[] reused */
  if (lane == 0) red[wave] = S;
  __syncthreads();
  if (tid == 0) out[v] = oz2_shift(bd, e, red[0] + red[1] + red[2] + red[3]);
}

/* Bound sums per vector with vectors adjacent (x[v + p*eld]), in the blocks
 * of absmax_strided_kernel; ex holds the exponents, sums must start at zero
 * (integer atomics: the result does not depend on the order). */
__global__ void __launch_bounds__(256)
oz2_sums_strided_kernel(const double *__restrict__ x, int64_t eld, int nvec, int len,
                        const int *__restrict__ ex, unsigned long long *__restrict__ sums) {
  __shared__ unsigned long long part[4][64];
  const int vl = threadIdx.x & 63, kq = threadIdx.x >> 6;
  const int v = blockIdx.x * 64 + vl;
  const int p0 = blockIdx.y * 256 + kq * 64;
  double sd = 0.0; /* at most 64 terms of at most 2^20: exact */
  if (v < nvec) {
    const int e = ex[v];
    if (e != OZ_NONFINITE) {
      const int p1 = p0 + 64 < len ? p0 + 64 : len;
      for (int p = p0; p < p1; ++p) sd += oz2_term_d(x[v + p * eld], e);
    }
  }
  unsigned long long S = (unsigned long long)sd;
  part[kq][vl] = S;
  __syncthreads();
  if (kq == 0 && v < nvec) {
#pragma unroll
    for (int q = 1; q < 4; ++q) S += part[q][vl];
    atomicAdd(sums + v, S);
  }
}

/* ex: exponent -> shift D = P - e, from the bound sums. */
__global__ void __launch_bounds__(256)
oz2_shift_from_sums_kernel(const unsigned long long *__restrict__ sums, int nvec, Oz2Bound bd,
                           int *__restrict__ ex) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nvec) ex[v] = oz2_shift(bd, ex[v], sums[v]);
}

/* 4 NW consecutive K entries of one vector per thread -> 4 NW bytes per
 * slice (vectors contiguous: x[v*vld + p]); a thread whose entries are all in
 * range and 16-byte aligned reads them with 16-byte loads.  Eight entries
 * (NW = 2) keep many waves per SIMD and the loads wide. */
template <int S, int NW>
__global__ void __launch_bounds__(256)
ozaki_split_kernel(const double *__restrict__ x, int64_t vld, int nvec, int len, int64_t kpad,
                   const int *__restrict__ ex, int8_t *__restrict__ out) {
  constexpr int EPT = 4 * NW;
  typedef int vec_t __attribute__((ext_vector_type(NW)));
  const int64_t nblk = kpad / EPT;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (tid >= nblk * nvec) return;
  const int64_t v = tid / nblk, blk = tid % nblk;
  const int e = ex[v];
  const double *xp = x + v * vld + blk * EPT;
  double xv[EPT];
  if (blk * EPT + EPT <= len && (reinterpret_cast<uintptr_t>(xp) & 15) == 0) {
#pragma unroll
    for (int i = 0; i < EPT / 2; ++i) {
      const d2_t t = reinterpret_cast<const d2_t *>(xp)[i];
      xv[2 * i] = t[0];
      xv[2 * i + 1] = t[1];
    }
  } else {
#pragma unroll
    for (int i = 0; i < EPT; ++i) xv[i] = blk * EPT + i < len ? xp[i] : 0.0;
  }
  vec_t dig[S];
#pragma unroll
  for (int t = 0; t < S; ++t) dig[t] = vec_t{};
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    double r = (e == OZ_NONFINITE) ? 0.0 : ldexp(xv[q], -e);
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const double y = r * 128.0;
      const double d = trunc(y);
      r = y - d;
      const unsigned byte = (unsigned)((int)d) & 0xffu;
      dig[t][q >> 2] |= (int)(byte << (8 * (q & 3)));
    }
  }
  int8_t *o = out + v * kpad + blk * EPT;
  const int64_t plane = (int64_t)nvec * kpad;
#pragma unroll
  for (int t = 0; t < S; ++t, o += plane) *reinterpret_cast<vec_t *>(o) = dig[t];
}

/* Same digits for vectors adjacent in memory (x[v + p*eld]), through an LDS
 * transpose so that both the reads (64 adjacent vectors per entry) and the
 * writes (128 contiguous bytes of one vector per slice) coalesce.  Block:
 * 32 vectors x 128 K entries.  LDS row of a vector keeps entry p at
 * (p % 16) * 8 + p / 16, so the 8 threads of one vector read 8 adjacent
 * words for each of their 16 entries. */
template <int S>
__global__ void __launch_bounds__(256)
ozaki_split_strided_kernel(const double *__restrict__ x, int64_t eld, int nvec,
                           int len, int64_t kpad, const int *__restrict__ ex,
                           int8_t *__restrict__ out) {
  constexpr int ROW = 129; /* 128 entries + 1 against bank conflicts */
  __shared__ double tile[32 * ROW];
  const int v0 = blockIdx.x * 32;
  const int64_t p0 = (int64_t)blockIdx.y * 128;
  {
    const int vl = threadIdx.x & 31, pr = threadIdx.x >> 5;
    const int v = v0 + vl;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int p = pr + 8 * j;
      const double val = (v < nvec && p0 + p < len) ? x[v + (p0 + p) * eld] : 0.0;
      tile[vl * ROW + (p % 16) * 8 + p / 16] = val;
    }
  }
  __syncthreads();
  const int vl = threadIdx.x >> 3, q = threadIdx.x & 7;
  const int v = v0 + vl;
  if (v >= nvec) return;
  const int e = ex[v];
  i4_t dig[S];
#pragma unroll
  for (int t = 0; t < S; ++t) dig[t] = i4_t{0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const double val = tile[vl * ROW + i * 8 + q];
    double r = (e == OZ_NONFINITE) ? 0.0 : ldexp(val, -e);
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const double y = r * 128.0;
      const double d = trunc(y);
      r = y - d;
      const unsigned byte = (unsigned)((int)d) & 0xffu;
      dig[t][i >> 2] |= (int)(byte << (8 * (i & 3)));
    }
  }
#pragma unroll
  for (int t = 0; t < S; ++t)
    *reinterpret_cast<i4_t *>(out + ((int64_t)t * nvec + v) * kpad + p0 + 16 * q) = dig[t];
}

/* ---- Ozaki-II residues ------------------------------------------------
 * The scaled integer a = rint(x * 2^D), D = P - e (|a| < 2^78; x * 2^D is
 * exact except below 2^-1022, where rint gives 0 either way; ties to even),
 * is split as a = h * 2^39 + l with h = trunc(a / 2^39): h and l carry a's
 * sign, |h|, |l| < 2^39, all exact in FP64.  For an odd modulus m,
 *   v = h * (2^39 mod m) + l  (|v| < 2^47, exact, v = a mod m)
 * and q = rint(v / m) is the exact nearest integer from one rounded product
 * v * (1/m): its error |v/m| 2^-52 stays below 1/(2m), the least distance of
 * v/m from a half-integer (2v - (2k+1)m is odd), so r = v - q m is the
 * centred residue in [-(m-1)/2, (m-1)/2].  Only r's low byte is stored, so
 * it comes from integer arithmetic modulo 2^8: q from the low word of
 * fma(v, 1/m, 1.5 * 2^52), v's low bits from the low words of h and l (each
 * read the same way), and r = v + q (2^24 - m) with 24-bit multiply-adds.
 * m = 256 takes l's low byte (256 divides 2^39). */
struct Oz2Val {
  double h, l;      /* a = h 2^39 + l */
  unsigned hlo, llo; /* h, l modulo 2^32 */
};

constexpr double OZ2_MAGIC = 0x1.8p52; /* x + 1.5 * 2^52: low word = rint(x) mod 2^32 */

__device__ __forceinline__ unsigned oz2_low_word(double x) {
  return (unsigned)__double_as_longlong(x + OZ2_MAGIC);
}

__device__ __forceinline__ Oz2Val oz2_val(double x, int D) {
  Oz2Val r;
  const double a = __builtin_rint(ldexp(x, D));
  r.h = __builtin_trunc(a * 0x1p-39);
  r.l = __builtin_fma(-r.h, 0x1p39, a);
  r.hlo = oz2_low_word(r.h);
  r.llo = oz2_low_word(r.l);
  return r;
}

constexpr unsigned pow2_mod(int e, int m) {
  unsigned r = 1u % (unsigned)m;
  for (int i = 0; i < e; ++i) r = (r * 2u) % (unsigned)m;
  return r;
}

/* a * b + c over the low 24 bits of a and b (v_mad_u32_u24, b scalar). */
__device__ __forceinline__ unsigned oz2_mad24(unsigned a, unsigned b, unsigned c) {
  unsigned r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}

/* Residue of a modulo OZ2_MODULI[J] in the low byte (higher bits unused). */
template <int J>
__device__ __forceinline__ unsigned oz2_residue_byte(const Oz2Val &a) {
  constexpr int m = OZ2_MODULI[J];
  if constexpr (m == 256) {
    return a.llo;
  } else {
    /* The constants sit in scalar registers, so that both products and both
     * multiply-adds are single three-operand instructions. */
    double c39d = (double)pow2_mod(39, m), rcp = 1.0 / (double)m;
    unsigned c39 = pow2_mod(39, m), mneg = (1u << 24) - (unsigned)m;
    asm("" : "+s"(c39d), "+s"(rcp), "+s"(c39), "+s"(mneg));
    const double v = __builtin_fma(a.h, c39d, a.l);
    const unsigned q = (unsigned)__double_as_longlong(__builtin_fma(v, rcp, OZ2_MAGIC));
    return oz2_mad24(q, mneg, oz2_mad24(a.hlo, c39, a.llo));
  }
}

/* Low bytes of b0..b3 -> one word, b0 lowest. */
__device__ __forceinline__ int oz2_pack4(unsigned b0, unsigned b1, unsigned b2, unsigned b3) {
  const unsigned lo = __builtin_amdgcn_perm(b1, b0, 0x0c0c0400u);
  const unsigned hi = __builtin_amdgcn_perm(b3, b2, 0x0c0c0400u);
  return (int)__builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

/* The residue bytes of four consecutive entries, packed into word W of each
 * of the nmod plane vectors (all indices compile-time constants). */
template <int J, int W, typename V>
__device__ __forceinline__ void oz2_encode4(const Oz2Val (&a)[4], int nmod,
                                            V (&res)[OZ2_MAX_MODULI]) {
  if (J < nmod)
    res[J][W] = oz2_pack4(oz2_residue_byte<J>(a[0]), oz2_residue_byte<J>(a[1]),
                          oz2_residue_byte<J>(a[2]), oz2_residue_byte<J>(a[3]));
  if constexpr (J + 1 < OZ2_MAX_MODULI) oz2_encode4<J + 1, W>(a, nmod, res);
}

/* nmod residue planes of 4 NW consecutive K entries of one vector per thread
 * (vectors contiguous: x[v*vld + p]); a thread whose entries are all in range
 * and 16-byte aligned reads them with 16-byte loads. */
template <int NW>
__global__ void __launch_bounds__(256)
ozaki2_split_kernel(const double *__restrict__ x, int64_t vld, int nvec, int len, int64_t kpad,
                    const int *__restrict__ ex, int nmod, int8_t *__restrict__ out) {
  constexpr int EPT = 4 * NW;
  typedef int vec_t __attribute__((ext_vector_type(NW)));
  const int64_t nblk = kpad / EPT;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (tid >= nblk * nvec) return;
  const int64_t v = tid / nblk, blk = tid % nblk;
  const int d = ex[v]; /* shift P - e */
  const double *xp = x + v * vld + blk * EPT;
  double xv[EPT];
  if (blk * EPT + EPT <= len && d != OZ_NONFINITE && (reinterpret_cast<uintptr_t>(xp) & 15) == 0) {
#pragma unroll
    for (int i = 0; i < EPT / 2; ++i) {
      const d2_t t = reinterpret_cast<const d2_t *>(xp)[i];
      xv[2 * i] = t[0];
      xv[2 * i + 1] = t[1];
    }
  } else {
#pragma unroll
    for (int i = 0; i < EPT; ++i)
      xv[i] = (blk * EPT + i < len && d != OZ_NONFINITE) ? xp[i] : 0.0;
  }
  vec_t res[OZ2_MAX_MODULI];
#pragma unroll
  for (int j = 0; j < OZ2_MAX_MODULI; ++j) res[j] = vec_t{};
  Oz2Val a[4];
#define GEMMX_OZ2_WORD(W)                                                                \
  if constexpr ((W) < NW) {                                                              \
    for (int q = 0; q < 4; ++q) a[q] = oz2_val(xv[4 * (W) + q], d);                      \
    oz2_encode4<0, W>(a, nmod, res);                                                     \
  }
  GEMMX_OZ2_WORD(0)
  GEMMX_OZ2_WORD(1)
  GEMMX_OZ2_WORD(2)
  GEMMX_OZ2_WORD(3)
#undef GEMMX_OZ2_WORD
  int8_t *o = out + v * kpad + blk * EPT;
  const int64_t plane = (int64_t)nvec * kpad;
#pragma unroll
  for (int j = 0; j < OZ2_MAX_MODULI; ++j, o += plane)
    if (j < nmod) *reinterpret_cast<vec_t *>(o) = res[j];
}

/* Bound pass and residue split in one kernel for contiguous vectors of at
 * most 256 * 8 * OZ2_FUSED_CHUNKS entries: a block of 256 threads per vector
 * reads it once, as chunks of 8 entries (chunk c = thread + 256 i, 16-byte
 * loads when aligned), derives the shift D = P - e exactly as
 * oz2_shift_contig_kernel does, and writes every chunk's residue words
 * (8 bytes per plane) from the registers, so the operand is read from HBM
 * once instead of twice. */
constexpr int OZ2_FUSED_CHUNKS = 4;

__global__ void __launch_bounds__(256)
ozaki2_shift_split_contig_kernel(const double *__restrict__ x, int64_t vld, int nvec, int len,
                                 int64_t kpad, Oz2Bound bd, int nmod, int *__restrict__ ex,
                                 int8_t *__restrict__ out) {
  constexpr int R = OZ2_FUSED_CHUNKS, EPT = 8;
  typedef int vec_t __attribute__((ext_vector_type(2)));
  __shared__ unsigned long long red[4];
  __shared__ int eshared, dshared;
  const int v = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double *p = x + (int64_t)v * vld;
  const int nchunks = (int)(kpad / EPT);
  double xv[R][EPT];
  unsigned long long mx = 0;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int c = tid + 256 * i;
    const double *cp = p + (int64_t)c * EPT;
    if (c * EPT + EPT <= len && (reinterpret_cast<uintptr_t>(cp) & 15) == 0) {
#pragma unroll
      for (int h = 0; h < EPT / 2; ++h) {
        const d2_t t = reinterpret_cast<const d2_t *>(cp)[h];
        xv[i][2 * h] = t[0];
        xv[i][2 * h + 1] = t[1];
      }
    } else {
#pragma unroll
      for (int q = 0; q < EPT; ++q) xv[i][q] = c * EPT + q < len ? cp[q] : 0.0;
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int q = 0; q < EPT; ++q)
      mx = max_u64(mx, (unsigned long long)__double_as_longlong(xv[i][q]) & 0x7fffffffffffffffull);
  for (int off = 32; off > 0; off >>= 1) mx = max_u64(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) red[wave] = mx;
  __syncthreads();
  if (tid == 0)
    eshared = exponent_of_absbits(max_u64(max_u64(red[0], red[1]), max_u64(red[2], red[3])));
  __syncthreads();
  const int e = eshared;
  double sd = 0.0;
  if (e != OZ_NONFINITE) {
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int q = 0; q < EPT; ++q) sd += oz2_term_d(xv[i][q], e);
  }
  unsigned long long S = (unsigned long long)sd;
  for (int off = 32; off > 0; off >>= 1) S += __shfl_xor(S, off, 64);
  __syncthreads(); /* red[] reused */
  if (lane == 0) red[wave] = S;
  __syncthreads();
  if (tid == 0) {
    const int d = oz2_shift(bd, e, red[0] + red[1] + red[2] + red[3]);
    ex[v] = d;
    dshared = d;
  }
  __syncthreads();
  const int d = dshared; /* shift P - e */
  const int64_t plane = (int64_t)nvec * kpad;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int c = tid + 256 * i;
    if (c >= nchunks) break;
    vec_t res[OZ2_MAX_MODULI];
#pragma unroll
    for (int j = 0; j < OZ2_MAX_MODULI; ++j) res[j] = vec_t{};
    Oz2Val a[4];
#define GEMMX_OZ2_WORD(W)                                                                \
    for (int q = 0; q < 4; ++q)                                                          \
      a[q] = oz2_val(d != OZ_NONFINITE ? xv[i][4 * (W) + q] : 0.0, d);                   \
    oz2_encode4<0, W>(a, nmod, res);
    GEMMX_OZ2_WORD(0)
    GEMMX_OZ2_WORD(1)
#undef GEMMX_OZ2_WORD
    int8_t *o = out + (int64_t)v * kpad + (int64_t)c * EPT;
#pragma unroll
    for (int j = 0; j < OZ2_MAX_MODULI; ++j, o += plane)
      if (j < nmod) *reinterpret_cast<vec_t *>(o) = res[j];
  }
}

/* The same for vectors adjacent in memory (x[v + p*eld]) through an LDS
 * tile of 32 vectors x 128 K entries: entry p of a vector sits in column
 * (p % EPT) TPV + p / EPT of its padded row, so that the TPV threads of a
 * vector, thread q taking entries q EPT .. q EPT + EPT - 1, read adjacent
 * columns.  The tile is loaded two vectors at a time (16-byte loads when x
 * and eld allow). */
template <int NW>
__global__ void __launch_bounds__(1024 / NW)
ozaki2_split_strided_kernel(const double *__restrict__ x, int64_t eld, int nvec, int len,
                            int64_t kpad, const int *__restrict__ ex, int nmod,
                            int8_t *__restrict__ out) {
  constexpr int EPT = 4 * NW, TPV = 128 / EPT, NT = 32 * TPV, ROW = 129;
  typedef int vec_t __attribute__((ext_vector_type(NW)));
  __shared__ double tile[32 * ROW];
  const int tid = threadIdx.x;
  const int v0 = blockIdx.x * 32;
  const int64_t p0 = (int64_t)blockIdx.y * 128;
  const bool aligned = (reinterpret_cast<uintptr_t>(x) & 15) == 0 && (eld & 1) == 0;
#pragma unroll
  for (int j = 0; j < 2048 / NT; ++j) {
    const int idx = tid + NT * j, vp = idx & 15, pr = idx >> 4;
    const int v = v0 + 2 * vp;
    const int64_t p = p0 + pr;
    double a0 = 0.0, a1 = 0.0;
    if (p < len) {
      const double *xp = x + v + p * eld;
      if (aligned && v + 1 < nvec) {
        const d2_t t = *reinterpret_cast<const d2_t *>(xp);
        a0 = t[0];
        a1 = t[1];
      } else {
        if (v < nvec) a0 = xp[0];
        if (v + 1 < nvec) a1 = xp[1];
      }
    }
    const int col = (pr % EPT) * TPV + pr / EPT;
    tile[(2 * vp) * ROW + col] = a0;
    tile[(2 * vp + 1) * ROW + col] = a1;
  }
  __syncthreads();
  const int vl = tid / TPV, q = tid % TPV;
  const int v = v0 + vl;
  if (v >= nvec) return;
  const int d = ex[v]; /* shift P - e */
  double xv[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) xv[i] = d != OZ_NONFINITE ? tile[vl * ROW + i * TPV + q] : 0.0;
  vec_t res[OZ2_MAX_MODULI];
#pragma unroll
  for (int j = 0; j < OZ2_MAX_MODULI; ++j) res[j] = vec_t{};
  Oz2Val a[4];
#define GEMMX_OZ2_WORD(W)                                                                \
  if constexpr ((W) < NW) {                                                              \
    for (int i = 0; i < 4; ++i) a[i] = oz2_val(xv[4 * (W) + i], d);                      \
    oz2_encode4<0, W>(a, nmod, res);                                                     \
  }
  GEMMX_OZ2_WORD(0)
  GEMMX_OZ2_WORD(1)
  GEMMX_OZ2_WORD(2)
  GEMMX_OZ2_WORD(3)
#undef GEMMX_OZ2_WORD
  int8_t *o = out + (int64_t)v * kpad + p0 + EPT * q;
  const int64_t plane = (int64_t)nvec * kpad;
#pragma unroll
  for (int j = 0; j < OZ2_MAX_MODULI; ++j, o += plane)
    if (j < nmod) *reinterpret_cast<vec_t *>(o) = res[j];
}

__device__ __forceinline__ unsigned bf16_rne_bits(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

/* 8 consecutive K entries of one vector per thread (vectors contiguous:
 * x[v*vld + p]) -> 16 bytes of hi and of lo.  out is [2][nvec][kpad] bf16 (hi
 * slice first). */
__global__ void __launch_bounds__(256)
bf16_split_kernel(const float *__restrict__ x, int64_t vld, int nvec, int len, int64_t kpad,
                  unsigned short *__restrict__ out) {
  const int64_t nblk = kpad / 8;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (tid >= nblk * nvec) return;
  const int64_t v = tid / nblk, blk = tid % nblk;
  float xv[8];
  const float *xp = x + v * vld + blk * 8;
  if (blk * 8 + 8 <= len && (reinterpret_cast<uintptr_t>(xp) & 15) == 0) {
    /* in range and aligned: two 16-byte loads */
    const f4_t a = reinterpret_cast<const f4_t *>(xp)[0], b = reinterpret_cast<const f4_t *>(xp)[1];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      xv[i] = a[i];
      xv[4 + i] = b[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) xv[i] = blk * 8 + i < len ? xp[i] : 0.0f;
  }
  i4_t hi, lo;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned h2[2], l2[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const float val = xv[2 * q + e];
      h2[e] = bf16_rne_bits(val);
      l2[e] = bf16_rne_bits(val - __uint_as_float(h2[e] << 16));
    }
    hi[q] = (int)(h2[0] | (h2[1] << 16));
    lo[q] = (int)(l2[0] | (l2[1] << 16));
  }
  *reinterpret_cast<i4_t *>(out + v * kpad + blk * 8) = hi;
  *reinterpret_cast<i4_t *>(out + ((int64_t)nvec + v) * kpad + blk * 8) = lo;
}

/* The same for vectors adjacent in memory (x[v + p*eld]) through an LDS tile
 * of 32 vectors x 64 K entries: the loads read 32 adjacent vectors per entry,
 * and the 8 threads of a vector (thread q taking entries 8q .. 8q + 7, kept in
 * LDS column (p % 8) * 8 + p / 8 of the vector's padded row) write 128
 * contiguous bytes of hi and of lo. */
__global__ void __launch_bounds__(256)
bf16_split_strided_kernel(const float *__restrict__ x, int64_t eld, int nvec, int len,
                          int64_t kpad, unsigned short *__restrict__ out) {
  constexpr int ROW = 65;
  __shared__ float tile[32 * ROW];
  const int tid = threadIdx.x;
  const int v0 = blockIdx.x * 32;
  const int64_t p0 = (int64_t)blockIdx.y * 64;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int idx = tid + 256 * j, vl = idx & 31, pr = idx >> 5;
    const int v = v0 + vl;
    const int64_t p = p0 + pr;
    const float val = (v < nvec && p < len) ? x[v + p * eld] : 0.0f;
    tile[vl * ROW + (pr % 8) * 8 + pr / 8] = val;
  }
  __syncthreads();
  const int vl = tid >> 3, q = tid & 7;
  const int v = v0 + vl;
  if (v >= nvec) return;
  i4_t hi, lo;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    unsigned h2[2], l2[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const float val = tile[vl * ROW + (2 * w + e) * 8 + q];
      h2[e] = bf16_rne_bits(val);
      l2[e] = bf16_rne_bits(val - __uint_as_float(h2[e] << 16));
    }
    hi[w] = (int)(h2[0] | (h2[1] << 16));
    lo[w] = (int)(l2[0] | (l2[1] << 16));
  }
  *reinterpret_cast<i4_t *>(out + (int64_t)v * kpad + p0 + 8 * q) = hi;
  *reinterpret_cast<i4_t *>(out + ((int64_t)nvec + v) * kpad + p0 + 8 * q) = lo;
}

template <int S>
hipError_t launch_ozaki_split(const double *x, int64_t vld, int nvec, int len, int64_t kpad,
                              const int *ex, int8_t *out, hipStream_t st) {
  constexpr int NW = 2;
  const int64_t threads = (kpad / (4 * NW)) * nvec;
  hipLaunchKernelGGL((ozaki_split_kernel<S, NW>), dim3((unsigned)((threads + 255) / 256)),
                     dim3(256), 0, st, x, vld, nvec, len, kpad, ex, out);
  return hipGetLastError();
}

template <int S>
hipError_t launch_split_strided(const double *x, int64_t eld, int nvec, int len,
                                int64_t kpad, const int *ex, int8_t *out, hipStream_t st) {
  hipLaunchKernelGGL((ozaki_split_strided_kernel<S>), dim3((unsigned)(((int64_t)nvec + 31) / 32), (unsigned)(kpad / 128)),
                     dim3(256), 0, st, x, eld, nvec, len, kpad, ex, out);
  return hipGetLastError();
}

hipError_t ozaki_split_strided(int s, const double *x, int64_t eld, int nvec, int len,
                               int64_t kpad, const int *ex, int8_t *out, hipStream_t st) {
  switch (s) {
  case 3: return launch_split_strided<3>(x, eld, nvec, len, kpad, ex, out, st);
  case 4: return launch_split_strided<4>(x, eld, nvec, len, kpad, ex, out, st);
  case 5: return launch_split_strided<5>(x, eld, nvec, len, kpad, ex, out, st);
  case 6: return launch_split_strided<6>(x, eld, nvec, len, kpad, ex, out, st);
  case 7: return launch_split_strided<7>(x, eld, nvec, len, kpad, ex, out, st);
  default: return launch_split_strided<8>(x, eld, nvec, len, kpad, ex, out, st);
  }
}

/* Contiguous vectors (x[v*vld + p]). */
hipError_t ozaki_split(int s, const double *x, int64_t vld, int nvec, int len, int64_t kpad,
                       const int *ex, int8_t *out, hipStream_t st) {
  switch (s) {
  case 3: return launch_ozaki_split<3>(x, vld, nvec, len, kpad, ex, out, st);
  case 4: return launch_ozaki_split<4>(x, vld, nvec, len, kpad, ex, out, st);
  case 5: return launch_ozaki_split<5>(x, vld, nvec, len, kpad, ex, out, st);
  case 6: return launch_ozaki_split<6>(x, vld, nvec, len, kpad, ex, out, st);
  case 7: return launch_ozaki_split<7>(x, vld, nvec, len, kpad, ex, out, st);
  default: return launch_ozaki_split<8>(x, vld, nvec, len, kpad, ex, out, st);
  }
}

/* Exponents and slices of op(A) rows (vectors = rows) or op(B) columns.
 * contiguous: vector v at x + v*ld (elements adjacent); else at x + v with
 * element stride ld. */

} // namespace

hipError_t ozaki_prepare(const double *x, int64_t ld, bool contiguous, int nvec,
                         int k, int64_t kpad, int s, int *ex, unsigned long long *mbits,
                         int8_t *slices, hipStream_t st) {
  if (contiguous) {
    hipLaunchKernelGGL(exp_contig_kernel, dim3((unsigned)(((int64_t)nvec + 3) / 4)), dim3(256), 0, st, x,
                       ld, nvec, k, ex);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return ozaki_split(s, x, ld, nvec, k, kpad, ex, slices, st);
  }
  hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)(((int64_t)nvec + 255) / 256)), dim3(256), 0, st, mbits,
                     nvec);
  hipLaunchKernelGGL(absmax_strided_kernel, dim3((unsigned)(((int64_t)nvec + 63) / 64), (unsigned)(((int64_t)k + 255) / 256)),
                     dim3(256), 0, st, x, ld, nvec, k, mbits);
  hipLaunchKernelGGL(exp_from_bits_kernel, dim3((unsigned)(((int64_t)nvec + 255) / 256)), dim3(256), 0, st,
                     mbits, nvec, ex);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return ozaki_split_strided(s, x, ld, nvec, k, kpad, ex, slices, st);
}

hipError_t ozaki2_prepare(const double *x, int64_t ld, bool contiguous, int nvec, int k,
                          int64_t kpad, int nmod, Oz2Bound bd, int *ex, unsigned long long *mbits,
                          unsigned long long *sums, int8_t *res, hipStream_t st) {
  /* 8 entries per thread (two words per plane): fewer registers, so 5 waves
   * per SIMD, measured faster than 16 for both forms */
  constexpr int NW = 2;
  if (contiguous && kpad <= 256 * 8 * OZ2_FUSED_CHUNKS) {
    hipLaunchKernelGGL(ozaki2_shift_split_contig_kernel, dim3(nvec), dim3(256), 0, st, x, ld,
                       nvec, k, kpad, bd, nmod, ex, res);
    return hipGetLastError();
  }
  if (contiguous) {
    hipLaunchKernelGGL(oz2_shift_contig_kernel, dim3(nvec), dim3(256), 0, st, x, ld, nvec, k,
                       bd, ex);
    const int64_t threads = (kpad / (4 * NW)) * nvec;
    hipLaunchKernelGGL(ozaki2_split_kernel<NW>, dim3((unsigned)((threads + 255) / 256)),
                       dim3(256), 0, st, x, ld, nvec, k, kpad, ex, nmod, res);
    return hipGetLastError();
  }
  const dim3 red((unsigned)(((int64_t)nvec + 63) / 64), (unsigned)(((int64_t)k + 255) / 256));
  hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)(((int64_t)nvec + 255) / 256)), dim3(256), 0, st, mbits,
                     nvec);
  hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)(((int64_t)nvec + 255) / 256)), dim3(256), 0, st, sums, nvec);
  hipLaunchKernelGGL(absmax_strided_kernel, red, dim3(256), 0, st, x, ld, nvec, k, mbits);
  hipLaunchKernelGGL(exp_from_bits_kernel, dim3((unsigned)(((int64_t)nvec + 255) / 256)), dim3(256), 0, st,
                     mbits, nvec, ex);
  hipLaunchKernelGGL(oz2_sums_strided_kernel, red, dim3(256), 0, st, x, ld, nvec, k, ex, sums);
  hipLaunchKernelGGL(oz2_shift_from_sums_kernel, dim3((unsigned)(((int64_t)nvec + 255) / 256)), dim3(256), 0, st,
                     sums, nvec, bd, ex);
  hipLaunchKernelGGL(ozaki2_split_strided_kernel<NW>,
                     dim3((unsigned)(((int64_t)nvec + 31) / 32), (unsigned)(kpad / 128)), dim3(1024 / NW), 0, st, x,
                     ld, nvec, k, kpad, ex, nmod, res);
  return hipGetLastError();
}

hipError_t bf16_prepare(const float *x, int64_t ld, bool contiguous, int nvec,
                        int k, int64_t kpad, unsigned short *out, hipStream_t st) {
  const int64_t threads = (kpad / 8) * nvec;
  const dim3 grid((unsigned)((threads + 255) / 256));
  if (contiguous)
    hipLaunchKernelGGL(bf16_split_kernel, grid, dim3(256), 0, st, x, ld, nvec, k, kpad, out);
  else /* kpad is a multiple of 64 */
    hipLaunchKernelGGL(bf16_split_strided_kernel, dim3((unsigned)(((int64_t)nvec + 31) / 32), (unsigned)(kpad / 64)),
                       dim3(256), 0, st, x, ld, nvec, k, kpad, out);
  return hipGetLastError();
}

} // namespace gemmx
