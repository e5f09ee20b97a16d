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
 * Ozaki-II (DGEMM): split into residues of P-bit scaled integers modulo
 * nmod pairwise-coprime moduli <= 256 -> one INT8 GEMM per modulus, its
 * INT32 sum stored modulo the modulus (K chunked so that kc * 128^2 < 2^31)
 * -> exact CRT reconstruction of the integer product here, rounded once to
 * FP64 and scaled; bit-identical to the oracle's exact integer restatement.
 *
 * split-BF16x3 (SGEMM): split -> pairs (lo,hi), (hi,lo), (hi,hi) in one FP32
 * accumulator -> C = fma(alpha, acc, beta*C) in the products epilogue.
 */
#include <string.h>

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

/* ---- Ozaki-II reconstruction ------------------------------------------
 * The residues c_j = X mod m_j of the exact integer X = sum_p a_ip b_pj give
 * X by the Chinese remainder theorem: with M = prod m_j, M_j = M / m_j,
 * w_j = M_j^-1 mod m_j and W_j = (w_j M_j) mod M, X = X' - q M with
 * X' = sum_j c_j W_j and q = rint(X' / M), since |X| <= 2^(L-1) < M/2 (the
 * vectors' bit counts).  X' (< nmod 2^8 M) is formed exactly with 24-bit
 * multiply-adds of the byte c_j into LW 16-bit limbs of W_j (limb sums below
 * 2^29).  q comes exactly from the top four limb sums in FP64: X'/M lies at
 * least (M - 2^L) / (2M) >= 2^-6.2 from a half-integer for every modulus
 * count (checked on the host), far above the estimate's error (< 2^-18).
 * Adding q (2^(16 LW) - M) limb by limb gives X in two's complement over
 * 16 LW bits; |X| is rounded to 53 bits (ties to even) by one conversion of
 * its normalised top 64 bits with the lower bits folded into a sticky bit,
 * and scaled by 2^-(D_i + D_j), D = P - e of row i and column j, with one
 * ldexp. */
constexpr int OZ2_LIMBS = 6;    /* 32-bit limbs: M < 2^171 */
constexpr int OZ2_LIMBS16 = 11; /* 16-bit limbs of M_j */

typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));

struct Oz2Crt {
  int nmod, L;    /* L = floor(log2 M) */
  Oz2Bound bound; /* the bit-count budget of the split (gemmx_slices.h) */
  /* W_j = (w_j M_j) mod M in 16-bit limbs, two moduli per word: limb l of
   * W_2p in the low half of W2[p][l], of W_2p+1 (0 past nmod) in the high */
  unsigned W2[OZ2_MAX_MODULI / 2][OZ2_LIMBS16];
  unsigned Mc16[OZ2_LIMBS16];                /* 2^(16 LW) - M, 16-bit limbs */
  double rMs;                                /* 2^(16 (LW - 4)) / M, rounded */
};

/* 16-bit limbs of W_j and of M for L = floor(log2 M). */
inline int oz2_lw(int L) { return (L + 16) / 16; }

/* Host: the constants of the first nmod moduli. */
bool oz2_constants(int nmod, Oz2Crt *c) {
  if (nmod < OZ2_MIN_MODULI || nmod > OZ2_MAX_MODULI) return false;
  memset(c, 0, sizeof *c);
  c->nmod = nmod;
  unsigned M[OZ2_LIMBS] = {1u}; /* M, 32-bit limbs */
  for (int j = 0; j < nmod; ++j) {
    unsigned long long carry = 0;
    for (int l = 0; l < OZ2_LIMBS; ++l) {
      const unsigned long long v = (unsigned long long)M[l] * OZ2_MODULI[j] + carry;
      M[l] = (unsigned)v;
      carry = v >> 32;
    }
  }
  for (int l = OZ2_LIMBS - 1; l >= 0; --l)
    if (M[l]) {
      c->L = 32 * l + 31 - __builtin_clz(M[l]);
      break;
    }
  for (int j = 0; j < nmod; ++j) {
    const int m = OZ2_MODULI[j];
    unsigned mj[OZ2_LIMBS];
    unsigned long long rem = 0;
    for (int l = OZ2_LIMBS - 1; l >= 0; --l) { /* M_j = M / m, exact */
      const unsigned long long cur = (rem << 32) | M[l];
      mj[l] = (unsigned)(cur / m);
      rem = cur % m;
    }
    unsigned long long r = 0; /* M_j mod m */
    for (int l = OZ2_LIMBS - 1; l >= 0; --l) r = ((r << 32) | mj[l]) % m;
    int inv = 0;
    for (int t = 1; t < m; ++t)
      if ((r * t) % m == 1) {
        inv = t;
        break;
      }
    if (!inv) return false;
    /* W_j = inv * M_j (< m M_j = M, so no reduction is needed) */
    unsigned wj[OZ2_LIMBS];
    unsigned long long carry = 0;
    for (int l = 0; l < OZ2_LIMBS; ++l) {
      const unsigned long long v = (unsigned long long)mj[l] * inv + carry;
      wj[l] = (unsigned)v;
      carry = v >> 32;
    }
    if (carry || (wj[OZ2_LIMBS - 1] >> 16)) return false; /* W_j < 2^176 */
    for (int l = 0; l < OZ2_LIMBS16; ++l)
      c->W2[j / 2][l] |= ((wj[l / 2] >> (16 * (l % 2))) & 0xffffu) << (16 * (j % 2));
  }
  const int lw = oz2_lw(c->L);
  /* 2^(16 lw) - M in 16-bit limbs: the complement of M's limbs plus one */
  unsigned carry16 = 1;
  for (int l = 0; l < lw; ++l) {
    const unsigned m16 = (M[l / 2] >> (16 * (l % 2))) & 0xffffu;
    const unsigned v = (~m16 & 0xffffu) + carry16;
    c->Mc16[l] = v & 0xffffu;
    carry16 = v >> 16;
  }
  double md = 0.0;
  for (int l = OZ2_LIMBS - 1; l >= 0; --l) md = md * 4294967296.0 + (double)M[l];
  c->rMs = ldexp(1.0 / md, 16 * (lw - 4));
  /* the split's bit counts keep |X| <= (1 - 2^-10) M / 2, so X'/M stays
   * 2^-11 or more from a half-integer, far above the estimate's error
   * (below 2^-18: the four top limb sums, M >= 2^(16 (LW - 1))) */
  c->bound = oz2_bound(nmod);
  return true;
}

/* X from the limb sums acc (X' = sum_l acc[l] 2^(16 l), each below 2^29),
 * rounded and scaled; then the alpha/beta update of C(i, j).  LW: 16-bit
 * limbs of the W_j (M < 2^(16 LW)), at least 4. */
template <int LW>
__device__ __forceinline__ void oz2_finish(unsigned (&acc)[LW], const Oz2Crt &crt, int da,
                                           int db, double *cp, double alpha, double beta) {
  constexpr int NW = (LW + 1) / 2; /* 32-bit words of X */
  /* q = rint(X' / M) from the top four limb sums */
  double xd = (double)acc[LW - 1];
#pragma unroll
  for (int l = LW - 2; l >= LW - 4; --l) xd = __builtin_fma(xd, 65536.0, (double)acc[l]);
  const unsigned q = (unsigned)__builtin_rint(xd * crt.rMs);
  /* X = X' + q (2^(16 LW) - M) modulo 2^(16 LW), carried into 32-bit words */
#pragma unroll
  for (int l = 0; l < LW; ++l) acc[l] += __umul24(q, crt.Mc16[l]);
  unsigned w[NW];
  unsigned carry = 0;
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    const unsigned lo = acc[2 * j] + carry;
    const unsigned hi = (2 * j + 1 < LW ? acc[2 * j + 1] : 0u) + (lo >> 16);
    w[j] = (lo & 0xffffu) | (hi << 16);
    carry = hi >> 16;
  }
  constexpr int TOP = (16 * LW - 1) & 31; /* sign bit within w[NW - 1] */
  const bool neg = (w[NW - 1] >> TOP) & 1u;
  if constexpr (TOP != 31) w[NW - 1] &= (2u << TOP) - 1u;
  if (neg) { /* |X| = 2^(16 LW) - w */
    unsigned c = 1;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      const unsigned nv = ~w[j] + c;
      c = (c && nv == 0u) ? 1u : 0u;
      w[j] = nv;
    }
    if constexpr (TOP != 31) w[NW - 1] &= (2u << TOP) - 1u;
  }
  /* top nonzero word k and the two below it */
  unsigned t0 = 0, t1 = 0, t2 = 0;
  int k = -1;
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    if (w[j] != 0u) {
      k = j;
      t0 = w[j];
      t1 = j >= 1 ? w[j - 1] : 0u;
      t2 = j >= 2 ? w[j - 2] : 0u;
    }
  }
  double scaled;
  if (da == OZ_NONFINITE || db == OZ_NONFINITE) {
    scaled = __builtin_nan("");
  } else if (k < 0) {
    scaled = 0.0;
  } else {
    /* |X| = (N 2^32 + rest) 2^(32 (k - 2) - z) + (words below k - 2) */
    unsigned under = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) under |= j + 3 <= k ? w[j] : 0u;
    const int z = __builtin_clz(t0);
    const unsigned long long top = ((unsigned long long)t0 << 32) | t1;
    unsigned long long N = z ? (top << z) | (t2 >> (32 - z)) : top;
    const unsigned rest = z ? t2 << z : t2;
    N |= (rest | under) != 0u ? 1ull : 0ull; /* sticky: below the rounding bit */
    const double mag = (double)N; /* one rounding, to nearest even */
    scaled = ldexp(mag, 32 * (k - 1) - z - da - db);
    scaled = neg ? -scaled : scaled;
  }
  *cp = __builtin_fma(alpha, scaled, beta * *cp);
}

/* Four consecutive rows of one column per thread: the residue planes are
 * Mp x N bytes with column stride Mp = M rounded up to a multiple of 4, so
 * each plane's bytes of the four come with one aligned 4-byte load (rows
 * M..Mp-1 are padding, skipped); all planes' loads are issued before the
 * first is used, and each W_j limb is read once for the four. */
template <int LW>
__global__ void __launch_bounds__(256)
ozaki2_recon_kernel(const uint8_t *__restrict__ R, const Oz2Crt crt,
                    const int *__restrict__ ea, const int *__restrict__ fb,
                    double *__restrict__ C, int64_t ldc, int M, int N, double alpha,
                    double beta) {
  constexpr int E = 4;
  const int Mp = (M + 3) & ~3;
  const int64_t total = (int64_t)Mp * N;
  const int64_t quads = total / E;
  /* one quad per thread (no grid-stride loop, whose invariant W_j loads the
   * compiler would keep in registers across iterations) */
  const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (u < quads) {
    const int64_t t0 = u * E;
    /* X'_e = sum_g c_g W_g < nmod 2^8 M */
    unsigned acc[E][LW];
#pragma unroll
    for (int q = 0; q < E; ++q)
#pragma unroll
      for (int l = 0; l < LW; ++l) acc[q][l] = 0u;
    /* every plane's bytes are loaded before the first is used, so the
     * loads overlap */
    unsigned c4[OZ2_MAX_MODULI];
#pragma unroll
    for (int g = 0; g < OZ2_MAX_MODULI; ++g) {
      c4[g] = 0;
      if (g < crt.nmod) {
        c4[g] = *reinterpret_cast<const unsigned *>(R + (int64_t)g * total + t0);
      }
    }
    /* two moduli per v_dot2_u32_u16: element q's residues of planes 2p
     * and 2p + 1 as the two 16-bit halves of one word (one byte permute),
     * times the paired limbs; each step adds below 2^25, so the limb sums
     * stay below nmod 2^24 < 2^29 as with one multiply-add per modulus */
#pragma unroll
    for (int p = 0; p < OZ2_MAX_MODULI / 2; ++p) {
      if (2 * p < crt.nmod) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
          /* bytes: [c4[2p] byte q, 0, c4[2p+1] byte q, 0] */
          const unsigned sel = (unsigned)q | 0x0c00u | ((4u + q) << 16) | 0x0c000000u;
          const u16x2_t cp = __builtin_bit_cast(u16x2_t, __builtin_amdgcn_perm(c4[2 * p + 1], c4[2 * p], sel));
#pragma unroll
          for (int l = 0; l < LW; ++l)
            acc[q][l] = __builtin_amdgcn_udot2(cp, __builtin_bit_cast(u16x2_t, crt.W2[p][l]), acc[q][l], false);
        }
      }
    }
    /* (i, j) of the first element; the four share column j */
    const int64_t j = t0 / Mp;
    const int i = (int)(t0 - j * Mp);
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (i + q < M)
        oz2_finish<LW>(acc[q], crt, ea[i + q], fb[j], C + (i + q) + j * ldc, alpha, beta);
  }
}

template <int LW>
hipError_t launch_recon_lw(const uint8_t *R, const Oz2Crt &crt, const int *ea, const int *fb,
                           const gemmx_problem *p, double beta, unsigned blocks,
                           hipStream_t st) {
  hipLaunchKernelGGL((ozaki2_recon_kernel<LW>), dim3(blocks), dim3(256), 0, st, R, crt, ea, fb,
                     (double *)p->c, p->ldc, p->m, p->n, p->alpha, beta);
  return hipGetLastError();
}

/* The reconstruction with as many 16-bit limbs as M needs (M < 2^(L+1)). */
hipError_t launch_recon(const uint8_t *R, const Oz2Crt &crt, const int *ea, const int *fb,
                        const gemmx_problem *p, double beta, unsigned blocks, hipStream_t st) {
  switch ((crt.L + 16) / 16) {
  case 4: return launch_recon_lw<4>(R, crt, ea, fb, p, beta, blocks, st);
  case 5: return launch_recon_lw<5>(R, crt, ea, fb, p, beta, blocks, st);
  case 6: return launch_recon_lw<6>(R, crt, ea, fb, p, beta, blocks, st);
  case 7: return launch_recon_lw<7>(R, crt, ea, fb, p, beta, blocks, st);
  case 8: return launch_recon_lw<8>(R, crt, ea, fb, p, beta, blocks, st);
  case 9: return launch_recon_lw<9>(R, crt, ea, fb, p, beta, blocks, st);
  case 10: return launch_recon_lw<10>(R, crt, ea, fb, p, beta, blocks, st);
  case 11: return launch_recon_lw<11>(R, crt, ea, fb, p, beta, blocks, st);
  default: return hipErrorInvalidValue;
  }
}

struct Oz2Layout {
  size_t ea, fb, mb, sq, as, bs, r8, total;
  int64_t kpad;
};
Oz2Layout oz2_layout(int m, int n, int k, int nmod) {
  Oz2Layout L;
  L.kpad = (int64_t)rup((size_t)k, SL_BK);
  L.ea = 0;
  L.fb = rup((size_t)m * 4, 256);
  L.mb = L.fb + rup((size_t)n * 4, 256);
  L.sq = L.mb + rup((size_t)(m > n ? m : n) * 8, 256);
  L.as = L.sq + rup((size_t)(m > n ? m : n) * 8, 256);
  L.bs = L.as + rup((size_t)nmod * m * L.kpad, 256);
  L.r8 = L.bs + rup((size_t)nmod * n * L.kpad, 256);
  L.total = L.r8 + rup((size_t)nmod * ((m + 3) & ~3) * n, 256); /* Mp x n per modulus */
  return L;
}

} // namespace

extern "C" size_t gemmx_ozaki2_workspace(int m, int n, int k, int moduli) {
  return oz2_layout(m, n, k, moduli).total;
}

extern "C" int gemmx_ozaki2_bits(int moduli, int k) {
  Oz2Crt c;
  if (k < 1 || !oz2_constants(moduli, &c)) return -1;
  return ozaki2_bits_for_sum(c.bound, (unsigned long long)k << 20); /* every t = 1024 */
}

extern "C" hipError_t gemmx_ozaki2_f64(const gemmx_problem *p, int nmod, void *work,
                                       hipStream_t st, hipEvent_t *ev) {
  Oz2Crt crt;
  if (!work || !oz2_constants(nmod, &crt)) return hipErrorInvalidValue;
  const Oz2Layout L = oz2_layout(p->m, p->n, p->k, nmod);
  char *w = (char *)work;
  int *ea = (int *)(w + L.ea), *fb = (int *)(w + L.fb);
  int8_t *as = (int8_t *)(w + L.as), *bs = (int8_t *)(w + L.bs);
  uint8_t *R = (uint8_t *)(w + L.r8);
  unsigned long long *mb = (unsigned long long *)(w + L.mb);
  unsigned long long *sq = (unsigned long long *)(w + L.sq);
  if (ev) (void)hipEventRecord(ev[0], st);
  hipError_t e = ozaki2_prepare((const double *)p->a, p->lda, p->ta != 0, p->m, p->k, L.kpad,
                                nmod, crt.bound, ea, mb, sq, as, st);
  if (e != hipSuccess) return e;
  e = ozaki2_prepare((const double *)p->b, p->ldb, p->tb == 0, p->n, p->k, L.kpad, nmod,
                     crt.bound, fb, mb, sq, bs, st);
  if (e != hipSuccess) return e;
  if (ev) (void)hipEventRecord(ev[1], st);
  /* group g: plane g of A times plane g of B, stored modulo m_g */
  SliceGroups G;
  G.ngroups = 0;
  for (int g = 0; g < nmod; ++g) {
    const int zero = 0;
    G.add(OZ2_MODULI[g], 1, &zero, &zero);
  }
  const int kp = (int)L.kpad;
  const int64_t total = (int64_t)((p->m + 3) & ~3) * p->n; /* residue plane bytes */
  const int64_t rblocks = (total + 4 * 256 - 1) / (4 * 256); /* four rows per thread */
  for (int kb0 = 0; kb0 < kp; kb0 += OZ2_KCHUNK) {
    const int kb1 = kb0 + OZ2_KCHUNK < kp ? kb0 + OZ2_KCHUNK : kp;
    e = slice_products(SP_I8_MOD, as, bs, p->m, p->n, L.kpad, kb0, kb1, G, R, NULL, 0, 0.f,
                       0.f, st);
    if (e != hipSuccess) return e;
    if (ev && kb0 + OZ2_KCHUNK >= kp) (void)hipEventRecord(ev[2], st);
    /* chunks after the first add onto C: beta = 1 leaves C as it is */
    e = launch_recon(R, crt, ea, fb, p, kb0 == 0 ? p->beta : 1.0, (unsigned)rblocks, st);
    if (e != hipSuccess) return e;
  }
  if (ev) (void)hipEventRecord(ev[3], st);
  return hipSuccess;
}

extern "C" int gemmx_ozaki_kchunk(int slices) {
  const int64_t lim = (int64_t)2147483647 / ((int64_t)slices * 127 * 127);
  return (int)(lim - lim % SL_BK);
}

extern "C" size_t gemmx_ozaki_workspace(int m, int n, int k, int slices) {
  return oz_layout(m, n, k, slices).total;
}

extern "C" hipError_t gemmx_ozaki_f64(const gemmx_problem *p, int s, void *work,
                                      hipStream_t st, hipEvent_t *ev) {
  if (s < 3 || s > 8 || !work) return hipErrorInvalidValue;
  const OzLayout L = oz_layout(p->m, p->n, p->k, s);
  char *w = (char *)work;
  int *ea = (int *)(w + L.ea), *fb = (int *)(w + L.fb);
  int8_t *as = (int8_t *)(w + L.as), *bs = (int8_t *)(w + L.bs);
  int *S = (int *)(w + L.sw);
  /* op(A) rows: transa N -> row i at a + i, stride lda; T -> contiguous.
   * op(B) columns: transb N -> contiguous at b + j*ldb; T -> stride ldb. */
  unsigned long long *mb = (unsigned long long *)(w + L.mb);
  if (ev) (void)hipEventRecord(ev[0], st);
  hipError_t e = ozaki_prepare((const double *)p->a, p->lda, p->ta != 0, p->m, p->k,
                               L.kpad, s, ea, mb, as, st);
  if (e != hipSuccess) return e;
  e = ozaki_prepare((const double *)p->b, p->ldb, p->tb == 0, p->n, p->k, L.kpad, s,
                    fb, mb, bs, st);
  if (e != hipSuccess) return e;
  if (ev) (void)hipEventRecord(ev[1], st);
  const SliceGroups groups = ozaki_groups(s);
  const int kc = gemmx_ozaki_kchunk(s); /* bytes, a multiple of SL_BK */
  const int kp = (int)L.kpad;
  const int64_t total = (int64_t)p->m * p->n;
  int64_t rblocks = (total + 255) / 256;
  if (rblocks > 65536) rblocks = 65536;
  for (int kb0 = 0; kb0 < kp; kb0 += kc) {
    const int kb1 = kb0 + kc < kp ? kb0 + kc : kp;
    e = slice_products(SP_I8_SUM, as, bs, p->m, p->n, L.kpad, kb0, kb1, groups, S, NULL, 0,
                       0.f, 0.f, st);
    if (e != hipSuccess) return e;
    if (ev && kb0 + kc >= kp) (void)hipEventRecord(ev[2], st);
    /* chunks after the first add onto C: beta = 1 leaves C as it is */
    hipLaunchKernelGGL(ozaki_recombine_kernel, dim3((unsigned)rblocks), dim3(256), 0, st,
                       S, s, ea, fb, (double *)p->c, p->ldc, p->m, p->n, p->alpha,
                       kb0 == 0 ? p->beta : 1.0);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (ev) (void)hipEventRecord(ev[3], st);
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
  return slice_products(SP_BF16, (const int8_t *)as, (const int8_t *)bs, p->m, p->n, kpad_bytes,
                        0, (int)kpad_bytes, G, NULL, (float *)p->c, p->ldc, (float)p->alpha,
                        (float)p->beta, st);
}
