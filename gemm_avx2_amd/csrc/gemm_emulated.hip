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
 * w_j = M_j^-1 mod m_j and W_j = (w_j M_j) mod M, X = (sum_j c_j W_j) mod M,
 * taken in (-M/2, M/2) (|X| < M/2 by the vectors' bit counts).  The sum
 * (< nmod 2^8 M < 2^185) is formed exactly with 24-bit multiply-adds of the
 * byte c_j into 16-bit limbs of W_j (limb sums stay below 2^29), carried
 * into 192-bit integers (32-bit limbs), reduced modulo M with an FP64
 * quotient and one correction, rounded to 53 bits (ties to even) and scaled
 * by 2^-(D_i + D_j), D = P - e of row i and column j, with one ldexp. */
constexpr int OZ2_LIMBS = 6;    /* 32-bit limbs: M < 2^171 */
constexpr int OZ2_LIMBS16 = 11; /* 16-bit limbs of M_j */

struct Oz2Crt {
  int nmod, L; /* L = floor(log2 M) */
  unsigned W16[OZ2_MAX_MODULI][OZ2_LIMBS16]; /* W_j = (w_j M_j) mod M, 16-bit limbs */
  unsigned M[OZ2_LIMBS], H[OZ2_LIMBS];        /* M and floor(M / 2) */
  double rMd;                                 /* 1 / M, rounded */
};

/* Host: the constants of the first nmod moduli. */
bool oz2_constants(int nmod, Oz2Crt *c) {
  if (nmod < OZ2_MIN_MODULI || nmod > OZ2_MAX_MODULI) return false;
  memset(c, 0, sizeof *c);
  c->nmod = nmod;
  c->M[0] = 1;
  for (int j = 0; j < nmod; ++j) {
    unsigned long long carry = 0;
    for (int l = 0; l < OZ2_LIMBS; ++l) {
      const unsigned long long v = (unsigned long long)c->M[l] * OZ2_MODULI[j] + carry;
      c->M[l] = (unsigned)v;
      carry = v >> 32;
    }
  }
  for (int l = OZ2_LIMBS - 1; l >= 0; --l)
    if (c->M[l]) {
      c->L = 32 * l + 31 - __builtin_clz(c->M[l]);
      break;
    }
  for (int j = 0; j < nmod; ++j) {
    const int m = OZ2_MODULI[j];
    unsigned mj[OZ2_LIMBS];
    unsigned long long rem = 0;
    for (int l = OZ2_LIMBS - 1; l >= 0; --l) { /* M_j = M / m, exact */
      const unsigned long long cur = (rem << 32) | c->M[l];
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
      c->W16[j][l] = (wj[l / 2] >> (16 * (l % 2))) & 0xffffu;
  }
  for (int l = 0; l < OZ2_LIMBS; ++l)
    c->H[l] = (c->M[l] >> 1) | (l + 1 < OZ2_LIMBS ? c->M[l + 1] << 31 : 0u);
  double md = 0.0;
  for (int l = OZ2_LIMBS - 1; l >= 0; --l) md = md * 4294967296.0 + (double)c->M[l];
  c->rMd = 1.0 / md;
  return true;
}

/* Bits [s, s+64) of an NL-limb number. */
template <int NL>
__device__ __forceinline__ unsigned long long bits64(const unsigned (&x)[NL], int s) {
  unsigned long long r = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const int pos = 32 * l - s; /* where limb l lands in the result */
    if (pos >= 64 || pos <= -32) continue;
    const unsigned long long v = x[l];
    r |= pos >= 0 ? v << pos : v >> -pos;
  }
  return r;
}

/* X from the carried 16-bit limb sums, reduced modulo M, rounded and scaled;
 * then the alpha/beta update of C(i, j).  LW: 16-bit limbs of the W_j (M <
 * 2^(16 LW)); the sum X' = sum_j c_j W_j < 22 * 2^8 * M needs NL = LW / 2 + 1
 * 32-bit limbs. */
template <int LW>
__device__ __forceinline__ void oz2_finish(const unsigned (&acc)[LW], const Oz2Crt &crt,
                                           int da, int db, double *cp, double alpha,
                                           double beta) {
  constexpr int NL = LW / 2 + 1;
  unsigned x[NL];
  {
    /* 16-bit digits of X': the limb sums (< 2^29) carried upward; the final
     * carry (< 2^14) is digit LW */
    unsigned d[2 * NL];
    unsigned carry = 0;
#pragma unroll
    for (int l = 0; l < LW; ++l) {
      const unsigned v = acc[l] + carry;
      d[l] = v & 0xffffu;
      carry = v >> 16;
    }
#pragma unroll
    for (int l = LW; l < 2 * NL; ++l) {
      d[l] = carry & 0xffffu;
      carry >>= 16;
    }
#pragma unroll
    for (int l = 0; l < NL; ++l) x[l] = d[2 * l] | (d[2 * l + 1] << 16);
  }
  /* X = X' mod M: subtract q M with q < nmod 2^8 from FP64 (off by at most
   * one), then correct */
  double xd = 0.0;
#pragma unroll
  for (int l = NL - 1; l >= 0; --l) xd = xd * 4294967296.0 + (double)x[l];
  const long long q = (long long)(xd * crt.rMd);
  long long borrow = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const long long d = (long long)x[l] - (long long)((unsigned long long)q * crt.M[l]) + borrow;
    x[l] = (unsigned)d;
    borrow = d >> 32; /* arithmetic: floor(d / 2^32) */
  }
  if (borrow < 0) { /* X < 0: add M */
    unsigned long long c2 = 0;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const unsigned long long v = (unsigned long long)x[l] + crt.M[l] + c2;
      x[l] = (unsigned)v;
      c2 = v >> 32;
    }
  }
  bool ge = true; /* X >= M: subtract M */
#pragma unroll
  for (int l = NL - 1; l >= 0; --l) {
    if (x[l] != crt.M[l]) {
      ge = x[l] > crt.M[l];
      break;
    }
  }
  if (ge) {
    long long b2 = 0;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const long long d = (long long)x[l] - (long long)crt.M[l] + b2;
      x[l] = (unsigned)d;
      b2 = d >> 32;
    }
  }
  /* symmetric range: X > M/2 stands for X - M */
  bool gt = false;
#pragma unroll
  for (int l = NL - 1; l >= 0; --l) {
    if (x[l] != crt.H[l]) {
      gt = x[l] > crt.H[l];
      break;
    }
  }
  const bool neg = gt;
  if (neg) {
    long long b3 = 0;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const long long d = (long long)crt.M[l] - (long long)x[l] + b3;
      x[l] = (unsigned)d;
      b3 = d >> 32;
    }
  }
  /* round |X| to 53 bits, ties to even */
  int top = -1;
#pragma unroll
  for (int l = NL - 1; l >= 0; --l)
    if (top < 0 && x[l]) top = 32 * l + 31 - __builtin_clz(x[l]);
  double mag = 0.0;
  int shift = 0;
  if (top >= 0) {
    unsigned long long mant;
    if (top <= 52) {
      mant = (unsigned long long)x[0] | ((unsigned long long)x[1] << 32);
    } else {
      shift = top - 52;
      mant = bits64(x, shift) & ((1ull << 53) - 1);
      const int hb = shift - 1;
      const bool half = (bits64(x, hb) & 1ull) != 0;
      bool sticky = false;
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        const int lo = 32 * l;
        if (lo >= hb) continue;
        const unsigned mask = hb - lo >= 32 ? 0xffffffffu : ((1u << (hb - lo)) - 1u);
        sticky |= (x[l] & mask) != 0;
      }
      if (half && (sticky || (mant & 1))) {
        ++mant;
        if (mant == (1ull << 53)) {
          mant >>= 1;
          ++shift;
        }
      }
    }
    mag = (double)mant;
  }
  double scaled;
  if (da == OZ_NONFINITE || db == OZ_NONFINITE) {
    scaled = __builtin_nan("");
  } else {
    scaled = ldexp(mag, shift - da - db);
    scaled = neg ? -scaled : scaled;
  }
  *cp = __builtin_fma(alpha, scaled, beta * *cp);
}

/* Four consecutive elements (column-major index t) per thread: each plane's
 * bytes of the four come with one load (VEC: m n a multiple of 4) or four
 * clamped byte loads, all planes' loads issued before the first is used, and
 * each W_j limb is read once for the four. */
template <int LW, bool VEC>
__global__ void __launch_bounds__(256)
ozaki2_recon_kernel(const uint8_t *__restrict__ R, const Oz2Crt crt,
                    const int *__restrict__ ea, const int *__restrict__ fb,
                    double *__restrict__ C, int64_t ldc, int M, int N, double alpha,
                    double beta) {
  constexpr int E = 4;
  const int64_t total = (int64_t)M * N;
  const int64_t quads = (total + E - 1) / E;
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
        const uint8_t *rp = R + (int64_t)g * total;
        if constexpr (VEC) {
          c4[g] = *reinterpret_cast<const unsigned *>(rp + t0);
        } else { /* clamped reads, bytes past the end taken as 0 */
#pragma unroll
          for (int q = 0; q < E; ++q) {
            const unsigned b = rp[t0 + q < total ? t0 + q : total - 1];
            c4[g] |= (t0 + q < total ? b : 0u) << (8 * q);
          }
        }
      }
    }
#pragma unroll
    for (int g = 0; g < OZ2_MAX_MODULI; ++g) {
      if (g < crt.nmod) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
          const unsigned cg = (c4[g] >> (8 * q)) & 0xffu;
#pragma unroll
          for (int l = 0; l < LW; ++l) acc[q][l] += __umul24(cg, crt.W16[g][l]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int64_t t = t0 + q;
      if (t < total) {
        const int64_t j = t / M, i = t - j * M;
        oz2_finish<LW>(acc[q], crt, ea[i], fb[j], C + i + j * ldc, alpha, beta);
      }
    }
  }
}

template <int LW>
hipError_t launch_recon_lw(const uint8_t *R, const Oz2Crt &crt, const int *ea, const int *fb,
                           const gemmx_problem *p, double beta, unsigned blocks,
                           hipStream_t st) {
  if (((int64_t)p->m * p->n) % 4 == 0)
    hipLaunchKernelGGL((ozaki2_recon_kernel<LW, true>), dim3(blocks), dim3(256), 0, st, R, crt,
                       ea, fb, (double *)p->c, p->ldc, p->m, p->n, p->alpha, beta);
  else
    hipLaunchKernelGGL((ozaki2_recon_kernel<LW, false>), dim3(blocks), dim3(256), 0, st, R, crt,
                       ea, fb, (double *)p->c, p->ldc, p->m, p->n, p->alpha, beta);
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
  L.total = L.r8 + rup((size_t)nmod * m * n, 256);
  return L;
}

} // namespace

extern "C" size_t gemmx_ozaki2_workspace(int m, int n, int k, int moduli) {
  return oz2_layout(m, n, k, moduli).total;
}

extern "C" int gemmx_ozaki2_bits(int moduli, int k) {
  Oz2Crt c;
  if (k < 1 || !oz2_constants(moduli, &c)) return -1;
  return ozaki2_bits_for_sum(c.L, (unsigned long long)k << 20); /* every t = 1024 */
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
                                nmod, crt.L, ea, mb, sq, as, st);
  if (e != hipSuccess) return e;
  e = ozaki2_prepare((const double *)p->b, p->ldb, p->tb == 0, p->n, p->k, L.kpad, nmod,
                     crt.L, fb, mb, sq, bs, st);
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
  const int64_t total = (int64_t)p->m * p->n;
  const int64_t rblocks = (total + 4 * 256 - 1) / (4 * 256); /* four elements per thread */
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
