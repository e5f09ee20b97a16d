/*
 * gemmx_tile.h -- one BM x BN tile of C = alpha*op(A)*op(B) + beta*C on the
 * FP64 / FP32 16x16x4 MFMA, shared by the single-GEMM kernels
 * (gemm_native.hip) and the grouped batch kernel (gemm_batch.hip).
 *
 * Replaces the reference's micro-kernels and their edge variants
 * (wjc404/GEMM_AVX2 src/kernel/gemm_kernel.S:117-238,
 * src/edge_kernel/gemm_edge_kernel.c:39-117): K is walked in BK slices, each
 * slice of op(A) and op(B) copied global->registers->LDS (double-buffered,
 * the next slice's loads in flight during the current slice's MFMAs).  The
 * LDS layout follows the global layout, so the global->LDS copy is a straight
 * copy of contiguous runs and the transpose is absorbed by how the MFMA
 * operands are read:
 *   op(A), transa N: LDS [k][m]     transa T: LDS [m][k]
 *   op(B), transb N: LDS [n][k]     transb T: LDS [k][n]
 * In MFMA step s of a slice, lane group g = lane/16 supplies k = g*BK/4+s,
 * identically for A and B; where LDS is K-contiguous, KV consecutive steps
 * are read with one 16-byte LDS read.
 *
 * FAST (interior tile, 16-byte loadable operands): a branch-free loop over
 * the K slices with an explicit issue order -- the next slice's global
 * loads, each step's fragment reads spread over the previous step's MFMAs,
 * the next slice's LDS writes spread over the last step.  A K remainder is
 * the loop's last slice, loaded as the BK positions ending at K with the
 * positions the previous slice already summed zeroed (no peeled iteration:
 * its extra registers made the 128x128 FP64 tile spill).  FAST with CLAMP: the same loop for
 * an edge tile, its loads past M or N clamped back into the operand
 * (TileCopy::load_clamp).  FAST with NARROW: the same loop for operands that
 * do not allow 16-byte loads, one load per element (TileCopy::load_narrow).
 * !FAST: every slice predicated.
 * The epilogue writes C = fma(alpha, acc, beta*C) for rows < M, columns < N
 * only (the reference's masked stores, src/kernel/gemm_kernel_unroll6.S:
 * 145-155); alpha moves from the reference's packed B (src/gemm_copy.c:65)
 * into this epilogue.  PARTIAL (split-K): the raw accumulators of one K
 * range are stored instead, and a reduction kernel applies alpha and beta.
 */
#ifndef GEMMX_TILE_H
#define GEMMX_TILE_H

#include <type_traits>

#include "gemmx_device.h"

namespace gemmx {

/* LDS elements the tile needs (both stages of A and B). */
template <typename T, bool TA, bool TB, int BM, int BN, int BK>
constexpr int tile_lds_elems() {
  return 2 * ((TA ? BM : BK) * ((TA ? BK : BM) + Vec16<T>::N) +
              (TB ? BK : BN) * ((TB ? BN : BK) + Vec16<T>::N));
}
/* LDS elements of a tile with KG K groups and WM x WN waves per group: the
 * staging buffers, or, if that is more, the accumulators groups 1..KG-1 hand
 * to group 0 followed by the epilogue regions of group 0's waves
 * (store_c_tile_lds). */
template <int BM, int BN, int WM, int WN, int KG>
constexpr int tile_lds_epi() {
  return (KG - 1) * BM * BN + WM * WN * epi_region_elems<BN / WN / 16>();
}
template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WM, int WN, int KG>
constexpr int tile_lds_need() {
  constexpr int st = tile_lds_elems<T, TA, TB, BM, BN, BK>();
  constexpr int ep = tile_lds_epi<BM, BN, WM, WN, KG>();
  return st > ep ? st : ep;
}
template <typename T, int BM, int BN, int BK, int WM, int WN>
constexpr int tile_lds_max() {
  constexpr int a = tile_lds_need<T, false, false, BM, BN, BK, WM, WN, 1>();
  constexpr int b = tile_lds_need<T, false, true, BM, BN, BK, WM, WN, 1>();
  constexpr int c = tile_lds_need<T, true, false, BM, BN, BK, WM, WN, 1>();
  constexpr int d = tile_lds_need<T, true, true, BM, BN, BK, WM, WN, 1>();
  return a > b ? (a > c ? (a > d ? a : d) : (c > d ? c : d)) : (b > c ? (b > d ? b : d) : (c > d ? c : d));
}

/* Work the caller overlaps with the first slice's global loads (default:
 * none); the grouped batch takes its next tile's ticket there. */
struct NoHook {
  __device__ __forceinline__ void operator()() const {}
};

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WM, int WN, int KVMAX,
          bool FAST, bool CLAMP = false, bool NARROW = false, bool PARTIAL = false,
          int DEPTH = 1, int KG = 1, bool EPI_LDS = true, typename Hook = NoHook>
__device__ __forceinline__ void gemm_tile(const T *__restrict__ A, const T *__restrict__ B,
                                          T *__restrict__ C, int M, int N, int K,
                                          int64_t lda, int64_t ldb, int64_t ldc, T alpha,
                                          T beta, int m0, int n0, int vecA, int vecB,
                                          T *lds, Hook after_first_loads = Hook()) {
  constexpr int NT = 64 * WM * WN * KG;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  constexpr int KS = BK / 4;            /* MFMA steps per slice */
  constexpr int KSG = KS / KG;          /* MFMA steps per slice of one K group */
  constexpr int VE = Vec16<T>::N;       /* elements per 16 bytes */
  /* MFMA steps read together where LDS is K-contiguous (more registers, fewer reads) */
  constexpr int KV = (VE < KSG ? VE : KSG) < KVMAX ? (VE < KSG ? VE : KSG) : KVMAX;
  constexpr int A_NL = TA ? BM : BK, A_LEN = TA ? BK : BM, A_LS = A_LEN + VE;
  constexpr int B_NL = TB ? BK : BN, B_LEN = TB ? BN : BK, B_LS = B_LEN + VE;
  static_assert(WTM % 16 == 0 && WTN % 16 == 0 && BK % 4 == 0 && KS % KG == 0 && KSG % KV == 0,
                "tile shape");
  typedef typename Mfma16<T>::acc_t acc_t;
  typedef T kvec_t __attribute__((ext_vector_type(KV)));

  // Stage pointers by arithmetic on `lds` (an array of pointers indexed at
  // run time would live in scratch and turn every LDS access into a flat one).
  auto sA = [&](int st) { return lds + st * (A_NL * A_LS); };
  auto sB = [&](int st) { return lds + 2 * A_NL * A_LS + st * (B_NL * B_LS); };

  /* KG > 1: KG groups of WM x WN waves share the tile and its LDS slices,
   * group q multiplying MFMA steps [q KSG, (q+1) KSG) of every slice; their
   * accumulators are summed in group order after the loop */
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wq = wave / (WM * WN), wt = wave % (WM * WN);
  const int wm = wt % WM, wn = wt / WM;
  const int li = lane & 15, kg = lane >> 4;

  TileCopy<T, A_NL, A_LEN, NT> ca;
  TileCopy<T, B_NL, B_LEN, NT> cb;
  auto abase = [&](int k0) { return TA ? A + k0 + m0 * lda : A + m0 + k0 * lda; };
  auto bbase = [&](int k0) { return TB ? B + n0 + k0 * ldb : B + k0 + n0 * ldb; };
  auto load_safe = [&](int k0) {
    if (!TA) ca.load(abase(k0), lda, K - k0, M - m0, vecA, tid);
    else     ca.load(abase(k0), lda, M - m0, K - k0, vecA, tid);
    if (!TB) cb.load(bbase(k0), ldb, N - n0, K - k0, vecB, tid);
    else     cb.load(bbase(k0), ldb, K - k0, N - n0, vecB, tid);
  };
  auto load_into = [&](auto &xa, auto &xb, int k0) {
    if (NARROW) {
      if (!TA) xa.template load_narrow<CLAMP>(abase(k0), lda, A_NL, M - m0, tid);
      else     xa.template load_narrow<CLAMP>(abase(k0), lda, M - m0, A_LEN, tid);
      if (!TB) xb.template load_narrow<CLAMP>(bbase(k0), ldb, N - n0, B_LEN, tid);
      else     xb.template load_narrow<CLAMP>(bbase(k0), ldb, B_NL, N - n0, tid);
    } else if (CLAMP) {
      if (!TA) xa.load_clamp(abase(k0), lda, A_NL, M - m0, tid);
      else     xa.load_clamp(abase(k0), lda, M - m0, A_LEN, tid);
      if (!TB) xb.load_clamp(bbase(k0), ldb, N - n0, B_LEN, tid);
      else     xb.load_clamp(bbase(k0), ldb, B_NL, N - n0, tid);
    } else {
      xa.load_vec(abase(k0), lda, tid);
      xb.load_vec(bbase(k0), ldb, tid);
    }
  };
  auto load_fast = [&](int k0) { load_into(ca, cb, k0); };
  auto store = [&](int buf) {
    ca.template store<A_LS>(sA(buf), tid);
    cb.template store<B_LS>(sB(buf), tid);
  };
  /* the stored slice starts `skip` positions before the first K position
   * not yet summed: those positions are written as zeros */
  auto store_skip_from = [&](const auto &xa, const auto &xb, int buf, int skip) {
    xa.template store_skip<A_LS, !TA>(sA(buf), tid, skip);
    xb.template store_skip<B_LS, TB>(sB(buf), tid, skip);
  };
  auto store_skip = [&](int buf, int skip) { store_skip_from(ca, cb, buf, skip); };

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = T(0);

  /* KV consecutive MFMA steps of one slice from LDS buffer `buf`. */
  auto compute = [&](int buf) {
    const T *as = sA(buf);
    const T *bs = sB(buf);
#pragma unroll
    for (int s0 = 0; s0 < KSG; s0 += KV) {
      const int kx = kg * KS + wq * KSG + s0;
      T af[KV][FM], bf[KV][FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int row = wm * WTM + i * 16 + li;
        if (TA) {
          const kvec_t v = *reinterpret_cast<const kvec_t *>(as + row * A_LS + kx);
#pragma unroll
          for (int v2 = 0; v2 < KV; ++v2) af[v2][i] = v[v2];
        } else {
#pragma unroll
          for (int v2 = 0; v2 < KV; ++v2) af[v2][i] = as[(kx + v2) * A_LS + row];
        }
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = wn * WTN + j * 16 + li;
        if (!TB) {
          const kvec_t v = *reinterpret_cast<const kvec_t *>(bs + col * B_LS + kx);
#pragma unroll
          for (int v2 = 0; v2 < KV; ++v2) bf[v2][j] = v[v2];
        } else {
#pragma unroll
          for (int v2 = 0; v2 < KV; ++v2) bf[v2][j] = bs[(kx + v2) * B_LS + col];
        }
      }
#pragma unroll
      for (int v2 = 0; v2 < KV; ++v2)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = Mfma16<T>::mma(af[v2][i], bf[v2][j], acc[i][j]);
    }
  };

  const int nfull = K / BK, nk = (K + BK - 1) / BK;
  /* Issue order of one branch-free iteration: the next slice's global
   * loads first, each MFMA step's fragment reads spread over the previous
   * step's MFMAs, the next slice's LDS writes spread over the last step. */
  auto issue_order = [&]() {
    constexpr int NLD = TileCopy<T, A_NL, A_LEN, NT>::PER + TileCopy<T, B_NL, B_LEN, NT>::PER;
    constexpr int NGL = NLD * (NARROW ? VE : 1); /* global loads per slice */
    constexpr int NR = (TA ? FM : FM * KV) + (TB ? FN * KV : FN); /* LDS reads per KV steps */
    constexpr int NMF = KV * FM * FN;                             /* MFMAs per KV steps */
    __builtin_amdgcn_sched_group_barrier(0x020, NGL, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, NR, 0);
#pragma unroll
    for (int s = KV; s < KSG; s += KV) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, NMF / NR, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, NMF - (NMF / NR) * NR, 0);
    }
#pragma unroll
    for (int w = 0; w < NLD; ++w) {
      __builtin_amdgcn_sched_group_barrier(0x008, NMF / NLD, 0);
      __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, NMF - (NMF / NLD) * NLD, 0);
  };
  if (FAST && DEPTH == 2 && nfull >= 1) {
    /* Two slices in flight: slice s + 2's global loads are issued while
     * slice s is multiplied, into the register stage that slice s vacated,
     * so a load has two iterations to land.  The loads are unconditional
     * (past the end they re-read the last slice): loads behind a branch make
     * the compiler wait for every outstanding load where the paths join. */
    TileCopy<T, A_NL, A_LEN, NT> ca1;
    TileCopy<T, B_NL, B_LEN, NT> cb1;
    auto slice_k = [&](int sl) {
      const int k0 = sl * BK;
      return k0 + BK <= K ? k0 : K - BK;
    };
    /* slice sl (LDS buffer P) multiplied, slice sl + 1 (register stage !P)
     * written to LDS buffer !P, slice sl + 2 loaded into stage P; LAST: the
     * stored slice may be the shifted last one */
    auto step = [&](auto par, auto last, int sl) {
      constexpr int P = decltype(par)::value;
      constexpr bool LAST = decltype(last)::value;
      auto &la = P ? ca1 : ca;
      auto &lb = P ? cb1 : cb;
      const auto &xa = P ? ca : ca1;
      const auto &xb = P ? cb : cb1;
      load_into(la, lb, slice_k(sl + 2 < nk ? sl + 2 : nk - 1));
      compute(P);
      if (LAST) {
        store_skip_from(xa, xb, P ^ 1, (sl + 1) * BK - slice_k(sl + 1));
      } else {
        xa.template store<A_LS>(sA(P ^ 1), tid);
        xb.template store<B_LS>(sB(P ^ 1), tid);
      }
      issue_order();
      __syncthreads();
    };
    typedef std::integral_constant<int, 0> I0;
    typedef std::integral_constant<int, 1> I1;
    typedef std::integral_constant<bool, false> F;
    typedef std::integral_constant<bool, true> L;
    load_fast(0);
    load_into(ca1, cb1, slice_k(nk > 1 ? 1 : 0));
    after_first_loads();
    store(0);
    __syncthreads();
    int sl = 0;
    /* every slice stored in the loop is a full one: only the last can be
     * shifted, and it is stored by one of the steps after the loop */
    for (; sl + 3 < nk; sl += 2) {
      step(I0(), F(), sl);
      step(I1(), F(), sl + 1);
    }
    if (sl + 1 < nk) {
      step(I0(), L(), sl);
      ++sl;
      if (sl + 1 < nk) {
        step(I1(), L(), sl);
        ++sl;
      }
    }
    compute(sl & 1);
  } else if (FAST && nfull >= 1) {
    load_fast(0);
    after_first_loads();
    store(0);
    __syncthreads();
    int kt = 0;
    /* the slices the loop stores (kt + 1 <= nk - 2) are full ones, stored
     * plainly: a select on loaded values in the loop lets the compiler
     * consume the loads early and wait for them */
    for (; kt + 2 < nk; ++kt) {
      load_fast((kt + 1) * BK);
      compute(kt & 1);
      store((kt + 1) & 1);
      issue_order();
      __syncthreads();
    }
    if (kt + 1 < nk) {
      /* the last slice (a K remainder) is loaded as the BK positions
       * ending at K, its first `skip` positions zeroed */
      const int k0 = (kt + 1) * BK;
      const int ks = k0 + BK <= K ? k0 : K - BK;
      load_fast(ks);
      compute(kt & 1);
      store_skip((kt + 1) & 1, k0 - ks);
      issue_order();
      __syncthreads();
      ++kt;
    }
    compute(kt & 1);
  } else {
    load_safe(0);
    after_first_loads();
    store(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      if (kt + 1 < nk) load_safe((kt + 1) * BK);
      compute(kt & 1);
      if (kt + 1 < nk) store((kt + 1) & 1);
      __syncthreads();
    }
  }

  /* K groups: groups 1..KG-1 leave their accumulators in LDS (the staging
   * buffers, free once every wave is past its last slice; the caller sizes
   * `lds` for (KG-1)*BM*BN elements too) in fragment order, and group 0
   * adds them in group order and goes on alone. */
  if constexpr (KG > 1) {
    __syncthreads();
    if (wq > 0) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            lds[((((wq - 1) * WM * WN + wt) * FM + i) * FN + j) * 256 + r * 64 + lane] =
                acc[i][j][r];
    }
    __syncthreads();
    if (wq > 0) return;
#pragma unroll
    for (int q = 1; q < KG; ++q)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc[i][j][r] += lds[((((q - 1) * WM * WN + wt) * FM + i) * FN + j) * 256 + r * 64 + lane];
  }

  /* PARTIAL (one K range of a split-K product): the raw accumulators go to
   * the BM*BN block C in the fragment order of the registers -- element
   * (((wave*FM + i)*FN + j)*4 + r)*64 + lane -- so every store is 64
   * consecutive elements; splitk_partial_coords maps them back. */
  if (PARTIAL) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          C[(((wt * FM + i) * FN + j) * 4 + r) * 64 + lane] = acc[i][j][r];
    return;
  }
  /* Epilogue: rows m..ldc-1 and columns >= n are never touched, matching the
   * reference's masked stores (src/kernel/gemm_kernel_unroll6.S:145-155).
   * EPI_LDS: C is read and written through the wave's LDS region
   * (store_c_tile_lds), after the last slice's reads of every wave (one K
   * group: a barrier; with K groups the regions lie past the handed-over
   * accumulators, and group 0's waves are past their own reads). */
  if constexpr (EPI_LDS) {
    if constexpr (KG == 1) __syncthreads();
    T *reg = lds + (KG - 1) * BM * BN + wt * epi_region_elems<FN>();
    store_c_tile_lds<T, FM, FN>(acc, C, M, N, ldc, alpha, beta, m0 + wm * WTM, n0 + wn * WTN,
                                li, kg, reg, lane);
  } else {
    store_c_tile<T, FM, FN>(acc, C, M, N, ldc, alpha, beta, m0 + wm * WTM, n0 + wn * WTN, li,
                            kg);
  }
}

/* Row and column, inside its BM x BN tile, of element e of a partial tile
 * stored by gemm_tile<..., PARTIAL> (the inverse of its fragment order). */
template <typename T, int BM, int BN, int WM, int WN>
__device__ __forceinline__ void splitk_partial_coords(int e, int *row, int *col) {
  constexpr int WTM = BM / WM, WTN = BN / WN, FM = WTM / 16, FN = WTN / 16;
  const int lane = e & 63, r = (e >> 6) & 3, f = e >> 8;
  const int j = f % FN, i = (f / FN) % FM, wave = f / (FN * FM);
  const int wm = wave % WM, wn = wave / WM;
  *row = wm * WTM + i * 16 + Mfma16<T>::row(lane >> 4, r);
  *col = wn * WTN + j * 16 + (lane & 15);
}

} // namespace gemmx
#endif
