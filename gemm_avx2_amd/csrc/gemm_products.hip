/*
 * gemm_products.hip -- slice-pair products on the INT8 and BF16 matrix cores
 * for the emulated forms (no reference counterpart; see gemm_emulated.hip).
 *
 * One workgroup computes one BM x BN tile of one group: it walks the group's
 * slice pairs and K stages, staging K-contiguous slice tiles through LDS
 * (double-buffered, the next stage's global loads in flight during the
 * current stage's MFMAs), and multiplies them with
 *   INT8: v_mfma_i32_16x16x64_i8   -> exact INT32 sums (Ozaki-I weight sums,
 *         or Ozaki-II residue products stored modulo their modulus)
 *   BF16: v_mfma_f32_16x16x32_bf16 -> FP32 sums
 * Both consume 64 K-bytes per lane-group step: lane l holds row l%16 and
 * bytes [16*(l/16), 16*(l/16)+16) of each 64-byte K chunk, for A and B
 * alike, so the hardware's order of those bytes cancels out.  D: lane l,
 * element r is row 4*(l/16)+r, column l%16.
 * Grid: groups x tiles, heaviest group first (the host orders the groups).
 */
#include <stdlib.h>

#include <atomic>

#include "gemmx_device.h"
#include "gemmx_slices.h"

namespace gemmx {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

template <bool BF16> struct SliceMfma;
template <> struct SliceMfma<false> {
  typedef i4_t acc_t;
  static __device__ __forceinline__ acc_t mma(i4_t a, i4_t b, acc_t c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
  }
};
template <> struct SliceMfma<true> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t mma(i4_t a, i4_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                   __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
  }
};

/* Global->LDS copy of an NL x LEN-int slice tile (NL lines, line stride ld
 * ints), 16 bytes per chunk.  Slices are zero-padded in K, so only lines can
 * fall outside the matrix; those are clamped to the last valid line (their
 * products land in rows/columns that are never stored).  No branches: every
 * chunk is one 16-byte load whatever the edge, so the loads of a stage stay
 * in flight together. */
template <int NL, int LEN, int NT>
struct SliceTileCopy {
  static constexpr int CPL = LEN / 4;
  static constexpr int PER = NL * CPL / NT;
  static_assert((NL * CPL) % NT == 0, "chunks must divide evenly over the threads");
  i4_t r[PER];

  __device__ __forceinline__ void load(const int *__restrict__ g, int64_t ld, int nl_valid,
                                       int tid) {
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const int ch = tid + c * NT, line = ch / CPL, off = (ch % CPL) * 4;
      const int ln = line < nl_valid ? line : nl_valid - 1;
      r[c] = *reinterpret_cast<const i4_t *>(g + ln * ld + off);
    }
  }
  template <int LS>
  __device__ __forceinline__ void store(int *s, int tid) const {
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const int ch = tid + c * NT, line = ch / CPL, off = (ch % CPL) * 4;
      *reinterpret_cast<i4_t *>(s + line * LS + off) = r[c];
    }
  }
};

/* Epilogue of one wave's FM x FN fragment grid (rows from r0, columns from
 * c0): BF16 -> C = fma(alpha, acc, beta*C) (store_c_tile); INT8 sums -> S slot gout[g];
 * INT8 residue products -> the sum modulo gout[g] as a byte in plane g. */
template <int KIND, int FM, int FN, typename acc_t>
__device__ __forceinline__ void store_tile(const acc_t (&acc)[FM][FN], int g, const SliceGroups &G,
                                           void *__restrict__ S, float *__restrict__ C,
                                           int64_t ldc, float alpha, float beta, int M, int N,
                                           int r0, int c0, int li, int kg) {
  constexpr bool BF16 = KIND == SP_BF16;
  if constexpr (KIND == SP_I8_MOD) {
    /* Residue of each INT32 sum v: v = hi 2^16 + lo (hi signed, lo < 2^16),
     * t = hi (2^16 mod m) + lo = v (mod m) with |t| < 2^23, exact in FP32;
     * q = floor(t / m) from one FP32 reciprocal is off by at most one (the
     * relative error of t * (1/m) is below 2^-22 and |t / m| < 2^16), and the
     * remainder is corrected into [0, m).  All integer / FP32 work: the
     * epilogue runs while the CU's matrix cores wait for the next tile. */
    const int mod = G.gout[g];
    const int c16 = 65536 % mod;
    const float rcp = 1.0f / (float)mod;
    /* residue planes are Mp x N bytes, column stride Mp = M rounded up to a
     * multiple of 4, so each lane's four rows are one aligned 4-byte store;
     * rows M..Mp-1 are padding that nothing reads */
    const int Mp = (M + 3) & ~3;
    uint8_t *const sbase = (uint8_t *)S + (int64_t)g * Mp * N + r0 + kg * 4 + (int64_t)(c0 + li) * Mp;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      if (c0 + j * 16 + li >= N) continue;
      uint8_t *const sj = sbase + (int64_t)(j * 16) * Mp;
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int row = r0 + i * 16 + kg * 4;
        unsigned bytes = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int v = acc[i][j][r];
          const int t = (v >> 16) * c16 + (v & 0xffff);
          const int q = (int)floorf((float)t * rcp);
          int rem = t - __mul24(q, mod);
          rem += rem < 0 ? mod : 0;
          rem -= rem >= mod ? mod : 0;
          bytes |= (unsigned)rem << (8 * r);
        }
        if (row < Mp) *reinterpret_cast<unsigned *>(sj + i * 16) = bytes;
      }
    }
    return;
  }
  if constexpr (BF16) {
    /* the BF16 MFMA's result rows are 4*(lane/16) + r, as FP32's */
    store_c_tile<float, FM, FN>(acc, C, M, N, ldc, alpha, beta, r0, c0, li, kg);
    return;
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = c0 + j * 16 + li;
      const int row = r0 + i * 16 + kg * 4;
      if (col >= N) continue;
      int *sp = (int *)S + (int64_t)G.gout[g] * M * N + row + (int64_t)col * M;
      if (row + 3 < M && (M & 3) == 0) {
        *reinterpret_cast<i4_t *>(sp) = *reinterpret_cast<const i4_t *>(&acc[i][j]);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row + r < M) sp[r] = (int)acc[i][j][r];
      }
    }
}

template <int KIND, int BM, int BN, int WM, int WN, int BKB, int PAD>
__global__ void __launch_bounds__(64 * WM * WN)
slice_products_kernel(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs,
                      int M, int N, int64_t kpad_bytes, int kb_begin, int kb_end,
                      SliceGroups G, int tiles_m, int tiles_n, void *__restrict__ S,
                      float *__restrict__ C, int64_t ldc, float alpha, float beta) {
  constexpr bool BF16 = KIND == SP_BF16;
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  constexpr int LEN = BKB / 4; /* ints per LDS line: one K stage of BKB bytes */
  constexpr int LS = LEN + PAD; /* PAD ints against bank conflicts */
  constexpr int KSTEPS = BKB / 64;
  typedef SliceMfma<BF16> MM;
  typedef typename MM::acc_t acc_t;
  extern __shared__ __align__(16) int smem[];
  /* LDS: A stage 0 | A stage 1 | B stage 0 | B stage 1 */
  auto sA = [&](int st) { return smem + st * (BM * LS); };
  auto sB = [&](int st) { return smem + 2 * BM * LS + st * (BN * LS); };

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, kg = lane >> 4;
  const int ntiles = tiles_m * tiles_n;
  const int g = blockIdx.x / ntiles;
  int tm, tn;
  tile_coords(blockIdx.x - g * ntiles, ntiles, tiles_m, tiles_n, &tm, &tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int64_t ld_i = kpad_bytes / 4;
  /* Scalar (uniform) group data; the pair of an iteration is unpacked with
   * shifts, so the loop issues no memory access besides the tile loads. */
  const int pc = G.gcnt[g];
  const unsigned tpack = G.tpack[g], upack = G.upack[g];

  SliceTileCopy<BM, LEN, NT> ca;
  SliceTileCopy<BN, LEN, NT> cb;
  const int kt_begin = kb_begin / BKB;
  const int nkt = (kb_end - kb_begin) / BKB;
  const int iters = pc * nkt;
  const int8_t *a_tile = As + (int64_t)m0 * kpad_bytes;
  const int8_t *b_tile = Bs + (int64_t)n0 * kpad_bytes;
  auto load_iter = [&](int q) {
    const int pl = q / nkt, kt = kt_begin + q - pl * nkt;
    /* residue products: group g multiplies plane g by plane g */
    const int t = KIND == SP_I8_MOD ? g : (tpack >> (4 * pl)) & 15;
    const int u = KIND == SP_I8_MOD ? g : (upack >> (4 * pl)) & 15;
    const int *ga = reinterpret_cast<const int *>(a_tile + (int64_t)t * M * kpad_bytes) + kt * LEN;
    const int *gb = reinterpret_cast<const int *>(b_tile + (int64_t)u * N * kpad_bytes) + kt * LEN;
    ca.load(ga, ld_i, M - m0, tid);
    cb.load(gb, ld_i, N - n0, tid);
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
    /* Branch-free body: the last iteration reloads its own stage into the
     * other buffer, which nobody reads any more. */
    load_iter(q + 1 < iters ? q + 1 : q);
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
    ca.template store<LS>(sA(cur ^ 1), tid);
    cb.template store<LS>(sB(cur ^ 1), tid);
    {
      /* Issue order: the stage's global loads, the first K step's fragment
       * reads, then every later K step's fragment reads spread over the
       * previous step's MFMAs, and the LDS writes of the next stage spread
       * over the last step's MFMAs. */
      constexpr int NR = FM + FN, NMF = FM * FN;
      constexpr int NW = SliceTileCopy<BM, LEN, NT>::PER + SliceTileCopy<BN, LEN, NT>::PER;
      __builtin_amdgcn_sched_group_barrier(0x020, NW, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, NR, 0);
#pragma unroll
      for (int ks = 1; ks < KSTEPS; ++ks) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          __builtin_amdgcn_sched_group_barrier(0x008, NMF / NR, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, NMF - (NMF / NR) * NR, 0);
      }
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        __builtin_amdgcn_sched_group_barrier(0x008, NMF / NW, 0);
        __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, NMF - (NMF / NW) * NW, 0);
    }
    __syncthreads();
  }

  store_tile<KIND, FM, FN>(acc, g, G, S, C, ldc, alpha, beta, M, N, m0 + wm * WTM,
                           n0 + wn * WTN, li, kg);
}

/* Direct-to-LDS form with 128-byte K stages in two LDS buffers (64 KB
 * each): one barrier per two MFMA k-steps.  A 128-byte row holds its eight
 * 16-byte chunks with chunk j at position j ^ ((row >> 1) & 7).  Iteration q:
 * the fragments of (stage q, k-step 0) are in registers; those of k-step 1
 * are read during k-step 0's MFMAs; after this wave's loads of stage q + 1
 * and its reads are complete, a barrier frees stage q's buffer for stage
 * q + 2, whose loads are issued; (stage q + 1, k-step 0) is read during
 * k-step 1's MFMAs. */
__device__ __forceinline__ int dl2_swz(int r) { return (r >> 1) & 7; }

template <int KIND>
__global__ void __launch_bounds__(512)
slice_products_dl2_kernel(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs,
                          int M, int N, int64_t kpad_bytes, int kb_begin, int kb_end,
                          SliceGroups G, int tiles_m, int tiles_n, void *__restrict__ S,
                          float *__restrict__ C, int64_t ldc, float alpha, float beta) {
  constexpr int BM = 256, BN = 256, BKB = 128, WM = 2, WN = 4, NW = WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  constexpr int STAGE = (BM + BN) * BKB;
  constexpr int LPA = BM * BKB / 1024 / NW, LPB = BN * BKB / 1024 / NW;
  constexpr bool BF16 = KIND == SP_BF16;
  typedef SliceMfma<BF16> MM;
  typedef typename MM::acc_t acc_t;
  extern __shared__ __align__(16) int smem[];
  char *const lds = reinterpret_cast<char *>(smem);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int li = lane & 15, kg = lane >> 4;
  const int ntiles = tiles_m * tiles_n;
  const int g = blockIdx.x / ntiles;
  int tm, tn;
  tile_coords(blockIdx.x - g * ntiles, ntiles, tiles_m, tiles_n, &tm, &tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int pc = G.gcnt[g];
  const unsigned tpack = G.tpack[g], upack = G.upack[g];
  const int kt_begin = kb_begin / BKB;
  const int nkt = (kb_end - kb_begin) / BKB;
  const int iters = pc * nkt;

  /* loads: LDS chunk c = (LP wave + u) 64 + lane, row c / 8, position c % 8
   * holding chunk (c % 8) ^ dl2_swz(row) (rows past the matrix clamped) */
  int64_t a_off[LPA], b_off[LPB];
#pragma unroll
  for (int u = 0; u < LPA; ++u) {
    const int c = (LPA * wave + u) * 64 + lane, r = c >> 3, j = (c & 7) ^ dl2_swz(r);
    a_off[u] = (r < M - m0 ? r : M - m0 - 1) * kpad_bytes + 16 * j;
  }
#pragma unroll
  for (int u = 0; u < LPB; ++u) {
    const int c = (LPB * wave + u) * 64 + lane, r = c >> 3, j = (c & 7) ^ dl2_swz(r);
    b_off[u] = (r < N - n0 ? r : N - n0 - 1) * kpad_bytes + 16 * j;
  }
  const int8_t *a_tile = As + (int64_t)m0 * kpad_bytes;
  const int8_t *b_tile = Bs + (int64_t)n0 * kpad_bytes;
  auto issue = [&](int q, int buf) {
    const int pl = q / nkt, kt = kt_begin + q - pl * nkt;
    const int t = KIND == SP_I8_MOD ? g : (tpack >> (4 * pl)) & 15;
    const int uu = KIND == SP_I8_MOD ? g : (upack >> (4 * pl)) & 15;
    const int8_t *ga = a_tile + (int64_t)t * M * kpad_bytes + kt * BKB;
    const int8_t *gb = b_tile + (int64_t)uu * N * kpad_bytes + kt * BKB;
    char *la = lds + buf * STAGE, *lb = la + BM * BKB;
#pragma unroll
    for (int u = 0; u < LPA; ++u)
      __builtin_amdgcn_global_load_lds((const void *)(ga + a_off[u]),
                                       (__attribute__((address_space(3))) void *)(la + (LPA * wave + u) * 1024),
                                       16, 0, 0);
#pragma unroll
    for (int u = 0; u < LPB; ++u)
      __builtin_amdgcn_global_load_lds((const void *)(gb + b_off[u]),
                                       (__attribute__((address_space(3))) void *)(lb + (LPB * wave + u) * 1024),
                                       16, 0, 0);
  };
  struct Frag {
    i4_t a[FM], b[FN];
  };
  auto read = [&](int buf, int ks, Frag &f) {
    const char *sa = lds + buf * STAGE, *sb = sa + BM * BKB;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int r = wm * WTM + i * 16 + li;
      f.a[i] = *reinterpret_cast<const i4_t *>(sa + r * BKB + 16 * ((ks * 4 + kg) ^ dl2_swz(r)));
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int r = wn * WTN + j * 16 + li;
      f.b[j] = *reinterpret_cast<const i4_t *>(sb + r * BKB + 16 * ((ks * 4 + kg) ^ dl2_swz(r)));
    }
  };

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  constexpr int NR = FM + FN, NMF = FM * FN, PER = NMF / NR;
  auto mmas = [&](const Frag &f) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = MM::mma(f.a[i], f.b[j], acc[i][j]);
  };

  issue(0, 0);
  if (iters > 1) issue(1, 1);
  if (iters > 1) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(LPA + LPB) : "memory");
  else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  asm volatile("s_barrier" ::: "memory");
  Frag f0, f1;
  read(0, 0, f0);
  for (int q = 0; q < iters; ++q) {
    const int buf = q & 1;
    /* k-step 1's fragments during k-step 0's MFMAs */
    read(buf, 1, f1);
    mmas(f0);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      __builtin_amdgcn_sched_group_barrier(0x008, PER, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, NMF - PER * NR, 0);
    /* stage q + 1 landed (this wave's loads), stage q fully read */
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
    asm volatile("s_barrier" ::: "memory");
    if (q + 2 < iters) issue(q + 2, buf);
    if (q + 1 < iters) read(buf ^ 1, 0, f0);
    mmas(f1);
    __builtin_amdgcn_sched_group_barrier(0x020, LPA + LPB, 0);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      __builtin_amdgcn_sched_group_barrier(0x008, PER, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, NMF - PER * NR, 0);
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");

  store_tile<KIND, FM, FN>(acc, g, G, S, C, ldc, alpha, beta, M, N, m0 + wm * WTM,
                           n0 + wn * WTN, li, kg);
}

template <int KIND>
hipError_t launch_dl2(const int8_t *as, const int8_t *bs, int m, int n, int64_t kpad, int kb0,
                      int kb1, const SliceGroups &G, void *S, float *C, int64_t ldc,
                      float alpha, float beta, hipStream_t st) {
  auto kern = slice_products_dl2_kernel<KIND>;
  const size_t lds = (size_t)2 * 512 * 128;
  hipError_t e = hipFuncSetAttribute((const void *)kern,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int tiles_m = (int)(((int64_t)m + 255) / 256), tiles_n = (int)(((int64_t)n + 255) / 256);
  hipLaunchKernelGGL(kern, dim3(tiles_m * tiles_n * G.ngroups), dim3(512), lds, st, as, bs, m,
                     n, kpad, kb0, kb1, G, tiles_m, tiles_n, S, C, ldc, alpha, beta);
  return hipGetLastError();
}

template <int KIND, int BM, int BN, int WM, int WN, int BKB, int PAD = 4>
hipError_t launch_cfg(const int8_t *as, const int8_t *bs, int m, int n, int64_t kpad,
                      int kb0, int kb1, const SliceGroups &G, void *S, float *C,
                      int64_t ldc, float alpha, float beta, hipStream_t st) {
  auto kern = slice_products_kernel<KIND, BM, BN, WM, WN, BKB, PAD>;
  const size_t lds = (size_t)2 * (BM + BN) * (BKB / 4 + PAD) * 4;
  hipError_t e = hipFuncSetAttribute((const void *)kern,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int tiles_m = (int)(((int64_t)m + BM - 1) / BM);
  const int tiles_n = (int)(((int64_t)n + BN - 1) / BN);
  hipLaunchKernelGGL(kern, dim3(tiles_m * tiles_n * G.ngroups), dim3(64 * WM * WN), lds,
                     st, as, bs, m, n, kpad, kb0, kb1, G, tiles_m, tiles_n, S, C, ldc,
                     alpha, beta);
  return hipGetLastError();
}

/* Product kernel: 1 (the default for every kind) = direct-to-LDS loads,
 * 128-byte K stages in two buffers, one barrier per two k-steps; 0 =
 * register-staged 128-byte stages with MFMAs interleaved with the fragment
 * reads and LDS writes, lines padded by 16 bytes.  Both: 256x256 tiles, 8
 * waves of 128x64.  Measured at 8192^3, 1 is 5-6% faster than 0 for
 * residue, slice and BF16 products alike.  Variants measured slower and
 * dropped: direct-to-LDS with 64-byte stages in four buffers, three stages
 * ahead (4.6% slower than 1) or two ahead, 4 waves of 128x128 with 128-byte
 * stages and direct-to-LDS with 4 waves of 128x128 (both spill), 256x128 and
 * 128x256 tiles (one wave per SIMD), two 256x128 blocks per CU, 32 bytes of
 * line padding, 64-byte stages with 8 waves and register staging, the
 * compiler's built-in GEMM interleaving, a bank-conflict-free
 * fragment-ordered direct-to-LDS layout (32-byte global runs: slower), and
 * raising the wave priority (s_setprio 1) over each k-step's MFMAs (5-8%
 * slower: the waves in their MFMAs starve the other's loads and reads).
 * gemmx_set_product_config (else GEMMX_SLICE_CFG, read at first use)
 * selects one for every kind. */
std::atomic<int> g_product_cfg{-2};

int product_cfg() {
  int cfg = g_product_cfg.load(std::memory_order_relaxed);
  if (cfg == -2) {
    const char *e = getenv("GEMMX_SLICE_CFG");
    const int v = (e && *e) ? atoi(e) : -1;
    cfg = (v == 0 || v == 1) ? v : -1;
    int expected = -2;
    if (!g_product_cfg.compare_exchange_strong(expected, cfg)) cfg = expected;
  }
  return cfg >= 0 ? cfg : 1;
}

template <int KIND>
hipError_t products(const int8_t *as, const int8_t *bs, int m, int n, int64_t kpad,
                    int kb0, int kb1, const SliceGroups &G, void *S, float *C,
                    int64_t ldc, float alpha, float beta, hipStream_t st) {
  if (product_cfg() == 0)
    return launch_cfg<KIND, 256, 256, 2, 4, 128>(as, bs, m, n, kpad, kb0, kb1, G, S, C, ldc,
                                                 alpha, beta, st);
  return launch_dl2<KIND>(as, bs, m, n, kpad, kb0, kb1, G, S, C, ldc, alpha, beta, st);
}

} // namespace

hipError_t slice_products(int kind, const int8_t *as, const int8_t *bs, int m, int n,
                          int64_t kpad_bytes, int kb0, int kb1, const SliceGroups &g,
                          void *S, float *C, int64_t ldc, float alpha, float beta,
                          hipStream_t st) {
  switch (kind) {
  case SP_BF16:
    return products<SP_BF16>(as, bs, m, n, kpad_bytes, kb0, kb1, g, S, C, ldc, alpha, beta,
                             st);
  case SP_I8_MOD:
    return products<SP_I8_MOD>(as, bs, m, n, kpad_bytes, kb0, kb1, g, S, C, ldc, alpha, beta,
                               st);
  default:
    return products<SP_I8_SUM>(as, bs, m, n, kpad_bytes, kb0, kb1, g, S, C, ldc, alpha, beta,
                               st);
  }
}

} // namespace gemmx

extern "C" int gemmx_set_product_config(int cfg) {
  if (cfg < -1 || cfg > 1) return -1;
  gemmx::g_product_cfg.store(cfg, std::memory_order_relaxed);
  return 0;
}
