/*
 * gemmx_device.h -- device-side building blocks shared by the gfx950 tile
 * kernels: MFMA wrappers, the global->LDS tile copy with the transpose folded
 * into the choice of LDS layout, and the XCD-aware tile order.
 *
 * Replaces the reference's packing routines (wjc404/GEMM_AVX2
 * src/gemm_copy.c:4-143, src/gemm_driver.c:34-79): instead of one thread
 * packing an A panel into a shared buffer while the others spin
 * (src/gemm_driver.c:23-33), every workgroup stages its own A and B tiles in
 * LDS, copying contiguous global runs straight into LDS lines.
 */
#ifndef GEMMX_DEVICE_H
#define GEMMX_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gemmx_index.h"

namespace gemmx {

typedef double d2_t __attribute__((ext_vector_type(2)));
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef float f16_t __attribute__((ext_vector_type(16)));
typedef int i4_t __attribute__((ext_vector_type(4)));
typedef int i16_t __attribute__((ext_vector_type(16)));

/* 16x16x4 MFMA for the native types.  Operand layout on gfx950 (checked on
 * the device with one-hot operands):
 *   A (16x4):  lane l holds A[l%16][l/16]
 *   B (4x16):  lane l holds B[l/16][l%16]
 *   D (16x16): lane l, element r holds D[row(l/16, r)][l%16] with
 *              FP32: row = 4*(l/16) + r    (four consecutive rows)
 *              FP64: row = (l/16) + 4*r    (rows interleaved by 4)         */
template <typename T> struct Mfma16;
template <> struct Mfma16<double> {
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int group, int r) { return group + 4 * r; }
};
template <> struct Mfma16<float> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int group, int r) { return 4 * group + r; }
};

template <typename T> struct Vec16;
template <> struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  enum { N = 2 };
};
template <> struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  enum { N = 4 };
};
template <> struct Vec16<int> {
  typedef int type __attribute__((ext_vector_type(4)));
  enum { N = 4 };
};

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

/* Epilogue of one wave's FM x FN grid of 16x16 MFMA fragments (row r of
 * fragment (i, j) at C row r0 + 16 i + Mfma16<T>::row(kg, r), column
 * c0 + 16 j + li): C = fma(alpha, acc, beta*C) for rows < M and columns < N
 * only, the reference's masked stores (wjc404/GEMM_AVX2
 * src/kernel/gemm_kernel_unroll6.S:145-155).  The C values of RC fragment
 * rows (at most 16 values per lane) are loaded together before any of their
 * stores, from addresses clamped into C, so the loads are unconditional and
 * in flight at once: a masked load-multiply-store per element made the
 * compiler wait for each load in turn, one memory round trip per element.
 * A clamped load may read an element an earlier batch already stored; it
 * only feeds a position that is not stored. */
template <typename T, int FM, int FN, typename acc_t>
__device__ __forceinline__ void store_c_tile(const acc_t (&acc)[FM][FN], T *__restrict__ C, int M,
                                             int N, int64_t ldc, T alpha, T beta, int r0, int c0,
                                             int li, int kg) {
  constexpr int PER_ROW = FN * 4;
  constexpr int RC0 = PER_ROW >= 16 ? 1 : 16 / PER_ROW;
  constexpr int RC = RC0 < FM ? RC0 : FM;
  static_assert(FM % RC == 0, "fragment rows per load batch");
#pragma unroll
  for (int i0 = 0; i0 < FM; i0 += RC) {
    T cv[RC][FN][4];
#pragma unroll
    for (int ii = 0; ii < RC; ++ii)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = c0 + j * 16 + li;
        const int64_t cc = col < N ? col : N - 1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + (i0 + ii) * 16 + Mfma16<T>::row(kg, r);
          cv[ii][j][r] = C[(row < M ? row : M - 1) + cc * ldc];
        }
      }
#pragma unroll
    for (int ii = 0; ii < RC; ++ii)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = c0 + j * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + (i0 + ii) * 16 + Mfma16<T>::row(kg, r);
          if (col < N && row < M)
            C[row + (int64_t)col * ldc] = fma_t(alpha, (T)acc[i0 + ii][j][r], beta * cv[ii][j][r]);
        }
      }
  }
}

/* The same epilogue with the C accesses coalesced through a per-wave LDS
 * region of FN*16 columns of EPI_CS elements (epi_region_elems): fragment
 * row i (16 rows of the wave's tile) is written there column by column,
 * then read back so that lane t holds row t % 16 of column t / 16 (and the
 * columns 4, 8, ... further on): every C load and store of a wave covers
 * whole 16-row column segments, 128 contiguous bytes (FP64) or 64 (FP32) per
 * 16 lanes, where the MFMA result layout puts 4 rows of 16 different columns
 * in one instruction.  The region is private to the wave, and a wave's LDS
 * operations complete in order, so no barrier is needed between its writes
 * and its reads; the caller makes sure no other wave still reads that LDS.
 * Addresses are clamped into C as in store_c_tile. */
constexpr int EPI_CS = 17;
template <int FN>
constexpr int epi_region_elems() {
  return FN * 16 * EPI_CS;
}
template <typename T, int FM, int FN, typename acc_t>
__device__ __forceinline__ void store_c_tile_lds(const acc_t (&acc)[FM][FN], T *__restrict__ C,
                                                 int M, int N, int64_t ldc, T alpha, T beta,
                                                 int r0, int c0, int li, int kg, T *reg,
                                                 int lane) {
  constexpr int PER = FN * 16 * 16 / 64; /* region elements per lane */
  const int er = lane & 15, ec = lane >> 4;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        reg[(j * 16 + li) * EPI_CS + Mfma16<T>::row(kg, r)] = (T)acc[i][j][r];
    const int row = r0 + i * 16 + er;
    const int64_t rr = row < M ? row : M - 1;
    T v[PER], cv[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      v[q] = reg[(ec + 4 * q) * EPI_CS + er];
      const int col = c0 + ec + 4 * q;
      cv[q] = C[rr + (int64_t)(col < N ? col : N - 1) * ldc];
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int col = c0 + ec + 4 * q;
      if (col < N && row < M) C[row + (int64_t)col * ldc] = fma_t(alpha, v[q], beta * cv[q]);
    }
  }
}

/* Copies an NL x LEN tile -- NL lines of LEN contiguous elements, line
 * stride `ld` in global memory -- into LDS lines of stride LS, in 16-byte
 * chunks.  Elements outside (nl_valid, len_valid) read as zero and are never
 * dereferenced.  The register stage `r` lets the next tile's global loads be
 * in flight while the current tile is multiplied. */
template <typename T, int NL, int LEN, int NT>
struct TileCopy {
  typedef typename Vec16<T>::type vec_t;
  static constexpr int VE = Vec16<T>::N;
  static constexpr int CPL = LEN / VE;
  static constexpr int NCH = NL * CPL;
  static constexpr int PER = NCH / NT;
  static_assert(LEN % VE == 0, "tile line must be a whole number of 16-byte chunks");
  static_assert(NCH % NT == 0, "tile chunks must divide evenly over the threads");
  vec_t r[PER];

  __device__ __forceinline__ void load(const T *__restrict__ g, int64_t ld,
                                       int nl_valid, int len_valid, bool vec_ok,
                                       int tid) {
    if (vec_ok && nl_valid >= NL && len_valid >= LEN) {
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const int ch = tid + c * NT, line = ch / CPL, off = (ch % CPL) * VE;
        r[c] = *reinterpret_cast<const vec_t *>(g + line * ld + off);
      }
    } else {
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const int ch = tid + c * NT, line = ch / CPL, off = (ch % CPL) * VE;
        vec_t v;
#pragma unroll
        for (int e = 0; e < VE; ++e)
          v[e] = (line < nl_valid && off + e < len_valid) ? g[line * ld + off + e] : T(0);
        r[c] = v;
      }
    }
  }

  /* Interior tile with aligned operands: every chunk is one 16-byte load. */
  __device__ __forceinline__ void load_vec(const T *__restrict__ g, int64_t ld, int tid) {
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const int ch = tid + c * NT, line = ch / CPL, off = (ch % CPL) * VE;
      r[c] = *reinterpret_cast<const vec_t *>(g + line * ld + off);
    }
  }

  /* A full-K slice of an edge tile with aligned operands: lines at or past
   * nl_valid re-read line nl_valid-1, and chunks at or past len_valid re-read
   * the last chunk holding a valid element (clamp_chunk), so every chunk is
   * still one 16-byte load.  Each load is the aligned 16-byte block around a
   * valid element (an aligned block never crosses a page); in the operand's
   * last line that block can reach up to VE-1 elements past the last valid
   * one, beyond the operand's minimal extent ld*(lines-1)+len.  What lands in
   * a position past nl_valid/len_valid feeds only rows or columns of C that
   * the epilogue does not store. */
  __device__ __forceinline__ void load_clamp(const T *__restrict__ g, int64_t ld, int nl_valid,
                                             int len_valid, int tid) {
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const ChunkAt at = clamp_chunk<CPL, VE>(tid + c * NT, nl_valid, len_valid);
      r[c] = *reinterpret_cast<const vec_t *>(g + at.line * ld + at.off);
    }
  }

  /* A full-K slice of operands that do not allow 16-byte loads: one load per
   * element (consecutive lanes still read consecutive elements of a line).
   * With CLAMP (edge tile) a line or element index past nl_valid/len_valid is
   * clamped to the last valid one, with the same effect as in load_clamp. */
  template <bool CLAMP>
  __device__ __forceinline__ void load_narrow(const T *__restrict__ g, int64_t ld, int nl_valid,
                                              int len_valid, int tid) {
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const int ch = tid + c * NT;
      const int line = CLAMP ? clamp_index(ch / CPL, nl_valid) : ch / CPL;
      const int off = (ch % CPL) * VE;
      const T *p = g + line * ld;
      vec_t v;
#pragma unroll
      for (int e = 0; e < VE; ++e) v[e] = p[CLAMP ? clamp_index(off + e, len_valid) : off + e];
      r[c] = v;
    }
  }

  /* store() for the last slice of the branch-free loop, which is loaded
   * shifted back so that it ends at K (every load in bounds): its first
   * `skip` K positions were summed by the previous slice and are written as
   * zeros, in both operands (so no 0 * Inf).  K runs along the lines
   * (KLINES: line index = K) or along each line.  The selects sit at the LDS
   * writes, where the loaded values are consumed anyway, so the loads stay
   * in flight over the slice's MFMAs. */
  template <int LS, bool KLINES>
  __device__ __forceinline__ void store_skip(T *s, int tid, int skip) const {
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const int ch = tid + c * NT, line = ch / CPL, off = (ch % CPL) * VE;
      vec_t v = r[c];
#pragma unroll
      for (int e = 0; e < VE; ++e) {
        const bool z = KLINES ? line < skip : off + e < skip;
        v[e] = z ? T(0) : v[e];
      }
      *reinterpret_cast<vec_t *>(s + line * LS + off) = v;
    }
  }

  template <int LS>
  __device__ __forceinline__ void store(T *s, int tid) const {
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const int ch = tid + c * NT, line = ch / CPL, off = (ch % CPL) * VE;
      *reinterpret_cast<vec_t *>(s + line * LS + off) = r[c];
    }
  }
};

/* Workgroups are dispatched round-robin over the 8 XCDs of MI355X, each with
 * its own L2.  Renumber them so that consecutive tile ids run on one XCD,
 * then walk tiles in groups of GROUP_M tile-rows so the workgroups that an
 * XCD runs together share A rows and B columns in its L2. */
__device__ __forceinline__ void tile_coords(int pid, int num, int tiles_m,
                                            int tiles_n, int *tm, int *tn) {
  constexpr int XCDS = 8, GROUP_M = 8;
  const int xcd = pid % XCDS, local = pid / XCDS;
  const int per = num / XCDS, rem = num % XCDS;
  pid = (xcd < rem) ? xcd * (per + 1) + local
                    : rem * (per + 1) + (xcd - rem) * per + local;
  const int in_group = GROUP_M * tiles_n;
  const int g = pid / in_group, first = g * GROUP_M;
  const int gsize = (tiles_m - first < GROUP_M) ? tiles_m - first : GROUP_M;
  const int r = pid % in_group;
  *tm = first + r % gsize;
  *tn = r / gsize;
}

} // namespace gemmx
#endif
