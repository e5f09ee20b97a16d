/*
 * gemmx_index.h -- index math of the clamped edge-tile loads, kept free of
 * HIP so that host code can check it (the device code uses these same
 * functions).
 */
#ifndef GEMMX_INDEX_H
#define GEMMX_INDEX_H

#ifdef __HIPCC__
#define GEMMX_HD __host__ __device__ __forceinline__
#else
#define GEMMX_HD inline
#endif

namespace gemmx {

struct ChunkAt {
  int line, off;
};

/* Chunk `ch` of a tile slice whose lines are CPL chunks of VE elements, in
 * an edge tile with nl_valid valid lines of len_valid valid elements each
 * (both >= 1): a line at or past nl_valid re-reads line nl_valid - 1, and a
 * chunk at or past the last chunk holding a valid element re-reads that
 * chunk.  Chunks inside the valid region keep their place.  Every chunk
 * read is the VE-aligned block around a valid element; in the last valid
 * line that block reaches up to VE - 1 - (len_valid - 1) % VE elements past
 * element len_valid - 1 (never into another aligned block, so never across a
 * page). */
template <int CPL, int VE>
GEMMX_HD ChunkAt clamp_chunk(int ch, int nl_valid, int len_valid) {
  const int line = ch / CPL, off = (ch % CPL) * VE;
  const int last_off = ((len_valid - 1) / VE) * VE;
  return ChunkAt{line < nl_valid - 1 ? line : nl_valid - 1, off < last_off ? off : last_off};
}

/* Line and element of a one-element-per-load edge tile (operands without
 * 16-byte loads): both clamped to the last valid one, so no read leaves the
 * operand's minimal extent. */
GEMMX_HD int clamp_index(int i, int n_valid) { return i < n_valid - 1 ? i : n_valid - 1; }

} // namespace gemmx

#endif
