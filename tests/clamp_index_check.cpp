// Host check of the edge-tile load index math (gemmx_index.h, used by
// TileCopy::load_clamp and load_narrow in gemmx_device.h).  For every tile
// line length the kernels instantiate and every valid extent of an edge tile
// it checks, against a buffer of exactly the operand's minimal extent
// ld*(lines-1)+len elements:
//   * chunks inside the valid region keep their place (no valid data moves);
//   * every clamped chunk lies in a valid line, starts at a VE-aligned offset
//     and holds at least one valid element;
//   * the only reads past the minimal extent are the last line's aligned
//     block tail, at most VE-1-(len-1)%VE elements, inside the same aligned
//     block as the last valid element;
//   * the one-element form never leaves the minimal extent.
// Prints "ok" and exits 0, or names the first failing case and exits 1.
#include <cstdio>

#include "gemmx_index.h"

using gemmx::ChunkAt;

static int fails = 0;

#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      if (fails++ < 10) {                     \
        std::printf("FAIL %s: ", #cond);      \
        std::printf(__VA_ARGS__);             \
        std::printf("\n");                    \
      }                                       \
    }                                         \
  } while (0)

template <int LEN, int VE>
static void lines_of(int nl_max) {
  constexpr int CPL = LEN / VE;
  for (int nl = 1; nl <= nl_max; ++nl)
    for (int len = 1; len <= LEN; ++len)
      for (int pad = 0; pad < 2 * VE; pad += VE) {
        const long ld = ((len + VE - 1) / VE) * VE + pad; // 16-byte loads need ld % VE == 0
        const long extent = ld * (nl - 1) + len;          // minimal BLAS extent
        const long block_end = ld * (nl - 1) + ((len + VE - 1) / VE) * VE;
        for (int ch = 0; ch < nl_max * CPL; ++ch) {
          const ChunkAt at = gemmx::clamp_chunk<CPL, VE>(ch, nl, len);
          const int line0 = ch / CPL, off0 = (ch % CPL) * VE;
          if (line0 < nl && off0 < len)
            CHECK(at.line == line0 && at.off == off0, "LEN %d VE %d nl %d len %d ch %d", LEN, VE,
                  nl, len, ch);
          CHECK(at.line >= 0 && at.line < nl && at.off >= 0 && at.off % VE == 0 && at.off < len,
                "LEN %d VE %d nl %d len %d ch %d -> line %d off %d", LEN, VE, nl, len, ch,
                at.line, at.off);
          const long first = at.line * ld + at.off, last = first + VE - 1;
          CHECK(first < extent, "chunk starts past the extent: LEN %d nl %d len %d", LEN, nl, len);
          if (last >= extent) {
            CHECK(at.line == nl - 1 && last < block_end &&
                      last - extent + 1 <= VE - 1 - (len - 1) % VE,
                  "over-read beyond the last aligned block: LEN %d VE %d nl %d len %d", LEN, VE,
                  nl, len);
            CHECK((extent - 1) / VE == last / VE,
                  "over-read leaves the last valid element's block: LEN %d VE %d nl %d len %d",
                  LEN, VE, nl, len);
          }
        }
        // one element per load: line and element clamped separately
        for (int l = 0; l < nl_max; ++l)
          for (int e = 0; e < LEN; ++e) {
            const long i = gemmx::clamp_index(l, nl) * ld + gemmx::clamp_index(e, len);
            CHECK(i >= 0 && i < extent, "narrow read outside: nl %d len %d l %d e %d", nl, len, l,
                  e);
            if (l < nl && e < len) CHECK(i == l * ld + e, "narrow valid element moved");
          }
      }
}

int main() {
  // FP64 (VE 2) and FP32 (VE 4) lines of the native, batch and split tiles
  lines_of<16, 2>(9);
  lines_of<32, 2>(9);
  lines_of<64, 2>(5);
  lines_of<128, 2>(3);
  lines_of<256, 2>(3);
  lines_of<16, 4>(9);
  lines_of<32, 4>(9);
  lines_of<64, 4>(5);
  lines_of<128, 4>(3);
  lines_of<256, 4>(3);
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
