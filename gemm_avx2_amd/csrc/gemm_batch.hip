/*
 * gemm_batch.hip -- grouped batch: many independent GEMMs of uneven shapes
 * and mixed transposes in one persistent launch.  No reference counterpart
 * (the reference computes one GEMM per dgemm_ call, wjc404/GEMM_AVX2
 * src/gemm_driver.c:118); every problem keeps dgemm_'s semantics, including
 * the beta != 1 rule and the alpha == 0 gate.
 *
 * The grid is as large as the device holds at once.  The host deals whole
 * problems, heaviest first, to eight queues, one per XCD (longest processing
 * time first: each problem to the queue with the least work), so the tiles
 * of a problem share the operand panels in one XCD's L2; a workgroup takes
 * the next tile of its XCD's queue with an atomic ticket, and when that queue
 * is empty, of the other queues in turn, so every wave reaches the exit
 * after at most eight empty queues.  The ticket for the next tile is taken
 * while a tile's first K slice is being loaded, right after those loads are
 * issued: the atomic's round trip overlaps the tile, and since vector memory
 * operations retire in order, an atomic issued before the loads would hold
 * back the first wait on them (and the workgroup's barriers with it), while
 * one issued after them does not.  A tile is computed by
 * the same gemm_tile as the single-GEMM kernels: the branch-free loop with
 * an edge tile's loads clamped back into the operand, in 16-byte chunks
 * where both operands allow them and one element per load otherwise,
 * instantiated for the four transpose pairs and chosen per problem.
 */
#include "gemmx_device.h"
#include "gemmx_internal.h"
#include "gemmx_tile.h"

using namespace gemmx;

namespace {

constexpr int BT_THREADS = 256; /* 4 waves in every configuration */

/* Tile configurations of the batch kernel (BM x BN x BK, waves WM x WN). */
template <int CFG> struct BatchCfg;
template <> struct BatchCfg<0> { static constexpr int BM = 64, BN = 64, BK = 16, WM = 2, WN = 2; };
template <> struct BatchCfg<1> { static constexpr int BM = 64, BN = 64, BK = 32, WM = 2, WN = 2; };
template <> struct BatchCfg<2> { static constexpr int BM = 128, BN = 64, BK = 16, WM = 2, WN = 2; };
template <> struct BatchCfg<3> { static constexpr int BM = 32, BN = 32, BK = 32, WM = 2, WN = 2; };

/* Zeroes the ticket counters of the XCD queues (a kernel, not a memset
 * node, so a captured graph replays it reliably). */
__global__ void batch_reset_kernel(unsigned *sched) {
  if (threadIdx.x < GEMMX_BATCH_QUEUES) sched[threadIdx.x * GEMMX_BATCH_CTR_STRIDE] = 0u;
}

/* Operand pointers read from the descriptors are generic; building them in
 * the global address space lets every load through them be a global (not
 * flat) load. */
template <typename T>
__device__ __forceinline__ T *as_global(T *p) {
  return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}

template <typename T, int CFG>
__global__ void __launch_bounds__(BT_THREADS, 2)
gemm_batch_kernel(const gemmx_batch_desc *__restrict__ desc,
                  const int *__restrict__ tile_prob, unsigned *__restrict__ sched,
                  gemmx_batch_queues Q) {
  typedef BatchCfg<CFG> G;
  static_assert(64 * G::WM * G::WN == BT_THREADS, "batch block size");
  __shared__ __align__(16) T lds[tile_lds_max<T, G::BM, G::BN, G::BK, G::WM, G::WN>()];
  __shared__ int ticket;
  /* sched: the queues' ticket counters, GEMMX_BATCH_CTR_STRIDE words apart
   * (each on a line of its own); Q: the queues' first tiles, passed by value
   * so that no tile waits on a load from a line the atomics keep busy.
   * Workgroups are dispatched round-robin over the XCDs, so blockIdx.x % 8 is
   * this workgroup's XCD: it takes the tiles of its XCD's queue (whole
   * problems, so their operand panels stay in that XCD's L2), then those of
   * the other queues. */
  const int x = (int)(blockIdx.x % GEMMX_BATCH_QUEUES);
  auto fetch = [&](int queue) {
    __syncthreads(); /* every wave has read `ticket` and left the LDS */
    if (threadIdx.x == 0)
      ticket = (int)atomicAdd(sched + queue * GEMMX_BATCH_CTR_STRIDE, 1u);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(ticket);
  };
  int y = x, round = 0;
  int t = fetch(y);
  for (;;) {
    /* an exhausted queue: move to the next one; after all of them, exit
     * (uniform: every wave of the workgroup leaves together) */
    while (t >= Q.qbeg[y + 1] - Q.qbeg[y]) {
      if (++round == GEMMX_BATCH_QUEUES) return;
      y = (x + round) % GEMMX_BATCH_QUEUES;
      t = fetch(y);
    }
    const int tt = Q.qbeg[y] + t;
    const int p = __builtin_amdgcn_readfirstlane(tile_prob[tt]);
    const gemmx_batch_desc &q = desc[p];
    const int local = tt - q.tile_begin;
    const int tm = local % q.tiles_m, tn = local / q.tiles_m;
    const int m0 = tm * G::BM, n0 = tn * G::BN;
    const T alpha = (T)q.alpha, beta = (T)q.beta;
    const T *A = as_global((const T *)q.a);
    const T *B = as_global((const T *)q.b);
    T *C = as_global((T *)q.c);
    const int f = q.flags;
    /* the next tile's ticket, taken by thread 0 while this tile's first
     * loads are in flight (issued after them, so waits on those loads do
     * not wait for the atomic) */
    int nxt = 0;
    auto take = [&]() {
      if (threadIdx.x == 0) nxt = (int)atomicAdd(sched + y * GEMMX_BATCH_CTR_STRIDE, 1u);
    };
    if (f & GEMMX_BF_SCALE) {
      take();
      /* alpha == 0 or k == 0: C = beta*C on this tile (beta != 1 here) */
      for (int e = threadIdx.x; e < G::BM * G::BN; e += BT_THREADS) {
        const int i = m0 + e % G::BM, j = n0 + e / G::BM;
        if (i < q.m && j < q.n) C[i + j * q.ldc] *= beta;
      }
    } else {
      const int va = (f & GEMMX_BF_VEC_A) != 0, vb = (f & GEMMX_BF_VEC_B) != 0;
      /* every tile takes the branch-free loop, an edge tile's loads past m
       * or n clamped back into the operand: 16-byte loads where both operands
       * allow them, else one element per load */
      const int fast = va && vb;
/* C through LDS in the epilogue for FP32 only: the FP64 kernel's extra
 * registers would cost it a workgroup per CU */
#define GEMMX_TILE(TA, TB, FAST)                                                        \
  gemm_tile<T, TA, TB, G::BM, G::BN, G::BK, G::WM, G::WN, 1, true, true, !(FAST), false, \
            1, 1, std::is_same<T, float>::value>(                                       \
      A, B, C, q.m, q.n, q.k, q.lda, q.ldb, q.ldc, alpha, beta, m0, n0, va, vb, lds, take)
      switch ((f & (GEMMX_BF_TA | GEMMX_BF_TB)) | (fast ? 4 : 0)) {
      case 0: GEMMX_TILE(false, false, false); break;
      case GEMMX_BF_TB: GEMMX_TILE(false, true, false); break;
      case GEMMX_BF_TA: GEMMX_TILE(true, false, false); break;
      case GEMMX_BF_TA | GEMMX_BF_TB: GEMMX_TILE(true, true, false); break;
      case 4: GEMMX_TILE(false, false, true); break;
      case 4 | GEMMX_BF_TB: GEMMX_TILE(false, true, true); break;
      case 4 | GEMMX_BF_TA: GEMMX_TILE(true, false, true); break;
      default: GEMMX_TILE(true, true, true); break;
      }
#undef GEMMX_TILE
    }
    /* the next tile of this queue: every wave is done with this tile's LDS
     * and with the previous ticket before thread 0 publishes the new one */
    __syncthreads();
    if (threadIdx.x == 0) ticket = nxt;
    __syncthreads();
    t = __builtin_amdgcn_readfirstlane(ticket);
  }
}

template <typename T, int CFG>
int batch_grid() {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gemm_batch_kernel<T, CFG>, BT_THREADS,
                                                   0) != hipSuccess)
    return 0;
  return cus * (per > 0 ? per : 1);
}

template <typename T, int CFG>
void launch(const gemmx_batch_desc *d, const int *tp, unsigned *sched,
            const gemmx_batch_queues &q, int grid, hipStream_t s) {
  hipLaunchKernelGGL((gemm_batch_kernel<T, CFG>), dim3(grid), dim3(BT_THREADS), 0, s, d, tp,
                     sched, q);
}

} // namespace

extern "C" int gemmx_batch_tile(int cfg, int *bm, int *bn) {
  switch (cfg) {
  case 0: *bm = BatchCfg<0>::BM; *bn = BatchCfg<0>::BN; return 0;
  case 1: *bm = BatchCfg<1>::BM; *bn = BatchCfg<1>::BN; return 0;
  case 2: *bm = BatchCfg<2>::BM; *bn = BatchCfg<2>::BN; return 0;
  case 3: *bm = BatchCfg<3>::BM; *bn = BatchCfg<3>::BN; return 0;
  default: return -1;
  }
}

extern "C" int gemmx_batch_grid(int es, int cfg) {
  switch (cfg) {
  case 0: return es == 8 ? batch_grid<double, 0>() : batch_grid<float, 0>();
  case 1: return es == 8 ? batch_grid<double, 1>() : batch_grid<float, 1>();
  case 2: return es == 8 ? batch_grid<double, 2>() : batch_grid<float, 2>();
  case 3: return es == 8 ? batch_grid<double, 3>() : batch_grid<float, 3>();
  default: return 0;
  }
}

extern "C" hipError_t gemmx_batch_launch(int es, int cfg, const gemmx_batch_desc *d_desc,
                                         const int *d_tile_prob, unsigned *d_sched,
                                         const gemmx_batch_queues *q, int grid, hipStream_t s) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, d_sched);
  const bool d = es == 8;
  switch (cfg) {
  case 0: (d ? launch<double, 0> : launch<float, 0>)(d_desc, d_tile_prob, d_sched, *q, grid, s); break;
  case 1: (d ? launch<double, 1> : launch<float, 1>)(d_desc, d_tile_prob, d_sched, *q, grid, s); break;
  case 2: (d ? launch<double, 2> : launch<float, 2>)(d_desc, d_tile_prob, d_sched, *q, grid, s); break;
  case 3: (d ? launch<double, 3> : launch<float, 3>)(d_desc, d_tile_prob, d_sched, *q, grid, s); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
