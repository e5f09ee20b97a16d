/*
 * gemmx_batch.c -- host side of the grouped batch extension (no reference
 * counterpart; each problem has the semantics of dgemm_/sgemm_,
 * wjc404/GEMM_AVX2 src/gemm_driver.c:118-197).
 *
 * A plan is built once from host arrays that describe the problems (device
 * pointers for the matrices): the arguments are checked, problems with work
 * are sorted heaviest first and dealt to one tile queue per XCD, their tiles
 * numbered queue by queue, the descriptors and the tile->problem map copied
 * to the device, and the launch (ticket reset + persistent grouped kernel,
 * the queue boundaries as a kernel argument) captured into a hipGraph.  Running the plan
 * replays the graph on the caller's stream.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <string.h>

#include "../../include/gemm_mi355x.h"
#include "gemmx_internal.h"
#include "gemmx_tune.h"

#define EXPORT __attribute__((visibility("default")))

struct gemmx_batch {
  int es, cfg, device, count, total_tiles, grid;
  double flops;
  gemmx_batch_desc *d_desc;
  int *d_tile_prob;
  unsigned *d_sched; /* the queues' ticket counters */
  gemmx_batch_queues queues;
  hipStream_t capture;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

typedef struct {
  double work;
  int idx;
} order_t;

static int by_work_desc(const void *a, const void *b) {
  const order_t *x = (const order_t *)a, *y = (const order_t *)b;
  if (x->work != y->work) return x->work < y->work ? 1 : -1;
  return x->idx - y->idx;
}

static void destroy(struct gemmx_batch *p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->capture) (void)hipStreamDestroy(p->capture);
  if (p->d_desc) (void)hipFree(p->d_desc);
  if (p->d_tile_prob) (void)hipFree(p->d_tile_prob);
  if (p->d_sched) (void)hipFree(p->d_sched);
  free(p);
}

static int notrans(char c) { return c == 'N' || c == 'n'; }

/* Batch kernel configuration per element size (index 0: FP64, 1: FP32): the
 * one set by *_x_batch_set_config, else GEMMX_BATCH_CFG from the
 * environment, else the tuning table's `batch` entry (measured by the
 * autotuner), else the default. */
static int g_batch_cfg[2] = {-1, -1};
static const int k_batch_default[2] = {0, 0};

static int batch_cfg(int es) {
  const int i = es == 8 ? 0 : 1;
  if (g_batch_cfg[i] >= 0) return g_batch_cfg[i];
  const char *e = getenv("GEMMX_BATCH_CFG");
  if (e && *e) {
    const int c = atoi(e);
    if (c >= 0 && c < GEMMX_BATCH_CONFIGS) return c;
  }
  const int t = gemmx_tune_pick_batch(es);
  return t >= 0 ? t : k_batch_default[i];
}

static int set_batch_cfg(int es, int cfg) {
  gemmx_clear_error();
  if (cfg < -1 || cfg >= GEMMX_BATCH_CONFIGS)
    return gemmx_report(GEMMX_ERR_ARG, "batch configuration %d not in -1..%d", cfg,
                        GEMMX_BATCH_CONFIGS - 1);
  g_batch_cfg[es == 8 ? 0 : 1] = cfg;
  return GEMMX_OK;
}

/* alpha/beta/a/b/c arrays hold doubles and double* (es 8) or floats and
 * float* (es 4). */
static int create(int es, void **out, int count, const char *transa, const char *transb,
                  const int *m, const int *n, const int *k, const void *alpha,
                  const void *const *a, const int *lda, const void *const *b,
                  const int *ldb, const void *beta, void *const *c, const int *ldc) {
  gemmx_clear_error();
  if (!out) return gemmx_report(GEMMX_ERR_ARG, "batch: NULL plan pointer");
  *out = NULL;
  if (count < 0) return gemmx_report(GEMMX_ERR_ARG, "batch: count %d < 0", count);
  if (count > 0 && (!transa || !transb || !m || !n || !k || !alpha || !a || !lda || !b ||
                    !ldb || !beta || !c || !ldc))
    return gemmx_report(GEMMX_ERR_ARG, "batch: NULL argument array");
  const int cfg = batch_cfg(es);
  int tbm = 0, tbn = 0;
  if (gemmx_batch_tile(cfg, &tbm, &tbn) != 0)
    return gemmx_report(GEMMX_ERR_ARG, "batch: unknown configuration %d", cfg);
  gemmx_batch_desc *h = (gemmx_batch_desc *)calloc(count > 0 ? count : 1, sizeof *h);
  order_t *ord = (order_t *)calloc(count > 0 ? count : 1, sizeof *ord);
  struct gemmx_batch *p = (struct gemmx_batch *)calloc(1, sizeof *p);
  int *tp = NULL;
  int rc = GEMMX_OK, used = 0;
  if (!h || !ord || !p) {
    rc = gemmx_report(GEMMX_ERR_ALLOC, "batch: host allocation failed");
    goto fail;
  }
  p->es = es;
  p->cfg = cfg;
  for (int i = 0; i < count; i++) {
    const double al = es == 8 ? ((const double *)alpha)[i] : (double)((const float *)alpha)[i];
    const double be = es == 8 ? ((const double *)beta)[i] : (double)((const float *)beta)[i];
    const int na = notrans(transa[i]), nb = notrans(transb[i]);
    const int ra = na ? m[i] : k[i], rb = nb ? k[i] : n[i];
    if (m[i] < 0 || n[i] < 0 || k[i] < 0) {
      rc = gemmx_report(GEMMX_ERR_ARG, "batch problem %d: negative dimension", i);
      goto fail;
    }
    if (m[i] == 0 || n[i] == 0) continue;
    const int product = al != 0.0 && k[i] > 0;
    if (!product && be == 1.0) continue; /* C unchanged */
    if (ldc[i] < (m[i] > 1 ? m[i] : 1) || !c[i]) {
      rc = gemmx_report(GEMMX_ERR_ARG, "batch problem %d: ldc=%d or C invalid", i, ldc[i]);
      goto fail;
    }
    if (product && (lda[i] < (ra > 1 ? ra : 1) || ldb[i] < (rb > 1 ? rb : 1) || !a[i] || !b[i])) {
      rc = gemmx_report(GEMMX_ERR_ARG, "batch problem %d: lda/ldb too small or NULL operand", i);
      goto fail;
    }
    gemmx_batch_desc *d = &h[used];
    d->a = a[i];
    d->b = b[i];
    d->c = c[i];
    d->lda = lda[i];
    d->ldb = ldb[i];
    d->ldc = ldc[i];
    d->alpha = al;
    d->beta = be;
    d->m = m[i];
    d->n = n[i];
    d->k = k[i];
    const int ve = 16 / es;
    d->flags = (na ? 0 : GEMMX_BF_TA) | (nb ? 0 : GEMMX_BF_TB) |
               ((((uintptr_t)a[i] % 16) == 0 && lda[i] % ve == 0) ? GEMMX_BF_VEC_A : 0) |
               ((((uintptr_t)b[i] % 16) == 0 && ldb[i] % ve == 0) ? GEMMX_BF_VEC_B : 0) |
               (product ? 0 : GEMMX_BF_SCALE);
    d->tiles_m = (int)(((long)m[i] + tbm - 1) / tbm);
    d->tiles_n = (int)(((long)n[i] + tbn - 1) / tbn);
    ord[used].work = product ? (double)k[i] * m[i] * n[i] : 0.0;
    ord[used].idx = used;
    if (product) p->flops += 2.0 * m[i] * n[i] * k[i];
    used++;
  }
  /* Heaviest first, each problem dealt whole to the XCD queue with the least
   * work so far (longest processing time first); a queue holds its problems'
   * tiles in that order, so long-K tiles are handed out first. */
  qsort(ord, used, sizeof *ord, by_work_desc);
  gemmx_batch_desc *sorted = (gemmx_batch_desc *)calloc(used > 0 ? used : 1, sizeof *sorted);
  int *queue_of = (int *)calloc(used > 0 ? used : 1, sizeof *queue_of);
  if (!sorted || !queue_of) {
    free(sorted);
    free(queue_of);
    rc = gemmx_report(GEMMX_ERR_ALLOC, "batch: host allocation failed");
    goto fail;
  }
  double qwork[GEMMX_BATCH_QUEUES] = {0};
  long qtiles[GEMMX_BATCH_QUEUES] = {0};
  long total = 0;
  for (int i = 0; i < used; i++) {
    sorted[i] = h[ord[i].idx];
    const long tiles = (long)sorted[i].tiles_m * sorted[i].tiles_n;
    int best = 0;
    for (int x = 1; x < GEMMX_BATCH_QUEUES; x++)
      if (qwork[x] < qwork[best]) best = x;
    /* a tile costs its K slices plus a fixed part (prologue, epilogue) */
    qwork[best] += (double)tiles * ((sorted[i].flags & GEMMX_BF_SCALE) ? 16.0
                                                                        : sorted[i].k + 64.0);
    qtiles[best] += tiles;
    queue_of[i] = best;
    total += tiles;
  }
  free(h);
  h = sorted;
  if (total > 0x7fffffffL) {
    free(queue_of);
    rc = gemmx_report(GEMMX_ERR_UNSUPPORTED, "batch: %ld tiles", total);
    goto fail;
  }
  int *const qbeg = p->queues.qbeg;
  qbeg[0] = 0;
  for (int x = 0; x < GEMMX_BATCH_QUEUES; x++) qbeg[x + 1] = qbeg[x] + (int)qtiles[x];
  p->count = used;
  p->total_tiles = (int)total;
  tp = (int *)malloc(sizeof(int) * (total > 0 ? total : 1));
  if (!tp) {
    free(queue_of);
    rc = gemmx_report(GEMMX_ERR_ALLOC, "batch: host allocation failed");
    goto fail;
  }
  {
    int fill[GEMMX_BATCH_QUEUES];
    memcpy(fill, qbeg, sizeof fill);
    for (int i = 0; i < used; i++) {
      const int tiles = h[i].tiles_m * h[i].tiles_n;
      h[i].tile_begin = fill[queue_of[i]];
      for (int t = 0; t < tiles; t++) tp[h[i].tile_begin + t] = i;
      fill[queue_of[i]] += tiles;
    }
  }
  free(queue_of);

  /* Arguments are checked before the device is looked up. */
  if (gemmx_current_device(&p->device) != GEMMX_OK) {
    rc = gemmx_last_code();
    goto fail;
  }
  if (gemmx_alloc_failure_injected()) {
    rc = gemmx_report(GEMMX_ERR_ALLOC, "batch: device allocation failure injected");
    goto fail;
  }
  hipError_t e;
  if ((e = hipMalloc((void **)&p->d_desc, sizeof *h * (used > 0 ? used : 1))) != hipSuccess ||
      (e = hipMalloc((void **)&p->d_tile_prob, sizeof(int) * (total > 0 ? total : 1))) != hipSuccess ||
      (e = hipMalloc((void **)&p->d_sched,
                     sizeof(unsigned) * GEMMX_BATCH_QUEUES * GEMMX_BATCH_CTR_STRIDE)) != hipSuccess) {
    (void)hipGetLastError();
    rc = gemmx_report(GEMMX_ERR_ALLOC, "batch: device allocation failed: %s", hipGetErrorString(e));
    goto fail;
  }
  if ((e = hipMemcpy(p->d_desc, h, sizeof *h * (used > 0 ? used : 1), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->d_tile_prob, tp, sizeof(int) * (total > 0 ? total : 1), hipMemcpyHostToDevice)) != hipSuccess) {
    rc = gemmx_report(GEMMX_ERR_LAUNCH, "batch: descriptor upload failed: %s", hipGetErrorString(e));
    goto fail;
  }
  p->grid = gemmx_batch_grid(es, cfg);
  if (p->grid <= 0) {
    rc = gemmx_report(GEMMX_ERR_LAUNCH, "batch: occupancy query failed");
    goto fail;
  }
  /* fewer workgroups than tiles are never needed; a multiple of the queue
   * count keeps every XCD's workgroups on its own queue */
  if (p->grid > p->total_tiles)
    p->grid = (p->total_tiles + GEMMX_BATCH_QUEUES - 1) / GEMMX_BATCH_QUEUES * GEMMX_BATCH_QUEUES;
  /* Capture ticket reset + persistent kernel into a graph (no memset nodes:
   * the ticket is cleared by a kernel). */
  if ((e = hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
    rc = gemmx_report(GEMMX_ERR_LAUNCH, "batch: graph capture failed: %s", hipGetErrorString(e));
    goto fail;
  }
  const hipError_t le = p->total_tiles > 0
                           ? gemmx_batch_launch(es, cfg, p->d_desc, p->d_tile_prob, p->d_sched,
                                                &p->queues, p->grid, p->capture)
                           : hipSuccess;
  e = hipStreamEndCapture(p->capture, &p->graph);
  if (le != hipSuccess) e = le;
  if (e != hipSuccess || (e = hipGraphInstantiate(&p->exec, p->graph, NULL, NULL, 0)) != hipSuccess) {
    rc = gemmx_report(GEMMX_ERR_LAUNCH, "batch: graph instantiation failed: %s", hipGetErrorString(e));
    goto fail;
  }
  free(h);
  free(ord);
  free(tp);
  *out = p;
  return GEMMX_OK;
fail:
  free(h);
  free(ord);
  free(tp);
  destroy(p);
  return rc;
}

static int run(void *plan, void *stream) {
  gemmx_clear_error();
  struct gemmx_batch *p = (struct gemmx_batch *)plan;
  if (!p) return gemmx_report(GEMMX_ERR_ARG, "batch: NULL plan");
  if (p->total_tiles == 0) return GEMMX_OK;
  int dev;
  if (gemmx_current_device(&dev) != GEMMX_OK) return gemmx_last_code();
  if (dev != p->device)
    return gemmx_report(GEMMX_ERR_ARG, "batch: plan built on device %d, current %d", p->device, dev);
  hipError_t e = hipGraphLaunch(p->exec, (hipStream_t)stream);
  if (e != hipSuccess)
    return gemmx_report(GEMMX_ERR_LAUNCH, "batch: graph launch failed: %s", hipGetErrorString(e));
  return GEMMX_OK;
}

EXPORT int dgemm_x_batch_create(void **plan, int count, const char *transa, const char *transb,
                                const int *m, const int *n, const int *k, const double *alpha,
                                const double *const *a, const int *lda,
                                const double *const *b, const int *ldb, const double *beta,
                                double *const *c, const int *ldc) {
  return create(8, plan, count, transa, transb, m, n, k, alpha, (const void *const *)a, lda,
                (const void *const *)b, ldb, beta, (void *const *)c, ldc);
}

EXPORT int sgemm_x_batch_create(void **plan, int count, const char *transa, const char *transb,
                                const int *m, const int *n, const int *k, const float *alpha,
                                const float *const *a, const int *lda, const float *const *b,
                                const int *ldb, const float *beta, float *const *c,
                                const int *ldc) {
  return create(4, plan, count, transa, transb, m, n, k, alpha, (const void *const *)a, lda,
                (const void *const *)b, ldb, beta, (void *const *)c, ldc);
}

EXPORT int dgemm_x_batch_run(void *plan, void *stream) { return run(plan, stream); }
EXPORT int sgemm_x_batch_run(void *plan, void *stream) { return run(plan, stream); }

EXPORT double dgemm_x_batch_flops(const void *plan) {
  return plan ? ((const struct gemmx_batch *)plan)->flops : 0.0;
}
EXPORT double sgemm_x_batch_flops(const void *plan) {
  return plan ? ((const struct gemmx_batch *)plan)->flops : 0.0;
}

EXPORT int dgemm_x_batch_set_config(int cfg) { return set_batch_cfg(8, cfg); }
EXPORT int sgemm_x_batch_set_config(int cfg) { return set_batch_cfg(4, cfg); }
/* The configuration a plan created now would use (no device needed). */
EXPORT int dgemm_x_batch_config(void) {
  gemmx_init();
  return batch_cfg(8);
}
EXPORT int sgemm_x_batch_config(void) {
  gemmx_init();
  return batch_cfg(4);
}

EXPORT int dgemm_x_batch_destroy(void *plan) {
  destroy((struct gemmx_batch *)plan);
  return GEMMX_OK;
}
EXPORT int sgemm_x_batch_destroy(void *plan) {
  destroy((struct gemmx_batch *)plan);
  return GEMMX_OK;
}
