/*
 * gemmx_stage.c -- host<->device staging of the Fortran-ABI path for
 * operands in pageable host memory (the reference's callers pass host
 * arrays, src/gemm_driver.c:118).
 *
 * A column-major block (cols columns of row_bytes bytes, leading dimension
 * src_ld bytes) is moved in chunks of whole columns through a ring of pinned
 * host buffers: the host copies chunk i into buffer i % NBUF with a small
 * pool of copy threads while the DMA engine moves earlier chunks, so host
 * copies and PCIe transfers overlap.  Device->host runs the same ring the
 * other way: the DMA of chunk i + NBUF - 1 is queued before the host copies
 * chunk i out.  Uploads and downloads have rings of their own (buffers, copy
 * stream, events), so the two directions overlap.  Pinned buffers are
 * allocated once per device and direction on first use (hipHostMalloc) and
 * kept until the process ends.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/gemm_mi355x.h"
#include "gemmx_internal.h"

#define NBUF 3
#define MAX_DEV 64
#define MAX_THREADS 16

/* ---------------------------------------------------------------- copy pool */
typedef struct {
  char *dst;
  const char *src;
  size_t dld, sld, row_bytes, cols;
} copy_job_t;

static pthread_once_t g_pool_once = PTHREAD_ONCE_INIT;
static pthread_mutex_t g_pool_lock = PTHREAD_MUTEX_INITIALIZER;
static pthread_mutex_t g_job_lock = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t g_job_cv = PTHREAD_COND_INITIALIZER;
static pthread_cond_t g_done_cv = PTHREAD_COND_INITIALIZER;
static int g_nthreads;         /* workers besides the calling thread */
static copy_job_t g_job;
static unsigned long g_job_seq; /* bumped per job */
static int g_parts, g_done;     /* parts of the current job, parts finished */

static void copy_part(const copy_job_t *j, int part, int parts) {
  const size_t c0 = j->cols * part / parts, c1 = j->cols * (part + 1) / parts;
  if (j->dld == j->row_bytes && j->sld == j->row_bytes) {
    memcpy(j->dst + c0 * j->dld, j->src + c0 * j->sld, (c1 - c0) * j->row_bytes);
    return;
  }
  for (size_t c = c0; c < c1; c++) memcpy(j->dst + c * j->dld, j->src + c * j->sld, j->row_bytes);
}

static void *worker(void *arg) {
  const int id = (int)(intptr_t)arg; /* 1 .. g_nthreads; the caller is part 0 */
  unsigned long seen = 0;
  for (;;) {
    pthread_mutex_lock(&g_job_lock);
    while (g_job_seq == seen) pthread_cond_wait(&g_job_cv, &g_job_lock);
    seen = g_job_seq;
    const copy_job_t j = g_job;
    const int parts = g_parts;
    pthread_mutex_unlock(&g_job_lock);
    if (id < parts) copy_part(&j, id, parts);
    pthread_mutex_lock(&g_job_lock);
    if (++g_done == g_nthreads) pthread_cond_signal(&g_done_cv);
    pthread_mutex_unlock(&g_job_lock);
  }
  return NULL;
}

static void pool_init(void) {
  const char *e = getenv("GEMMX_COPY_THREADS");
  int n = e && *e ? atoi(e) : 8;
  if (n < 1) n = 1;
  if (n > MAX_THREADS) n = MAX_THREADS;
  g_nthreads = 0;
  for (int i = 1; i < n; i++) {
    pthread_t t;
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    pthread_attr_setdetachstate(&attr, PTHREAD_CREATE_DETACHED);
    if (pthread_create(&t, &attr, worker, (void *)(intptr_t)(g_nthreads + 1)) == 0) g_nthreads++;
    pthread_attr_destroy(&attr);
  }
}

/* dst[c*dld + r] = src[c*sld + r] for r < row_bytes, c < cols, split over the
 * pool (one job at a time per process). */
static void parallel_copy(char *dst, size_t dld, const char *src, size_t sld, size_t row_bytes,
                          size_t cols) {
  pthread_once(&g_pool_once, pool_init);
  const size_t bytes = row_bytes * cols;
  if (g_nthreads == 0 || bytes < ((size_t)1 << 20)) {
    copy_job_t j = {dst, src, dld, sld, row_bytes, cols};
    copy_part(&j, 0, 1);
    return;
  }
  pthread_mutex_lock(&g_pool_lock);
  const int parts = (size_t)(g_nthreads + 1) < cols ? g_nthreads + 1 : (int)cols;
  pthread_mutex_lock(&g_job_lock);
  g_job = (copy_job_t){dst, src, dld, sld, row_bytes, cols};
  g_parts = parts;
  g_done = 0;
  g_job_seq++;
  pthread_cond_broadcast(&g_job_cv);
  pthread_mutex_unlock(&g_job_lock);
  copy_part(&g_job, 0, parts);
  pthread_mutex_lock(&g_job_lock);
  while (g_done < g_nthreads) pthread_cond_wait(&g_done_cv, &g_job_lock);
  pthread_mutex_unlock(&g_job_lock);
  pthread_mutex_unlock(&g_pool_lock);
}

/* ---------------------------------------------------------------- ring */
typedef struct {
  void *buf[NBUF];
  hipEvent_t ev[NBUF]; /* the last DMA from or into buf[i] */
  hipEvent_t start;    /* ordering marker on the compute stream */
  size_t bytes;
  hipStream_t copy;
  int ready;
} ring_t;

/* One ring per direction and device (0 uploads, 1 downloads): each has its
 * own copy stream, so a download can run while uploads are in flight (PCIe
 * carries both directions at once). */
static ring_t g_ring[2][MAX_DEV];

size_t gemmx_stage_chunk_bytes(void) {
  const char *e = getenv("GEMMX_STAGE_CHUNK_MB");
  long mb = e && *e ? atol(e) : 32;
  if (mb < 1) mb = 1;
  if (mb > 1024) mb = 1024;
  return (size_t)mb << 20;
}

/* The device's ring with buffers of at least `bytes`; caller holds the
 * device lock. */
static int ring_get(int dev, int dir, size_t bytes, ring_t **out) {
  ring_t *r = &g_ring[dir][dev];
  if (!r->ready) {
    if (hipStreamCreateWithFlags(&r->copy, hipStreamNonBlocking) != hipSuccess)
      return gemmx_report(GEMMX_ERR_LAUNCH, "staging: stream creation failed");
    for (int i = 0; i < NBUF; i++)
      if (hipEventCreateWithFlags(&r->ev[i], hipEventDisableTiming) != hipSuccess)
        return gemmx_report(GEMMX_ERR_LAUNCH, "staging: event creation failed");
    if (hipEventCreateWithFlags(&r->start, hipEventDisableTiming) != hipSuccess)
      return gemmx_report(GEMMX_ERR_LAUNCH, "staging: event creation failed");
    r->ready = 1;
  }
  if (r->bytes < bytes) {
    for (int i = 0; i < NBUF; i++) {
      if (r->buf[i]) {
        (void)hipEventSynchronize(r->ev[i]);
        (void)hipHostFree(r->buf[i]);
        r->buf[i] = NULL;
      }
    }
    r->bytes = 0;
    for (int i = 0; i < NBUF; i++) {
      if (hipHostMalloc(&r->buf[i], bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        for (int j = 0; j < i; j++) {
          (void)hipHostFree(r->buf[j]);
          r->buf[j] = NULL;
        }
        return gemmx_report(GEMMX_ERR_ALLOC, "staging: pinned allocation of %zu bytes failed",
                            bytes);
      }
    }
    r->bytes = bytes;
  }
  *out = r;
  return GEMMX_OK;
}

/* Columns per chunk for a block of row_bytes-long columns (0: one column is
 * larger than a chunk). */
static size_t cols_per_chunk(size_t row_bytes, size_t chunk) {
  return row_bytes ? chunk / row_bytes : 0;
}

int gemmx_stage_h2d(int dev, void *dst, size_t dld, const void *src, size_t sld,
                    size_t row_bytes, size_t cols, hipStream_t compute, hipEvent_t after) {
  const size_t chunk = gemmx_stage_chunk_bytes();
  const size_t cpc = cols_per_chunk(dld, chunk);
  hipError_t e;
  if (cpc == 0 || row_bytes * cols < ((size_t)4 << 20)) {
    /* small blocks, or columns longer than a chunk: one runtime copy */
    e = hipMemcpy2DAsync(dst, dld, src, sld, row_bytes, cols, hipMemcpyHostToDevice, compute);
    return e == hipSuccess ? GEMMX_OK
                           : gemmx_report(GEMMX_ERR_LAUNCH, "copy to the device failed: %s",
                                          hipGetErrorString(e));
  }
  ring_t *r;
  int rc = ring_get(dev, 0, chunk, &r);
  if (rc != GEMMX_OK) return rc;
  /* the copy stream starts after `after`, or else after the work already
   * queued on the compute stream (the destination may still be in use) */
  if (!after && (e = hipEventRecord(r->start, compute)) != hipSuccess)
    return gemmx_report(GEMMX_ERR_LAUNCH, "staging: stream ordering failed: %s",
                        hipGetErrorString(e));
  if ((e = hipStreamWaitEvent(r->copy, after ? after : r->start, 0)) != hipSuccess)
    return gemmx_report(GEMMX_ERR_LAUNCH, "staging: stream ordering failed: %s",
                        hipGetErrorString(e));
  size_t i = 0;
  for (size_t c0 = 0; c0 < cols; c0 += cpc, i++) {
    const size_t nc = cols - c0 < cpc ? cols - c0 : cpc;
    const int b = (int)(i % NBUF);
    /* buf[b] may still feed a DMA of this or an earlier block */
    if ((e = hipEventSynchronize(r->ev[b])) != hipSuccess) break;
    parallel_copy((char *)r->buf[b], dld, (const char *)src + c0 * sld, sld, row_bytes, nc);
    /* a block narrower than its leading dimension (a row panel) must leave
     * the device rows outside it alone: copy row_bytes per column */
    e = row_bytes < dld
            ? hipMemcpy2DAsync((char *)dst + c0 * dld, dld, r->buf[b], dld, row_bytes, nc,
                               hipMemcpyHostToDevice, r->copy)
            : hipMemcpyAsync((char *)dst + c0 * dld, r->buf[b], nc * dld, hipMemcpyHostToDevice,
                             r->copy);
    if (e != hipSuccess ||
        (e = hipEventRecord(r->ev[b], r->copy)) != hipSuccess)
      break;
  }
  if (e == hipSuccess) {
    /* the compute stream waits for the last copy (the copies are in order) */
    const int last = (int)((i + NBUF - 1) % NBUF);
    e = hipStreamWaitEvent(compute, r->ev[last], 0);
  }
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(r->copy);
    return gemmx_report(GEMMX_ERR_LAUNCH, "staged copy to the device failed: %s",
                        hipGetErrorString(e));
  }
  return GEMMX_OK;
}

int gemmx_stage_d2h(int dev, void *dst, size_t dld, const void *src, size_t sld,
                    size_t row_bytes, size_t cols, hipStream_t compute, hipEvent_t after) {
  const size_t chunk = gemmx_stage_chunk_bytes();
  const size_t cpc = cols_per_chunk(sld, chunk);
  hipError_t e;
  if (cpc == 0 || row_bytes * cols < ((size_t)4 << 20)) {
    e = hipMemcpy2DAsync(dst, dld, src, sld, row_bytes, cols, hipMemcpyDeviceToHost, compute);
    if (e == hipSuccess) e = hipStreamSynchronize(compute);
    return e == hipSuccess ? GEMMX_OK
                           : gemmx_report(GEMMX_ERR_LAUNCH, "copy to the host failed: %s",
                                          hipGetErrorString(e));
  }
  ring_t *r;
  int rc = ring_get(dev, 1, chunk, &r);
  if (rc != GEMMX_OK) return rc;
  const size_t nchunks = (cols + cpc - 1) / cpc;
  /* the copies start after `after`, or else after the computation queued on
   * the compute stream; the copy stream is in order, so they also follow
   * every earlier DMA of the ring */
  if (!after && (e = hipEventRecord(r->start, compute)) != hipSuccess)
    return gemmx_report(GEMMX_ERR_LAUNCH, "staging: stream ordering failed: %s",
                        hipGetErrorString(e));
  if ((e = hipStreamWaitEvent(r->copy, after ? after : r->start, 0)) != hipSuccess)
    return gemmx_report(GEMMX_ERR_LAUNCH, "staging: stream ordering failed: %s",
                        hipGetErrorString(e));
  /* queue the first NBUF chunks, then for each chunk: wait, copy out, queue
   * the chunk NBUF ahead into the freed buffer */
  size_t queued = 0;
  for (; queued < nchunks && queued < NBUF && e == hipSuccess; queued++) {
    const size_t c0 = queued * cpc, nc = cols - c0 < cpc ? cols - c0 : cpc;
    const int b = (int)(queued % NBUF);
    e = hipMemcpyAsync(r->buf[b], (const char *)src + c0 * sld, nc * sld,
                       hipMemcpyDeviceToHost, r->copy);
    if (e == hipSuccess) e = hipEventRecord(r->ev[b], r->copy);
  }
  for (size_t i = 0; i < nchunks && e == hipSuccess; i++) {
    const size_t c0 = i * cpc, nc = cols - c0 < cpc ? cols - c0 : cpc;
    const int b = (int)(i % NBUF);
    if ((e = hipEventSynchronize(r->ev[b])) != hipSuccess) break;
    parallel_copy((char *)dst + c0 * dld, dld, (const char *)r->buf[b], sld, row_bytes, nc);
    if (queued < nchunks) {
      const size_t q0 = queued * cpc, qn = cols - q0 < cpc ? cols - q0 : cpc;
      e = hipMemcpyAsync(r->buf[b], (const char *)src + q0 * sld, qn * sld,
                         hipMemcpyDeviceToHost, r->copy);
      if (e == hipSuccess) e = hipEventRecord(r->ev[b], r->copy);
      queued++;
    }
  }
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(r->copy);
    return gemmx_report(GEMMX_ERR_LAUNCH, "staged copy to the host failed: %s",
                        hipGetErrorString(e));
  }
  return GEMMX_OK;
}
