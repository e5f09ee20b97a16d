/*
 * gemmx_api.c -- host C layer of DGEMM.so / SGEMM.so.
 *
 * Replaces the reference's API/driver layer, GEMM_DRIVER (wjc404/GEMM_AVX2
 * src/gemm_driver.c:118-197): argument decoding (:121-127), the beta rule
 * (:133, C_MULT_BETA :106-116), the alpha gate (:134) and the per-call scratch
 * (:135-142).  The OpenMP team, the N-split and the spin barrier
 * (:23-33, :128-160) have no host counterpart here: the tile kernels schedule
 * themselves over the GPU's compute units.
 *
 * Additions of this layer: pointer classification (host operands are staged
 * to device memory and C is copied back; device operands are used in place),
 * argument checking with error reporting, process-wide arithmetic mode
 * (native / emulated), and the device-pointer extension entry points.
 */
#define _GNU_SOURCE
#include <limits.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/gemm_mi355x.h"
#include "gemmx_internal.h"
#include "gemmx_tune.h"

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Errors: per-thread last error, printed once per error code per process.  */
/* ------------------------------------------------------------------------ */
static __thread int t_err;
static __thread char t_msg[320];
static pthread_mutex_t g_report_lock = PTHREAD_MUTEX_INITIALIZER;
static unsigned g_reported;

static int set_error_v(int code, const char *fmt, va_list ap) {
  vsnprintf(t_msg, sizeof t_msg, fmt, ap);
  t_err = code;
  const char *q = getenv("GEMMX_QUIET");
  if (!(q && *q && *q != '0')) {
    pthread_mutex_lock(&g_report_lock);
    if (!(g_reported & (1u << code))) {
      g_reported |= 1u << code;
      fprintf(stderr, "gemm_mi355x: %s\n", t_msg);
    }
    pthread_mutex_unlock(&g_report_lock);
  }
  return code;
}

static int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  set_error_v(code, fmt, ap);
  va_end(ap);
  return code;
}

static void clear_error(void) {
  t_err = GEMMX_OK;
  t_msg[0] = 0;
}

/* The same, for the other host sources of the library (gemmx_batch.c). */
int gemmx_report(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  set_error_v(code, fmt, ap);
  va_end(ap);
  return code;
}
void gemmx_clear_error(void) { clear_error(); }
int gemmx_last_code(void) { return t_err; }

static int hip_fail(hipError_t e, const char *what) {
  return set_error(e == hipErrorOutOfMemory ? GEMMX_ERR_ALLOC : GEMMX_ERR_LAUNCH,
                   "%s failed: %s", what, hipGetErrorString(e));
}

/* ------------------------------------------------------------------------ */
/* Process state: lazy HIP init, per-device scratch and staging stream.     */
/* ------------------------------------------------------------------------ */
#define GEMMX_MAX_DEV 64
typedef struct {
  pthread_mutex_t lock; /* held for the whole of a call that uses scratch */
  void *scratch;
  size_t scratch_bytes;
  int pool_tried;      /* stream-ordered pool of split-K partials (dev_call) */
  hipMemPool_t pool;
} devstate_t;

static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static int g_ndev;
static hipError_t g_init_err;
static devstate_t g_dev[GEMMX_MAX_DEV];
static int g_dmode = GEMMX_MODE_NATIVE, g_dslices = 8;
static int g_smode = GEMMX_MODE_NATIVE;

static void init_once(void) {
  for (int i = 0; i < GEMMX_MAX_DEV; i++) pthread_mutex_init(&g_dev[i].lock, NULL);
  const char *dm = getenv("GEMMX_DGEMM_MODE");
  if (dm && strcmp(dm, "auto") == 0) {
    g_dmode = GEMMX_MODE_AUTO;
  } else if (dm && strncmp(dm, "ozaki2", 6) == 0) {
    g_dmode = GEMMX_MODE_OZAKI2;
    g_dslices = 20;
    if (dm[6] == ':') {
      int s = atoi(dm + 7);
      if (s >= 8 && s <= 22) g_dslices = s;
    }
  } else if (dm && strncmp(dm, "ozaki", 5) == 0) {
    g_dmode = GEMMX_MODE_OZAKI;
    if (dm[5] == ':') {
      int s = atoi(dm + 6);
      if (s >= 3 && s <= 8) g_dslices = s;
    }
  }
  const char *sm = getenv("GEMMX_SGEMM_MODE");
  if (sm && strcmp(sm, "bf16x3") == 0) g_smode = GEMMX_MODE_BF16X3;
  if (sm && strcmp(sm, "auto") == 0) g_smode = GEMMX_MODE_AUTO;
  g_init_err = hipGetDeviceCount(&g_ndev);
  if (g_init_err != hipSuccess) {
    g_ndev = 0;
    (void)hipGetLastError();
  }
  /* the tuning table is checked against the arch of device 0 */
  hipDeviceProp_t prop;
  const int have_arch = g_ndev > 0 && hipGetDeviceProperties(&prop, 0) == hipSuccess;
  gemmx_tune_init(have_arch ? prop.gcnArchName : NULL);
}

static int current_device(int *dev) {
  pthread_once(&g_once, init_once);
  if (g_ndev <= 0)
    return set_error(GEMMX_ERR_NODEVICE, "no HIP device available (%s)",
                     hipGetErrorString(g_init_err));
  hipError_t e = hipGetDevice(dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (*dev < 0 || *dev >= GEMMX_MAX_DEV)
    return set_error(GEMMX_ERR_UNSUPPORTED, "device ordinal %d out of range", *dev);
  return GEMMX_OK;
}

int gemmx_current_device(int *dev) { return current_device(dev); }
void gemmx_init(void) { pthread_once(&g_once, init_once); }

/* GEMMX_FAIL_ALLOC=1 makes the library's device allocations fail as if the
 * device were out of memory (failure-path testing; read at every attempt). */
int gemmx_alloc_failure_injected(void) {
  const char *fail = getenv("GEMMX_FAIL_ALLOC");
  return fail && *fail && *fail != '0';
}

/* Grows the device's scratch to `bytes`; caller holds g_dev[dev].lock and
 * has no work in flight on the old scratch. */
static int get_scratch(int dev, size_t bytes, void **out) {
  devstate_t *d = &g_dev[dev];
  if (bytes > d->scratch_bytes) {
    if (gemmx_alloc_failure_injected())
      return set_error(GEMMX_ERR_ALLOC, "device scratch of %zu bytes: allocation failure injected", bytes);
    if (d->scratch) {
      hipError_t e = hipFree(d->scratch);
      d->scratch = NULL;
      d->scratch_bytes = 0;
      if (e != hipSuccess) return hip_fail(e, "hipFree(scratch)");
    }
    size_t want = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
    hipError_t e = hipMalloc(&d->scratch, want);
    if (e != hipSuccess) {
      d->scratch = NULL;
      (void)hipGetLastError();
      return set_error(GEMMX_ERR_ALLOC, "hipMalloc of %zu bytes of device scratch failed: %s",
                       want, hipGetErrorString(e));
    }
    d->scratch_bytes = want;
  }
  *out = d->scratch;
  return GEMMX_OK;
}

static int is_device_ptr(const void *p) {
  if (!p) return 0;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
         attr.type == hipMemoryTypeUnified;
}

/* ------------------------------------------------------------------------ */
/* Phase timing of the emulated DGEMMs (per thread, off by default).        */
/* ------------------------------------------------------------------------ */
static __thread int t_timing;
static __thread int t_ev_dev = -1;
static __thread int t_ev_used;
static __thread hipEvent_t t_ev[4];

/* The calling thread's four events on the current device when timing is on,
 * else NULL. */
static hipEvent_t *phase_events(void) {
  if (!t_timing) return NULL;
  int dev;
  if (hipGetDevice(&dev) != hipSuccess) return NULL;
  if (t_ev_dev != dev) {
    if (t_ev_dev >= 0)
      for (int i = 0; i < 4; i++) (void)hipEventDestroy(t_ev[i]);
    t_ev_dev = -1;
    for (int i = 0; i < 4; i++)
      if (hipEventCreate(&t_ev[i]) != hipSuccess) {
        for (int j = 0; j < i; j++) (void)hipEventDestroy(t_ev[j]);
        return NULL;
      }
    t_ev_dev = dev;
  }
  t_ev_used = 1;
  return t_ev;
}

/* ------------------------------------------------------------------------ */
/* Arguments.                                                               */
/* ------------------------------------------------------------------------ */
static int notrans(char c) { return c == 'N' || c == 'n'; }

typedef struct {
  int es; /* element size: 8 or 4 */
  char ta, tb;
  int m, n, k;
  int64_t lda, ldb, ldc;
  double alpha, beta;
  const void *a, *b;
  void *c;
} call_t;

static int validate(const call_t *x) {
  const int ra = notrans(x->ta) ? x->m : x->k;
  const int rb = notrans(x->tb) ? x->k : x->n;
  if (x->m < 0 || x->n < 0 || x->k < 0)
    return set_error(GEMMX_ERR_ARG, "negative dimension (m=%d n=%d k=%d)", x->m, x->n, x->k);
  if (x->m == 0 || x->n == 0) return GEMMX_OK;
  if (x->ldc < (x->m > 1 ? x->m : 1))
    return set_error(GEMMX_ERR_ARG, "ldc=%lld < max(1,m=%d)", (long long)x->ldc, x->m);
  if (!x->c) return set_error(GEMMX_ERR_ARG, "C is NULL");
  if (x->alpha != 0.0 && x->k > 0) {
    if (x->lda < (ra > 1 ? ra : 1))
      return set_error(GEMMX_ERR_ARG, "lda=%lld < max(1,%d)", (long long)x->lda, ra);
    if (x->ldb < (rb > 1 ? rb : 1))
      return set_error(GEMMX_ERR_ARG, "ldb=%lld < max(1,%d)", (long long)x->ldb, rb);
    if (!x->a || !x->b) return set_error(GEMMX_ERR_ARG, "A or B is NULL");
  }
  return GEMMX_OK;
}

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

/* Launches the arithmetic for device-resident operands (already checked). */
static int launch_device(const gemmx_problem *p, int es, int mode, int slices,
                         void *work, hipStream_t s) {
  hipError_t e;
  if (p->alpha == 0.0 || p->k == 0) {
    if (p->beta == 1.0) return GEMMX_OK;
    e = es == 8 ? gemmx_scale_f64((double *)p->c, p->ldc, p->m, p->n, p->beta, p->batch,
                                  p->sc, s)
                : gemmx_scale_f32((float *)p->c, p->ldc, p->m, p->n, (float)p->beta, p->batch,
                                  p->sc, s);
    return e == hipSuccess ? GEMMX_OK : hip_fail(e, "scale kernel launch");
  }
  if (es == 8 && mode == GEMMX_MODE_OZAKI) {
    e = gemmx_ozaki_f64(p, slices, work, s, phase_events());
    return e == hipSuccess ? GEMMX_OK : hip_fail(e, "ozaki kernel launch");
  }
  if (es == 8 && mode == GEMMX_MODE_OZAKI2) {
    e = gemmx_ozaki2_f64(p, slices, work, s, phase_events());
    return e == hipSuccess ? GEMMX_OK : hip_fail(e, "ozaki2 kernel launch");
  }
  if (es == 4 && mode == GEMMX_MODE_BF16X3) {
    e = gemmx_bf16x3_f32(p, work, s);
    return e == hipSuccess ? GEMMX_OK : hip_fail(e, "bf16x3 kernel launch");
  }
  const int cfg = p->batch > 1
                      ? gemmx_tune_pick_strided(es, p->ta, p->tb, p->m, p->n, p->k, p->batch)
                      : gemmx_tune_pick(es, p->ta, p->tb, p->m, p->n, p->k);
  /* split-K needs the partial tiles' workspace; without one: one pass */
  const int split = (p->batch > 1 || !work)
                        ? 1 : gemmx_tune_pick_split(es, p->ta, p->tb, p->m, p->n, p->k, cfg);
  e = es == 8 ? gemmx_native_f64(p, cfg, split, work, s)
              : gemmx_native_f32(p, cfg, split, work, s);
  return e == hipSuccess ? GEMMX_OK : hip_fail(e, "native kernel launch");
}

/* Bytes of split-K partial tiles a native call needs (0: one pass). */
static size_t native_ws(int es, int ta, int tb, int m, int n, int k) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  const int cfg = gemmx_tune_pick(es, ta, tb, m, n, k);
  const int split = gemmx_tune_pick_split(es, ta, tb, m, n, k, cfg);
  return gemmx_native_workspace(es, cfg, m, n, k, split);
}

/* Device scratch of a call: the emulated forms' workspace, or the native
 * path's split-K partial tiles. */
static size_t call_ws(const call_t *x, int mode, int slices) {
  if (x->alpha == 0.0 || x->k == 0 || x->m == 0 || x->n == 0) return 0;
  if (mode == GEMMX_MODE_NATIVE || (x->es == 4 && mode != GEMMX_MODE_BF16X3) ||
      (x->es == 8 && mode == GEMMX_MODE_BF16X3))
    return native_ws(x->es, !notrans(x->ta), !notrans(x->tb), x->m, x->n, x->k);
  if (x->es == 8 && mode == GEMMX_MODE_OZAKI)
    return gemmx_ozaki_workspace(x->m, x->n, x->k, slices);
  if (x->es == 8 && mode == GEMMX_MODE_OZAKI2)
    return gemmx_ozaki2_workspace(x->m, x->n, x->k, slices);
  if (x->es == 4 && mode == GEMMX_MODE_BF16X3)
    return gemmx_bf16x3_workspace(x->m, x->n, x->k);
  return 0;
}

/* Column panels of C for a native call with host B or C: about eight
 * panels, multiples of 256 columns; 0 (no panels) for emulated arithmetic,
 * narrow C or less than 64 MB of B and C to move. */
#define GEMMX_MAX_PANELS 16
static int panel_cols(const call_t *x, int mode, int need_ab, int host_b, int host_c) {
  if (mode != GEMMX_MODE_NATIVE || !need_ab || !(host_b || host_c) || x->n < 512) return 0;
  const size_t moved = ((host_b ? (size_t)x->k : 0) + (host_c ? (size_t)x->m : 0)) *
                       (size_t)x->n * (size_t)x->es;
  if (moved < ((size_t)64 << 20)) return 0;
  int np = (x->n + 7) / 8;
  np = (np + 255) / 256 * 256;
  return np < x->n ? np : 0;
}

/* A native call with host operands, pipelined over column panels of C: A is
 * staged first; then for each panel j the panel of op(B) and of C are staged
 * and its GEMM is queued, and the previous panel's C is copied back after an
 * event recorded behind that panel's GEMM, so the transfers of panel j + 1
 * and j - 1 run while the GPU computes panel j.  p holds the device
 * operands (staged buffers for host operands: A of leading dimension p->lda,
 * op(B) columns j0.. at column j0 of B (transb N) or row j0 (transb T)). */
static int panelled_call(const call_t *x, const gemmx_problem *p, int dev, int np, int host_a,
                         int host_b, int host_c, hipStream_t s) {
  const size_t es = (size_t)x->es;
  const int np_count = (x->n + np - 1) / np;
  if (np_count > GEMMX_MAX_PANELS) return set_error(GEMMX_ERR_UNSUPPORTED, "too many panels");
  hipEvent_t start = NULL, done[GEMMX_MAX_PANELS] = {0};
  int rc = GEMMX_OK, made = 0;
  hipError_t e = hipEventCreateWithFlags(&start, hipEventDisableTiming);
  for (; e == hipSuccess && made < np_count; made++)
    e = hipEventCreateWithFlags(&done[made], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(start, s);
  if (e != hipSuccess) {
    rc = hip_fail(e, "panel events");
    goto end;
  }
  const int ra = notrans(x->ta) ? x->m : x->k, ca = notrans(x->ta) ? x->k : x->m;
  if (host_a && (rc = gemmx_stage_h2d(dev, (void *)p->a, (size_t)p->lda * es, x->a,
                                      (size_t)x->lda * es, (size_t)ra * es, (size_t)ca, s,
                                      start)) != GEMMX_OK)
    goto end;
  for (int j = 0; j <= np_count && rc == GEMMX_OK; j++) {
    if (j < np_count) {
      const int j0 = j * np, nj = x->n - j0 < np ? x->n - j0 : np;
      /* op(B) columns j0 .. j0 + nj - 1: columns of B, or rows of B^T */
      const size_t boff_d = p->tb ? (size_t)j0 * es : (size_t)j0 * (size_t)p->ldb * es;
      const size_t boff_h = p->tb ? (size_t)j0 * es : (size_t)j0 * (size_t)x->ldb * es;
      if (host_b &&
          (rc = p->tb ? gemmx_stage_h2d(dev, (char *)p->b + boff_d, (size_t)p->ldb * es,
                                        (const char *)x->b + boff_h, (size_t)x->ldb * es,
                                        (size_t)nj * es, (size_t)x->k, s, start)
                      : gemmx_stage_h2d(dev, (char *)p->b + boff_d, (size_t)p->ldb * es,
                                        (const char *)x->b + boff_h, (size_t)x->ldb * es,
                                        (size_t)x->k * es, (size_t)nj, s, start)) != GEMMX_OK)
        break;
      if (host_c &&
          (rc = gemmx_stage_h2d(dev, (char *)p->c + (size_t)j0 * (size_t)p->ldc * es,
                                (size_t)p->ldc * es,
                                (const char *)x->c + (size_t)j0 * (size_t)x->ldc * es,
                                (size_t)x->ldc * es, (size_t)x->m * es, (size_t)nj, s,
                                start)) != GEMMX_OK)
        break;
      gemmx_problem q = *p;
      q.n = nj;
      q.b = (const char *)p->b + boff_d;
      q.c = (char *)p->c + (size_t)j0 * (size_t)p->ldc * es;
      if ((rc = launch_device(&q, x->es, GEMMX_MODE_NATIVE, 0, NULL, s)) != GEMMX_OK) break;
      if ((e = hipEventRecord(done[j], s)) != hipSuccess) {
        rc = hip_fail(e, "panel event");
        break;
      }
    }
    if (j >= 1 && host_c) { /* panel j - 1 is complete: copy it back */
      const int i0 = (j - 1) * np, ni = x->n - i0 < np ? x->n - i0 : np;
      rc = gemmx_stage_d2h(dev, (char *)x->c + (size_t)i0 * (size_t)x->ldc * es,
                           (size_t)x->ldc * es,
                           (const char *)p->c + (size_t)i0 * (size_t)p->ldc * es,
                           (size_t)p->ldc * es, (size_t)x->m * es, (size_t)ni, s, done[j - 1]);
    }
  }
end:
  e = hipStreamSynchronize(s);
  if (e != hipSuccess && rc == GEMMX_OK) rc = hip_fail(e, "GEMM execution");
  for (int i = 0; i < made; i++)
    if (done[i]) (void)hipEventDestroy(done[i]);
  if (start) (void)hipEventDestroy(start);
  return rc;
}

/* The Fortran-ABI path: synchronous, host or device operands. */
static void fortran_call(call_t *x, int mode, int slices) {
  clear_error();
  if (validate(x) != GEMMX_OK) return;
  if (x->m == 0 || x->n == 0) return;
  if (mode == GEMMX_MODE_AUTO) /* per-shape arithmetic from the tuning table */
    mode = x->es == 8 ? gemmx_tune_pick_emulated(!notrans(x->ta), !notrans(x->tb), x->m, x->n,
                                                 x->k, &slices)
                      : gemmx_tune_pick_emulated32(!notrans(x->ta), !notrans(x->tb), x->m,
                                                   x->n, x->k);
  int dev;
  if (current_device(&dev) != GEMMX_OK) return;

  const int need_ab = x->alpha != 0.0 && x->k > 0;
  const int ra = notrans(x->ta) ? x->m : x->k, ca = notrans(x->ta) ? x->k : x->m;
  const int rb = notrans(x->tb) ? x->k : x->n, cb = notrans(x->tb) ? x->n : x->k;
  const int host_a = need_ab && !is_device_ptr(x->a);
  const int host_b = need_ab && !is_device_ptr(x->b);
  const int host_c = !is_device_ptr(x->c);
  const size_t es = (size_t)x->es;
  /* Staged operands get a leading dimension rounded to 64 bytes so the tile
   * kernels can use 16-byte loads. */
  const size_t al = 64 / es;
  const size_t lda2 = host_a ? round_up((size_t)ra, al) : (size_t)x->lda;
  const size_t ldb2 = host_b ? round_up((size_t)rb, al) : (size_t)x->ldb;
  const size_t ldc2 = host_c ? round_up((size_t)x->m, al) : (size_t)x->ldc;
  const size_t bytes_a = host_a ? round_up(lda2 * ca * es, 256) : 0;
  const size_t bytes_b = host_b ? round_up(ldb2 * cb * es, 256) : 0;
  const size_t bytes_c = host_c ? round_up(ldc2 * x->n * es, 256) : 0;
  const size_t bytes_w = round_up(call_ws(x, mode, slices), 256);
  const size_t total = bytes_a + bytes_b + bytes_c + bytes_w;

  gemmx_problem p = {.ta = !notrans(x->ta), .tb = !notrans(x->tb), .m = x->m,
                     .n = x->n, .k = x->k, .lda = (int64_t)lda2,
                     .ldb = (int64_t)ldb2, .ldc = (int64_t)ldc2,
                     .alpha = x->alpha, .beta = x->beta, .a = x->a, .b = x->b,
                     .c = x->c};
  /* The null (default) stream orders this call after work the caller left
   * on blocking streams; the call waits for it before returning. */
  hipStream_t s = NULL;
  devstate_t *d = &g_dev[dev];
  pthread_mutex_lock(&d->lock);
  char *base = NULL;
  if (total > 0 && get_scratch(dev, total, (void **)&base) != GEMMX_OK) goto out;
  char *pa = base, *pb = pa + bytes_a, *pc = pb + bytes_b, *pw = pc + bytes_c;
  hipError_t e;
  const int np = panel_cols(x, mode, need_ab, host_b, host_c);
  if (np > 0) {
    if (host_a) p.a = pa;
    if (host_b) p.b = pb;
    if (host_c) p.c = pc;
    (void)panelled_call(x, &p, dev, np, host_a, host_b, host_c, s);
    goto out;
  }
  /* Host operands go through the pinned staging ring (gemmx_stage.c). */
  if (host_a) {
    if (gemmx_stage_h2d(dev, pa, lda2 * es, x->a, (size_t)x->lda * es, (size_t)ra * es,
                        (size_t)ca, s, NULL) != GEMMX_OK)
      goto out;
    p.a = pa;
  }
  if (host_b) {
    if (gemmx_stage_h2d(dev, pb, ldb2 * es, x->b, (size_t)x->ldb * es, (size_t)rb * es,
                        (size_t)cb, s, NULL) != GEMMX_OK)
      goto out;
    p.b = pb;
  }
  if (host_c) {
    /* C is always read: beta == 0 still multiplies (NaN propagates). */
    if (gemmx_stage_h2d(dev, pc, ldc2 * es, x->c, (size_t)x->ldc * es, (size_t)x->m * es,
                        (size_t)x->n, s, NULL) != GEMMX_OK)
      goto out;
    p.c = pc;
  }
  if (launch_device(&p, x->es, mode, slices, bytes_w ? pw : NULL, s) != GEMMX_OK) {
    (void)hipStreamSynchronize(s);
    goto out;
  }
  if (host_c && gemmx_stage_d2h(dev, x->c, (size_t)x->ldc * es, pc, ldc2 * es,
                                 (size_t)x->m * es, (size_t)x->n, s, NULL) != GEMMX_OK) {
    (void)hipStreamSynchronize(s);
    goto out;
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) hip_fail(e, "GEMM execution");
out:
  pthread_mutex_unlock(&d->lock);
}

/* ------------------------------------------------------------------------ */
/* Drop-in entry points (src/gemm_driver.c:118).                            */
/* ------------------------------------------------------------------------ */
EXPORT void dgemm_(const char *transa, const char *transb, const int *m,
                   const int *n, const int *k, const double *alpha,
                   const double *a, const int *lda, const double *b,
                   const int *ldb, const double *beta, double *c,
                   const int *ldc) {
  if (!transa || !transb || !m || !n || !k || !alpha || !lda || !ldb || !beta || !ldc) {
    set_error(GEMMX_ERR_ARG, "dgemm_: NULL scalar argument");
    return;
  }
  call_t x = {8, *transa, *transb, *m, *n, *k, *lda, *ldb, *ldc, *alpha, *beta, a, b, c};
  pthread_once(&g_once, init_once);
  fortran_call(&x, g_dmode, g_dslices);
}

EXPORT void sgemm_(const char *transa, const char *transb, const int *m,
                   const int *n, const int *k, const float *alpha,
                   const float *a, const int *lda, const float *b,
                   const int *ldb, const float *beta, float *c,
                   const int *ldc) {
  if (!transa || !transb || !m || !n || !k || !alpha || !lda || !ldb || !beta || !ldc) {
    set_error(GEMMX_ERR_ARG, "sgemm_: NULL scalar argument");
    return;
  }
  call_t x = {4, *transa, *transb, *m, *n, *k, *lda, *ldb, *ldc, *alpha, *beta, a, b, c};
  pthread_once(&g_once, init_once);
  fortran_call(&x, g_smode, 0);
}

/* ------------------------------------------------------------------------ */
/* Complex GEMM: zgemm_/cgemm_ (four real products) and zgemm3m_/cgemm3m_   */
/* (three), the level-3 complex siblings kept outside the reference's build */
/* (OpenBLAS-like_implementation/zgemm_kernel_4x2_haswell.c,                 */
/* zgemm3m_kernel_4x4_haswell.c and their c* forms).  BLAS argument order,   */
/* complex scalars and elements as (re, im) pairs, the transpose rule of     */
/* dgemm_ ('N'/'n' none, anything else transposes) with 'C'/'c' also         */
/* conjugating, and dgemm_'s beta rule (beta != 1 scales C first, also 0).   */
/* ------------------------------------------------------------------------ */
static void complex_call(int es, char ta_c, char tb_c, int m, int n, int k, const double *alpha,
                         const void *a, int lda, const void *b, int ldb, const double *beta,
                         void *c, int ldc, int three) {
  clear_error();
  call_t x = {es, ta_c, tb_c, m, n, k, lda, ldb, ldc, alpha[0], beta[0], a, b, c};
  /* the scalar check of validate reads alpha != 0: either part counts */
  if (alpha[1] != 0.0) x.alpha = 1.0;
  if (validate(&x) != GEMMX_OK) return;
  if (m == 0 || n == 0) return;
  int dev;
  if (current_device(&dev) != GEMMX_OK) return;
  const int product = (alpha[0] != 0.0 || alpha[1] != 0.0) && k > 0;
  const int ta = !notrans(ta_c), tb = !notrans(tb_c);
  const int ra = ta ? k : m, ca = ta ? m : k, rb = tb ? n : k, cb = tb ? k : n;
  const size_t ce = 2 * (size_t)es; /* bytes per complex element */
  const int host_a = product && !is_device_ptr(a), host_b = product && !is_device_ptr(b);
  const int host_c = !is_device_ptr(c);
  const size_t lda2 = host_a ? (size_t)ra : (size_t)lda, ldb2 = host_b ? (size_t)rb : (size_t)ldb;
  const size_t ldc2 = host_c ? (size_t)m : (size_t)ldc;
  const size_t bytes_a = host_a ? round_up(lda2 * ca * ce, 256) : 0;
  const size_t bytes_b = host_b ? round_up(ldb2 * cb * ce, 256) : 0;
  const size_t bytes_c = host_c ? round_up(ldc2 * (size_t)n * ce, 256) : 0;
  const size_t bytes_w = product ? round_up(gemmx_complex_workspace(es, ta, tb, m, n, k, three), 256) : 0;
  const size_t total = bytes_a + bytes_b + bytes_c + bytes_w;
  gemmx_cproblem p = {.ta = ta, .tb = tb, .conja = ta_c == 'C' || ta_c == 'c',
                      .conjb = tb_c == 'C' || tb_c == 'c', .m = m, .n = n, .k = k,
                      .lda = (int64_t)lda2, .ldb = (int64_t)ldb2, .ldc = (int64_t)ldc2,
                      .alpha = {alpha[0], alpha[1]}, .beta = {beta[0], beta[1]},
                      .a = a, .b = b, .c = c};
  hipStream_t s = NULL;
  devstate_t *d = &g_dev[dev];
  pthread_mutex_lock(&d->lock);
  char *base = NULL;
  if (total > 0 && get_scratch(dev, total, (void **)&base) != GEMMX_OK) goto out;
  char *pa = base, *pb = pa + bytes_a, *pc = pb + bytes_b, *pw = pc + bytes_c;
  hipError_t e;
  if (host_a) {
    if (gemmx_stage_h2d(dev, pa, lda2 * ce, a, (size_t)lda * ce, (size_t)ra * ce, (size_t)ca,
                        s, NULL) != GEMMX_OK)
      goto out;
    p.a = pa;
  }
  if (host_b) {
    if (gemmx_stage_h2d(dev, pb, ldb2 * ce, b, (size_t)ldb * ce, (size_t)rb * ce, (size_t)cb,
                        s, NULL) != GEMMX_OK)
      goto out;
    p.b = pb;
  }
  if (host_c) {
    if (gemmx_stage_h2d(dev, pc, ldc2 * ce, c, (size_t)ldc * ce, (size_t)m * ce, (size_t)n,
                        s, NULL) != GEMMX_OK)
      goto out;
    p.c = pc;
  }
  e = gemmx_complex(es, &p, three, pw, s);
  if (e != hipSuccess) {
    hip_fail(e, "complex GEMM launch");
    (void)hipStreamSynchronize(s);
    goto out;
  }
  if (host_c && gemmx_stage_d2h(dev, c, (size_t)ldc * ce, pc, ldc2 * ce, (size_t)m * ce,
                                 (size_t)n, s, NULL) != GEMMX_OK) {
    (void)hipStreamSynchronize(s);
    goto out;
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) hip_fail(e, "complex GEMM execution");
out:
  pthread_mutex_unlock(&d->lock);
}

#define GEMMX_COMPLEX_ENTRY(NAME, T, ES, THREE)                                                   \
  EXPORT void NAME(const char *transa, const char *transb, const int *m, const int *n,          \
                   const int *k, const T *alpha, const T *a, const int *lda, const T *b,        \
                   const int *ldb, const T *beta, T *c, const int *ldc) {                       \
    if (!transa || !transb || !m || !n || !k || !alpha || !lda || !ldb || !beta || !ldc) {     \
      clear_error();                                                                            \
      set_error(GEMMX_ERR_ARG, #NAME ": NULL scalar argument");                                 \
      return;                                                                                   \
    }                                                                                           \
    const double al[2] = {alpha[0], alpha[1]}, be[2] = {beta[0], beta[1]};                      \
    pthread_once(&g_once, init_once);                                                           \
    complex_call(ES, *transa, *transb, *m, *n, *k, al, a, *lda, b, *ldb, be, c, *ldc, THREE);   \
  }
GEMMX_COMPLEX_ENTRY(zgemm_, double, 8, 0)
GEMMX_COMPLEX_ENTRY(zgemm3m_, double, 8, 1)
GEMMX_COMPLEX_ENTRY(cgemm_, float, 4, 0)
GEMMX_COMPLEX_ENTRY(cgemm3m_, float, 4, 1)
#undef GEMMX_COMPLEX_ENTRY

/* ------------------------------------------------------------------------ */
/* Triangular multiply and solve: dtrmm_/dtrsm_ and strmm_/strsm_, the      */
/* level-3 siblings kept outside the reference's build                      */
/* (OpenBLAS-like_implementation/trmm_kernel_4x8_simplified.c,               */
/* strsm_kernel_8x4_haswell_{LN,LT,RN,RT}.c).  BLAS argument order; side    */
/* 'L'/'l' = op(A) on the left, else right; uplo 'U'/'u' = upper stored     */
/* triangle, else lower; transa as for dgemm_; diag 'U'/'u' = unit.  alpha  */
/* = 0 sets B to zero without reading A or B, as BLAS specifies.            */
/* ------------------------------------------------------------------------ */
static void tri_call(int es, int solve, char side, char uplo, char transa, char diag, int m,
                     int n, double alpha, const void *a, int lda, void *b, int ldb) {
  clear_error();
  const int left = side == 'L' || side == 'l';
  const int k = left ? m : n;
  if (m < 0 || n < 0) {
    set_error(GEMMX_ERR_ARG, "negative dimension (m=%d n=%d)", m, n);
    return;
  }
  if (m == 0 || n == 0) return;
  if (ldb < (m > 1 ? m : 1)) {
    set_error(GEMMX_ERR_ARG, "ldb=%d < max(1,m=%d)", ldb, m);
    return;
  }
  if (!b) {
    set_error(GEMMX_ERR_ARG, "B is NULL");
    return;
  }
  if (alpha != 0.0) {
    if (lda < (k > 1 ? k : 1)) {
      set_error(GEMMX_ERR_ARG, "lda=%d < max(1,%d)", lda, k);
      return;
    }
    if (!a) {
      set_error(GEMMX_ERR_ARG, "A is NULL");
      return;
    }
  }
  int dev;
  if (current_device(&dev) != GEMMX_OK) return;
  const int host_a = alpha != 0.0 && !is_device_ptr(a), host_b = !is_device_ptr(b);
  const size_t e8 = (size_t)es;
  const size_t lda2 = host_a ? (size_t)k : (size_t)lda, ldb2 = host_b ? (size_t)m : (size_t)ldb;
  const size_t bytes_a = host_a ? round_up(lda2 * k * e8, 256) : 0;
  const size_t bytes_b = host_b ? round_up(ldb2 * (size_t)n * e8, 256) : 0;
  const size_t bytes_w = alpha != 0.0 ? round_up(gemmx_tri_workspace(es, solve, left, m, n), 256) : 0;
  const size_t total = bytes_a + bytes_b + bytes_w;
  gemmx_tproblem p = {.left = left, .upper = uplo == 'U' || uplo == 'u', .trans = !notrans(transa),
                      .unit = diag == 'U' || diag == 'u', .m = m, .n = n,
                      .lda = (int64_t)lda2, .ldb = (int64_t)ldb2, .alpha = alpha, .a = a,
                      .b = b};
  hipStream_t s = NULL;
  devstate_t *d = &g_dev[dev];
  pthread_mutex_lock(&d->lock);
  char *base = NULL;
  if (total > 0 && get_scratch(dev, total, (void **)&base) != GEMMX_OK) goto out;
  char *pa = base, *pb = pa + bytes_a, *pw = pb + bytes_b;
  hipError_t e;
  if (host_a) {
    if (gemmx_stage_h2d(dev, pa, lda2 * e8, a, (size_t)lda * e8, (size_t)k * e8, (size_t)k,
                        s, NULL) != GEMMX_OK)
      goto out;
    p.a = pa;
  }
  if (host_b) {
    if (gemmx_stage_h2d(dev, pb, ldb2 * e8, b, (size_t)ldb * e8, (size_t)m * e8, (size_t)n,
                        s, NULL) != GEMMX_OK)
      goto out;
    p.b = pb;
  }
  e = gemmx_tri(es, solve, &p, pw, s);
  if (e != hipSuccess) {
    hip_fail(e, solve ? "triangular solve launch" : "triangular multiply launch");
    (void)hipStreamSynchronize(s);
    goto out;
  }
  if (host_b && gemmx_stage_d2h(dev, b, (size_t)ldb * e8, pb, ldb2 * e8, (size_t)m * e8,
                                 (size_t)n, s, NULL) != GEMMX_OK) {
    (void)hipStreamSynchronize(s);
    goto out;
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) hip_fail(e, solve ? "triangular solve" : "triangular multiply");
out:
  pthread_mutex_unlock(&d->lock);
}

#define GEMMX_TRI_ENTRY(NAME, T, ES, SOLVE)                                                      \
  EXPORT void NAME(const char *side, const char *uplo, const char *transa, const char *diag,   \
                   const int *m, const int *n, const T *alpha, const T *a, const int *lda,     \
                   T *b, const int *ldb) {                                                     \
    if (!side || !uplo || !transa || !diag || !m || !n || !alpha || !lda || !ldb) {            \
      clear_error();                                                                           \
      set_error(GEMMX_ERR_ARG, #NAME ": NULL scalar argument");                                \
      return;                                                                                  \
    }                                                                                          \
    pthread_once(&g_once, init_once);                                                          \
    tri_call(ES, SOLVE, *side, *uplo, *transa, *diag, *m, *n, (double)*alpha, a, *lda, b,      \
             *ldb);                                                                            \
  }
GEMMX_TRI_ENTRY(dtrmm_, double, 8, 0)
GEMMX_TRI_ENTRY(dtrsm_, double, 8, 1)
GEMMX_TRI_ENTRY(strmm_, float, 4, 0)
GEMMX_TRI_ENTRY(strsm_, float, 4, 1)
#undef GEMMX_TRI_ENTRY

/* ILP64 forms: the same calls with 64-bit Fortran integers (callers built
 * with -fdefault-integer-8 and the "_64_" suffix of ILP64 BLAS builds).
 * Leading dimensions are used at full width; m, n and k must stay below
 * 2^31 (the kernels' tile grid). */
static int ilp64_dims(call_t *x, const int64_t *m, const int64_t *n, const int64_t *k) {
  if (*m > INT_MAX || *n > INT_MAX || *k > INT_MAX || *m < INT_MIN || *n < INT_MIN ||
      *k < INT_MIN)
    return set_error(GEMMX_ERR_ARG, "dimension beyond 2^31-1 (m=%lld n=%lld k=%lld)",
                     (long long)*m, (long long)*n, (long long)*k);
  x->m = (int)*m;
  x->n = (int)*n;
  x->k = (int)*k;
  return GEMMX_OK;
}

EXPORT void dgemm_64_(const char *transa, const char *transb, const int64_t *m,
                      const int64_t *n, const int64_t *k, const double *alpha,
                      const double *a, const int64_t *lda, const double *b,
                      const int64_t *ldb, const double *beta, double *c,
                      const int64_t *ldc) {
  clear_error();
  if (!transa || !transb || !m || !n || !k || !alpha || !lda || !ldb || !beta || !ldc) {
    set_error(GEMMX_ERR_ARG, "dgemm_64_: NULL scalar argument");
    return;
  }
  call_t x = {8, *transa, *transb, 0, 0, 0, *lda, *ldb, *ldc, *alpha, *beta, a, b, c};
  if (ilp64_dims(&x, m, n, k) != GEMMX_OK) return;
  pthread_once(&g_once, init_once);
  fortran_call(&x, g_dmode, g_dslices);
}

EXPORT void sgemm_64_(const char *transa, const char *transb, const int64_t *m,
                      const int64_t *n, const int64_t *k, const float *alpha,
                      const float *a, const int64_t *lda, const float *b,
                      const int64_t *ldb, const float *beta, float *c,
                      const int64_t *ldc) {
  clear_error();
  if (!transa || !transb || !m || !n || !k || !alpha || !lda || !ldb || !beta || !ldc) {
    set_error(GEMMX_ERR_ARG, "sgemm_64_: NULL scalar argument");
    return;
  }
  call_t x = {4, *transa, *transb, 0, 0, 0, *lda, *ldb, *ldc, *alpha, *beta, a, b, c};
  if (ilp64_dims(&x, m, n, k) != GEMMX_OK) return;
  pthread_once(&g_once, init_once);
  fortran_call(&x, g_smode, 0);
}

/* ------------------------------------------------------------------------ */
/* Extensions.                                                              */
/* ------------------------------------------------------------------------ */
EXPORT int dgemm_x_set_mode(int mode, int slices) {
  pthread_once(&g_once, init_once);
  if (mode == GEMMX_MODE_NATIVE || mode == GEMMX_MODE_AUTO) {
    g_dmode = mode;
    return GEMMX_OK;
  }
  if ((mode == GEMMX_MODE_OZAKI && slices >= 3 && slices <= 8) ||
      (mode == GEMMX_MODE_OZAKI2 && slices >= 8 && slices <= 22)) {
    g_dmode = mode;
    g_dslices = slices;
    return GEMMX_OK;
  }
  return set_error(GEMMX_ERR_ARG, "dgemm_x_set_mode: mode %d slices %d", mode, slices);
}

EXPORT int sgemm_x_set_mode(int mode) {
  pthread_once(&g_once, init_once);
  if (mode == GEMMX_MODE_NATIVE || mode == GEMMX_MODE_BF16X3 || mode == GEMMX_MODE_AUTO) {
    g_smode = mode;
    return GEMMX_OK;
  }
  return set_error(GEMMX_ERR_ARG, "sgemm_x_set_mode: mode %d", mode);
}

/* Tile configuration of the native kernels: cfg >= 0 forces one for every
 * call, -1 returns to the tuning table.  Table reload: path or NULL. */
EXPORT int dgemm_x_set_tile_config(int cfg) {
  pthread_once(&g_once, init_once);
  if (gemmx_tune_force(8, cfg) != 0) return set_error(GEMMX_ERR_ARG, "tile config %d", cfg);
  return GEMMX_OK;
}
EXPORT int sgemm_x_set_tile_config(int cfg) {
  pthread_once(&g_once, init_once);
  if (gemmx_tune_force(4, cfg) != 0) return set_error(GEMMX_ERR_ARG, "tile config %d", cfg);
  return GEMMX_OK;
}
/* Split-K factor of the native kernels: split >= 1 forces it for every
 * call (1 = one pass), -1 returns to the tuning table / heuristic. */
EXPORT int dgemm_x_set_split_config(int split) {
  pthread_once(&g_once, init_once);
  if (gemmx_tune_force_split(8, split) != 0)
    return set_error(GEMMX_ERR_ARG, "split config %d", split);
  return GEMMX_OK;
}
EXPORT int sgemm_x_set_split_config(int split) {
  pthread_once(&g_once, init_once);
  if (gemmx_tune_force_split(4, split) != 0)
    return set_error(GEMMX_ERR_ARG, "split config %d", split);
  return GEMMX_OK;
}
/* "major.minor" of the ROCm release the library was built with (the key a
 * tuning table must carry). */
EXPORT const char *dgemm_x_build_rocm(void) { return gemmx_tune_rocm(); }
EXPORT const char *sgemm_x_build_rocm(void) { return gemmx_tune_rocm(); }
/* Product kernel of the emulated forms: 1 direct-to-LDS (default), 0
 * register-staged, -1 back to GEMMX_SLICE_CFG / the default. */
EXPORT int dgemm_x_set_product_config(int cfg) {
  if (gemmx_set_product_config(cfg) != 0)
    return set_error(GEMMX_ERR_ARG, "product config %d", cfg);
  return GEMMX_OK;
}
EXPORT int sgemm_x_set_product_config(int cfg) {
  if (gemmx_set_product_config(cfg) != 0)
    return set_error(GEMMX_ERR_ARG, "product config %d", cfg);
  return GEMMX_OK;
}
EXPORT int dgemm_x_load_tuning(const char *path) {
  pthread_once(&g_once, init_once);
  return gemmx_tune_load(path);
}
EXPORT int sgemm_x_load_tuning(const char *path) {
  pthread_once(&g_once, init_once);
  return gemmx_tune_load(path);
}

/* The arithmetic dgemm_ in the automatic mode would use for these
 * arguments (GEMMX_MODE_*), with the slice or modulus count in *count.
 * Needs no device. */
EXPORT int dgemm_x_tune_lookup_emulated(char transa, char transb, int m, int n, int k,
                                        int *count) {
  pthread_once(&g_once, init_once);
  int c = 0;
  const int mode = gemmx_tune_pick_emulated(!notrans(transa), !notrans(transb), m, n, k, &c);
  if (count) *count = c;
  return mode;
}

/* The arithmetic sgemm_ in the automatic mode would use for these
 * arguments (GEMMX_MODE_NATIVE or GEMMX_MODE_BF16X3).  Needs no device. */
EXPORT int sgemm_x_tune_lookup_emulated(char transa, char transb, int m, int n, int k) {
  pthread_once(&g_once, init_once);
  return gemmx_tune_pick_emulated32(!notrans(transa), !notrans(transb), m, n, k);
}

/* The tile configuration a call with these arguments would use (forced
 * configuration, else the table entry of its shape bucket, else the size
 * heuristic).  Needs no device. */
static int tune_lookup(int es, char transa, char transb, int m, int n, int k) {
  pthread_once(&g_once, init_once);
  if (m < 0 || n < 0 || k < 0) return set_error(GEMMX_ERR_ARG, "negative dimension"), -1;
  return gemmx_tune_pick(es, !notrans(transa), !notrans(transb), m, n, k);
}
EXPORT int dgemm_x_tune_lookup(char transa, char transb, int m, int n, int k) {
  return tune_lookup(8, transa, transb, m, n, k);
}
EXPORT int sgemm_x_tune_lookup(char transa, char transb, int m, int n, int k) {
  return tune_lookup(4, transa, transb, m, n, k);
}
/* The tile configuration of a strided batch of `batch` such problems. */
static int strided_lookup(int es, char transa, char transb, int m, int n, int k, int batch) {
  pthread_once(&g_once, init_once);
  if (m < 0 || n < 0 || k < 0 || batch < 0)
    return set_error(GEMMX_ERR_ARG, "negative dimension or batch count"), -1;
  return gemmx_tune_pick_strided(es, !notrans(transa), !notrans(transb), m, n, k, batch);
}
EXPORT int dgemm_x_tune_lookup_strided(char transa, char transb, int m, int n, int k,
                                       int batch) {
  return strided_lookup(8, transa, transb, m, n, k, batch);
}
EXPORT int sgemm_x_tune_lookup_strided(char transa, char transb, int m, int n, int k,
                                       int batch) {
  return strided_lookup(4, transa, transb, m, n, k, batch);
}
/* The number of K ranges a native call with these arguments runs (1 = one
 * pass), and the bytes of partial tiles they need.  Needs no device. */
static int split_lookup(int es, char transa, char transb, int m, int n, int k, size_t *bytes) {
  const int cfg = tune_lookup(es, transa, transb, m, n, k);
  if (cfg < 0) return -1;
  const int ta = !notrans(transa), tb = !notrans(transb);
  const int split = gemmx_tune_pick_split(es, ta, tb, m, n, k, cfg);
  if (bytes) *bytes = gemmx_native_workspace(es, cfg, m, n, k, split);
  return gemmx_native_splits(es, cfg, k, split);
}
EXPORT int dgemm_x_tune_lookup_split(char transa, char transb, int m, int n, int k,
                                     size_t *bytes) {
  return split_lookup(8, transa, transb, m, n, k, bytes);
}
EXPORT int sgemm_x_tune_lookup_split(char transa, char transb, int m, int n, int k,
                                     size_t *bytes) {
  return split_lookup(4, transa, transb, m, n, k, bytes);
}

EXPORT int dgemm_x_last_error(const char **msg) {
  if (msg) *msg = t_msg;
  return t_err;
}
EXPORT int sgemm_x_last_error(const char **msg) {
  if (msg) *msg = t_msg;
  return t_err;
}

/* The arithmetic launch_device runs for an element size and mode. */
static int launch_mode(int es, int mode) {
  if (es == 8 && (mode == GEMMX_MODE_OZAKI || mode == GEMMX_MODE_OZAKI2)) return mode;
  if (es == 4 && mode == GEMMX_MODE_BF16X3) return mode;
  return GEMMX_MODE_NATIVE;
}

/* `bytes` of device memory ordered on stream s, from a pool of the device
 * that keeps up to 256 MB between calls; NULL when the pool is unavailable
 * or the allocation fails (the caller then runs without it). */
static void *stream_alloc(int dev, size_t bytes, hipStream_t s) {
  devstate_t *d = &g_dev[dev];
  pthread_mutex_lock(&d->lock);
  if (!d->pool_tried) {
    d->pool_tried = 1;
    hipMemPoolProps props;
    memset(&props, 0, sizeof props);
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    if (hipMemPoolCreate(&d->pool, &props) == hipSuccess) {
      uint64_t keep = (uint64_t)256 << 20;
      (void)hipMemPoolSetAttribute(d->pool, hipMemPoolAttrReleaseThreshold, &keep);
    } else {
      d->pool = NULL;
      (void)hipGetLastError();
    }
  }
  hipMemPool_t pool = d->pool;
  pthread_mutex_unlock(&d->lock);
  void *w = NULL;
  if (gemmx_alloc_failure_injected()) return NULL;
  if (!pool || hipMallocFromPoolAsync(&w, bytes, pool, s) != hipSuccess) {
    (void)hipGetLastError();
    return NULL;
  }
  return w;
}

static int dev_call(call_t *x, int mode, int slices, void *work, size_t work_bytes,
                    void *stream) {
  clear_error();
  if (validate(x) != GEMMX_OK) return t_err;
  if (x->m == 0 || x->n == 0) return GEMMX_OK;
  int dev;
  if (current_device(&dev) != GEMMX_OK) return t_err;
  gemmx_problem p = {.ta = !notrans(x->ta), .tb = !notrans(x->tb), .m = x->m,
                     .n = x->n, .k = x->k, .lda = x->lda, .ldb = x->ldb,
                     .ldc = x->ldc, .alpha = x->alpha, .beta = x->beta,
                     .a = x->a, .b = x->b, .c = x->c};
  hipStream_t s = (hipStream_t)stream;
  const size_t need = call_ws(x, mode, slices);
  if (need == 0 || work) {
    if (need > 0 && work_bytes < need)
      return set_error(GEMMX_ERR_ARG, "workspace of %zu bytes < %zu needed", work_bytes, need);
    return launch_device(&p, x->es, mode, slices, work, s);
  }
  if (launch_mode(x->es, mode) == GEMMX_MODE_NATIVE) {
    /* split-K partial tiles: stream-ordered memory of the device's pool, so
     * the call stays asynchronous; without it the product runs in one pass */
    void *w = stream_alloc(dev, need, s);
    const int rc = launch_device(&p, x->es, mode, slices, w, s);
    if (w) (void)hipFreeAsync(w, s);
    return rc;
  }
  /* No caller workspace: borrow the device scratch for the duration. */
  devstate_t *d = &g_dev[dev];
  pthread_mutex_lock(&d->lock);
  void *w = NULL;
  int rc = get_scratch(dev, need, &w);
  if (rc == GEMMX_OK) {
    rc = launch_device(&p, x->es, mode, slices, w, s);
    hipError_t e = hipStreamSynchronize(s);
    if (rc == GEMMX_OK && e != hipSuccess) rc = hip_fail(e, "GEMM execution");
  }
  pthread_mutex_unlock(&d->lock);
  return rc;
}

EXPORT int dgemm_x_dev(char transa, char transb, int m, int n, int k,
                       double alpha, const double *a, int lda, const double *b,
                       int ldb, double beta, double *c, int ldc, void *stream) {
  call_t x = {8, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return dev_call(&x, GEMMX_MODE_NATIVE, 0, NULL, 0, stream);
}

EXPORT int sgemm_x_dev(char transa, char transb, int m, int n, int k,
                       float alpha, const float *a, int lda, const float *b,
                       int ldb, float beta, float *c, int ldc, void *stream) {
  call_t x = {4, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return dev_call(&x, GEMMX_MODE_NATIVE, 0, NULL, 0, stream);
}

/* The allocation-free form of dgemm_x_dev / sgemm_x_dev: the same kernels
 * and split choice, the split-K partial tiles in the library's static device
 * scratch instead of the stream-ordered pool; no device allocation or free
 * of any kind.  A call whose partial tiles exceed the scratch is refused. */
static int noalloc_call(call_t *x, void *stream) {
  clear_error();
  if (validate(x) != GEMMX_OK) return t_err;
  if (x->m == 0 || x->n == 0) return GEMMX_OK;
  int dev;
  if (current_device(&dev) != GEMMX_OK) return t_err;
  gemmx_problem p = {.ta = !notrans(x->ta), .tb = !notrans(x->tb), .m = x->m,
                     .n = x->n, .k = x->k, .lda = x->lda, .ldb = x->ldb,
                     .ldc = x->ldc, .alpha = x->alpha, .beta = x->beta,
                     .a = x->a, .b = x->b, .c = x->c};
  const size_t need = call_ws(x, GEMMX_MODE_NATIVE, 0);
  void *w = NULL;
  if (need > 0) {
    if (need > (size_t)GEMMX_NOALLOC_SCRATCH)
      return set_error(GEMMX_ERR_UNSUPPORTED,
                       "%cgemm_x_dev_noalloc: the split-K partial tiles of this call need %zu "
                       "bytes, more than the %u bytes of static scratch (use %cgemm_x_dev)",
                       x->es == 8 ? 'd' : 's', need, (unsigned)GEMMX_NOALLOC_SCRATCH,
                       x->es == 8 ? 'd' : 's');
    const hipError_t e = gemmx_noalloc_scratch(dev, &w);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  }
  return launch_device(&p, x->es, GEMMX_MODE_NATIVE, 0, w, (hipStream_t)stream);
}

EXPORT int dgemm_x_dev_noalloc(char transa, char transb, int m, int n, int k, double alpha,
                               const double *a, int lda, const double *b, int ldb,
                               double beta, double *c, int ldc, void *stream) {
  call_t x = {8, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return noalloc_call(&x, stream);
}

EXPORT int sgemm_x_dev_noalloc(char transa, char transb, int m, int n, int k, float alpha,
                               const float *a, int lda, const float *b, int ldb, float beta,
                               float *c, int ldc, void *stream) {
  call_t x = {4, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return noalloc_call(&x, stream);
}

/* Strided batch on device pointers, native arithmetic: `batch` problems of
 * one shape in one set of launches (grid row per problem). */
static int strided_call(call_t *x, int64_t sa, int64_t sb, int64_t sc, int batch,
                        void *stream) {
  clear_error();
  if (batch < 0) return set_error(GEMMX_ERR_ARG, "batch count %d < 0", batch);
  if (validate(x) != GEMMX_OK) return t_err;
  if (batch == 0 || x->m == 0 || x->n == 0) return GEMMX_OK;
  if (sa < 0 || sb < 0 || sc < 0)
    return set_error(GEMMX_ERR_ARG, "negative stride (%lld, %lld, %lld)", (long long)sa,
                     (long long)sb, (long long)sc);
  if (batch > 1 && sc < x->ldc * (int64_t)x->n)
    return set_error(GEMMX_ERR_ARG, "stride_c=%lld < ldc*n=%lld: the C matrices overlap",
                     (long long)sc, (long long)(x->ldc * (int64_t)x->n));
  int dev;
  if (current_device(&dev) != GEMMX_OK) return t_err;
  gemmx_problem p = {.ta = !notrans(x->ta), .tb = !notrans(x->tb), .m = x->m,
                     .n = x->n, .k = x->k, .lda = x->lda, .ldb = x->ldb,
                     .ldc = x->ldc, .alpha = x->alpha, .beta = x->beta,
                     .a = x->a, .b = x->b, .c = x->c, .batch = batch,
                     .sa = sa, .sb = sb, .sc = sc};
  return launch_device(&p, x->es, GEMMX_MODE_NATIVE, 0, NULL, (hipStream_t)stream);
}

EXPORT int dgemm_x_strided_batched_dev(char transa, char transb, int m, int n, int k,
                                       double alpha, const double *a, int lda,
                                       int64_t stride_a, const double *b, int ldb,
                                       int64_t stride_b, double beta, double *c, int ldc,
                                       int64_t stride_c, int batch, void *stream) {
  call_t x = {8, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return strided_call(&x, stride_a, stride_b, stride_c, batch, stream);
}

EXPORT int sgemm_x_strided_batched_dev(char transa, char transb, int m, int n, int k,
                                       float alpha, const float *a, int lda,
                                       int64_t stride_a, const float *b, int ldb,
                                       int64_t stride_b, float beta, float *c, int ldc,
                                       int64_t stride_c, int batch, void *stream) {
  call_t x = {4, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return strided_call(&x, stride_a, stride_b, stride_c, batch, stream);
}

EXPORT size_t dgemm_x_ozaki_workspace(int m, int n, int k, int slices) {
  if (m <= 0 || n <= 0 || k <= 0 || slices < 3 || slices > 8) return 0;
  return gemmx_ozaki_workspace(m, n, k, slices);
}

EXPORT int dgemm_x_ozaki_dev(char transa, char transb, int m, int n, int k,
                             double alpha, const double *a, int lda,
                             const double *b, int ldb, double beta, double *c,
                             int ldc, int slices, void *work, size_t work_bytes,
                             void *stream) {
  if (slices < 3 || slices > 8) {
    clear_error();
    return set_error(GEMMX_ERR_ARG, "slices=%d outside 3..8", slices);
  }
  call_t x = {8, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return dev_call(&x, GEMMX_MODE_OZAKI, slices, work, work_bytes, stream);
}

EXPORT int dgemm_x_set_phase_timing(int on) {
  pthread_once(&g_once, init_once);
  t_timing = on != 0;
  t_ev_used = 0;
  return GEMMX_OK;
}

EXPORT int dgemm_x_phase_times(float *ms) {
  clear_error();
  if (!ms) return set_error(GEMMX_ERR_ARG, "phase_times: NULL output");
  if (!t_ev_used || t_ev_dev < 0)
    return set_error(GEMMX_ERR_ARG, "phase_times: no timed emulated call on this thread");
  hipError_t e = hipEventSynchronize(t_ev[3]);
  if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
  for (int i = 0; i < 3; i++) {
    e = hipEventElapsedTime(&ms[i], t_ev[i], t_ev[i + 1]);
    if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
  }
  return GEMMX_OK;
}

EXPORT size_t dgemm_x_ozaki2_workspace(int m, int n, int k, int moduli) {
  if (m <= 0 || n <= 0 || k <= 0 || moduli < 8 || moduli > 22) return 0;
  return gemmx_ozaki2_workspace(m, n, k, moduli);
}

EXPORT int dgemm_x_ozaki2_bits(int moduli, int k) {
  if (k <= 0) return -1;
  return gemmx_ozaki2_bits(moduli, k);
}

EXPORT int dgemm_x_ozaki2_dev(char transa, char transb, int m, int n, int k,
                              double alpha, const double *a, int lda,
                              const double *b, int ldb, double beta, double *c,
                              int ldc, int moduli, void *work, size_t work_bytes,
                              void *stream) {
  if (moduli < 8 || moduli > 22) {
    clear_error();
    return set_error(GEMMX_ERR_ARG, "moduli=%d outside 8..22", moduli);
  }
  call_t x = {8, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return dev_call(&x, GEMMX_MODE_OZAKI2, moduli, work, work_bytes, stream);
}

EXPORT size_t sgemm_x_bf16x3_workspace(int m, int n, int k) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  return gemmx_bf16x3_workspace(m, n, k);
}

EXPORT int sgemm_x_bf16x3_dev(char transa, char transb, int m, int n, int k,
                              float alpha, const float *a, int lda,
                              const float *b, int ldb, float beta, float *c,
                              int ldc, void *work, size_t work_bytes,
                              void *stream) {
  call_t x = {4, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta, a, b, c};
  return dev_call(&x, GEMMX_MODE_BF16X3, 0, work, work_bytes, stream);
}
