/*
 * gemmx_tune.c -- run-time tuning table (see gemmx_tune.h).
 *
 * The reference fixes its blocking at compile time (dgemm_tune.h:6-10,
 * sgemm_tune.h:7-11, derived in src/gemm_set_parameters.h:54-61).  This layer
 * looks the tile configuration up per call, so one build serves every shape
 * and the autotuner can change the choice without a rebuild.
 */
#define _GNU_SOURCE
#include "gemmx_tune.h"

#include <dlfcn.h>
#include <hip/hip_version.h>
#include <libgen.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/gemm_mi355x.h"
#include "gemmx_internal.h"

#define NB 12 /* buckets 4..15 */
static signed char g_table[2][4][NB][NB][NB]; /* [f64/f32][trans][m][n][k] */
/* split-K factor per bucket: 0 = not in the table (the heuristic decides) */
static signed char g_split[2][4][NB][NB][NB];
/* DGEMM arithmetic per bucket: 0 = no entry, else (mode + 1) << 8 | count */
static short g_emu[4][NB][NB][NB];
/* SGEMM arithmetic per bucket: 0 = no entry, else mode + 1 */
static signed char g_emu32[4][NB][NB][NB];
/* strided-batch tile configuration per bucket: 0 = no entry, else cfg + 1 */
static signed char g_strided[2][4][NB][NB][NB];
/* grouped-batch kernel configuration per element size: 0 = no entry, else cfg + 1 */
static signed char g_batch[2];
static int g_entries;
static int g_force[2] = {-1, -1}; /* [f64/f32]: one configuration for every call */
static int g_force_split[2] = {0, 0}; /* [f64/f32]: one split factor for every call (0: off) */
static char g_path[1024];
static char g_arch[64];           /* current device arch, "" = unknown */

/* The table is measured at powers of two, so a dimension takes the nearest
 * one on a log scale: bucket b covers [2^b/sqrt(2), 2^b*sqrt(2)), i.e.
 * b = floor(log2(2*dim^2) / 2). */
int gemmx_tune_bucket(int dim) {
  uint64_t v = dim > 0 ? 2 * (uint64_t)dim * (uint64_t)dim : 1;
  int f = -1;
  while (v) {
    v >>= 1;
    f++;
  }
  int b = f >> 1;
  if (b < 4) b = 4;
  if (b > 15) b = 15;
  return b;
}

static int trans_index(const char *t) {
  if (strcmp(t, "NN") == 0) return 0;
  if (strcmp(t, "NT") == 0) return 1;
  if (strcmp(t, "TN") == 0) return 2;
  if (strcmp(t, "TT") == 0) return 3;
  return -1;
}

/* The ROCm (HIP) release the library was compiled with, "major.minor": a
 * table measured with kernels from another compiler is not used. */
static const char *build_rocm(void) {
  static char v[32];
  if (!v[0]) snprintf(v, sizeof v, "%d.%d", HIP_VERSION_MAJOR, HIP_VERSION_MINOR);
  return v;
}

static void clear_table(void) {
  memset(g_table, 0, sizeof g_table);
  memset(g_split, 0, sizeof g_split);
  memset(g_emu, 0, sizeof g_emu);
  memset(g_emu32, 0, sizeof g_emu32);
  memset(g_strided, 0, sizeof g_strided);
  memset(g_batch, 0, sizeof g_batch);
  g_entries = 0;
  g_path[0] = 0;
}

/* Reads a table; on a format, arch or ROCm mismatch, or without a rocm
 * directive, nothing is kept. */
static int load_table(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  static signed char staged[2][4][NB][NB][NB];
  static signed char staged_split[2][4][NB][NB][NB];
  static short staged_emu[4][NB][NB][NB];
  static signed char staged_emu32[4][NB][NB][NB];
  static signed char staged_strided[2][4][NB][NB][NB];
  signed char staged_batch[2] = {0, 0};
  memset(staged, 0, sizeof staged);
  memset(staged_split, 0, sizeof staged_split);
  memset(staged_emu, 0, sizeof staged_emu);
  memset(staged_emu32, 0, sizeof staged_emu32);
  memset(staged_strided, 0, sizeof staged_strided);
  char line[256];
  int format = 0, entries = 0, ok = 1, rocm = 0;
  while (ok && fgets(line, sizeof line, f)) {
    char *hash = strchr(line, '#');
    if (hash) *hash = 0;
    char w0[32], w1[64];
    if (sscanf(line, "%31s %63s", w0, w1) == 2) {
      if (strcmp(w0, "format") == 0) {
        format = atoi(w1);
        if (format != GEMMX_TUNE_FORMAT) ok = 0;
        continue;
      }
      if (strcmp(w0, "arch") == 0) {
        if (g_arch[0] && strcmp(w1, g_arch) != 0) ok = 0;
        continue;
      }
      if (strcmp(w0, "rocm") == 0) {
        if (strcmp(w1, build_rocm()) != 0) ok = 0;
        rocm = 1;
        continue;
      }
    }
    char dt[8], tr[8], sc[16];
    int mb, nb, kb, cfg, split = 0;
    if (sscanf(line, "batch %7s %d", dt, &cfg) == 2) {
      if (format != GEMMX_TUNE_FORMAT) {
        ok = 0;
        break;
      }
      const int d = strcmp(dt, "f64") == 0 ? 0 : strcmp(dt, "f32") == 0 ? 1 : -1;
      if (d < 0 || cfg < 0 || cfg >= GEMMX_BATCH_CONFIGS) continue;
      staged_batch[d] = (signed char)(cfg + 1);
      entries++;
      continue;
    }
    if (sscanf(line, "strided %7s %7s %d %d %d %d", dt, tr, &mb, &nb, &kb, &cfg) == 6) {
      if (format != GEMMX_TUNE_FORMAT) {
        ok = 0;
        break;
      }
      const int d = strcmp(dt, "f64") == 0 ? 0 : strcmp(dt, "f32") == 0 ? 1 : -1;
      const int t = trans_index(tr);
      if (d < 0 || t < 0 || mb < 4 || mb > 15 || nb < 4 || nb > 15 || kb < 4 || kb > 15 ||
          cfg < 0 || cfg >= GEMMX_NCFG_COUNT)
        continue;
      staged_strided[d][t][mb - 4][nb - 4][kb - 4] = (signed char)(cfg + 1);
      entries++;
      continue;
    }
    if (sscanf(line, "emu32 %7s %d %d %d %15s", tr, &mb, &nb, &kb, sc) == 5) {
      if (format != GEMMX_TUNE_FORMAT) {
        ok = 0;
        break;
      }
      const int t = trans_index(tr);
      const int mode = strcmp(sc, "native") == 0   ? GEMMX_MODE_NATIVE
                       : strcmp(sc, "bf16x3") == 0 ? GEMMX_MODE_BF16X3
                                                   : -1;
      if (t < 0 || mode < 0 || mb < 4 || mb > 15 || nb < 4 || nb > 15 || kb < 4 || kb > 15)
        continue;
      staged_emu32[t][mb - 4][nb - 4][kb - 4] = (signed char)(mode + 1);
      entries++;
      continue;
    }
    if (sscanf(line, "emu %7s %d %d %d %15s %d", tr, &mb, &nb, &kb, sc, &cfg) == 6) {
      if (format != GEMMX_TUNE_FORMAT) {
        ok = 0;
        break;
      }
      const int t = trans_index(tr);
      const int mode = strcmp(sc, "native") == 0   ? GEMMX_MODE_NATIVE
                       : strcmp(sc, "ozaki") == 0  ? GEMMX_MODE_OZAKI
                       : strcmp(sc, "ozaki2") == 0 ? GEMMX_MODE_OZAKI2
                                                   : -1;
      const int cmin = mode == GEMMX_MODE_OZAKI ? 3 : mode == GEMMX_MODE_OZAKI2 ? 8 : 0;
      const int cmax = mode == GEMMX_MODE_OZAKI ? 8 : mode == GEMMX_MODE_OZAKI2 ? 22 : 0;
      if (t < 0 || mode < 0 || mb < 4 || mb > 15 || nb < 4 || nb > 15 || kb < 4 || kb > 15 ||
          cfg < cmin || cfg > cmax)
        continue;
      staged_emu[t][mb - 4][nb - 4][kb - 4] = (short)(((mode + 1) << 8) | cfg);
      entries++;
      continue;
    }
    const int got = sscanf(line, "%7s %7s %d %d %d %d %d", dt, tr, &mb, &nb, &kb, &cfg, &split);
    if (got != 6 && got != 7) continue;
    if (format != GEMMX_TUNE_FORMAT) {
      ok = 0; /* entries before (or without) a format directive */
      break;
    }
    int d = strcmp(dt, "f64") == 0 ? 0 : strcmp(dt, "f32") == 0 ? 1 : -1;
    int t = trans_index(tr);
    if (d < 0 || t < 0 || mb < 4 || mb > 15 || nb < 4 || nb > 15 || kb < 4 || kb > 15) continue;
    if (cfg < 0 || cfg >= GEMMX_NCFG_COUNT) continue;
    if (got == 7 && (split < 1 || split > GEMMX_MAX_SPLIT)) continue;
    staged[d][t][mb - 4][nb - 4][kb - 4] = (signed char)(cfg + 1);
    staged_split[d][t][mb - 4][nb - 4][kb - 4] = (signed char)(got == 7 ? split : 0);
    entries++;
  }
  fclose(f);
  if (!ok || format != GEMMX_TUNE_FORMAT || !rocm) return -1;
  memcpy(g_table, staged, sizeof g_table);
  memcpy(g_split, staged_split, sizeof g_split);
  memcpy(g_emu, staged_emu, sizeof g_emu);
  memcpy(g_emu32, staged_emu32, sizeof g_emu32);
  memcpy(g_strided, staged_strided, sizeof g_strided);
  memcpy(g_batch, staged_batch, sizeof g_batch);
  g_entries = entries;
  snprintf(g_path, sizeof g_path, "%s", path);
  return entries;
}

void gemmx_tune_init(const char *arch) {
  clear_table();
  g_arch[0] = 0;
  if (arch) {
    /* "gfx950:sramecc+:xnack-" -> "gfx950" */
    snprintf(g_arch, sizeof g_arch, "%s", arch);
    char *colon = strchr(g_arch, ':');
    if (colon) *colon = 0;
  }
  const char *force = getenv("GEMMX_NATIVE_CFG");
  if (force && *force) {
    int c = atoi(force);
    if (c >= 0 && c < GEMMX_NCFG_COUNT) g_force[0] = g_force[1] = c;
  }
  const char *fsplit = getenv("GEMMX_NATIVE_SPLIT");
  if (fsplit && *fsplit) {
    int c = atoi(fsplit);
    if (c >= 1 && c <= GEMMX_MAX_SPLIT) g_force_split[0] = g_force_split[1] = c;
  }
  const char *env = getenv("GEMMX_TUNE_FILE");
  if (env && *env) {
    (void)load_table(env);
    return;
  }
  Dl_info info;
  if (dladdr((void *)&gemmx_tune_init, &info) && info.dli_fname) {
    char buf[1024];
    snprintf(buf, sizeof buf, "%s", info.dli_fname);
    char path[1100];
    snprintf(path, sizeof path, "%s/gemm_mi355x_tune.txt", dirname(buf));
    (void)load_table(path);
  }
}

int gemmx_tune_entries(void) { return g_entries; }

int gemmx_tune_force(int es, int cfg) {
  if (cfg < -1 || cfg >= GEMMX_NCFG_COUNT) return -1;
  g_force[es == 8 ? 0 : 1] = cfg;
  return 0;
}

int gemmx_tune_force_split(int es, int split) {
  if (split != -1 && (split < 1 || split > GEMMX_MAX_SPLIT)) return -1;
  g_force_split[es == 8 ? 0 : 1] = split < 0 ? 0 : split;
  return 0;
}

const char *gemmx_tune_rocm(void) { return build_rocm(); }

int gemmx_tune_load(const char *path) {
  if (!path || !*path) {
    clear_table();
    return 0;
  }
  return load_table(path);
}

const char *gemmx_tune_path(void) { return g_path; }

int gemmx_tune_pick_emulated(int ta, int tb, int m, int n, int k, int *count) {
  const int t = (ta ? 2 : 0) + (tb ? 1 : 0);
  const int v = g_emu[t][gemmx_tune_bucket(m) - 4][gemmx_tune_bucket(n) - 4]
                     [gemmx_tune_bucket(k) - 4];
  if (!v) {
    *count = 0;
    return GEMMX_MODE_NATIVE;
  }
  *count = v & 0xff;
  return (v >> 8) - 1;
}

int gemmx_tune_pick_emulated32(int ta, int tb, int m, int n, int k) {
  const int t = (ta ? 2 : 0) + (tb ? 1 : 0);
  const int v = g_emu32[t][gemmx_tune_bucket(m) - 4][gemmx_tune_bucket(n) - 4]
                       [gemmx_tune_bucket(k) - 4];
  return v ? v - 1 : GEMMX_MODE_NATIVE;
}

int gemmx_tune_pick(int es, int ta, int tb, int m, int n, int k) {
  const int d = es == 8 ? 0 : 1, t = (ta ? 2 : 0) + (tb ? 1 : 0);
  if (g_force[d] >= 0) return g_force[d];
  const int c = g_table[d][t][gemmx_tune_bucket(m) - 4][gemmx_tune_bucket(n) - 4]
                       [gemmx_tune_bucket(k) - 4];
  if (c > 0) return c - 1;
  /* Heuristic: keep at least ~2 workgroups per CU of the 256 on MI355X. */
  const long tiles128 = (((long)m + 127) / 128) * (((long)n + 127) / 128);
  return tiles128 >= 512 ? GEMMX_NCFG_128x128 : GEMMX_NCFG_64x64;
}

int gemmx_tune_pick_strided(int es, int ta, int tb, int m, int n, int k, int batch) {
  const int d = es == 8 ? 0 : 1, t = (ta ? 2 : 0) + (tb ? 1 : 0);
  if (g_force[d] >= 0) return g_force[d];
  /* A strided entry is measured on batches that fill the GPU many times
   * over; a batch below one 128x128 tile per CU (256 CUs on MI355X) keeps
   * the single problem's choice. */
  const long tiles128 = (((long)m + 127) / 128) * (((long)n + 127) / 128);
  if (batch > 1 && tiles128 * batch >= 256) {
    const int c = g_strided[d][t][gemmx_tune_bucket(m) - 4][gemmx_tune_bucket(n) - 4]
                           [gemmx_tune_bucket(k) - 4];
    if (c > 0) return c - 1;
  }
  return gemmx_tune_pick(es, ta, tb, m, n, k);
}

int gemmx_tune_pick_batch(int es) { return g_batch[es == 8 ? 0 : 1] - 1; }

int gemmx_tune_pick_split(int es, int ta, int tb, int m, int n, int k, int cfg) {
  const int d = es == 8 ? 0 : 1, t = (ta ? 2 : 0) + (tb ? 1 : 0);
  int bm, bn;
  if (gemmx_native_tile(cfg, &bm, &bn) != 1) return 1; /* no split-K kernels */
  if (g_force_split[d] >= 1) return g_force_split[d];
  const int v = g_split[d][t][gemmx_tune_bucket(m) - 4][gemmx_tune_bucket(n) - 4]
                       [gemmx_tune_bucket(k) - 4];
  if (v > 0 && g_force[d] < 0) return v;
  /* Heuristic: about 1024 workgroups (four per CU of the 256 on MI355X) when
   * the tiles alone fill fewer than two per CU, K ranges of at least 256, and
   * at most GEMMX_SPLIT_SCRATCH bytes of partial tiles. */
  const long tiles = (((long)m + bm - 1) / bm) * (((long)n + bn - 1) / bn);
  if (tiles < 1 || tiles >= 512) return 1;
  long sp = 1024 / tiles;
  if (sp > k / 256) sp = k / 256;
  if (sp > GEMMX_MAX_SPLIT) sp = GEMMX_MAX_SPLIT;
  while (sp > 1 && (size_t)sp * (size_t)tiles * (size_t)bm * (size_t)bn * (size_t)es >
                       (size_t)GEMMX_SPLIT_SCRATCH)
    sp--;
  return sp < 2 ? 1 : (int)sp;
}
