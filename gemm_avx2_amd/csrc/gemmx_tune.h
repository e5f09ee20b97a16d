/*
 * gemmx_tune.h -- run-time tuning table.
 *
 * Replaces the reference's compile-time tuning headers dgemm_tune.h /
 * sgemm_tune.h and their clamping in src/gemm_set_parameters.h:10-61 (edited
 * by hand and rebuilt, README.md:13-15).  Here the block shape is a run-time
 * choice among the compiled tile configurations (GEMMX_NCFG_*), read from a
 * text table written by the autotuner (gemm_avx2_amd/autotune.py).
 *
 * Table file: GEMMX_TUNE_FILE, else gemm_mi355x_tune.txt beside the loaded
 * library.  '#' starts a comment.  Directives:
 *     format 2            table format version (required, first directive)
 *     arch <gfxNNN>       the table is ignored on a device of another arch
 *     rocm <major.minor>  required: the table is ignored by a library built
 *                         with another ROCm (HIP) release, whose compiler
 *                         makes other code of the same kernels
 * and one entry per line:
 *     <dtype:f64|f32> <trans:NN|NT|TN|TT> <mbucket> <nbucket> <kbucket> <cfg> [split]
 * where a bucket is round(log2(dim)) clamped to 4..15, cfg is a
 * configuration id and split (1..GEMMX_MAX_SPLIT, optional) the number of K
 * ranges of a split-K launch (without it a heuristic decides), or, for the
 * arithmetic of DGEMM in the automatic mode,
 *     emu <trans> <mbucket> <nbucket> <kbucket> <native|ozaki|ozaki2> <count>
 * (count: slices 3..8 for ozaki, moduli 8..22 for ozaki2, 0 for native), or,
 * for the grouped-batch kernel of one element type,
 *     batch <dtype:f64|f32> <cfg>
 * A later entry for the same key replaces an earlier one.  Missing entries
 * fall back to a size heuristic (tile) or native arithmetic.
 */
#ifndef GEMMX_TUNE_H
#define GEMMX_TUNE_H

#ifdef __cplusplus
extern "C" {
#endif

#define GEMMX_TUNE_FORMAT 2
#define GEMMX_MAX_SPLIT 64
/* The split heuristic keeps the partial tiles of one call within this many
 * bytes. */
#define GEMMX_SPLIT_SCRATCH (64u << 20)

/* `arch` is the current device's architecture name ("gfx950"), or NULL when
 * there is no device (then a table's arch directive is not checked). */
void gemmx_tune_init(const char *arch);
int gemmx_tune_pick(int es, int ta, int tb, int m, int n, int k);
/* Split-K factor for a native call with tile configuration cfg (1 = one
 * pass): a forced factor, else the table's, else a heuristic. */
int gemmx_tune_pick_split(int es, int ta, int tb, int m, int n, int k, int cfg);
/* Forces one split factor (1..GEMMX_MAX_SPLIT) for every native call of the
 * element size, -1 returns to the table; 0 on success, -1 otherwise. */
int gemmx_tune_force_split(int es, int split);
/* "major.minor" of the ROCm release the library was built with. */
const char *gemmx_tune_rocm(void);
/* Arithmetic of a DGEMM in the automatic mode: GEMMX_MODE_NATIVE, _OZAKI or
 * _OZAKI2, with the slice or modulus count in *count. */
int gemmx_tune_pick_emulated(int ta, int tb, int m, int n, int k, int *count);
/* Arithmetic of an SGEMM in the automatic mode: GEMMX_MODE_NATIVE or
 * GEMMX_MODE_BF16X3 (the table's `emu32` entry of the shape's bucket). */
int gemmx_tune_pick_emulated32(int ta, int tb, int m, int n, int k);
int gemmx_tune_bucket(int dim);
/* Grouped-batch kernel configuration of the table for element size es (8 or
 * 4), or -1 when it has none. */
int gemmx_tune_pick_batch(int es);
/* Tile configuration of a strided batch of `batch` problems of one shape: a
 * forced configuration, else the table's `strided` entry of the bucket when
 * the batch holds at least 256 tiles of 128x128, else gemmx_tune_pick. */
int gemmx_tune_pick_strided(int es, int ta, int tb, int m, int n, int k, int batch);
/* Number of entries loaded from the table file (0 when none). */
int gemmx_tune_entries(void);
/* Forces one configuration for every call of the given element size (8 or
 * 4); -1 returns to the table.  Returns 0, or -1 for an unknown cfg. */
int gemmx_tune_force(int es, int cfg);
/* Loads (or, with NULL, clears) a table at run time; returns entries read,
 * or -1 when the file cannot be read or has a different format or arch. */
int gemmx_tune_load(const char *path);
/* Path of the table that was loaded, or "" when none. */
const char *gemmx_tune_path(void);

#ifdef __cplusplus
}
#endif
#endif
