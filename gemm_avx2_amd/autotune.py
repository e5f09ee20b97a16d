"""Autotuner: measures every native tile configuration on a set of shapes
and writes the tuning table the libraries read at first use.

It takes the place of the reference's hand-edited compile-time tuning
headers (wjc404/GEMM_AVX2 dgemm_tune.h:6-10, sgemm_tune.h:7-11, README.md:13-15:
edit UNROLL_N / block sizes, rebuild, time with an external benchmark).  Here
the choice is a run-time table keyed by (element type, transposes, shape
bucket), bucket = round(log2(dim)) clamped to 4..15, so one build serves every
shape and retuning needs no rebuild.

    python -m gemm_avx2_amd.autotune [--out PATH] [--dtypes d,s]
        [--trans NN,NT,TN,TT] [--set quick|full] [--reps N] [--resume]
        [--emulated on|off] [--strided on|off] [--batch on|off]

For each (dtype, trans, shape) every configuration runs on device-resident
U[-1,1) operands (alpha = 1.5, beta = -0.5); the fastest by median event time
is written (configuration 0 in one pass unless another choice is more than 1%
faster), after its result is checked against configuration 0's on the same
inputs (|difference| within 64 k u (|alpha| k + |beta|), the bound for
U[-1,1) inputs).  For shapes with min(m, n, k) >= 1024 it also picks the
arithmetic of the automatic modes: for dgemm_, Ozaki-II with the smallest
modulus count or Ozaki-I with the smallest slice count whose error is no
larger than native FP64's on phi in {0, 1, 4} inputs (relative Frobenius
error, and half of native's largest componentwise error, over 2048 sampled
entries on two seeds), whichever is fastest, when that beats native (an
`emu` entry); for sgemm_, split-BF16x3 under the same rule against native
FP32 (an `emu32` entry); the tile configuration of strided batches of
m = n = k problems (64 to 1024, about 2^34 flops per batch) per element type
and transpose (a `strided` entry, used for batches of at least 256 tiles of
128x128); and the grouped-batch kernel configuration per element type on a
mixed batch (a `batch` entry).
With --resume, keys already
in the output table are kept and not measured again.  The table is written in
sorted key order with the measured times as comments.
"""
from __future__ import annotations

import argparse
import os
import platform
import re
import sys
import time

FORMAT = 2
TILE_CONFIGS = 10  # compiled native tile configurations (GEMMX_TILE_CONFIGS)
DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib",
                           "gemm_mi355x_tune.txt")
TRANS = ("NN", "NT", "TN", "TT")
ALPHA, BETA = 1.5, -0.5

# Shapes per set; every shape fills the table entry of its bucket.
SHAPES = {
    "quick": [(s, s, s) for s in (256, 1024, 4096, 8192)],
    "full": [(s, s, s) for s in (128, 256, 512, 1024, 2048, 4096, 8192, 16384)]
    + [(8192, 8192, 256), (8192, 8192, 1024), (8192, 256, 8192), (256, 8192, 8192),
       (16384, 1024, 1024), (1024, 16384, 1024), (4096, 4096, 16384), (2048, 2048, 8192),
       (65536, 256, 256), (512, 512, 65536)],
}


def bucket(dim: int) -> int:
    """round(log2(dim)) clamped to 4..15, as the libraries compute it: the
    power of two nearest dim on a log scale, floor(log2(2*dim^2) / 2)."""
    d = max(int(dim), 1)
    b = ((2 * d * d).bit_length() - 1) >> 1
    return min(max(b, 4), 15)


def key_of(dtype: str, trans: str, m: int, n: int, k: int):
    return ("f64" if dtype == "d" else "f32", trans, bucket(m), bucket(n), bucket(k))


# ------------------------------------------------------------------ table I/O
_ENTRY = re.compile(r"^\s*(f64|f32)\s+(NN|NT|TN|TT)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)"
                    r"(?:\s+(\d+))?\s*(#.*)?$")
MAX_SPLIT = 64  # GEMMX_MAX_SPLIT
SPLIT_CONFIGS = (0, 1, 5, 7, 8, 9)  # native tile configurations with split-K kernels
_EMU = re.compile(r"^\s*emu\s+(NN|NT|TN|TT)\s+(\d+)\s+(\d+)\s+(\d+)\s+(native|ozaki|ozaki2)"
                  r"\s+(\d+)\s*(#.*)?$")
_EMU_RANGE = {"native": (0, 0), "ozaki": (3, 8), "ozaki2": (8, 22)}
_EMU32 = re.compile(r"^\s*emu32\s+(NN|NT|TN|TT)\s+(\d+)\s+(\d+)\s+(\d+)\s+(native|bf16x3)"
                    r"\s*(#.*)?$")
_BATCH = re.compile(r"^\s*batch\s+(f64|f32)\s+(\d+)\s*(#.*)?$")
_STRIDED = re.compile(r"^\s*strided\s+(f64|f32)\s+(NN|NT|TN|TT)\s+(\d+)\s+(\d+)\s+(\d+)"
                      r"\s+(\d+)\s*(#.*)?$")
BATCH_CONFIGS = 4  # compiled grouped-batch kernel configurations (GEMMX_BATCH_CONFIGS)


def emu_key(trans, m, n, k):
    return ("emu", trans, bucket(m), bucket(n), bucket(k))


def emu32_key(trans, m, n, k):
    return ("emu32", trans, bucket(m), bucket(n), bucket(k))


def read_table(path):
    """(entries, comments, arch or None, rocm or None); later lines replace
    earlier ones, like the libraries' reader.  A tile entry's value is its
    configuration, or (configuration, split) when the line gives a split."""
    entries, notes, arch, rocm = {}, {}, None, None
    if not os.path.exists(path):
        return entries, notes, arch, rocm
    with open(path) as f:
        for line in f:
            body = line.split("#", 1)[0].split()
            if len(body) == 2 and body[0] == "format" and int(body[1]) != FORMAT:
                raise ValueError(f"{path}: table format {body[1]}, expected {FORMAT}")
            if len(body) == 2 and body[0] == "arch":
                arch = body[1]
                continue
            if len(body) == 2 and body[0] == "rocm":
                rocm = body[1]
                continue
            m = _BATCH.match(line)
            if m:
                key = ("batch", m.group(1))
                if 0 <= int(m.group(2)) < BATCH_CONFIGS:
                    entries[key] = int(m.group(2))
                    notes[key] = (m.group(3) or "").lstrip("# ").rstrip()
                continue
            m = _STRIDED.match(line)
            if m:
                key = ("strided", m.group(1), m.group(2), int(m.group(3)), int(m.group(4)),
                       int(m.group(5)))
                if all(4 <= b <= 15 for b in key[3:]) and 0 <= int(m.group(6)) < TILE_CONFIGS:
                    entries[key] = int(m.group(6))
                    notes[key] = (m.group(7) or "").lstrip("# ").rstrip()
                continue
            m = _EMU32.match(line)
            if m:
                key = ("emu32", m.group(1), int(m.group(2)), int(m.group(3)), int(m.group(4)))
                if all(4 <= b <= 15 for b in key[2:]):
                    entries[key] = m.group(5)
                    notes[key] = (m.group(6) or "").lstrip("# ").rstrip()
                continue
            m = _EMU.match(line)
            if m:
                key = ("emu", m.group(1), int(m.group(2)), int(m.group(3)), int(m.group(4)))
                scheme, cnt = m.group(5), int(m.group(6))
                lo, hi = _EMU_RANGE[scheme]
                if all(4 <= b <= 15 for b in key[2:]) and lo <= cnt <= hi:
                    entries[key] = (scheme, cnt)
                    notes[key] = (m.group(7) or "").lstrip("# ").rstrip()
                continue
            m = _ENTRY.match(line)
            if m:
                key = (m.group(1), m.group(2), int(m.group(3)), int(m.group(4)), int(m.group(5)))
                cfg = int(m.group(6))
                split = int(m.group(7)) if m.group(7) is not None else None
                # the libraries skip entries outside the buckets or configurations
                if not all(4 <= b <= 15 for b in key[2:]) or not 0 <= cfg < TILE_CONFIGS:
                    continue
                if split is not None and not 1 <= split <= MAX_SPLIT:
                    continue
                entries[key] = cfg if split is None else (cfg, split)
                notes[key] = (m.group(8) or "").lstrip("# ").rstrip()
    return entries, notes, arch, rocm


def write_table(path, entries, notes, arch, header=(), rocm=None):
    """Writes the table in sorted key order (the same entries always give
    the same file).  rocm: the release key, by default that of the libraries
    this package loads."""
    if rocm is None:
        import gemm_avx2_amd as gx
        rocm = gx.build_rocm()
    lines = ["# gemm_mi355x tuning table, written by gemm_avx2_amd.autotune"]
    lines += [f"# {h}" for h in header]
    lines += [f"format {FORMAT}"]
    if arch:
        lines += [f"arch {arch}"]
    lines += [f"rocm {rocm}"]
    lines += ["# dtype trans mbucket nbucket kbucket cfg [split]",
              "# emu trans mbucket nbucket kbucket scheme count   (DGEMM arithmetic, auto mode)",
              "# emu32 trans mbucket nbucket kbucket scheme   (SGEMM arithmetic, auto mode)",
              "# strided dtype trans mbucket nbucket kbucket cfg   (strided batches that fill the GPU)",
              "# batch dtype cfg   (grouped-batch kernel)"]
    order = {"f64": 0, "f32": 1, "emu": 2, "emu32": 3}
    tiles = [k for k in entries if k[0] not in ("batch", "strided")]
    for key in sorted(tiles, key=lambda k: (order[k[0]], TRANS.index(k[1]), k[2], k[3], k[4])):
        note = notes.get(key, "")
        val = entries[key]
        if isinstance(val, tuple):
            val = f"{val[0]} {val[1]}"
        lines.append(f"{key[0]} {key[1]} {key[2]} {key[3]} {key[4]} {val}"
                     + (f"  # {note}" if note else ""))
    strided = [k for k in entries if k[0] == "strided"]
    for key in sorted(strided, key=lambda k: (k[1] != "f64", TRANS.index(k[2]), k[3:])):
        note = notes.get(key, "")
        lines.append("strided " + " ".join(str(v) for v in key[1:]) + f" {entries[key]}"
                     + (f"  # {note}" if note else ""))
    for key in sorted(k for k in entries if k[0] == "batch"):
        note = notes.get(key, "")
        lines.append(f"batch {key[1]} {entries[key]}" + (f"  # {note}" if note else ""))
    tmp = path + ".tmp"
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(tmp, "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(tmp, path)


def pending(wanted, entries, resume):
    """(dtype, trans, m, n, k) still to measure, one per table key: all of
    them, or with resume those whose key is not in the table yet."""
    out, seen = [], set()
    for w in wanted:
        key = key_of(*w)
        if key in seen or (resume and key in entries):
            continue
        seen.add(key)
        out.append(w)
    return out


# ------------------------------------------------------------------ measurement
def _median(xs):
    xs = sorted(xs)
    n = len(xs)
    return xs[n // 2] if n % 2 else 0.5 * (xs[n // 2 - 1] + xs[n // 2])


TILE_DIMS = ((128, 128), (64, 64), (256, 128), (128, 128), (128, 128), (128, 128), (256, 128),
             (64, 64), (64, 64), (64, 64))
# K slice depth of each configuration's FP64 and FP32 kernels (split-K ranges
# are whole slices)
TILE_BK = {"d": (16, 16, 16, 32, 16, 16, 16, 16, 32, 32),
           "s": (16, 16, 16, 32, 16, 32, 16, 32, 64, 32)}


def split_candidates(cfg, m, n, k, es):
    """Split factors worth timing for configuration cfg: powers of two from
    2 while the tiles and their K ranges give at most 4096 workgroups, each
    K range at least 64 deep, and the partial tiles fit in 256 MiB; none for
    a configuration without split-K kernels or with enough tiles already."""
    if cfg not in SPLIT_CONFIGS:
        return []
    bm, bn = TILE_DIMS[cfg]
    tiles = -(-m // bm) * -(-n // bn)
    if tiles >= 1024:
        return []
    out, sp = [], 2
    while (sp <= MAX_SPLIT and tiles * sp <= 4096 and k // sp >= 64
           and sp * tiles * bm * bn * es <= (256 << 20)):
        out.append(sp)
        sp *= 2
    return out


def measure(torch, gx, dtype, trans, m, n, k, reps):
    """Median event time (ms) of every configuration in one pass and, for
    the configurations with split-K kernels, of every candidate split factor
    (split_candidates), keyed (cfg, split); which results agree with
    configuration 0's one-pass result; and the fastest agreeing (cfg, split)."""
    tdt = torch.float64 if dtype == "d" else torch.float32
    u = 2.0 ** -53 if dtype == "d" else 2.0 ** -24
    es = 8 if dtype == "d" else 4
    ra, ca = (m, k) if trans[0] == "N" else (k, m)
    rb, cb = (k, n) if trans[1] == "N" else (n, k)
    A = torch.empty(ra * ca, dtype=tdt, device="cuda")
    B = torch.empty(rb * cb, dtype=tdt, device="cuda")
    C0 = torch.empty(m * n, dtype=tdt, device="cuda")
    gx.fill(A, ra, ca, ra, 97, 0)
    gx.fill(B, rb, cb, rb, 97, 1)
    gx.fill(C0, m, n, m, 97, 2)
    C = torch.empty_like(C0)
    dev = gx.dgemm_dev if dtype == "d" else gx.sgemm_dev
    stream = torch.cuda.current_stream()

    def call():
        dev(trans[0], trans[1], m, n, k, ALPHA, A, ra, B, rb, BETA, C, m, stream=stream)

    times, ok, ref = {}, {}, None
    bound = 64 * k * u * (abs(ALPHA) * k + abs(BETA))
    runs = [(c, 1) for c in range(gx.TILE_CONFIGS)]
    runs += [(c, sp) for c in SPLIT_CONFIGS for sp in split_candidates(c, m, n, k, es)]
    try:
        for cfg, sp in runs:
            gx.set_tile_config(dtype, cfg)
            gx.set_split_config(dtype, sp)
            C.copy_(C0)
            call()
            torch.cuda.synchronize()
            if ref is None:
                ref = C.clone()
                ok[(cfg, sp)] = True
            else:
                ok[(cfg, sp)] = bool((C - ref).abs().max().item() <= 2 * bound)
            ev = []
            for _ in range(reps):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                call()
                e1.record(stream)
                e1.synchronize()
                ev.append(e0.elapsed_time(e1))
            times[(cfg, sp)] = _median(ev)
    finally:
        gx.set_tile_config(dtype, -1)
        gx.set_split_config(dtype, -1)
    best = min((t, key) for key, t in times.items() if ok[key])[1]
    # a choice within TILE_MARGIN of configuration 0 in one pass is a tie
    # (one measurement's noise, and the leading dimension can tip it the
    # other way): configuration 0 is kept
    if times[(0, 1)] * TILE_MARGIN <= times[best]:
        best = (0, 1)
    return times, ok, best


TILE_MARGIN = 0.99


def _times_note(m, n, k, times, ok):
    return f"{m}x{n}x{k} ms: " + " ".join(
        (f"c{c}" if sp == 1 else f"c{c}s{sp}") + f"={tm:.3f}" + ("" if ok[(c, sp)] else "!")
        for (c, sp), tm in sorted(times.items()))


# ------------------------------------------------------------------ strided batch
STRIDED_SIZES = (64, 128, 256, 512, 1024)  # m = n = k of the measured strided batches
STRIDED_FLOP = 2.0 ** 34  # the count is chosen for about this many flops (at least 64 problems)


def strided_count(m, n, k):
    return max(64, int(STRIDED_FLOP // (2 * m * n * k)))


def measure_strided(torch, gx, dtype, trans, m, n, k, reps, count=None):
    """Median event time (ms) of a strided batch of `count` problems (operands
    back to back) per tile configuration; which results agree with
    configuration 0's; and the fastest agreeing configuration."""
    count = count or strided_count(m, n, k)
    tdt = torch.float64 if dtype == "d" else torch.float32
    u = 2.0 ** -53 if dtype == "d" else 2.0 ** -24
    ra, ca = (m, k) if trans[0] == "N" else (k, m)
    rb, cb = (k, n) if trans[1] == "N" else (n, k)
    A = torch.empty(count * ra * ca, dtype=tdt, device="cuda")
    B = torch.empty(count * rb * cb, dtype=tdt, device="cuda")
    C0 = torch.empty(count * m * n, dtype=tdt, device="cuda")
    gx.fill(A, ra, count * ca, ra, 97, 0)
    gx.fill(B, rb, count * cb, rb, 97, 1)
    gx.fill(C0, m, count * n, m, 97, 2)
    C = torch.empty_like(C0)
    sb = gx.dgemm_strided_batched_dev if dtype == "d" else gx.sgemm_strided_batched_dev
    stream = torch.cuda.current_stream()

    def call():
        sb(trans[0], trans[1], m, n, k, ALPHA, A, ra, ra * ca, B, rb, rb * cb, BETA, C, m,
           m * n, count, stream=stream)

    times, ok, ref = {}, {}, None
    bound = 64 * k * u * (abs(ALPHA) * k + abs(BETA))
    try:
        for cfg in range(gx.TILE_CONFIGS):
            gx.set_tile_config(dtype, cfg)
            C.copy_(C0)
            call()
            torch.cuda.synchronize()
            if ref is None:
                ref = C.clone()
                ok[cfg] = True
            else:
                ok[cfg] = bool((C - ref).abs().max().item() <= 2 * bound)
            times[cfg] = _timed(torch, call, reps, stream)
    finally:
        gx.set_tile_config(dtype, -1)
    del A, B, C0, C, ref
    best = min((t, c) for c, t in times.items() if ok[c])[1]
    return times, ok, best, count


# ------------------------------------------------------------------ grouped batch
BATCH_MIX = (4096, 16, 1024)  # problems; m, n, k log-uniform integers in [lo, hi]


def measure_batch(torch, gx, dtype, reps, mix=BATCH_MIX, seed=11):
    """Median event time (ms) of a grouped-batch plan per kernel
    configuration on a mixed batch (m, n, k log-uniform, transposes uniform
    over {N, T}, alpha ~ U[0.5, 2], beta ~ U[-1, 1], ld = stored rows), and
    the configuration batch_pick chooses among those whose C agrees with
    configuration 0's (each result from the same C, within
    64 k u (|alpha| |A||B| + |beta C|) for U[-1,1) operands)."""
    import math
    import numpy as np
    count, lo, hi = mix
    rng = np.random.default_rng(seed)
    tdt = torch.float64 if dtype == "d" else torch.float32
    u = 2.0 ** -53 if dtype == "d" else 2.0 ** -24
    shapes, offs = [], [0, 0, 0]
    for _ in range(count):
        m, n, k = (int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
                   for _ in range(3))
        ta, tb = "NT"[rng.integers(2)], "NT"[rng.integers(2)]
        shapes.append((ta, tb, m, n, k, float(rng.uniform(0.5, 2.0)),
                       float(rng.uniform(-1.0, 1.0)), tuple(offs)))
        for j, size in enumerate((m * k, k * n, m * n)):
            offs[j] += (size + 1) // 2 * 2  # 16-byte aligned operands
    bufs = [torch.empty(max(o, 1), dtype=tdt, device="cuda") for o in offs]
    for j, b in enumerate(bufs):
        gx.fill(b, b.numel(), 1, b.numel(), seed, j)
    c0 = bufs[2].clone()
    es = bufs[0].element_size()
    specs = [(ta, tb, m, n, k, alpha, bufs[0].data_ptr() + oa * es, m if ta == "N" else k,
              bufs[1].data_ptr() + ob * es, k if tb == "N" else n, beta,
              bufs[2].data_ptr() + oc * es, m)
             for ta, tb, m, n, k, alpha, beta, (oa, ob, oc) in shapes]
    kmax = max(s[4] for s in specs)
    bound = 64 * kmax * u * (2.0 * kmax + 1.0)
    stream = torch.cuda.current_stream()
    times, ok, ref = [], [], None
    try:
        for cfg in range(BATCH_CONFIGS):
            gx.set_batch_config(dtype, cfg)
            plan = gx.BatchPlan(dtype, specs)
            bufs[2].copy_(c0)
            plan.run(stream)
            torch.cuda.synchronize()
            if ref is None:
                ref = bufs[2].clone()
                ok.append(True)
            else:
                ok.append(bool((bufs[2] - ref).abs().max().item() <= bound))
            ev = []
            for _ in range(reps):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                plan.run(stream)
                e1.record(stream)
                e1.synchronize()
                ev.append(e0.elapsed_time(e1))
            times.append(_median(ev))
            plan.destroy()
    finally:
        gx.set_batch_config(dtype, -1)
    return times, ok, batch_pick(times, ok)


BATCH_MARGIN = 0.97  # another batch configuration must beat configuration 0 by 3%


def batch_pick(times, ok):
    """The batch configuration to record: the fastest agreeing one when it
    is at least 3% faster than configuration 0 (the default 64x64x16 tile),
    else 0.  Runs of one plan vary by about 2% between replays, so a smaller
    lead is not a measured difference."""
    t, c = min((t, c) for c, t in enumerate(times) if ok[c])
    return c if t < BATCH_MARGIN * times[0] else 0


# ------------------------------------------------------------------ emulated DGEMM
PHIS = (0.0, 1.0, 4.0)
MODULI = tuple(range(14, 23))
EMU_SEEDS = (41, 43)
EMU_SAMPLES = 2048
# The largest componentwise error over a few thousand sampled entries varies
# between inputs of one distribution by up to a factor of about 1.5 (measured
# on 8192^3, phi = 4), so an emulated form must stay a factor 2 below native's
# on that measure to match it on other inputs too.
EMU_COMP_MARGIN = 0.5


def _two_prod(a, b):
    """Dekker/Veltkamp TwoProduct: a*b = p + e exactly."""
    c = 134217729.0

    def split(x):
        t = c * x
        hi = t - (t - x)
        return hi, x - hi
    p = a * b
    ah, al = split(a)
    bh, bl = split(b)
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def dot2_pairs(rows, cols, ri, cj, alpha=1.0, c1=0.0):
    """alpha * (rows[ri[t]] . cols[cj[t]]) + c1[t] for every pair t with Dot2
    (TwoProduct and TwoSum: as accurate as twice the working precision, then
    rounded once), vectorised over the pairs; also sum |a||b| per pair.
    rows (nr, k) and cols (nc, k) are float64 numpy arrays."""
    import numpy as np
    X = np.ascontiguousarray(rows[ri].T)  # (k, count)
    Y = np.ascontiguousarray(cols[cj].T)
    s = np.broadcast_to(np.asarray(c1, dtype=np.float64), (X.shape[1],)).copy()
    c = np.zeros_like(s)
    ab = np.zeros_like(s)
    for p in range(X.shape[0]):
        h, e = _two_prod(X[p], Y[p] * alpha)
        t = s + h
        z = t - s
        c += ((s - (t - z)) + (h - z)) + e
        s = t
        ab += np.abs(X[p] * Y[p])
    return s + c, ab


def sample_errors(got, exact, ab):
    """Relative Frobenius error and the largest error relative to
    sum |a||b| (componentwise) over sampled entries."""
    import numpy as np
    err = np.abs(np.asarray(got, dtype=np.float64) - exact)
    return {"rel_fro": float(np.linalg.norm(err) / max(np.linalg.norm(exact), 1e-300)),
            "max_comp": float(np.max(err / (ab + 1e-300)))}


SLICES = tuple(range(3, 9))


def _timed(torch, fn, reps, stream):
    fn()
    ev = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        ev.append(e0.elapsed_time(e1))
    return _median(ev)


def _qualify(torch, gx, trans, m, n, k, A, B, C, native, forms, seeds, samples):
    """The names in `forms` (name -> launch) whose error at `samples` random
    entries (alpha 1, beta 0, against Dot2 values of the same inputs) is no
    larger than native's in relative Frobenius norm and at most
    EMU_COMP_MARGIN times native's largest error relative to sum |a||b|, for
    entries (u - 0.5) exp(phi g), phi in PHIS, on every seed; also, per name,
    the largest ratio of its error to native's on each measure over the cases
    it was checked on (a name drops out at its first failing case).  A
    non-finite error never qualifies; a non-finite native error is an error."""
    import numpy as np
    ra, ca = (m, k) if trans[0] == "N" else (k, m)
    rb, cb = (k, n) if trans[1] == "N" else (n, k)
    qual = set(forms)
    ratio = {name: [0.0, 0.0] for name in forms}
    for seed in seeds:
        rng = np.random.default_rng(seed)
        is_, js = rng.integers(0, m, samples), rng.integers(0, n, samples)
        ui, ri = np.unique(is_, return_inverse=True)
        uj, cj = np.unique(js, return_inverse=True)
        tui, tuj = torch.from_numpy(ui).cuda(), torch.from_numpy(uj).cuda()
        flat = torch.from_numpy(is_.astype(np.int64) + js.astype(np.int64) * m).cuda()
        for phi in PHIS:
            gx.fill(A, ra, ca, ra, seed, 0, 1, phi)
            gx.fill(B, rb, cb, rb, seed, 1, 1, phi)
            torch.cuda.synchronize()
            # op(A) rows: A(i, :) of an m x k (N) or A(:, i) of a k x m (T) array
            rows = A.view(k, m)[:, tui].T if trans[0] == "N" else A.view(m, k)[tui]
            cols = B.view(n, k)[tuj] if trans[1] == "N" else B.view(k, n)[:, tuj].T
            # FP32 entries: their products are exact in FP64, Dot2 sums them
            exact, ab = dot2_pairs(rows.contiguous().double().cpu().numpy(),
                                   cols.contiguous().double().cpu().numpy(), ri, cj)

            def err(fn):
                # beta = 0 still multiplies C (dgemm_'s rule), so C must hold
                # no NaN left over from earlier allocations
                C.zero_()
                fn()
                torch.cuda.synchronize()
                return sample_errors(C[flat].double().cpu().numpy(), exact, ab)
            en = err(native)
            if not (np.isfinite(en["rel_fro"]) and np.isfinite(en["max_comp"])):
                raise RuntimeError(f"autotune: native {trans} {m}x{n}x{k} error is not finite "
                                   f"(seed {seed}, phi {phi}): {en}")
            for name in sorted(qual):
                e = err(forms[name])
                for i, what in enumerate(("rel_fro", "max_comp")):
                    q = e[what] / max(en[what], 1e-300)
                    ratio[name][i] = max(ratio[name][i], q if np.isfinite(q) else np.inf)
                # written so that a NaN error fails
                if not (e["rel_fro"] <= en["rel_fro"] and
                        e["max_comp"] <= EMU_COMP_MARGIN * en["max_comp"]):
                    qual.discard(name)
    return qual, ratio


def _ratio_note(r):
    return f"error x{r[0]:.2f} / x{r[1]:.2f} of native"


def measure_emulated(torch, gx, trans, m, n, k, reps, moduli=MODULI, seeds=EMU_SEEDS,
                     samples=EMU_SAMPLES, slices=SLICES):
    """DGEMM arithmetic for one shape: native FP64, Ozaki-II with the
    smallest qualifying modulus count or Ozaki-I with the smallest qualifying
    slice count (_qualify: error no larger than native's), whichever of them
    is fastest, when it is faster than native.  Returns (scheme, count,
    note)."""
    ra, ca = (m, k) if trans[0] == "N" else (k, m)
    rb, cb = (k, n) if trans[1] == "N" else (n, k)
    A = torch.empty(ra * ca, dtype=torch.float64, device="cuda")
    B = torch.empty(rb * cb, dtype=torch.float64, device="cuda")
    C = torch.empty(m * n, dtype=torch.float64, device="cuda")
    wbytes = max(gx.dgemm_ozaki2_workspace(m, n, k, max(moduli)) if moduli else 0,
                 gx.dgemm_ozaki_workspace(m, n, k, max(slices)) if slices else 0)
    ws = torch.empty(wbytes, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def native():
        gx.dgemm_dev(trans[0], trans[1], m, n, k, 1.0, A, ra, B, rb, 0.0, C, m, stream=stream)

    def ozaki2(nm):
        gx.dgemm_ozaki2_dev(trans[0], trans[1], m, n, k, 1.0, A, ra, B, rb, 0.0, C, m, nm, ws,
                            ws.numel(), stream=stream)

    def ozaki(s):
        gx.dgemm_ozaki_dev(trans[0], trans[1], m, n, k, 1.0, A, ra, B, rb, 0.0, C, m, s, ws,
                           ws.numel(), stream=stream)

    forms = {("ozaki2", nm): (lambda nm=nm: ozaki2(nm)) for nm in moduli}
    forms.update({("ozaki", s): (lambda s=s: ozaki(s)) for s in slices})
    qual, ratio = _qualify(torch, gx, trans, m, n, k, A, B, C, native, forms, seeds, samples)
    t_nat = _timed(torch, native, reps, stream)
    best, t_best = ("native", 0), t_nat
    parts = [f"{m}x{n}x{k} native {t_nat:.3f} ms"]
    for scheme, tag, what in (("ozaki2", "n", "modulus"), ("ozaki", "s", "slice")):
        counts = sorted(c for sc, c in qual if sc == scheme)
        if not counts:
            if (moduli if scheme == "ozaki2" else slices):
                parts.append(f"no {what} count matches its error")
            continue
        t = _timed(torch, forms[(scheme, counts[0])], reps, stream)
        parts.append(f"{scheme} {tag}{counts[0]} {t:.3f} ms, "
                     + _ratio_note(ratio[(scheme, counts[0])]))
        if t < t_best:
            best, t_best = (scheme, counts[0]), t
    return best[0], best[1], "; ".join(parts)


def measure_emulated32(torch, gx, trans, m, n, k, reps, seeds=EMU_SEEDS, samples=EMU_SAMPLES):
    """SGEMM arithmetic for one shape: split-BF16x3 when its error is no
    larger than native FP32's (_qualify) and it is faster, else native.
    Returns (scheme, note)."""
    ra, ca = (m, k) if trans[0] == "N" else (k, m)
    rb, cb = (k, n) if trans[1] == "N" else (n, k)
    A = torch.empty(ra * ca, dtype=torch.float32, device="cuda")
    B = torch.empty(rb * cb, dtype=torch.float32, device="cuda")
    C = torch.empty(m * n, dtype=torch.float32, device="cuda")
    ws = torch.empty(gx.sgemm_bf16x3_workspace(m, n, k), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def native():
        gx.sgemm_dev(trans[0], trans[1], m, n, k, 1.0, A, ra, B, rb, 0.0, C, m, stream=stream)

    def bf16x3():
        gx.sgemm_bf16x3_dev(trans[0], trans[1], m, n, k, 1.0, A, ra, B, rb, 0.0, C, m, ws,
                            ws.numel(), stream=stream)

    qual, ratio = _qualify(torch, gx, trans, m, n, k, A, B, C, native, {"bf16x3": bf16x3},
                           seeds, samples)
    t_nat = _timed(torch, native, reps, stream)
    t_b = _timed(torch, bf16x3, reps, stream)
    note = (f"{m}x{n}x{k} native {t_nat:.3f} ms, bf16x3 {t_b:.3f} ms, "
            + _ratio_note(ratio["bf16x3"])
            + ("" if qual else "; bf16x3 does not match its error"))
    return ("bf16x3" if qual and t_b < t_nat else "native"), note


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--dtypes", default="d,s")
    ap.add_argument("--trans", default=",".join(TRANS))
    ap.add_argument("--set", default="quick", choices=sorted(SHAPES))
    ap.add_argument("--shapes", default=None, help="m,n,k;m,n,k;... instead of --set")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--resume", action="store_true")
    ap.add_argument("--emulated", choices=["on", "off"], default="on",
                    help="also choose the arithmetic per bucket for shapes with min(m, n, k) >= "
                         "1024: DGEMM native, Ozaki-II or Ozaki-I; SGEMM native or BF16x3")
    ap.add_argument("--batch", choices=["on", "off"], default="on",
                    help="also choose the grouped-batch kernel configuration per element type "
                         "on a mix of %d problems with m, n, k log-uniform in [%d, %d]" % BATCH_MIX)
    ap.add_argument("--strided", choices=["on", "off"], default="on",
                    help="also choose the tile configuration of strided batches of m = n = k "
                         "in %s per dtype and transpose, each batch about 2^34 flops"
                         % (STRIDED_SIZES,))
    args = ap.parse_args(argv)

    import torch
    import gemm_avx2_amd as gx

    if not torch.cuda.is_available():
        raise SystemExit("autotune: no GPU visible")
    props = torch.cuda.get_device_properties(0)
    arch = getattr(props, "gcnArchName", "unknown").split(":")[0]
    shapes = ([tuple(int(v) for v in s.split(",")) for s in args.shapes.split(";")]
              if args.shapes else SHAPES[args.set])
    entries, notes, old_arch, old_rocm = (read_table(args.out) if args.resume
                                          else ({}, {}, None, None))
    if old_arch and old_arch != arch:
        raise SystemExit(f"autotune: {args.out} is for {old_arch}, this device is {arch}")
    if old_rocm and old_rocm != gx.build_rocm():
        raise SystemExit(f"autotune: {args.out} is for ROCm {old_rocm}, the libraries are built "
                         f"with {gx.build_rocm()}")
    header = [f"device: {props.name}, {props.multi_processor_count} CUs, arch {arch}",
              f"libraries built with ROCm {gx.build_rocm()}; torch {torch.__version__} "
              f"(HIP {torch.version.hip}); host: {platform.node()}",
              f"written: {time.strftime('%Y-%m-%d %H:%M:%S')}; set: {args.shapes or args.set}; "
              f"reps: {args.reps}"]
    wanted = []
    for d in args.dtypes.split(","):
        for t in args.trans.split(","):
            for (m, n, k) in shapes:
                wanted.append((d, t, m, n, k))
    todo = pending(wanted, entries, args.resume)
    for i, (d, t, m, n, k) in enumerate(todo):
        times, ok, best = measure(torch, gx, d, t, m, n, k, args.reps)
        key = key_of(d, t, m, n, k)
        entries[key] = best
        notes[key] = _times_note(m, n, k, times, ok)
        write_table(args.out, entries, notes, arch, header)  # after every shape: resumable
        print(f"[{i + 1}/{len(todo)}] {key[0]} {t} {m}x{n}x{k}: cfg {best[0]} split {best[1]} "
              f"({times[best]:.3f} ms; cfg0 {times[(0, 1)]:.3f} ms)", flush=True)
    for d in (args.dtypes.split(",") if args.emulated == "on" else ()):
        keyf = emu_key if d == "d" else emu32_key
        etodo, seen = [], set()
        for t in args.trans.split(","):
            for (m, n, k) in shapes:
                key = keyf(t, m, n, k)
                if min(m, n, k) < 1024 or key in seen or (args.resume and key in entries):
                    continue
                seen.add(key)
                etodo.append((t, m, n, k))
        for i, (t, m, n, k) in enumerate(etodo):
            key = keyf(t, m, n, k)
            if d == "d":
                scheme, cnt, note = measure_emulated(torch, gx, t, m, n, k, args.reps)
                entries[key] = (scheme, cnt)
            else:
                scheme, note = measure_emulated32(torch, gx, t, m, n, k, args.reps)
                entries[key], cnt = scheme, ""
            notes[key] = note
            write_table(args.out, entries, notes, arch, header)
            print(f"[{key[0]} {i + 1}/{len(etodo)}] {t} {m}x{n}x{k}: {scheme} {cnt} ({note})",
                  flush=True)
    for d in (args.dtypes.split(",") if args.strided == "on" else ()):
        for t in args.trans.split(","):
            for s in STRIDED_SIZES:
                key = ("strided", "f64" if d == "d" else "f32", t, bucket(s), bucket(s), bucket(s))
                if args.resume and key in entries:
                    continue
                times, ok, best, count = measure_strided(torch, gx, d, t, s, s, s, args.reps)
                entries[key] = best
                notes[key] = f"{count} x {s}^3 ms: " + " ".join(
                    f"c{c}={tm:.3f}" + ("" if ok[c] else "!") for c, tm in sorted(times.items()))
                write_table(args.out, entries, notes, arch, header)
                print(f"[strided] {key[1]} {t} {count}x{s}^3: cfg {best} ({times[best]:.3f} ms; "
                      f"cfg0 {times[0]:.3f} ms)", flush=True)
    if args.batch == "on":
        for d in args.dtypes.split(","):
            key = ("batch", "f64" if d == "d" else "f32")
            if args.resume and key in entries:
                continue
            times, ok, best = measure_batch(torch, gx, d, args.reps)
            entries[key] = best
            notes[key] = ("mix of %d problems, m, n, k in [%d, %d] ms: " % BATCH_MIX) + " ".join(
                f"c{c}={tm:.3f}" + ("" if ok[c] else "!") for c, tm in enumerate(times))
            write_table(args.out, entries, notes, arch, header)
            print(f"[batch] {key[1]}: cfg {best} ({times[best]:.3f} ms; cfg0 {times[0]:.3f} ms)",
                  flush=True)
    print(f"autotune: {len(entries)} entries in {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
