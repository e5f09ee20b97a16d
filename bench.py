#!/usr/bin/env python3
"""Benchmark of the MI355X DGEMM/SGEMM hot path.

Metric: "DGEMM/SGEMM TFLOP/s 8192^3: emulated vs native at matched FP64
error".  One step is one 8192^3 DGEMM through the library's device-pointer
entry point (dgemm_x_dev, or dgemm_x_ozaki_dev for the emulated form) with
inputs already resident in HBM:
    transa = transb = 'N', m = n = k = 8192, lda = ldb = ldc = 8200,
    alpha = 1.5, beta = -0.5, A/B/C ~ U[-1,1) from the counter-based generator.
TFLOP/s counts the algorithmic 2*m*n*k flops for every form (slice products
of the emulated form are not counted).  The headline `value` is the faster of
native FP64 MFMA and emulated (Ozaki-I INT8) DGEMM among the forms whose
error against an exact product is no larger than native FP64's on the same
inputs; both are reported, with their errors, under "dgemm".

Also reported (not the headline): SGEMM native/BF16x3 and the other transpose
pairs at 8192^3, the rooflines, and the reference's AVX2/OpenMP DGEMM timed on
this host's cores on a bounded slice of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0, "i8": 5000.0}  # dense, TFLOP/s (TOP/s)
HBM_PEAK = 8000.0  # GB/s
M = N = K = 8192
LD = 8200
ALPHA, BETA = 1.5, -0.5


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--profile-only",
                    choices=["native", "ozaki", "sgemm", "bf16x3", "batch", "copy"],
                    help="run just one form (for rocprofv3)")
    return ap.parse_args()


# ---------------------------------------------------------------- timing
def time_launches(torch, fn, steps, warmup):
    """Runs fn() warmup times, then times `steps` calls bracketed by
    synchronisation; returns (total_ms_wall, [per-launch ms from HIP events]).
    Events are recorded on torch's current stream, the stream fn launches on."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    return wall, [a.elapsed_time(b) for a, b in evs]


# ---------------------------------------------------------------- exact reference
def _two_prod(a, b):
    """Dekker/Veltkamp TwoProduct in numpy: a*b = p + e exactly."""
    c = 134217729.0
    def split(x):
        t = c * x
        hi = t - (t - x)
        return hi, x - hi
    p = a * b
    ah, al = split(a)
    bh, bl = split(b)
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
    return p, e


def exact_entries(arows, bcols, c0, alpha, beta):
    """Correctly rounded alpha*sum(a*b) + beta*c0 per sampled entry (math.fsum
    over the exact products); arows (r, k), bcols (k, c) -> (r, c)."""
    out = np.empty((arows.shape[0], bcols.shape[1]))
    for i in range(arows.shape[0]):
        for j in range(bcols.shape[1]):
            p, e = _two_prod(arows[i], bcols[:, j])
            s = math.fsum(np.concatenate([p, e]))
            # alpha*s rounded once more, then beta*c0: error ~1 ulp, far below
            # the errors measured here
            out[i, j] = math.fsum([alpha * s, beta * c0[i, j]])
    return out


def sampled_error(torch, dA, dB, dC, c0_samples, rows, cols, ta="N", tb="N"):
    """Max componentwise and relative-Frobenius error of C (after one call
    from the original C) at the sampled rows x cols."""
    A2 = dA.view(-1, LD)  # (cols of stored A, LD)
    B2 = dB.view(-1, LD)
    r = torch.tensor(rows, device="cuda")
    c = torch.tensor(cols, device="cuda")
    arows = A2[:, r].T.contiguous()[:, :K].double().cpu().numpy()    # (len rows, K)
    bcols = B2[c, :K].T.contiguous().double().cpu().numpy()         # (K, len cols)
    got = dC.view(-1, LD)[c][:, r].T.double().cpu().numpy()          # (rows, cols)
    ref = exact_entries(arows, bcols, c0_samples, ALPHA, BETA)
    scale = np.abs(arows) @ np.abs(bcols) * abs(ALPHA) + abs(BETA) * np.abs(c0_samples)
    err = np.abs(got - ref)
    return {"max_comp": float(np.max(err / scale)),
            "rel_fro": float(np.linalg.norm(err) / np.linalg.norm(ref))}


# ---------------------------------------------------------------- CPU baseline
def cpu_baseline():
    """The reference's own dgemm_ (compiled from its sources into oracle/_ref)
    on this host's cores, over a bounded slice of the workload: m = k = 8192,
    n = 1024 (one eighth of the 8192^3 flops), ld 8200, alpha 1.5, beta -0.5.
    Falls back to the oracle's Dot2 port on a smaller slice."""
    import ctypes
    import oracle
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    fn = oracle.ref("d")
    if fn is not None:
        m, n, k = M, 1024, K
        kind, sample = "reference", f"dgemm_ NN m=k=8192 n=1024 ld=8200 (1/8 of 8192^3), best of 2"
    else:
        m, n, k = 1024, 128, 1024
        kind, sample = "port", "oracle Dot2 port, NN 1024x128x1024, best of 2"
        fn = oracle.lib().oracle_dgemm
    A = oracle.fill(m, k, LD, 0, 0)
    B = oracle.fill(k, n, LD, 0, 1)
    C = oracle.fill(m, n, LD, 0, 2)
    I = ctypes.c_int
    best = float("inf")
    for _ in range(2):
        t0 = time.perf_counter()
        fn(b"N", b"N", ctypes.byref(I(m)), ctypes.byref(I(n)), ctypes.byref(I(k)),
           ctypes.byref(ctypes.c_double(ALPHA)), A.ctypes.data, ctypes.byref(I(LD)),
           B.ctypes.data, ctypes.byref(I(LD)), ctypes.byref(ctypes.c_double(BETA)),
           C.ctypes.data, ctypes.byref(I(LD)))
        best = min(best, time.perf_counter() - t0)
    return {"value": round(2.0 * m * n * k / best / 1e12, 4), "unit": "TFLOP/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", cores)) if kind == "reference" else cores,
            "kind": kind, "sample": sample, "seconds": round(best, 3)}


# ---------------------------------------------------------------- HBM traffic
PROFILES = os.path.join(ROOT, "profiles")


def profiled_traffic(summary):
    """HBM bytes per launch of a kernel, as recorded in its rocprofv3
    summary under profiles/ (FETCH_SIZE and WRITE_SIZE runs scaled by the
    plain-copy calibration); None when the summary is absent."""
    try:
        with open(os.path.join(PROFILES, summary)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


# ---------------------------------------------------------------- grouped batch
BATCH_COUNT, BATCH_LO, BATCH_HI = 4096, 16, 1024


def batch_workload(torch, gx, dtype, stream, seed=11):
    """4096 independent GEMMs, m, n, k log-uniform in [16, 1024], transa and
    transb uniform over {N, T}, alpha = 1.5, beta = -0.5, operands packed
    back to back (ld = stored rows) in three device buffers.  Returns the
    plan and the single-call argument lists of the same problems."""
    rng = np.random.default_rng(seed)
    tdt = torch.float64 if dtype == "d" else torch.float32
    shapes = []
    offs = [0, 0, 0]
    for _ in range(BATCH_COUNT):
        m, n, k = (int(round(math.exp(rng.uniform(math.log(BATCH_LO), math.log(BATCH_HI)))))
                   for _ in range(3))
        ta, tb = "NT"[rng.integers(2)], "NT"[rng.integers(2)]
        shapes.append((ta, tb, m, n, k, tuple(offs)))
        # 16-byte aligned starts for every operand
        for j, size in enumerate((m * k, k * n, m * n)):
            offs[j] += (size + 1) // 2 * 2
    bufs = [torch.empty(max(o, 1), dtype=tdt, device="cuda") for o in offs]
    for j, b in enumerate(bufs):
        gx.fill(b, b.numel(), 1, b.numel(), seed, j, stream=stream)
    es = bufs[0].element_size()
    specs = []
    for ta, tb, m, n, k, (oa, ob, oc) in shapes:
        ra = m if ta == "N" else k
        rb = k if tb == "N" else n
        specs.append((ta, tb, m, n, k, ALPHA, bufs[0].data_ptr() + oa * es, ra,
                      bufs[1].data_ptr() + ob * es, rb, BETA, bufs[2].data_ptr() + oc * es, m))
    return gx.BatchPlan(dtype, specs), specs, bufs


def bench_batch(torch, gx, stream, steps, warmup):
    out = {"problems": BATCH_COUNT, "sizes": f"log-uniform [{BATCH_LO}, {BATCH_HI}]",
           "trans": "mixed N/T"}
    for d, key in (("d", "f64"), ("s", "f32")):
        plan, specs, bufs = batch_workload(torch, gx, d, stream)
        _, per = time_launches(torch, lambda: plan.run(stream), steps, warmup)
        ms = float(np.mean(per))
        single = gx.dgemm_dev if d == "d" else gx.sgemm_dev

        def loop():
            for sp in specs:
                single(*sp, stream=stream)
        _, per_l = time_launches(torch, loop, 2, 1)
        ms_l = float(np.mean(per_l))
        out[key] = {"tflops": round(plan.flops / ms / 1e9, 2), "ms": round(ms, 3),
                    "gflop": round(plan.flops / 1e9, 2),
                    "single_call_loop_tflops": round(plan.flops / ms_l / 1e9, 2)}
        plan.destroy()
        del bufs
    return out


# ---------------------------------------------------------------- main
def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(every box has one MI355X; sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    import gemm_avx2_amd as gx

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    flops = 2.0 * M * N * K

    def alloc(dt):
        return torch.empty(LD * N, dtype=dt, device=dev)

    # FP64 operands for the headline
    dA, dB, dC = alloc(torch.float64), alloc(torch.float64), alloc(torch.float64)
    gx.fill(dA, M, K, LD, 1, 0, stream=stream)
    gx.fill(dB, K, N, LD, 1, 1, stream=stream)
    gx.fill(dC, M, N, LD, 1, 2, stream=stream)
    C0 = dC.clone()
    torch.cuda.synchronize()

    def native():
        gx.dgemm_dev("N", "N", M, N, K, ALPHA, dA, LD, dB, LD, BETA, dC, LD, stream=stream)

    slices = int(os.environ.get("GEMMX_BENCH_SLICES", "8"))
    ws = torch.empty(gx.dgemm_ozaki_workspace(M, N, K, slices), dtype=torch.uint8, device=dev)

    def ozaki(s=slices):
        gx.dgemm_ozaki_dev("N", "N", M, N, K, ALPHA, dA, LD, dB, LD, BETA, dC, LD, s,
                           ws, ws.numel(), stream=stream)

    if args.profile_only == "copy":
        # a plain device copy of one 8200 x 8192 FP64 operand (537 MB read,
        # 537 MB written): the calibration for the HBM byte counters
        dst = torch.empty_like(dA)
        wall, per = time_launches(torch, lambda: dst.copy_(dA), args.steps, args.warmup)
        print(json.dumps({"profile_only": "copy", "ms": per, "bytes": dA.numel() * 8}))
        return
    if args.profile_only == "batch":
        del ws, dA, dB, dC, C0
        plan, _, bufs = batch_workload(torch, gx, "d", stream)
        wall, per = time_launches(torch, lambda: plan.run(stream), args.steps, args.warmup)
        print(json.dumps({"profile_only": "batch", "ms": per, "gflop": plan.flops / 1e9}))
        return
    if args.profile_only:
        if args.profile_only in ("native", "ozaki"):
            fn = native if args.profile_only == "native" else ozaki
        else:
            del ws, dA, dB, dC, C0
            sA, sB, sC = alloc(torch.float32), alloc(torch.float32), alloc(torch.float32)
            for t, r, c, i in ((sA, M, K, 0), (sB, K, N, 1), (sC, M, N, 2)):
                gx.fill(t, r, c, LD, 1, i, stream=stream)
            if args.profile_only == "sgemm":
                def fn():
                    gx.sgemm_dev("N", "N", M, N, K, ALPHA, sA, LD, sB, LD, BETA, sC, LD,
                                 stream=stream)
            else:
                wsb = torch.empty(gx.sgemm_bf16x3_workspace(M, N, K), dtype=torch.uint8,
                                  device=dev)

                def fn():
                    gx.sgemm_bf16x3_dev("N", "N", M, N, K, ALPHA, sA, LD, sB, LD, BETA, sC,
                                        LD, wsb, wsb.numel(), stream=stream)
        wall, per = time_launches(torch, fn, args.steps, args.warmup)
        print(json.dumps({"profile_only": args.profile_only, "ms": per}))
        return

    result = {}
    # ---- headline candidates: native and emulated DGEMM
    wall_n, per_n = time_launches(torch, native, args.steps, args.warmup)
    ms_n = wall_n / args.steps
    ev_n = float(np.mean(per_n))
    # accuracy on sampled entries, one call from the original C
    rng = np.random.default_rng(0)
    rows = sorted(rng.choice(M, 12, replace=False).tolist())
    cols = sorted(rng.choice(N, 12, replace=False).tolist())
    c0s = C0.view(-1, LD)[torch.tensor(cols, device=dev)][:, torch.tensor(rows, device=dev)].T
    c0s = c0s.double().cpu().numpy()
    dC.copy_(C0)
    native()
    torch.cuda.synchronize()
    err_n = sampled_error(torch, dA, dB, dC, c0s, rows, cols)
    dgemm = {"native": {"tflops": flops / ev_n / 1e9, "ms_event": ev_n, "ms_wall": ms_n,
                        "error_vs_exact": err_n}}
    best_name, best_ms = "native", ev_n
    if not args.quick:
        wall_o, per_o = time_launches(torch, ozaki, max(3, args.steps // 2), 1)
        ev_o = float(np.mean(per_o))
        dC.copy_(C0)
        ozaki()
        torch.cuda.synchronize()
        err_o = sampled_error(torch, dA, dB, dC, c0s, rows, cols)
        dgemm[f"ozaki_s{slices}"] = {"tflops": flops / ev_o / 1e9, "ms_event": ev_o,
                                     "error_vs_exact": err_o}
        if err_o["rel_fro"] <= err_n["rel_fro"] and ev_o < best_ms:
            best_name, best_ms = f"ozaki_s{slices}", ev_o
    result["dgemm"] = dgemm

    # ---- the other transpose pairs and SGEMM (reported, not the headline)
    if not args.quick:
        trans = {}
        for ta, tb in (("N", "T"), ("T", "N"), ("T", "T")):
            def f(ta=ta, tb=tb):
                gx.dgemm_dev(ta, tb, M, N, K, ALPHA, dA, LD, dB, LD, BETA, dC, LD, stream=stream)
            _, per = time_launches(torch, f, 3, 1)
            trans[ta + tb] = round(flops / float(np.mean(per)) / 1e9, 2)
        result["dgemm_native_trans_tflops"] = trans
        del ws
        sA, sB, sC = alloc(torch.float32), alloc(torch.float32), alloc(torch.float32)
        gx.fill(sA, M, K, LD, 1, 0, stream=stream)
        gx.fill(sB, K, N, LD, 1, 1, stream=stream)
        gx.fill(sC, M, N, LD, 1, 2, stream=stream)
        sg = {}
        for ta, tb in (("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")):
            def f(ta=ta, tb=tb):
                gx.sgemm_dev(ta, tb, M, N, K, ALPHA, sA, LD, sB, LD, BETA, sC, LD, stream=stream)
            _, per = time_launches(torch, f, 3, 1)
            sg["native_" + ta + tb] = round(flops / float(np.mean(per)) / 1e9, 2)

        wsb = torch.empty(gx.sgemm_bf16x3_workspace(M, N, K), dtype=torch.uint8, device=dev)

        def fb():
            gx.sgemm_bf16x3_dev("N", "N", M, N, K, ALPHA, sA, LD, sB, LD, BETA, sC, LD,
                                wsb, wsb.numel(), stream=stream)
        _, per = time_launches(torch, fb, 3, 1)
        sg["bf16x3_NN"] = round(flops / float(np.mean(per)) / 1e9, 2)
        result["sgemm_tflops"] = sg
        del sA, sB, sC, wsb
        result["batch"] = bench_batch(torch, gx, stream, max(3, args.steps // 2), 1)

    value = flops / best_ms / 1e9
    ms_step = ms_n if best_name == "native" else best_ms
    line = {
        "metric": "DGEMM/SGEMM TFLOP/s 8192^3: emulated vs native at matched FP64 error",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (counter-based U[-1,1) generator, seed 1)",
        "config": {"workload": f"DGEMM {best_name} NN m=n=k=8192 ld=8200 alpha=1.5 beta=-0.5",
                   "m": M, "n": N, "k": K, "ld": LD, "alpha": ALPHA, "beta": BETA,
                   "mode": best_name},
        "roofline": {"bound": "mfma", "achieved": round(flops / ev_n / 1e9, 3),
                     "peak": PEAK["f64"], "unit": "TFLOP/s",
                     "frac": round(flops / ev_n / 1e9 / PEAK["f64"], 4),
                     "traffic": profiled_traffic("dgemm_native_nn_8192.md"),
                     "kernel": "gemm_native_kernel<double> (native FP64 MFMA)"},
    }
    line.update(result)
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # reported, never fatal for the GPU number
            line["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
