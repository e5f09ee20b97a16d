"""Grouped batch (dgemm_x_batch_* / sgemm_x_batch_*): many independent GEMMs
of uneven shapes and mixed transposes in one persistent launch, replayed from
a hipGraph.  The reference has no batch entry point (one GEMM per call,
wjc404/GEMM_AVX2 src/gemm_driver.c:118), so every problem of a batch is held
to the single-call bar: componentwise against the oracle (oracle_dgemm,
oracle_sgemm) with the tolerance of the native tests,
  |C - C_exact| <= 64 * k * u * (|alpha| |op(A)||op(B)| + |beta C0|),
ld padding untouched, and the reference's beta/alpha rules (beta != 1 scales
C first so beta == 0 keeps NaN, alpha == 0 or k == 0 reads neither A nor B).
Graph replays are checked one by one with eager work between them.
"""
import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_util import U32, U64, abs_bound, from_device, max_comp, stored_shape, to_device

pytestmark = pytest.mark.gpu


def _tol(dt, k):
    return 64 * max(k, 1) * (U64 if dt == np.float64 else U32)


def _problems(rng, count, lo, hi, dt):
    """Random problems: log-uniform sizes in [lo, hi], every trans spelling,
    odd leading dimensions, alpha == 0 / k == 0 / beta == 0 / beta == 1 cases,
    NaN in C where beta == 0 and in the ld padding of A and B."""
    out = []
    for i in range(count):
        m, n, k = (int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(3))
        ta = "NnTtC"[rng.integers(5)]
        tb = "NnTtC"[rng.integers(5)]
        alpha = float(rng.choice([1.0, -0.75, 2.5, 0.0])) if i % 7 else 1.0
        beta = float(rng.choice([0.0, 1.0, -0.5, 1.25]))
        if i % 11 == 5:
            k = 0
        pad = int(rng.integers(0, 4))
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        lda, ldb, ldc = max(ra, 1) + pad, max(rb, 1) + pad, m + pad
        A = np.asfortranarray(rng.uniform(-1, 1, (lda, ca)).astype(dt))
        B = np.asfortranarray(rng.uniform(-1, 1, (ldb, cb)).astype(dt))
        A[ra:] = np.nan  # ld padding: a value read from it must not reach C
        B[rb:] = np.nan
        C = np.asfortranarray(rng.uniform(-1, 1, (ldc, n)).astype(dt))
        C[m:] = 777.0
        if beta == 0.0 and m > 1 and n > 1:
            C[1, 1] = np.nan
        out.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, A=A, B=B, C=C,
                        lda=lda, ldb=ldb, ldc=ldc, off=int(rng.integers(0, 2))))
    return out


def _expected(p):
    R = p["C"].copy(order="F")
    oracle.gemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["A"], p["lda"], p["B"],
                p["ldb"], p["beta"], R, p["ldc"])
    return R


def _upload(p, torch, dt):
    # off = 1 shifts the operands by one element: not 16-byte aligned
    pre = np.zeros(p["off"], dtype=dt)
    p["dA"] = to_device(np.concatenate([pre, np.ravel(p["A"], order="F")]), torch)[p["off"]:]
    p["dB"] = to_device(np.concatenate([pre, np.ravel(p["B"], order="F")]), torch)[p["off"]:]
    p["dC"] = to_device(p["C"], torch)


def _spec(p):
    return (p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["dA"], p["lda"], p["dB"],
            p["ldb"], p["beta"], p["dC"], p["ldc"])


def _check(p, R, G, dt):
    m, n, k = p["m"], p["n"], p["k"]
    what = (p["ta"], p["tb"], m, n, k, p["alpha"], p["beta"], p["off"])
    assert np.all(G[m:] == 777.0), ("ld padding written", what)
    if m == 0 or n == 0:
        return
    nan_r = np.isnan(R[:m])
    assert np.array_equal(nan_r, np.isnan(G[:m])), ("NaN pattern", what)
    bound = abs_bound(p["A"], p["B"], np.nan_to_num(p["C"]), p["ta"], p["tb"], m, n, k,
                      p["alpha"], p["beta"])
    g, r = np.where(nan_r, 0, G[:m]), np.where(nan_r, 0, R[:m])
    assert max_comp(g, r, bound) <= _tol(dt, k), what


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_batch_matches_oracle_and_replays(cuda, dt):
    import torch
    rng = np.random.default_rng(101 if dt == np.float64 else 202)
    probs = _problems(rng, 160, 1, 300, dt)
    for p in probs:
        _upload(p, torch, dt)
    plan = gx.BatchPlan("d" if dt == np.float64 else "s", [_spec(p) for p in probs])
    want = sum(2.0 * p["m"] * p["n"] * p["k"] for p in probs if p["alpha"] != 0 and p["k"] > 0)
    assert plan.flops == pytest.approx(want)
    exp = [_expected(p) for p in probs]
    initial = [p["dC"].clone() for p in probs]
    stream = torch.cuda.Stream()
    for replay in range(3):
        # eager work between replays: restore every C, then replay the graph
        for p, c0 in zip(probs, initial):
            p["dC"].copy_(c0)
        torch.cuda.synchronize()
        if replay == 1:
            with torch.cuda.stream(stream):
                plan.run(stream)
            stream.synchronize()
        else:
            plan.run()
            torch.cuda.synchronize()
        for p, R in zip(probs, exp):
            _check(p, R, from_device(p["dC"], p["C"].shape), dt)
    plan.destroy()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_every_batch_config(cuda, dt):
    """Each compiled batch tile configuration on 60 mixed problems."""
    import torch
    d = "d" if dt == np.float64 else "s"
    try:
        for cfg in range(gx.BATCH_CONFIGS):
            gx.set_batch_config(d, cfg)
            rng = np.random.default_rng(300 + cfg)
            probs = _problems(rng, 60, 1, 200, dt)
            for p in probs:
                _upload(p, torch, dt)
            plan = gx.BatchPlan(d, [_spec(p) for p in probs])
            plan.run()
            torch.cuda.synchronize()
            for p in probs:
                _check(p, _expected(p), from_device(p["dC"], p["C"].shape), dt)
            plan.destroy()
    finally:
        gx.set_batch_config(d, -1)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_batch_agrees_with_single_calls_at_full_size(cuda, dt):
    """4096 problems with sizes log-uniform in [16, 1024] and mixed
    transposes, as in the batch benchmark: every problem against the same
    problem through the single-call device entry point (same tolerance), and
    32 of them in full against the oracle."""
    import torch
    d = "d" if dt == np.float64 else "s"
    tdt = torch.float64 if dt == np.float64 else torch.float32
    rng = np.random.default_rng(7)
    count = 4096
    specs, singles, picks = [], [], set(rng.choice(count, 32, replace=False).tolist())
    host = {}
    for i in range(count):
        m, n, k = (int(round(np.exp(rng.uniform(np.log(16), np.log(1024))))) for _ in range(3))
        ta, tb = "NT"[rng.integers(2)], "NT"[rng.integers(2)]
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = torch.empty(ra * ca, dtype=tdt, device="cuda")
        B = torch.empty(rb * cb, dtype=tdt, device="cuda")
        C = torch.empty(m * n, dtype=tdt, device="cuda")
        gx.fill(A, ra, ca, ra, 5, 3 * i)
        gx.fill(B, rb, cb, rb, 5, 3 * i + 1)
        gx.fill(C, m, n, m, 5, 3 * i + 2)
        C2 = C.clone()
        specs.append((ta, tb, m, n, k, 1.5, A, ra, B, rb, -0.5, C, m))
        singles.append((ta, tb, m, n, k, 1.5, A, ra, B, rb, -0.5, C2, m))
        if i in picks:
            host[i] = (A.cpu().numpy().reshape((ra, ca), order="F").copy(order="F"),
                       B.cpu().numpy().reshape((rb, cb), order="F").copy(order="F"),
                       C.cpu().numpy().reshape((m, n), order="F").copy(order="F"))
    plan = gx.BatchPlan(d, specs)
    plan.run()
    dev = gx.dgemm_dev if dt == np.float64 else gx.sgemm_dev
    for s in singles:
        dev(*s)
    torch.cuda.synchronize()
    for i, (s, t) in enumerate(zip(specs, singles)):
        ta, tb, m, n, k = s[:5]
        G = s[11].cpu().numpy().reshape((m, n), order="F")
        S = t[11].cpu().numpy().reshape((m, n), order="F")
        if i in picks:
            A, B, C0 = host[i]
            R = C0.copy(order="F")
            oracle.gemm(ta, tb, m, n, k, 1.5, A, A.shape[0], B, B.shape[0], -0.5, R, m)
            bound = abs_bound(A, B, C0, ta, tb, m, n, k, 1.5, -0.5)
            assert max_comp(G, R, bound) <= _tol(dt, k), (i, ta, tb, m, n, k)
            assert max_comp(S, R, bound) <= _tol(dt, k), (i, ta, tb, m, n, k)
        # |G - S| within twice the single-call bound; inputs are U[-1, 1), so
        # |alpha| |op(A)||op(B)| + |beta C0| <= 1.5 k + 0.5 entrywise
        assert np.max(np.abs(G - S)) <= 2 * _tol(dt, k) * (1.5 * k + 0.5), (i, ta, tb, m, n, k)
    plan.destroy()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_batch_empty_and_dropped(cuda, dt):
    """An empty batch and a batch whose problems are all no-ops (m == 0, or
    alpha == 0 with beta == 1) build and run without touching C."""
    import torch
    d = "d" if dt == np.float64 else "s"
    plan = gx.BatchPlan(d, [])
    plan.run()
    plan.destroy()
    C = to_device(np.full((4, 4), 3.0, dtype=dt, order="F"), torch)
    A = to_device(np.ones((4, 4), dtype=dt, order="F"), torch)
    plan = gx.BatchPlan(d, [("N", "N", 0, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4),
                            ("N", "N", 4, 4, 4, 0.0, A, 4, A, 4, 1.0, C, 4),
                            ("T", "N", 4, 4, 0, 1.0, A, 4, A, 4, 1.0, C, 4)])
    assert plan.flops == 0.0
    plan.run()
    torch.cuda.synchronize()
    assert torch.all(C == 3.0)
    plan.destroy()
