"""Parity of the native gfx950 DGEMM/SGEMM against the oracle and the
reference's golden vectors, through the C ABI (dgemm_/sgemm_ and the
device-pointer extension).

Tolerances (the reference's own error against an exact product is ~5e-16
relative Frobenius for FP64 and ~3e-7 for FP32, so these leave room for a
different summation order, not for a wrong term):
  componentwise |C - C_exact| <= 64 * k * u * (|alpha| |op(A)||op(B)| + |beta C0|)
  relative Frobenius          <= 1e-14 (FP64), 1e-5 (FP32)
with u = 2^-53 (FP64) or 2^-24 (FP32).
"""
import ctypes

import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_util import (U32, U64, abs_bound, case_inputs, from_device, load_fixtures,
                       max_comp, rel_fro, stored_shape, to_device)

pytestmark = pytest.mark.gpu
FIX = load_fixtures()
TRANS = [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")]


def _tol(dtype, k):
    return 64 * max(k, 1) * (U64 if dtype in ("f64", np.float64) else U32)


def _fro_tol(dtype):
    return 1e-14 if dtype in ("f64", np.float64) else 1e-5


def _gemm_host(dt, *args):
    (gx.dgemm if dt == np.float64 else gx.sgemm)(*args)


def _gemm_dev(dt, *args, stream=None):
    (gx.dgemm_dev if dt == np.float64 else gx.sgemm_dev)(*args, stream=stream)


@pytest.mark.parametrize("case", [c for c in FIX if c["group"] == "small"],
                         ids=lambda c: f"{c['dtype']}-{c['transa']}{c['transb']}-"
                                       f"{c['m']}x{c['n']}x{c['k']}-a{c['alpha']}b{c['beta']}")
def test_small_fixtures_host_pointers(cuda, case):
    """Fortran ABI with host memory: staged to the GPU and back."""
    dt = np.float64 if case["dtype"] == "f64" else np.float32
    A, B, C = case_inputs(case, oracle)
    C0 = C.copy(order="F")
    m, n, k = case["m"], case["n"], case["k"]
    _gemm_host(dt, case["transa"], case["transb"], m, n, k, case["alpha"], A, case["lda"], B,
               case["ldb"], case["beta"], C, case["ldc"])
    ref = case["values"].reshape((m, n), order="F")
    assert np.all(np.isnan(C[m:, :n])), "ld padding of C was written"
    if case["alpha"] == 0.0 or k == 0:
        np.testing.assert_array_equal(C[:m, :n].astype(np.float64), ref)
        return
    R = C0.copy(order="F")
    oracle.gemm(case["transa"], case["transb"], m, n, k, case["alpha"], A, case["lda"], B,
                case["ldb"], case["beta"], R, case["ldc"])
    bound = abs_bound(A, B, C0, case["transa"], case["transb"], m, n, k, case["alpha"],
                      case["beta"])
    assert max_comp(C[:m, :n], R[:m, :n], bound) <= _tol(dt, k)
    assert max_comp(C[:m, :n], ref, bound) <= _tol(dt, k)


@pytest.mark.parametrize("case", [c for c in FIX if c["group"] == "medium"],
                         ids=lambda c: f"{c['dtype']}-{c['transa']}{c['transb']}-"
                                       f"{c['m']}x{c['n']}x{c['k']}")
def test_medium_fixtures_device_pointers(cuda, case):
    import torch
    dt = np.float64 if case["dtype"] == "f64" else np.float32
    A, B, C = case_inputs(case, oracle)
    C0 = C.copy(order="F")
    m, n, k = case["m"], case["n"], case["k"]
    dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
    _gemm_dev(dt, case["transa"], case["transb"], m, n, k, case["alpha"], dA, case["lda"], dB,
              case["ldb"], case["beta"], dC, case["ldc"])
    torch.cuda.synchronize()
    G = from_device(dC, C.shape)
    R = C0.copy(order="F")
    oracle.gemm(case["transa"], case["transb"], m, n, k, case["alpha"], A, case["lda"], B,
                case["ldb"], case["beta"], R, case["ldc"])
    bound = abs_bound(A, B, C0, case["transa"], case["transb"], m, n, k, case["alpha"],
                      case["beta"])
    assert max_comp(G[:m, :n], R[:m, :n], bound) <= _tol(dt, k)
    assert rel_fro(G[:m, :n], R[:m, :n]) <= _fro_tol(dt)
    assert max_comp(G[case["is"], case["js"]], case["values"],
                    bound[case["is"], case["js"]]) <= _tol(dt, k)
    assert np.all(np.isnan(G[m:, :n]))


def _device_case(case, torch):
    """Device buffers of a large fixture case, drawn on the GPU."""
    dt = torch.float64 if case["dtype"] == "f64" else torch.float32
    m, n, k = case["m"], case["n"], case["k"]
    ra, ca = stored_shape(case["transa"], m, k)
    rb, cb = stored_shape(case["transb"], k, n)
    s, d, p = case["seed"], case["dist"], case["param"]
    dA = torch.empty(case["lda"] * ca, dtype=dt, device="cuda")
    dB = torch.empty(case["ldb"] * cb, dtype=dt, device="cuda")
    dC = torch.empty(case["ldc"] * n, dtype=dt, device="cuda")
    gx.fill(dA, ra, ca, case["lda"], s, 0, d, p)
    gx.fill(dB, rb, cb, case["ldb"], s, 1, d, p)
    gx.fill(dC, m, n, case["ldc"], s, 2, d, p)
    return dA, dB, dC


@pytest.mark.parametrize("case", [c for c in FIX if c["group"] in ("nn1024", "sq8192",
                                                                   "ragged8192")],
                         ids=lambda c: f"{c['group']}-{c['dtype']}-{c['transa']}{c['transb']}")
def test_large_fixtures_sampled(cuda, case):
    import torch
    dt = np.float64 if case["dtype"] == "f64" else np.float32
    m, n, k, ldc = case["m"], case["n"], case["k"], case["ldc"]
    print(f"case {case['group']} {case['dtype']} {case['transa']}{case['transb']} {m}x{n}x{k}",
          flush=True)
    dA, dB, dC = _device_case(case, torch)
    _gemm_dev(dt, case["transa"], case["transb"], m, n, k, case["alpha"], dA, case["lda"], dB,
              case["ldb"], case["beta"], dC, ldc)
    torch.cuda.synchronize()
    flat = torch.from_numpy((case["is"] + case["js"].astype(np.int64) * ldc)).to("cuda")
    got = dC[flat].double().cpu().numpy()
    exp, ab = oracle.sampled(dt, case["transa"], case["transb"], m, n, k, case["alpha"],
                             case["beta"], case["seed"], case["is"], case["js"])
    bound = abs(case["alpha"]) * ab + abs(case["beta"]) + 1e-300
    assert np.max(np.abs(got - exp) / bound) <= _tol(dt, k)
    assert np.max(np.abs(got - case["values"]) / bound) <= _tol(dt, k)
    # rows m..ldc-1 keep the generator's NaN
    pad = dC.view(n, ldc)[:, m:]
    assert bool(torch.isnan(pad).all()) if ldc > m else True
    del dA, dB, dC
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_ragged_shapes_and_lds(cuda, dt, ta, tb):
    """Odd sizes around every tile edge, odd leading dimensions (no 16-byte
    loads), unaligned operand offsets."""
    import torch
    rng = np.random.default_rng(17)
    for (m, n, k, pad) in [(1, 1, 1, 0), (15, 17, 3, 1), (63, 65, 129, 3), (129, 127, 17, 1),
                           (200, 70, 300, 0), (257, 129, 65, 2), (33, 300, 1000, 5)]:
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        lda, ldb, ldc = ra + pad, rb + pad, m + pad
        A = np.asfortranarray(rng.uniform(-1, 1, (lda, ca)).astype(dt))
        B = np.asfortranarray(rng.uniform(-1, 1, (ldb, cb)).astype(dt))
        C = np.asfortranarray(rng.uniform(-1, 1, (ldc, n)).astype(dt))
        C[m:] = 777.0
        R = C.copy(order="F")
        oracle.gemm(ta, tb, m, n, k, -1.25, A, lda, B, ldb, 0.5, R, ldc)
        bound = abs_bound(A, B, C, ta, tb, m, n, k, -1.25, 0.5)
        # device pointers with a one-element offset: misaligned for 16-byte loads
        dA = to_device(np.concatenate([[0], np.ravel(A, order="F")]).astype(dt), torch)
        dB = to_device(np.concatenate([[0], np.ravel(B, order="F")]).astype(dt), torch)
        dC = to_device(C, torch)
        _gemm_dev(dt, ta, tb, m, n, k, -1.25, dA[1:], lda, dB[1:], ldb, 0.5, dC, ldc)
        torch.cuda.synchronize()
        G = from_device(dC, C.shape)
        assert max_comp(G[:m], R[:m], bound) <= _tol(dt, k), (m, n, k, pad)
        assert np.all(G[m:] == 777.0), "ld padding written"
        # host pointers for the same problem
        H = C.copy(order="F")
        _gemm_host(dt, ta, tb, m, n, k, -1.25, A, lda, B, ldb, 0.5, H, ldc)
        assert max_comp(H[:m], R[:m], bound) <= _tol(dt, k), (m, n, k, pad)
        assert np.all(H[m:] == 777.0)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_semantics(cuda, dt):
    """beta != 1 scales first (beta == 0 keeps NaN/Inf of C), alpha == 0 reads
    neither A nor B, k == 0 and m == 0 behave as in the reference, any
    transpose character other than N/n transposes."""
    rng = np.random.default_rng(19)
    m, n, k = 37, 29, 41
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dt))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(dt))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
    C[3, 4] = np.nan
    C[5, 6] = np.inf
    X = C.copy(order="F")
    _gemm_host(dt, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, X, m)
    assert np.isnan(X[3, 4]) and np.isnan(X[5, 6]) and np.isfinite(X[0, 0])
    An = A.copy(order="F")
    An[0, 0] = np.nan
    for beta in (0.0, 1.0, 2.0):
        X = C.copy(order="F")
        _gemm_host(dt, "N", "N", m, n, k, 0.0, An, m, B, k, beta, X, m)
        np.testing.assert_array_equal(X, C * dt(beta) if beta != 1.0 else C)
        X = C.copy(order="F")
        _gemm_host(dt, "N", "N", m, n, 0, 1.0, An, m, B, k, beta, X, m)
        np.testing.assert_array_equal(X, C * dt(beta) if beta != 1.0 else C)
    X = C.copy(order="F")
    _gemm_host(dt, "N", "N", 0, n, k, 1.0, A, 1, B, k, 0.0, X, 1)
    np.testing.assert_array_equal(X, C)
    base = np.zeros((m, n), dtype=dt, order="F")
    _gemm_host(dt, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, base, m)
    At = np.asfortranarray(A.T)
    for ch in ("T", "t", "C", "c", "x"):
        X = np.zeros((m, n), dtype=dt, order="F")
        _gemm_host(dt, ch, "n", m, n, k, 1.0, At, k, B, k, 0.0, X, m)
        np.testing.assert_array_equal(X, base)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_known_answer(cuda, dt):
    """C[i,j] = 12*(5+i%2)*(20+j%8) (dgemm_kernel_4x4_haswell.c:472-492)."""
    n = 12
    idx = np.arange(n * n)
    A = np.asfortranarray((5.0 + idx % 2).reshape((n, n), order="F").astype(dt))
    B = np.asfortranarray((20.0 + (idx // n) % 8).reshape((n, n), order="F").astype(dt))
    C = np.zeros((n, n), dtype=dt, order="F")
    _gemm_host(dt, "N", "N", n, n, n, 1.0, A, n, B, n, 1.0, C, n)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    np.testing.assert_array_equal(C, (12.0 * (5 + i % 2) * (20 + j % 8)).astype(dt))


def test_device_generator_matches_oracle(cuda):
    import torch
    for dt, ndt in ((torch.float64, np.float64), (torch.float32, np.float32)):
        t = torch.empty(13 * 7, dtype=dt, device="cuda")
        gx.fill(t, 11, 7, 13, 42, 1)
        torch.cuda.synchronize()
        host = oracle.fill(11, 7, 13, 42, 1, ndt)
        got = t.cpu().numpy().reshape((13, 7), order="F")
        np.testing.assert_array_equal(got[:11], host[:11])
        assert np.all(np.isnan(got[11:]))


def test_errors_leave_c_untouched(cuda):
    C = np.ones((4, 4), order="F")
    A = np.ones((4, 4), order="F")
    gx.dgemm("N", "N", 4, 4, 4, 1.0, A, 3, A, 4, 0.0, C, 4, check=False)  # lda < m
    code, msg = gx.last_error("d")
    assert code == gx.ERR_ARG and "lda" in msg
    np.testing.assert_array_equal(C, 1.0)
    gx.dgemm("N", "N", 4, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4)
    assert gx.last_error("d")[0] == gx.OK
    np.testing.assert_array_equal(C, 4.0)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_every_tile_config(cuda, dt):
    """Each compiled tile configuration, forced through the C ABI, on shapes
    that hit interior tiles (branch-free path), edge tiles, a K remainder and
    all transposes."""
    import torch
    d = "d" if dt == np.float64 else "s"
    rng = np.random.default_rng(23)
    try:
        for cfg in range(gx.TILE_CONFIGS):
            gx.set_tile_config(d, cfg)
            for (m, n, k) in [(256, 256, 64), (300, 270, 77), (512, 129, 48)]:
                for ta, tb in TRANS:
                    ra, ca = stored_shape(ta, m, k)
                    rb, cb = stored_shape(tb, k, n)
                    A = np.asfortranarray(rng.uniform(-1, 1, (ra, ca)).astype(dt))
                    B = np.asfortranarray(rng.uniform(-1, 1, (rb, cb)).astype(dt))
                    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
                    R = C.copy(order="F")
                    oracle.gemm(ta, tb, m, n, k, 0.75, A, ra, B, rb, -1.0, R, m)
                    dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
                    _gemm_dev(dt, ta, tb, m, n, k, 0.75, dA, ra, dB, rb, -1.0, dC, m)
                    torch.cuda.synchronize()
                    G = from_device(dC, C.shape)
                    bound = abs_bound(A, B, C, ta, tb, m, n, k, 0.75, -1.0)
                    assert max_comp(G, R, bound) <= _tol(dt, k), (cfg, m, n, k, ta, tb)
    finally:
        gx.set_tile_config(d, -1)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_edge_tiles_with_clamped_loads(cuda, dt, ta, tb):
    """Every tile runs the branch-free loop, an edge tile's loads past m or n
    clamped back into the operand: with 16-byte loads (aligned base, ld a
    multiple of four elements; odd sizes whose last 16-byte chunk straddles
    the end of a column) and with one load per element (odd ld); m and n one
    past, one short of and far from every tile edge, for every tile
    configuration.  The ld padding of A and B holds NaN, so a value read from
    it that reached a stored element would show; rows m..ldc-1 of C stay
    untouched."""
    import torch
    d = "d" if dt == np.float64 else "s"
    rng = np.random.default_rng(29)
    try:
        for cfg in range(gx.TILE_CONFIGS):
            gx.set_tile_config(d, cfg)
            for (m, n, k, wide) in [(131, 70, 77, 1), (257, 383, 40, 1), (1, 129, 33, 1),
                                    (129, 1, 16, 1), (300, 260, 19, 1), (63, 127, 35, 1),
                                    (131, 70, 77, 0), (257, 383, 40, 0), (128, 129, 33, 0),
                                    (64, 64, 16, 0)]:
                ra, ca = stored_shape(ta, m, k)
                rb, cb = stored_shape(tb, k, n)
                if wide:  # 16-byte loadable
                    lda, ldb = (ra + 3) // 4 * 4 + 4, (rb + 3) // 4 * 4
                else:  # odd leading dimensions
                    lda, ldb = ra + 1 + (ra & 1), rb + 1 + (rb & 1)
                ldc = m + 3
                A = np.full((lda, ca), np.nan, dtype=dt, order="F")
                B = np.full((ldb, cb), np.nan, dtype=dt, order="F")
                A[:ra] = rng.uniform(-1, 1, (ra, ca))
                B[:rb] = rng.uniform(-1, 1, (rb, cb))
                C = np.full((ldc, n), 777.0, dtype=dt, order="F")
                C[:m] = rng.uniform(-1, 1, (m, n))
                R = C.copy(order="F")
                oracle.gemm(ta, tb, m, n, k, 1.5, A, lda, B, ldb, -0.5, R, ldc)
                dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
                _gemm_dev(dt, ta, tb, m, n, k, 1.5, dA, lda, dB, ldb, -0.5, dC, ldc)
                torch.cuda.synchronize()
                G = from_device(dC, C.shape)
                bound = abs_bound(A[:ra], B[:rb], C[:m], ta, tb, m, n, k, 1.5, -0.5)
                assert max_comp(G[:m], R[:m], bound) <= _tol(dt, k), (cfg, m, n, k, lda, ldb)
                assert np.all(G[m:] == 777.0), (cfg, m, n, k, lda, ldb)
    finally:
        gx.set_tile_config(d, -1)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_host_pointer_staging_ring(cuda, dt, monkeypatch):
    """Host operands above the staging threshold go through the ring of
    pinned chunks (1 MB chunks here, so A, B and C each take many chunks and
    reuse the buffers of the previous operand); odd leading dimensions and
    transposes; rows m..ldc-1 of C untouched."""
    monkeypatch.setenv("GEMMX_STAGE_CHUNK_MB", "1")
    rng = np.random.default_rng(31)
    for (ta, tb, m, n, k, pad) in [("N", "N", 1500, 1300, 1100, 3), ("T", "T", 1030, 1700, 1200, 1)]:
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = np.asfortranarray(rng.uniform(-1, 1, (ra + pad, ca)).astype(dt))
        B = np.asfortranarray(rng.uniform(-1, 1, (rb + pad, cb)).astype(dt))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)).astype(dt))
        C[m:] = 321.0
        C0 = C.copy(order="F")
        R = C.copy(order="F")
        oracle.gemm(ta, tb, m, n, k, 0.5, A, ra + pad, B, rb + pad, 2.0, R, m + pad)
        _gemm_host(dt, ta, tb, m, n, k, 0.5, A, ra + pad, B, rb + pad, 2.0, C, m + pad)
        bound = abs_bound(A, B, C0, ta, tb, m, n, k, 0.5, 2.0)
        assert max_comp(C[:m], R[:m], bound) <= _tol(dt, k), (ta, tb)
        assert np.all(C[m:] == 321.0)


@pytest.mark.parametrize("chunk_mb", ["32", "1"])
def test_host_pointer_column_panels(cuda, chunk_mb, monkeypatch):
    """Native calls moving 64 MB or more of host B and C run in column panels
    of C (multiples of 256 columns, about eight): the panel's op(B) columns
    and C columns are staged while the previous panel computes, and each C
    panel is copied back behind its own GEMM.  Columns at every panel edge
    (and the ragged last panel) against the oracle, for NN, NT (op(B)
    columns are rows of B) and TN in single precision; rows m..ldc-1 of C
    untouched.  With 1 MB chunks the staging ring's buffers alternate between
    uploads and downloads of different panels."""
    monkeypatch.setenv("GEMMX_STAGE_CHUNK_MB", chunk_mb)
    rng = np.random.default_rng(41)
    for (dt, ta, tb, m, n, k, pad) in [(np.float64, "N", "N", 300, 2300, 4000, 3),
                                       (np.float64, "N", "T", 400, 2100, 4200, 1),
                                       (np.float32, "T", "N", 700, 2600, 7000, 2)]:
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = np.asfortranarray(rng.uniform(-1, 1, (ra + pad, ca)).astype(dt))
        B = np.asfortranarray(rng.uniform(-1, 1, (rb + pad, cb)).astype(dt))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)).astype(dt))
        C[m:] = 321.0
        C0 = C.copy(order="F")
        _gemm_host(dt, ta, tb, m, n, k, 0.5, A, ra + pad, B, rb + pad, -1.5, C, m + pad)
        assert np.all(C[m:] == 321.0)
        np_cols = -(-((n + 7) // 8) // 256) * 256
        cols = sorted({c for j in range(0, n, np_cols) for c in (j - 1, j, j + 1) if 0 <= c < n}
                      | {n - 1} | set(rng.integers(0, n, 6).tolist()))
        Bs = np.asfortranarray(B[:k, cols] if tb == "N" else B[cols, :k])
        Cs = np.asfortranarray(C0[:m, cols])
        R = Cs.copy(order="F")
        oracle.gemm(ta, tb, m, len(cols), k, 0.5, A, ra + pad, Bs, Bs.shape[0], -1.5, R, m)
        bound = abs_bound(A, Bs, Cs, ta, tb, m, len(cols), k, 0.5, -1.5)
        assert max_comp(C[:m, cols], R, bound) <= _tol(dt, k), (dt, ta, tb)


def test_host_pointer_column_panels_mixed_operands(cuda):
    """Column panels with some operands already on the device: A on the device
    with B and C host arrays (NT), and A host, B on the device, C host (NN,
    only C moves enough to take panels).  Sampled columns at every panel edge
    against the oracle; the device operands are left unchanged."""
    import torch
    rng = np.random.default_rng(43)
    for (ta, tb, m, n, k, dev_a) in [("N", "T", 300, 2300, 4000, True),
                                     ("N", "N", 2000, 4200, 300, False)]:
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = np.asfortranarray(rng.uniform(-1, 1, (ra, ca)))
        B = np.asfortranarray(rng.uniform(-1, 1, (rb, cb)))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + 1, n)))
        C[m:] = 321.0
        C0 = C.copy(order="F")
        a_arg = to_device(A, torch) if dev_a else A
        b_arg = B if dev_a else to_device(B, torch)
        gx.dgemm(ta, tb, m, n, k, 1.25, a_arg, ra, b_arg, rb, 0.5, C, m + 1)
        torch.cuda.synchronize()
        dev_op, host_op = (a_arg, A) if dev_a else (b_arg, B)
        np.testing.assert_array_equal(from_device(dev_op, host_op.shape), host_op)
        assert np.all(C[m:] == 321.0)
        np_cols = -(-((n + 7) // 8) // 256) * 256
        cols = sorted({c for j in range(0, n, np_cols) for c in (j - 1, j) if 0 <= c < n} | {n - 1})
        Bs = np.asfortranarray(B[:k, cols] if tb == "N" else B[cols, :k])
        Cs = np.asfortranarray(C0[:m, cols])
        R = Cs.copy(order="F")
        oracle.gemm(ta, tb, m, len(cols), k, 1.25, A, ra, Bs, Bs.shape[0], 0.5, R, m)
        bound = abs_bound(A, Bs, Cs, ta, tb, m, len(cols), k, 1.25, 0.5)
        assert max_comp(C[:m, cols], R, bound) <= _tol(np.float64, k), (ta, tb)


def test_concurrent_calls_from_host_threads(cuda):
    """dgemm_/sgemm_ from four host threads at once (ctypes releases the GIL):
    column-panel host calls, small host calls and device-pointer calls
    interleave on one device; every result matches the oracle (sampled
    columns for the large ones) and no call reports an error."""
    import threading
    import torch
    rng = np.random.default_rng(47)
    jobs = []
    for t in range(4):
        for (dt, m, n, k) in [(np.float64, 300, 2300, 4000), (np.float32, 130, 70, 90),
                              (np.float64, 97, 65, 130)]:
            A = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dt))
            B = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(dt))
            C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
            jobs.append((t, dt, m, n, k, A, B, C, C.copy(order="F"), (m * n) % 2 == 0))
    errors = []

    def worker(tid):
        try:
            for (t, dt, m, n, k, A, B, C, C0, on_dev) in jobs:
                if t != tid:
                    continue
                if on_dev and m * n < 10**5:
                    dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C0, torch)
                    _gemm_host(dt, "N", "N", m, n, k, 0.75, dA, m, dB, k, 0.5, dC, m)
                    torch.cuda.synchronize()
                    C[:] = from_device(dC, C.shape)
                else:
                    _gemm_host(dt, "N", "N", m, n, k, 0.75, A, m, B, k, 0.5, C, m)
                code, msg = gx.last_error("d" if dt == np.float64 else "s")
                if code != gx.OK:
                    errors.append(msg)
        except Exception as exc:  # reported below, in the main thread
            errors.append(repr(exc))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads) and not errors, errors
    for (t, dt, m, n, k, A, B, C, C0, on_dev) in jobs:
        cols = sorted(set(rng.integers(0, n, 5).tolist()) | {0, n - 1})
        Bs, Cs = np.asfortranarray(B[:, cols]), np.asfortranarray(C0[:, cols])
        R = Cs.copy(order="F")
        oracle.gemm("N", "N", m, len(cols), k, 0.75, A, m, Bs, k, 0.5, R, m)
        bound = abs_bound(A, Bs, Cs, "N", "N", m, len(cols), k, 0.75, 0.5)
        assert max_comp(C[:, cols], R, bound) <= _tol(dt, k), (t, dt, m, n, k)
