"""Split-K of the native kernels on the GPU, through the C ABI.

A product with few C tiles divides K over several grid rows writing partial
tiles, which a second kernel sums in range order and combines with alpha and
beta*C (in place of the reference's N-split over all threads of every call,
src/gemm_driver.c:128-151).  Checked against the oracle with the native
tolerance (componentwise 64 k u (|alpha||A||B| + |beta C|), relative Frobenius
1e-14 / 1e-5), for every transpose pair, tile configuration with split-K
kernels, forced split factors (K ranges that do not divide K, a last range
shorter than one slice), ragged edges, padded leading dimensions (rows past m
stay untouched), the semantics cases (beta = 0 keeps NaN of C, alpha = 0),
host and device operands, a non-default stream, repeatability, and at full
size on 512 x 512 x 65536 and 1024^3 with the split the library picks.
"""
import numpy as np
import pytest

import gemm_avx2_amd as gx
from gemm_avx2_amd import autotune as at
import oracle
from gemm_util import (U32, U64, abs_bound, from_device, max_comp, rel_fro, stored_shape,
                       to_device)

pytestmark = pytest.mark.gpu
TRANS = [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")]


def _u(dt):
    return U64 if dt == np.float64 else U32


def _fro(dt):
    return 1e-14 if dt == np.float64 else 1e-5


def _inputs(rng, dt, ta, tb, m, n, k, pad):
    ra, ca = stored_shape(ta, m, k)
    rb, cb = stored_shape(tb, k, n)
    A = np.asfortranarray(rng.uniform(-1, 1, (ra + pad, ca)).astype(dt))
    B = np.asfortranarray(rng.uniform(-1, 1, (rb + pad, cb)).astype(dt))
    C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)).astype(dt))
    C[m:] = np.nan
    return A, B, C, ra + pad, rb + pad, m + pad


def _dev(dt, torch, *args, stream=None):
    (gx.dgemm_dev if dt == np.float64 else gx.sgemm_dev)(*args, stream=stream)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_forced_splits_match_the_oracle(cuda, dt, ta, tb):
    import torch
    d = "d" if dt == np.float64 else "s"
    rng = np.random.default_rng(7)
    try:
        for (m, n, k, pad) in [(130, 97, 3001, 3), (64, 64, 4096, 0), (257, 31, 1000, 1)]:
            A, B, C, lda, ldb, ldc = _inputs(rng, dt, ta, tb, m, n, k, pad)
            R = C.astype(np.float64, order="F") if dt == np.float32 else C.copy(order="F")
            oracle.gemm(ta, tb, m, n, k, 1.5, A.astype(R.dtype, order="F"), lda,
                        B.astype(R.dtype, order="F"), ldb, -0.5, R, ldc)
            bound = abs_bound(A, B, C, ta, tb, m, n, k, 1.5, -0.5)
            dA, dB = to_device(A, torch), to_device(B, torch)
            for cfg in at.SPLIT_CONFIGS:
                gx.set_tile_config(d, cfg)
                for split in (2, 3, 7, 16):
                    print(f"case {d} {ta}{tb} {m}x{n}x{k} cfg {cfg} split {split}", flush=True)
                    gx.set_split_config(d, split)
                    assert gx.tune_lookup_split(d, ta, tb, m, n, k)[0] > 1
                    dC = to_device(C, torch)
                    _dev(dt, torch, ta, tb, m, n, k, 1.5, dA, lda, dB, ldb, -0.5, dC, ldc)
                    torch.cuda.synchronize()
                    G = from_device(dC, C.shape)
                    assert max_comp(G[:m], R[:m], bound) <= 64 * k * _u(dt), (cfg, split)
                    assert rel_fro(G[:m], R[:m]) <= _fro(dt), (cfg, split)
                    assert np.all(np.isnan(G[m:])), "rows past m written"
    finally:
        gx.set_tile_config(d, -1)
        gx.set_split_config(d, -1)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_split_semantics_and_repeatability(cuda, dt):
    import torch
    d = "d" if dt == np.float64 else "s"
    rng = np.random.default_rng(8)
    m, n, k = 70, 50, 2000
    A, B, C, lda, ldb, ldc = _inputs(rng, dt, "N", "N", m, n, k, 0)
    dA, dB = to_device(A, torch), to_device(B, torch)
    try:
        gx.set_tile_config(d, 1)
        gx.set_split_config(d, 5)
        # beta = 0 still multiplies C: NaN in C stays NaN, other entries are alpha*AB
        Cn = C.copy(order="F")
        Cn[3, 4] = np.nan
        dC = to_device(Cn, torch)
        _dev(dt, torch, "N", "N", m, n, k, 1.0, dA, lda, dB, ldb, 0.0, dC, ldc)
        torch.cuda.synchronize()
        G = from_device(dC, C.shape)
        assert np.isnan(G[3, 4]) and np.isfinite(np.delete(G.ravel(order="F"), 3 + 4 * m)).all()
        # alpha = 0: C = beta*C, no product
        dC = to_device(C, torch)
        _dev(dt, torch, "N", "N", m, n, k, 0.0, dA, lda, dB, ldb, 2.0, dC, ldc)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_device(dC, C.shape), C * dt(2.0))
        # the same call twice gives the same bits (partials summed in range order)
        outs = []
        for _ in range(2):
            dC = to_device(C, torch)
            _dev(dt, torch, "N", "N", m, n, k, 1.25, dA, lda, dB, ldb, 0.5, dC, ldc)
            torch.cuda.synchronize()
            outs.append(from_device(dC, C.shape))
        np.testing.assert_array_equal(outs[0], outs[1])
        # on a non-default stream, queued behind other work of that stream
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            dC = to_device(C, torch)
            _dev(dt, torch, "N", "N", m, n, k, 1.25, dA, lda, dB, ldb, 0.5, dC, ldc,
                 stream=s)
        s.synchronize()
        np.testing.assert_array_equal(from_device(dC, C.shape), outs[0])
        # host operands through the Fortran-ABI entry (partials in the library's scratch)
        Ch = C.copy(order="F")
        (gx.dgemm if dt == np.float64 else gx.sgemm)("N", "N", m, n, k, 1.25, A, lda, B, ldb,
                                                     0.5, Ch, ldc)
        np.testing.assert_array_equal(Ch, outs[0])
    finally:
        gx.set_tile_config(d, -1)
        gx.set_split_config(d, -1)


@pytest.mark.parametrize("ta,tb", TRANS)
@pytest.mark.parametrize("m,n,k,force", [(512, 512, 65536, -1), (1024, 1024, 1024, 4)])
def test_library_split_at_full_size(cuda, ta, tb, m, n, k, force):
    """FP64 split-K at full size, 2048 sampled entries against Dot2 values of
    the same inputs: 512 x 512 x 65536 with the split the library picks, and
    1024^3 (which the tuning table runs in one pass) with 4 ranges forced."""
    import torch
    gx.set_split_config("d", force)
    splits, _ = gx.tune_lookup_split("d", ta, tb, m, n, k)
    print(f"case dgemm {ta}{tb} {m}x{n}x{k} splits {splits}", flush=True)
    try:
        assert splits > 1
        _full_size_case(torch, ta, tb, m, n, k)
    finally:
        gx.set_split_config("d", -1)


def _full_size_case(torch, ta, tb, m, n, k):
    ra, ca = stored_shape(ta, m, k)
    rb, cb = stored_shape(tb, k, n)
    A = torch.empty(ra * ca, dtype=torch.float64, device="cuda")
    B = torch.empty(rb * cb, dtype=torch.float64, device="cuda")
    C = torch.empty(m * n, dtype=torch.float64, device="cuda")
    gx.fill(A, ra, ca, ra, 12, 0)
    gx.fill(B, rb, cb, rb, 12, 1)
    gx.fill(C, m, n, m, 12, 2)
    C0 = C.clone()
    gx.dgemm_dev(ta, tb, m, n, k, 1.5, A, ra, B, rb, -0.5, C, m)
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    is_, js = rng.integers(0, m, 2048), rng.integers(0, n, 2048)
    ui, ri = np.unique(is_, return_inverse=True)
    uj, cj = np.unique(js, return_inverse=True)
    tui, tuj = torch.from_numpy(ui).cuda(), torch.from_numpy(uj).cuda()
    rows = A.view(k, m)[:, tui].T if ta == "N" else A.view(m, k)[tui]
    cols = B.view(n, k)[tuj] if tb == "N" else B.view(k, n)[:, tuj].T
    flat = torch.from_numpy(is_.astype(np.int64) + js.astype(np.int64) * m).cuda()
    c0 = C0[flat].cpu().numpy()
    exact, ab = oracle.dot2_pairs(rows.contiguous().cpu().numpy(), cols.contiguous().cpu().numpy(),
                                  ri, cj, 1.5, -0.5 * c0)
    got = C[flat].cpu().numpy()
    bound = 1.5 * ab + 0.5 * np.abs(c0)
    assert np.max(np.abs(got - exact) / bound) <= 64 * k * U64
    assert np.linalg.norm(got - exact) / np.linalg.norm(exact) <= 1e-14
