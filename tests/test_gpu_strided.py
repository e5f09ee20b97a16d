"""The BLAS-style caller interfaces next to the Fortran symbols: the strided
batch on device pointers (dgemm_x_strided_batched_dev / sgemm_...) and the
ILP64 Fortran forms (dgemm_64_ / sgemm_64_), through the C ABI.

Every problem of a batch is checked against the oracle (the CPU restatement
of src/gemm_driver.c:118) with the native tolerances of test_gpu_native:
componentwise |C - C_oracle| <= 64 k u (|alpha||op(A)||op(B)| + |beta C0|).
The ILP64 form runs the same kernels as dgemm_, so it must agree bit for bit.
"""
import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_util import U32, U64, abs_bound, max_comp, stored_shape

pytestmark = pytest.mark.gpu
TRANS = [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")]


def _tol(dt, k):
    return 64 * max(k, 1) * (U64 if dt == np.float64 else U32)


def _batched(dt):
    return gx.dgemm_strided_batched_dev if dt == np.float64 else gx.sgemm_strided_batched_dev


def _run_batch(torch, dt, ta, tb, m, n, k, alpha, beta, batch, pad_ld, gap, share_a=False,
               seed=0):
    """Builds `batch` problems in three flat buffers (gap elements between
    consecutive problems, NaN in C's gaps and padding rows), runs the strided
    batch and returns (inputs, result) as numpy."""
    rng = np.random.default_rng(seed)
    ra, ca = stored_shape(ta, m, k)
    rb, cb = stored_shape(tb, k, n)
    lda, ldb, ldc = ra + pad_ld, rb + pad_ld, m + pad_ld
    sa = 0 if share_a else lda * ca + gap
    sb, sc = ldb * cb + gap, ldc * n + gap
    na = (sa * (batch - 1) if not share_a else 0) + lda * ca
    A = rng.uniform(-1, 1, na).astype(dt)
    B = rng.uniform(-1, 1, sb * (batch - 1) + ldb * cb).astype(dt)
    C = np.full(sc * (batch - 1) + ldc * n, np.nan, dtype=dt)
    for i in range(batch):
        blk = C[i * sc:i * sc + ldc * n].reshape((ldc, n), order="F")
        blk[:m] = rng.uniform(-1, 1, (m, n))
    dA, dB, dC = (torch.from_numpy(x.copy()).cuda() for x in (A, B, C))
    _batched(dt)(ta, tb, m, n, k, alpha, dA, lda, sa, dB, ldb, sb, beta, dC, ldc, sc, batch)
    torch.cuda.synchronize()
    return (A, B, C, lda, ldb, ldc, sa, sb, sc), dC.cpu().numpy()


def _check_batch(dt, ta, tb, m, n, k, alpha, beta, batch, inputs, G, problems=None):
    A, B, C, lda, ldb, ldc, sa, sb, sc = inputs
    ra, ca = stored_shape(ta, m, k)
    rb, cb = stored_shape(tb, k, n)
    for i in (range(batch) if problems is None else problems):
        Ai = np.asfortranarray(A[i * sa:i * sa + lda * ca].reshape((lda, ca), order="F"))
        Bi = np.asfortranarray(B[i * sb:i * sb + ldb * cb].reshape((ldb, cb), order="F"))
        C0 = np.asfortranarray(C[i * sc:i * sc + ldc * n].reshape((ldc, n), order="F"))
        R = C0.copy(order="F")
        oracle.gemm(ta, tb, m, n, k, alpha, Ai, lda, Bi, ldb, beta, R, ldc)
        Gi = G[i * sc:i * sc + ldc * n].reshape((ldc, n), order="F")
        assert np.all(np.isnan(Gi[m:])), f"problem {i}: ld padding of C written"
        bound = abs_bound(Ai, Bi, C0, ta, tb, m, n, k, alpha, beta)
        assert max_comp(Gi[:m], R[:m], bound) <= _tol(dt, k), f"problem {i}"
        if i + 1 < batch and sc > ldc * n:
            assert np.all(np.isnan(G[i * sc + ldc * n:(i + 1) * sc])), f"gap after {i} written"


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_strided_batch_matches_the_oracle_per_problem(cuda, dt, ta, tb):
    import torch
    for (m, n, k, batch, pad, gap) in ((33, 29, 31, 5, 3, 7), (128, 128, 64, 3, 0, 0),
                                       (200, 130, 70, 4, 8, 16)):
        inputs, G = _run_batch(torch, dt, ta, tb, m, n, k, 1.5, -0.5, batch, pad, gap,
                               seed=m + n)
        _check_batch(dt, ta, tb, m, n, k, 1.5, -0.5, batch, inputs, G)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_strided_batch_shared_operand_and_single_problem(cuda, dt):
    """stride_a = 0 shares one A across the batch; batch 1 is dgemm_x_dev,
    bit for bit (same kernels, same configuration)."""
    import torch
    inputs, G = _run_batch(torch, dt, "N", "T", 70, 50, 90, 1.0, 0.0, 6, 2, 0, share_a=True)
    _check_batch(dt, "N", "T", 70, 50, 90, 1.0, 0.0, 6, inputs, G)
    inputs, G = _run_batch(torch, dt, "T", "N", 96, 64, 80, 2.0, 1.0, 1, 0, 0, seed=3)
    A, B, C, lda, ldb, ldc = inputs[:6]
    dA, dB, dC = (torch.from_numpy(x.copy()).cuda() for x in (A, B, C))
    (gx.dgemm_dev if dt == np.float64 else gx.sgemm_dev)("T", "N", 96, 64, 80, 2.0, dA, lda,
                                                         dB, ldb, 1.0, dC, ldc)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dC.cpu().numpy(), G)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N")])
def test_strided_batch_that_fills_the_gpu_takes_its_table_entry(cuda, dt, ta, tb):
    """64 problems of 256^3 hold 256 tiles of 128x128, so the batch takes the
    shipped table's `strided` entry of its bucket (when it has one) instead
    of the single problem's choice; its result, and that of every tile
    configuration forced on the same batch, matches the oracle on sampled
    problems."""
    import os
    import torch
    from gemm_avx2_amd import autotune as at
    d = "d" if dt == np.float64 else "s"
    m = n = k = 256
    batch = 64
    entries = at.read_table(os.path.join(gx.LIB_DIR, "gemm_mi355x_tune.txt"))[0]
    key = ("strided", "f64" if d == "d" else "f32", ta + tb, 8, 8, 8)
    want = entries.get(key, gx.tune_lookup(d, ta, tb, m, n, k))
    assert gx.tune_lookup_strided(d, ta, tb, m, n, k, batch) == want
    inputs, G = _run_batch(torch, dt, ta, tb, m, n, k, 1.5, -0.5, batch, 2, 5, seed=17)
    _check_batch(dt, ta, tb, m, n, k, 1.5, -0.5, batch, inputs, G, problems=[0, 1, 31, 63])
    for cfg in range(gx.TILE_CONFIGS):
        gx.set_tile_config(d, cfg)
        try:
            inputs, G = _run_batch(torch, dt, ta, tb, m, n, k, 1.5, -0.5, batch, 2, 5, seed=17)
        finally:
            gx.set_tile_config(d, -1)
        _check_batch(dt, ta, tb, m, n, k, 1.5, -0.5, batch, inputs, G, problems=[0, 63])


def test_strided_batch_beyond_one_grid_row_launch(cuda):
    """More problems than one launch's grid rows (65535): every problem, the
    last of the first launch and the first of the second included, is
    computed once."""
    import torch
    m = n = k = 2
    batch = 70000
    inputs, G = _run_batch(torch, np.float32, "N", "N", m, n, k, 1.0, 0.0, batch, 0, 0, seed=9)
    A, B, C, lda, ldb, ldc, sa, sb, sc = inputs
    a = A.reshape(batch, k, m).transpose(0, 2, 1).astype(np.float64)  # column-major blocks
    b = B.reshape(batch, n, k).transpose(0, 2, 1).astype(np.float64)
    want = np.matmul(a, b)  # [batch, m, n]
    got = G.reshape(batch, n, m).transpose(0, 2, 1).astype(np.float64)
    assert np.max(np.abs(got - want)) <= 64 * k * U32 * 4
    _check_batch(np.float32, "N", "N", m, n, k, 1.0, 0.0, batch, inputs, G,
                 problems=[0, 65534, 65535, 65536, batch - 1])


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_strided_batch_scale_only_semantics(cuda, dt):
    """alpha = 0 or k = 0: C = beta C per problem, beta = 0 included (NaN in
    C stays NaN, as in dgemm_); beta = 1 leaves C as it is."""
    import torch
    m, n, batch, ldc = 5, 4, 3, 7
    sc = ldc * n + 2
    C = np.full(sc * (batch - 1) + ldc * n, np.nan, dtype=dt)
    vals = np.arange(batch * m * n, dtype=dt).reshape(batch, n, m) + 1
    for i in range(batch):
        C[i * sc:i * sc + ldc * n].reshape((ldc, n), order="F")[:m] = vals[i].T
    C[sc + 1] = np.nan  # problem 1, element (1, 0)
    for alpha, k, beta in ((0.0, 3, 2.0), (1.0, 0, 0.0), (0.0, 0, 1.0)):
        dC = torch.from_numpy(C.copy()).cuda()
        dA = torch.zeros(64, dtype=torch.float64 if dt == np.float64 else torch.float32).cuda()
        _batched(dt)("N", "N", m, n, k, alpha, dA, m, 0, dA, max(k, 1), 0, beta, dC, ldc, sc,
                     batch)
        torch.cuda.synchronize()
        G = dC.cpu().numpy()
        for i in range(batch):
            gi = G[i * sc:i * sc + ldc * n].reshape((ldc, n), order="F")
            ci = C[i * sc:i * sc + ldc * n].reshape((ldc, n), order="F")
            want = ci[:m] * beta if beta != 1.0 else ci[:m]
            np.testing.assert_array_equal(gi[:m], want)
            assert np.all(np.isnan(gi[m:]))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_ilp64_form_is_dgemm_bit_for_bit(cuda, dt, ta, tb):
    """dgemm_64_ / sgemm_64_ on host arrays give exactly dgemm_ / sgemm_'s
    result (same kernels) and the oracle's within tolerance."""
    rng = np.random.default_rng(31)
    m, n, k = 97, 65, 130
    ra, ca = stored_shape(ta, m, k)
    rb, cb = stored_shape(tb, k, n)
    A = np.asfortranarray(rng.uniform(-1, 1, (ra + 3, ca)).astype(dt))
    B = np.asfortranarray(rng.uniform(-1, 1, (rb + 1, cb)).astype(dt))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (m + 2, n)).astype(dt))
    C0[m:] = np.nan
    C32, C64, R = C0.copy(order="F"), C0.copy(order="F"), C0.copy(order="F")
    f32, f64 = (gx.dgemm, gx.dgemm64) if dt == np.float64 else (gx.sgemm, gx.sgemm64)
    f32(ta, tb, m, n, k, 1.5, A, ra + 3, B, rb + 1, -0.5, C32, m + 2)
    f64(ta, tb, m, n, k, 1.5, A, ra + 3, B, rb + 1, -0.5, C64, m + 2)
    np.testing.assert_array_equal(C64, C32)
    oracle.gemm(ta, tb, m, n, k, 1.5, A, ra + 3, B, rb + 1, -0.5, R, m + 2)
    bound = abs_bound(A, B, C0, ta, tb, m, n, k, 1.5, -0.5)
    assert max_comp(C64[:m], R[:m], bound) <= _tol(dt, k)
    assert np.all(np.isnan(C64[m:]))
