"""Parity of the emulated forms on the GPU.

Ozaki-I (dgemm_x_ozaki_dev): every INT8 slice product is exact and every
FP64 step is the same operation as in the oracle's restatement, so the GPU
result must be bit-identical to oracle.ozaki_gemm for every transpose pair,
slice count and ragged shape.

split-BF16x3 (sgemm_x_bf16x3_dev): the split is exact integer work, the three
products accumulate in FP32 inside the MFMA in an order the oracle cannot
follow, so the GPU is compared with the oracle's exactly-accumulated BF16x3
within FP32 accumulation error (64*k*2^-24 componentwise), and with the FP64
product within the split error (3*2^-16 componentwise).
"""
import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_util import U32, abs_bound, from_device, max_comp, rel_fro, stored_shape, to_device

pytestmark = pytest.mark.gpu
TRANS = [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")]


def _ozaki_gpu(torch, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s):
    dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
    gx.dgemm_ozaki_dev(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, s)
    torch.cuda.synchronize()
    return from_device(dC, C.shape)


@pytest.mark.parametrize("s", [3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_ozaki_bit_exact(cuda, ta, tb, s):
    import torch
    rng = np.random.default_rng(100 + s)
    for (m, n, k, pad) in [(1, 1, 1, 0), (17, 9, 33, 1), (130, 129, 200, 2), (64, 256, 513, 0)]:
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = np.asfortranarray(rng.uniform(-1, 1, (ra + pad, ca)) * np.exp(2 * rng.standard_normal((ra + pad, ca))))
        B = np.asfortranarray(rng.uniform(-1, 1, (rb + pad, cb)))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)))
        C[m:] = 555.0
        R = C.copy(order="F")
        oracle.ozaki_gemm(ta, tb, m, n, k, 1.5, A, ra + pad, B, rb + pad, -0.5, R, m + pad, s)
        G = _ozaki_gpu(torch, ta, tb, m, n, k, 1.5, A, ra + pad, B, rb + pad, -0.5, C, m + pad, s)
        np.testing.assert_array_equal(G, R)


def test_ozaki_nonfinite_and_semantics(cuda):
    import torch
    m, n, k = 5, 4, 6
    A = np.ones((m, k), order="F")
    A[2, 3] = np.inf
    B = np.ones((k, n), order="F")
    C = np.zeros((m, n), order="F")
    G = _ozaki_gpu(torch, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, 6)
    assert np.all(np.isnan(G[2])) and np.all(G[[0, 1, 3, 4]] == 6.0)
    # alpha = 0: C = beta*C without touching A, B
    C = np.full((m, n), 3.0, order="F")
    G = _ozaki_gpu(torch, "N", "N", m, n, k, 0.0, A, m, B, k, 2.0, C, m, 6)
    assert np.all(G == 6.0)


def test_ozaki_accuracy_vs_dot2(cuda):
    """Error against the Dot2 product shrinks with the slice count and at 8
    slices reaches FP64-level error for well-scaled inputs."""
    import torch
    rng = np.random.default_rng(5)
    m, n, k = 96, 80, 700
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    R = np.zeros((m, n), order="F")
    oracle.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m)
    errs = []
    for s in range(3, 9):
        G = _ozaki_gpu(torch, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, np.zeros((m, n), order="F"), m, s)
        errs.append(rel_fro(G, R))
    assert all(b < a for a, b in zip(errs, errs[1:-1])), errs
    assert errs[-1] < 1e-15, errs


@pytest.mark.parametrize("ta,tb", TRANS)
def test_bf16x3(cuda, ta, tb):
    import torch
    rng = np.random.default_rng(31)
    for (m, n, k, pad) in [(1, 1, 1, 0), (19, 33, 45, 3), (130, 129, 257, 1), (128, 256, 512, 0)]:
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = np.asfortranarray(rng.uniform(-1, 1, (ra + pad, ca)).astype(np.float32))
        B = np.asfortranarray(rng.uniform(-1, 1, (rb + pad, cb)).astype(np.float32))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)).astype(np.float32))
        C[m:] = 555.0
        R3 = C.copy(order="F")
        oracle.bf16x3_gemm(ta, tb, m, n, k, 1.5, A, ra + pad, B, rb + pad, -0.5, R3, m + pad)
        R64 = C.astype(np.float64, order="F")
        oracle.gemm(ta, tb, m, n, k, 1.5, A.astype(np.float64, order="F"), ra + pad,
                    B.astype(np.float64, order="F"), rb + pad, -0.5, R64, m + pad)
        dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
        gx.sgemm_bf16x3_dev(ta, tb, m, n, k, 1.5, dA, ra + pad, dB, rb + pad, -0.5, dC, m + pad)
        torch.cuda.synchronize()
        G = from_device(dC, C.shape)
        bound = abs_bound(A, B, C, ta, tb, m, n, k, 1.5, -0.5)
        assert max_comp(G[:m], R3[:m], bound) <= 64 * k * U32, (m, n, k)
        assert max_comp(G[:m], R64[:m], bound) <= 3 * 2.0 ** -16 + 64 * k * U32, (m, n, k)
        assert np.all(G[m:] == 555.0)


def _ozaki2_gpu(torch, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nmod):
    dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
    gx.dgemm_ozaki2_dev(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, nmod)
    torch.cuda.synchronize()
    return from_device(dC, C.shape)


@pytest.mark.parametrize("nmod", [8, 12, 16, 20, 22])
@pytest.mark.parametrize("ta,tb", TRANS)
def test_ozaki2_bit_exact(cuda, ta, tb, nmod):
    """Residue split, INT8 residue products and CRT reconstruction against the
    oracle's exact 256-bit integer restatement: identical bits."""
    import torch
    rng = np.random.default_rng(200 + nmod)
    for (m, n, k, pad) in [(1, 1, 1, 0), (17, 9, 33, 1), (130, 129, 200, 2), (64, 256, 513, 0),
                           (257, 31, 129, 3)]:
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = np.asfortranarray(rng.uniform(-1, 1, (ra + pad, ca)) * np.exp(4 * rng.standard_normal((ra + pad, ca))))
        B = np.asfortranarray(rng.uniform(-1, 1, (rb + pad, cb)) * np.exp(2 * rng.standard_normal((rb + pad, cb))))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)))
        C[m:] = 555.0
        R = C.copy(order="F")
        oracle.ozaki2_gemm(ta, tb, m, n, k, 1.5, A, ra + pad, B, rb + pad, -0.5, R, m + pad, nmod)
        G = _ozaki2_gpu(torch, ta, tb, m, n, k, 1.5, A, ra + pad, B, rb + pad, -0.5, C, m + pad, nmod)
        np.testing.assert_array_equal(G, R)


@pytest.mark.parametrize("nmod", list(range(8, 23)))
def test_ozaki2_bit_exact_every_modulus_count(cuda, nmod):
    """Every modulus count (so every limb count of the reconstruction,
    4..11 sixteen-bit limbs) against the oracle, bit for bit: heavy-tailed
    entries, and rows and columns of equal-sign entries at the top of their
    binade, whose dot products come closest to the CRT range bound."""
    import torch
    rng = np.random.default_rng(300 + nmod)
    m, n, k = 40, 24, 300
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * np.exp(4 * rng.standard_normal((m, k))))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)) * np.exp(2 * rng.standard_normal((k, n))))
    top = np.nextafter(1.0, 0.0)
    A[:4] = top * 2.0 ** 5
    A[4:8] = -top
    B[:, :3] = top * 2.0 ** -3
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    R = C.copy(order="F")
    oracle.ozaki2_gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.5, R, m, nmod)
    G = _ozaki2_gpu(torch, "N", "N", m, n, k, 1.0, A, m, B, k, 0.5, C, m, nmod)
    np.testing.assert_array_equal(G, R)


def test_both_product_kernels(cuda):
    """The register-staged (0) and direct-to-LDS (1, default) product kernels:
    Ozaki-I and Ozaki-II bit-identical to the oracle, BF16x3 within FP32
    accumulation error of it and bit-identical between the two kernels (the
    same MFMA sequence per element), on ragged shapes that leave partial
    tiles and K stages."""
    import torch
    rng = np.random.default_rng(77)
    try:
        bf = {}
        for c in (0, 1):
            gx.set_product_config("d", c)
            gx.set_product_config("s", c)
            for (m, n, k) in [(300, 270, 333), (17, 520, 129)]:
                A = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * np.exp(2 * rng.standard_normal((m, k))))
                B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
                C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
                for s in (4, 7):
                    R = C.copy(order="F")
                    oracle.ozaki_gemm("N", "N", m, n, k, 1.25, A, m, B, k, 0.5, R, m, s)
                    G = _ozaki_gpu(torch, "N", "N", m, n, k, 1.25, A, m, B, k, 0.5, C, m, s)
                    np.testing.assert_array_equal(G, R)
                for nmod in (12, 16):
                    R = C.copy(order="F")
                    oracle.ozaki2_gemm("N", "N", m, n, k, 1.25, A, m, B, k, 0.5, R, m, nmod)
                    G = _ozaki2_gpu(torch, "N", "N", m, n, k, 1.25, A, m, B, k, 0.5, C, m, nmod)
                    np.testing.assert_array_equal(G, R)
            m, n, k = 260, 200, 300
            r2 = np.random.default_rng(78)
            A = np.asfortranarray(r2.uniform(-1, 1, (k, m)).astype(np.float32))
            B = np.asfortranarray(r2.uniform(-1, 1, (k, n)).astype(np.float32))
            C = np.asfortranarray(r2.uniform(-1, 1, (m, n)).astype(np.float32))
            R3 = C.copy(order="F")
            oracle.bf16x3_gemm("T", "N", m, n, k, 1.5, A, k, B, k, -0.5, R3, m)
            dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
            gx.sgemm_bf16x3_dev("T", "N", m, n, k, 1.5, dA, k, dB, k, -0.5, dC, m)
            torch.cuda.synchronize()
            G = from_device(dC, C.shape)
            bound = abs_bound(A, B, C, "T", "N", m, n, k, 1.5, -0.5)
            assert max_comp(G, R3, bound) <= 64 * k * U32
            bf[c] = G
        np.testing.assert_array_equal(bf[0], bf[1])
    finally:
        gx.set_product_config("d", -1)
        gx.set_product_config("s", -1)


@pytest.mark.parametrize("nmod", [8, 12, 16, 19, 22])
def test_ozaki2_rounding_ties_and_word_positions(cuda, nmod):
    """Integer products that need rounding to 53 bits: odd 54-bit products are
    exact ties (to even), (2^54 - 1) rounds up across a binade, and the
    other pairs give 55-bit and short results, of both signs, plus zeros.
    The modulus counts place the reconstructed integer in different 32-bit
    words.  The oracle is pinned to Python's correctly rounded int -> float
    conversion where the split keeps the integers exact (P >= 53 with
    k = 1, nmod >= 16), and the device must match the oracle bit for bit."""
    import torch
    a = [1, 3, 5, 7, -3, 9, 2**26 + 1, -(2**25 + 3)]
    b = [(2**53 + 1) // 3, (2**54 - 1) // 3, 2**52 + 1, 2**53 - 1, -(2**52 + 3),
         4503599627370497, 1, 0]
    A = np.asfortranarray(np.array(a, dtype=np.float64).reshape(8, 1))
    B = np.asfortranarray(np.array(b, dtype=np.float64).reshape(1, 8))
    C = np.zeros((8, 8), order="F")
    R = C.copy(order="F")
    oracle.ozaki2_gemm("N", "N", 8, 8, 1, 1.0, A, 8, B, 1, 0.0, R, 8, nmod)
    if nmod >= 16:
        want = np.array([[float(x * y) for y in b] for x in a])
        np.testing.assert_array_equal(R, want)
    G = _ozaki2_gpu(torch, "N", "N", 8, 8, 1, 1.0, A, 8, B, 1, 0.0, C, 8, nmod)
    np.testing.assert_array_equal(G, R)
    # the same pairs as two-term sums with a cancelling partner
    A2 = np.asfortranarray(np.hstack([A, -A]))
    B2 = np.asfortranarray(np.vstack([B, B - 2.0]))
    R2 = np.zeros((8, 8), order="F")
    oracle.ozaki2_gemm("N", "N", 8, 8, 2, 1.0, A2, 8, B2, 2, 0.0, R2, 8, nmod)
    G2 = _ozaki2_gpu(torch, "N", "N", 8, 8, 2, 1.0, A2, 8, B2, 2, 0.0, np.zeros((8, 8), order="F"), 8, nmod)
    np.testing.assert_array_equal(G2, R2)


def test_ozaki2_chunks_subnormals_and_semantics(cuda):
    import torch
    rng = np.random.default_rng(9)
    # K beyond one 65536 chunk
    m, n, k = 3, 5, 65536 + 300
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    R = C.copy(order="F")
    oracle.ozaki2_gemm("N", "N", m, n, k, 0.75, A, m, B, k, 2.0, R, m, 20)
    G = _ozaki2_gpu(torch, "N", "N", m, n, k, 0.75, A, m, B, k, 2.0, C, m, 20)
    np.testing.assert_array_equal(G, R)
    # tiny and subnormal entries, results near the bottom of the range
    A = np.asfortranarray(rng.uniform(-1, 1, (40, 50)) * 2.0 ** -1000)
    A[3, :5] = 5e-324
    B = np.asfortranarray(rng.uniform(-1, 1, (50, 30)) * 2.0 ** -60)
    C = np.zeros((40, 30), order="F")
    R = C.copy(order="F")
    oracle.ozaki2_gemm("N", "N", 40, 30, 50, 1.0, A, 40, B, 50, 0.0, R, 40, 16)
    G = _ozaki2_gpu(torch, "N", "N", 40, 30, 50, 1.0, A, 40, B, 50, 0.0, C, 40, 16)
    np.testing.assert_array_equal(G, R)
    # Inf in a row gives NaN in that row; alpha = 0 reads neither operand
    A = np.ones((5, 6), order="F")
    A[2, 3] = np.inf
    B = np.ones((6, 4), order="F")
    G = _ozaki2_gpu(torch, "N", "N", 5, 4, 6, 1.0, A, 5, B, 6, 0.0, np.zeros((5, 4), order="F"), 5, 12)
    assert np.all(np.isnan(G[2])) and np.all(G[[0, 1, 3, 4]] == 6.0)
    G = _ozaki2_gpu(torch, "N", "N", 5, 4, 6, 0.0, A, 5, B, 6, 2.0, np.full((5, 4), 3.0, order="F"), 5, 12)
    assert np.all(G == 6.0)
    with pytest.raises(gx.GemmError):
        gx.dgemm_ozaki2_dev("N", "N", 5, 4, 6, 1.0, 0, 5, 0, 6, 0.0, 0, 5, 7)


def test_ozaki2_accuracy_matches_native_at_20_moduli(cuda):
    """On entries with a wide exponent range ((u-0.5) exp(4g)) the error
    against Dot2 falls with the modulus count and at 20 moduli is no larger
    than native FP64's."""
    import torch
    rng = np.random.default_rng(6)
    m, n, k = 96, 80, 1000
    A = np.asfortranarray((rng.uniform(0, 1, (m, k)) - 0.5) * np.exp(4 * rng.standard_normal((m, k))))
    B = np.asfortranarray((rng.uniform(0, 1, (k, n)) - 0.5) * np.exp(4 * rng.standard_normal((k, n))))
    R = np.zeros((m, n), order="F")
    oracle.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m)
    dA, dB = to_device(A, torch), to_device(B, torch)
    dC = to_device(np.zeros((m, n), order="F"), torch)
    gx.dgemm_dev("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    torch.cuda.synchronize()
    native = rel_fro(from_device(dC, (m, n)), R)
    errs = []
    for nmod in (10, 14, 18, 20, 22):
        G = _ozaki2_gpu(torch, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0,
                        np.zeros((m, n), order="F"), m, nmod)
        errs.append(rel_fro(G, R))
    assert all(b <= a for a, b in zip(errs, errs[1:])), errs
    assert errs[3] <= native, (errs, native)
