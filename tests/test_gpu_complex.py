"""Complex GEMM (zgemm_, cgemm_ and the three-product zgemm3m_, cgemm3m_)
through the C ABI against the oracle's complex restatement.

Tolerances, componentwise on the real and imaginary parts, with
S = |alpha| |op(A)| |op(B)| + |beta| |C0| (complex magnitudes):
  four products: 128 k u S   (two real products per part, each within the
                              native 64 k u, plus the complex combine);
  three products: 512 k u S  (the sum product (Ar + Ai)(Br + Bi) is up to
                              twice |A||B| and cancels against the others).
u = 2^-53 (double complex) or 2^-24 (single complex).
"""
import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_util import U32, U64, stored_shape

pytestmark = pytest.mark.gpu
CTRANS = ["N", "T", "C"]
FORMS = {("z", 0): gx.zgemm, ("z", 1): gx.zgemm3m, ("c", 0): gx.cgemm, ("c", 1): gx.cgemm3m}


def _cplx(rng, shape, dt):
    return np.asfortranarray((rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
                             .astype(dt))


def _cop(X, t):
    return X if t in "Nn" else (X.conj().T if t in "Cc" else X.T)


def _check(G, R, A, B, C0, ta, tb, m, n, k, alpha, beta, dt, three):
    u = U64 if dt == np.complex128 else U32
    oa = np.abs(_cop(A[: (m if ta == "N" else k)], ta)[:m, :k]).astype(np.float64)
    ob = np.abs(_cop(B[: (k if tb == "N" else n)], tb)[:k, :n]).astype(np.float64)
    S = abs(alpha) * (oa @ ob) + abs(beta) * np.abs(C0[:m, :n]).astype(np.float64)
    tol = (512 if three else 128) * max(k, 1) * u
    for part in (np.real, np.imag):
        err = np.abs(part(G[:m, :n]).astype(np.float64) - part(R[:m, :n]).astype(np.float64))
        assert np.all(err <= tol * S + 1e-300), float(np.max(err / (S + 1e-300)))


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("three", [0, 1])
@pytest.mark.parametrize("ta", CTRANS)
@pytest.mark.parametrize("tb", CTRANS)
def test_complex_gemm_matches_the_oracle(cuda, prec, three, ta, tb):
    dt = np.complex128 if prec == "z" else np.complex64
    rng = np.random.default_rng(700 + 3 * CTRANS.index(ta) + CTRANS.index(tb))
    alpha, beta = 1.5 - 0.5j, -0.5 + 0.25j
    for (m, n, k, pad) in ((33, 29, 31, 3), (130, 70, 200, 0), (1, 1, 1, 1)):
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A, B = _cplx(rng, (ra + pad, ca), dt), _cplx(rng, (rb + 1, cb), dt)
        C = _cplx(rng, (m + pad, n), dt)
        C[m:] = np.nan
        C0, R = C.copy(order="F"), C.copy(order="F")
        FORMS[(prec, three)](ta, tb, m, n, k, alpha, A, ra + pad, B, rb + 1, beta, C, m + pad)
        oracle.complex_gemm(ta, tb, m, n, k, alpha, A, ra + pad, B, rb + 1, beta, R, m + pad)
        assert np.all(np.isnan(C[m:])), "ld padding of C written"
        _check(C, R, A, B, C0, ta, tb, m, n, k, alpha, beta, dt, three)


@pytest.mark.parametrize("prec", ["z", "c"])
def test_complex_gemm_device_operands_and_larger_shape(cuda, prec):
    """Device-resident complex tensors (used in place), 256 x 192 x 320."""
    import torch
    dt = np.complex128 if prec == "z" else np.complex64
    rng = np.random.default_rng(710)
    m, n, k = 256, 192, 320
    A, B, C = _cplx(rng, (k, m), dt), _cplx(rng, (k, n), dt), _cplx(rng, (m, n), dt)
    dA, dB, dC = (torch.from_numpy(np.ravel(x, order="F").copy()).cuda() for x in (A, B, C))
    FORMS[(prec, 0)]("C", "N", m, n, k, 0.5 + 2j, dA, k, dB, k, 1 + 0j, dC, m)
    torch.cuda.synchronize()
    G = dC.cpu().numpy().reshape((m, n), order="F")
    R = C.copy(order="F")
    oracle.complex_gemm("C", "N", m, n, k, 0.5 + 2j, A, k, B, k, 1 + 0j, R, m)
    _check(G, R, A, B, C, "C", "N", m, n, k, 0.5 + 2j, 1 + 0j, dt, 0)


@pytest.mark.parametrize("prec", ["z", "c"])
@pytest.mark.parametrize("three", [0, 1])
def test_complex_gemm_semantics(cuda, prec, three):
    """beta != 1 scales C first, also beta = 0 (NaN stays NaN); alpha = 0 or
    k = 0 leaves beta C; alpha = 0 with beta = 1 leaves C untouched; m = 0
    is a no-op -- exactly the oracle's values."""
    dt = np.complex128 if prec == "z" else np.complex64
    rng = np.random.default_rng(720)
    f = FORMS[(prec, three)]
    m, n, k = 5, 4, 6
    A, B = _cplx(rng, (m, k), dt), _cplx(rng, (k, n), dt)
    C = _cplx(rng, (m, n), dt)
    C[2, 1] = complex(np.nan, 1.0)
    G = C.copy(order="F")
    f("N", "N", m, n, k, 1 + 0j, A, m, B, k, 0j, G, m)
    assert np.isnan(G[2, 1].real) and np.isnan(G[2, 1].imag)
    for alpha, kk, beta in ((0j, k, 2 - 1j), (1 + 1j, 0, 0.5j), (0j, k, 1 + 0j)):
        G, R = C.copy(order="F"), C.copy(order="F")
        f("N", "N", m, n, kk, alpha, A, m, B, k, beta, G, m)
        oracle.complex_gemm("N", "N", m, n, kk, alpha, A, m, B, k, beta, R, m)
        np.testing.assert_array_equal(G, R)
    G = C.copy(order="F")
    f("N", "N", 0, n, k, 1 + 0j, A, 1, B, k, 0j, G, m)
    np.testing.assert_array_equal(G, C)
