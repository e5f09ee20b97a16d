"""Triangular multiply and solve (dtrmm_/dtrsm_, strmm_/strsm_) through the
C ABI against the oracle's restatements, every side/uplo/transpose/diagonal
combination, padded leading dimensions, host and device operands.

A is well conditioned (diagonal 2..3, off-diagonal U[-1,1)/k), so the solve
error is a few k u: the tolerance is 64 k u relative to max |X| (u = 2^-53
or 2^-24) for TRSM, and TRMM is checked like a GEMM (64 k u relative to
|alpha| |op(A)| |B|).
"""
import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_util import U32, U64

pytestmark = pytest.mark.gpu
TRI_CASES = [(s, u, t, d) for s in "LR" for u in "UL" for t in "NTC" for d in "NU"]
FN = {("d", 0): gx.dtrmm, ("d", 1): gx.dtrsm, ("s", 0): gx.strmm, ("s", 1): gx.strsm}


def _system(rng, side, m, n, pad, dt):
    k = m if side == "L" else n
    A = np.asfortranarray((rng.uniform(-1, 1, (k + pad, k)) / max(k, 1)).astype(dt))
    A[np.arange(k), np.arange(k)] = (2 + rng.uniform(0, 1, k)).astype(dt)
    B = np.asfortranarray(rng.uniform(-1, 1, (m + pad, n)).astype(dt))
    B[m:] = np.nan
    return A, B, k


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("side,uplo,trans,diag", TRI_CASES)
def test_triangular_ops_match_the_oracle(cuda, prec, side, uplo, trans, diag):
    dt = np.float64 if prec == "d" else np.float32
    u = U64 if prec == "d" else U32
    rng = np.random.default_rng(900 + TRI_CASES.index((side, uplo, trans, diag)))
    for (m, n, pad) in ((1, 1, 0), (70, 45, 3), (130, 257, 1), (64, 128, 0)):
        A, B, k = _system(rng, side, m, n, pad, dt)
        # the other triangle (and a unit diagonal) must not be read
        other = np.tril(np.ones(A.shape, bool), -1) if uplo == "U" else np.triu(
            np.ones(A.shape, bool), 1)
        other[k:] = False
        Ad = A.copy(order="F")
        Ad[other] = np.nan
        if diag == "U":
            Ad[np.arange(k), np.arange(k)] = np.nan
        for solve in (0, 1):
            G, R = B.copy(order="F"), B.copy(order="F")
            FN[(prec, solve)](side, uplo, trans, diag, m, n, 1.5, Ad, k + pad, G, m + pad)
            oracle.tri(solve, side, uplo, trans, diag, m, n, 1.5, A, k + pad, R, m + pad)
            assert np.all(np.isnan(G[m:])), "ld padding of B written"
            err = np.max(np.abs(G[:m].astype(np.float64) - R[:m]))
            if solve:
                assert err <= 64 * k * u * max(np.max(np.abs(R[:m])), 1e-300), (m, n, err)
            else:
                T = np.abs(np.triu(A[:k]) if uplo == "U" else np.tril(A[:k])).astype(np.float64)
                if diag == "U":
                    np.fill_diagonal(T, 1.0)
                opT = T.T if trans != "N" else T
                Bm = np.abs(B[:m]).astype(np.float64)
                scale = 1.5 * (opT @ Bm if side == "L" else Bm @ opT)
                assert np.all(np.abs(G[:m] - R[:m]) <= 64 * k * u * scale + 1e-300), (m, n)


@pytest.mark.parametrize("prec", ["d", "s"])
def test_triangular_device_operands_alpha_zero_and_round_trip(cuda, prec):
    """Device tensors used in place; alpha = 0 writes zeros (NaN in B and A
    not read); TRSM undoes TRMM on a well-conditioned triangle."""
    import torch
    dt = np.float64 if prec == "d" else np.float32
    u = U64 if prec == "d" else U32
    rng = np.random.default_rng(950)
    m, n = 200, 96
    A, B, k = _system(rng, "R", m, n, 0, dt)
    dA = torch.from_numpy(np.ravel(A, order="F").copy()).cuda()
    dB = torch.from_numpy(np.ravel(B, order="F").copy()).cuda()
    FN[(prec, 0)]("R", "L", "T", "N", m, n, 2.0, dA, k, dB, m)
    FN[(prec, 1)]("R", "L", "T", "N", m, n, 0.5, dA, k, dB, m)
    torch.cuda.synchronize()
    back = dB.cpu().numpy().reshape((m, n), order="F")
    assert np.max(np.abs(back - B)) <= 256 * k * u
    Z = np.full((m, n), np.nan, dtype=dt, order="F")
    FN[(prec, 1)]("L", "U", "N", "N", m, n, 0.0, np.full((m, m), np.nan, dtype=dt, order="F"), m,
                  Z, m)
    assert np.all(Z == 0.0)
