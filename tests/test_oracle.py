"""Pins the CPU oracle before anything is checked against it.

The oracle (oracle/gemm_oracle.c) restates the reference's dgemm_/sgemm_
semantics (wjc404/GEMM_AVX2 src/gemm_driver.c:118-197) with Dot2
accumulation.  It is checked here against
  * the golden vectors written by the reference itself (tests/golden/),
  * the closed-form 12x12x12 known-answer test of
    OpenBLAS-like_implementation/dgemm_kernel_4x4_haswell.c:472-492,
  * brute force in long double on tiny shapes,
  * invariants that a dropped term, a wrong sign, a transposed operand or a
    wrong beta rule would break,
and, when oracle/_ref/ is built, against the reference on extra shapes.
The emulated restatements (Ozaki-I, Ozaki-II, split-BF16x3) are pinned by
exact integer cases, by exact rational arithmetic (Python's Fraction: the
scaled integers, the rounding, the correctly rounded product when the inputs
fit), by a library BF16 rounding routine and by their error against Dot2
shrinking as the slice or modulus count grows.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from gemm_util import (U32, U64, abs_bound, case_inputs, load_fixtures, max_comp, op,
                       rel_fro, stored_shape)

FIX = load_fixtures()


def _tol(dtype, k):
    u = U64 if dtype == "f64" else U32
    return 64 * max(k, 1) * u


# --------------------------------------------------------------- fixtures
@pytest.mark.parametrize("case", [c for c in FIX if c["group"] == "small"],
                         ids=lambda c: f"{c['dtype']}-{c['transa']}{c['transb']}-"
                                       f"{c['m']}x{c['n']}x{c['k']}-a{c['alpha']}b{c['beta']}")
def test_oracle_matches_reference_small(case):
    A, B, C = case_inputs(case, oracle)
    C0 = C.copy(order="F")
    m, n, k = case["m"], case["n"], case["k"]
    oracle.gemm(case["transa"], case["transb"], m, n, k, case["alpha"], A, case["lda"], B,
                case["ldb"], case["beta"], C, case["ldc"])
    ref = case["values"].reshape((m, n), order="F")
    assert case["pad_touched"] == 0
    assert np.all(np.isnan(C[m:, :n])), "oracle wrote into the ld padding"
    if case["alpha"] == 0.0 or k == 0:
        np.testing.assert_array_equal(C[:m, :n].astype(np.float64), ref)
        return
    bound = abs_bound(A, B, C0, case["transa"], case["transb"], m, n, k, case["alpha"],
                      case["beta"])
    assert max_comp(C[:m, :n], ref, bound) <= _tol(case["dtype"], k)


@pytest.mark.parametrize("case", [c for c in FIX if c["group"] == "medium"],
                         ids=lambda c: f"{c['dtype']}-{c['transa']}{c['transb']}-"
                                       f"{c['m']}x{c['n']}x{c['k']}")
def test_oracle_matches_reference_medium(case):
    A, B, C = case_inputs(case, oracle)
    C0 = C.copy(order="F")
    m, n, k = case["m"], case["n"], case["k"]
    oracle.gemm(case["transa"], case["transb"], m, n, k, case["alpha"], A, case["lda"], B,
                case["ldb"], case["beta"], C, case["ldc"])
    got = C[case["is"], case["js"]].astype(np.float64)
    bound = abs_bound(A, B, C0, case["transa"], case["transb"], m, n, k, case["alpha"],
                      case["beta"])[case["is"], case["js"]]
    assert max_comp(got, case["values"], bound) <= _tol(case["dtype"], k)
    full = C[:m, :n].astype(np.float64)
    assert abs(np.linalg.norm(full) - case["fro"]) <= 1e-6 * case["fro"]


@pytest.mark.parametrize("case", [c for c in FIX if c["group"] not in ("small", "medium")],
                         ids=lambda c: f"{c['group']}-{c['dtype']}-{c['transa']}{c['transb']}"
                                       f"-p{c['param']}")
def test_oracle_sampled_matches_reference_large(case):
    dt = np.float64 if case["dtype"] == "f64" else np.float32
    exp, ab = oracle.sampled(dt, case["transa"], case["transb"], case["m"], case["n"],
                             case["k"], case["alpha"], case["beta"], case["seed"], case["is"],
                             case["js"], case["dist"], case["param"])
    bound = abs(case["alpha"]) * ab + abs(case["beta"]) + 1e-300
    err = np.max(np.abs(exp - case["values"]) / bound)
    assert err <= _tol(case["dtype"], case["k"])


def test_fixture_inventory():
    groups = {c["group"] for c in FIX}
    assert {"small", "medium", "nn1024", "sq8192", "ragged8192", "emu8192"} <= groups
    sq = [c for c in FIX if c["group"] == "sq8192"]
    assert {(c["dtype"], c["transa"] + c["transb"]) for c in sq} == {
        (d, t) for d in ("f64", "f32") for t in ("NN", "NT", "TN", "TT")}
    assert all(c["lda"] == c["ldb"] == c["ldc"] == 8200 for c in sq)
    assert all(c["pad_touched"] == 0 for c in FIX)


# --------------------------------------------------------------- closed forms
def test_known_answer_12x12x12():
    """OpenBLAS-like_implementation/dgemm_kernel_4x4_haswell.c:472-492:
    A[idx]=5+idx%2, B[idx]=20+(idx/ldb)%8, C=0, alpha=beta=1, so
    C[i,j] = 12*(5+i%2)*(20+j%8)."""
    n = 12
    idx = np.arange(n * n)
    A = np.asfortranarray((5.0 + idx % 2).reshape((n, n), order="F"))
    B = np.asfortranarray((20.0 + (idx // n) % 8).reshape((n, n), order="F"))
    C = np.zeros((n, n), order="F")
    oracle.gemm("N", "N", n, n, n, 1.0, A, n, B, n, 1.0, C, n)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    np.testing.assert_array_equal(C, 12.0 * (5 + i % 2) * (20 + j % 8))
    Cs = np.zeros((n, n), dtype=np.float32, order="F")
    oracle.gemm("N", "N", n, n, n, 1.0, A.astype(np.float32), n, B.astype(np.float32), n, 1.0,
                Cs, n)
    np.testing.assert_array_equal(Cs, (12.0 * (5 + i % 2) * (20 + j % 8)).astype(np.float32))


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (2.0, -0.5), (-0.25, 1.0), (0.5, 3.0)])
def test_dot2_matches_long_double(ta, tb, alpha, beta):
    rng = np.random.default_rng(7)
    for m, n, k in [(1, 1, 1), (3, 5, 7), (9, 4, 33), (17, 13, 64)]:
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = np.asfortranarray(rng.uniform(-1, 1, (ra + 2, ca)))
        B = np.asfortranarray(rng.uniform(-1, 1, (rb + 1, cb)))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + 3, n)))
        C1, C2 = C.copy(order="F"), C.copy(order="F")
        oracle.gemm(ta, tb, m, n, k, alpha, A, ra + 2, B, rb + 1, beta, C1, m + 3)
        oracle.longdouble_gemm(ta, tb, m, n, k, alpha, A, ra + 2, B, rb + 1, beta, C2, m + 3)
        bound = abs_bound(A, B, C, ta, tb, m, n, k, alpha, beta)
        assert max_comp(C1[:m], C2[:m], bound) <= 4 * U64
        np.testing.assert_array_equal(C1[m:], C[m:])  # padding untouched


def test_ones_give_k():
    for k in (1, 7, 288, 1000):
        A = np.ones((5, k), order="F")
        B = np.ones((k, 6), order="F")
        C = np.full((5, 6), np.nan, order="F")
        oracle.gemm("N", "N", 5, 6, k, 1.0, A, 5, B, k, 0.0, C, 5)
        # beta = 0 multiplies NaN: the reference's rule keeps the NaN
        assert np.all(np.isnan(C))
        C = np.zeros((5, 6), order="F")
        oracle.gemm("N", "N", 5, 6, k, 1.0, A, 5, B, k, 0.0, C, 5)
        np.testing.assert_array_equal(C, float(k))


def test_beta_rule_alpha_zero_k_zero():
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.uniform(-1, 1, (4, 4)))
    A[0, 0] = np.nan  # never read when alpha == 0
    B = np.asfortranarray(rng.uniform(-1, 1, (4, 4)))
    C = np.asfortranarray(rng.uniform(-1, 1, (4, 4)))
    for beta in (0.0, 1.0, -2.5):
        C1 = C.copy(order="F")
        oracle.gemm("N", "N", 4, 4, 4, 0.0, A, 4, B, 4, beta, C1, 4)
        np.testing.assert_array_equal(C1, C * beta if beta != 1.0 else C)
        C2 = C.copy(order="F")
        oracle.gemm("N", "N", 4, 4, 0, 1.0, A, 4, B, 4, beta, C2, 4)
        np.testing.assert_array_equal(C2, C * beta if beta != 1.0 else C)
    C3 = C.copy(order="F")
    oracle.gemm("N", "N", 0, 4, 4, 1.0, A, 4, B, 4, 0.0, C3, 4)
    np.testing.assert_array_equal(C3, C)


def test_transpose_chars_and_identity():
    """'N'/'n' are no-transpose, anything else transposes; op(A) stored
    transposed gives bit-identical results; (AB)^T = B^T A^T."""
    rng = np.random.default_rng(11)
    m, n, k = 13, 9, 21
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    base = np.zeros((m, n), order="F")
    oracle.gemm("N", "N", m, n, k, 1.5, A, m, B, k, 0.0, base, m)
    At, Bt = np.asfortranarray(A.T), np.asfortranarray(B.T)
    for ta, tb, a, lda, b, ldb in [("n", "n", A, m, B, k), ("T", "N", At, k, B, k),
                                   ("C", "n", At, k, B, k), ("N", "t", A, m, Bt, n),
                                   ("x", "c", At, k, Bt, n)]:
        C = np.zeros((m, n), order="F")
        oracle.gemm(ta, tb, m, n, k, 1.5, a, lda, b, ldb, 0.0, C, m)
        np.testing.assert_array_equal(C, base)
    # C^T = op(B)^T op(A)^T: swap roles, compare within rounding
    CT = np.zeros((n, m), order="F")
    oracle.gemm("T", "T", n, m, k, 1.5, B, k, A, m, 0.0, CT, n)
    np.testing.assert_allclose(CT.T, base, rtol=0, atol=4 * U64 * np.abs(base).max())


def test_linearity_in_alpha():
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.uniform(-1, 1, (8, 30)))
    B = np.asfortranarray(rng.uniform(-1, 1, (30, 7)))
    C1 = np.zeros((8, 7), order="F")
    C2 = np.zeros((8, 7), order="F")
    oracle.gemm("N", "N", 8, 7, 30, 0.75, A, 8, B, 30, 0.0, C1, 8)
    oracle.gemm("N", "N", 8, 7, 30, 1.5, A, 8, B, 30, 0.0, C2, 8)
    np.testing.assert_array_equal(2.0 * C1, C2)  # scaling by 2 is exact


def test_reference_agrees_on_extra_shapes():
    """Extra shapes straight against the reference when oracle/_ref is built."""
    if oracle.ref("d") is None:
        pytest.skip("reference not built (oracle/_ref)")
    rng = np.random.default_rng(13)
    for dt, dts in ((np.float64, "f64"), (np.float32, "f32")):
        for (m, n, k) in [(1, 1, 289), (8, 96, 288), (100, 3, 5), (70, 193, 337)]:
            for ta in "NT":
                for tb in "NT":
                    ra, ca = stored_shape(ta, m, k)
                    rb, cb = stored_shape(tb, k, n)
                    A = np.asfortranarray(rng.uniform(-1, 1, (ra, ca)).astype(dt))
                    B = np.asfortranarray(rng.uniform(-1, 1, (rb, cb)).astype(dt))
                    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
                    C1, C2 = C.copy(order="F"), C.copy(order="F")
                    oracle.gemm(ta, tb, m, n, k, 1.5, A, ra, B, rb, -0.5, C1, m)
                    assert oracle.ref_gemm(ta, tb, m, n, k, 1.5, A, ra, B, rb, -0.5, C2, m)
                    bound = abs_bound(A, B, C, ta, tb, m, n, k, 1.5, -0.5)
                    assert max_comp(C1, C2, bound) <= _tol(dts, k)


# --------------------------------------------------------------- generator
def test_generator_values():
    L = oracle.lib()
    xs = np.array([L.oracle_uniform_f64(0, 0, i) for i in range(20000)])
    assert xs.min() >= -1.0 and xs.max() < 1.0
    assert abs(xs.mean()) < 0.02 and abs(xs.var() - 1 / 3) < 0.01
    # exactly (x>>11) * 2^-53 * 2 - 1: every value is a multiple of 2^-52
    assert np.all(np.ldexp(xs + 1.0, 52) == np.floor(np.ldexp(xs + 1.0, 52)))
    fs = np.array([L.oracle_uniform_f32(0, 0, i) for i in range(20000)], dtype=np.float64)
    assert np.all(np.ldexp(fs + 1.0, 23) == np.floor(np.ldexp(fs + 1.0, 23)))
    # the FP32 draw uses the top 24 bits of the same counter as the FP64 draw
    for i in range(100):
        x = L.oracle_counter_u64(5, 1, i)
        assert L.oracle_uniform_f64(5, 1, i) == (x >> 11) * 2.0 ** -53 * 2 - 1
        assert L.oracle_uniform_f32(5, 1, i) == np.float32((x >> 40) * 2.0 ** -24 * 2 - 1)
    # streams differ by seed and by matrix id
    assert L.oracle_counter_u64(0, 0, 0) != L.oracle_counter_u64(0, 1, 0)
    assert L.oracle_counter_u64(0, 0, 0) != L.oracle_counter_u64(1, 0, 0)
    buf = oracle.fill(3, 2, 5, 9, 0)
    assert np.all(np.isnan(buf[3:])) and not np.any(np.isnan(buf[:3]))
    assert buf[1, 1] == L.oracle_uniform_f64(9, 0, 1 + 1 * 3)


# --------------------------------------------------------------- Ozaki-I
def test_ozaki_exponent_and_digits():
    L = oracle.lib()
    x = np.array([0.75, -3.0, 1e-3], dtype=np.float64)
    e = L.oracle_ozaki_exponent(x.ctypes.data, 3, 1)
    assert e == 2  # max 3.0 in [2^1, 2^2)
    z = np.zeros(4)
    assert L.oracle_ozaki_exponent(z.ctypes.data, 4, 1) == 0
    z[2] = np.inf
    assert L.oracle_ozaki_exponent(z.ctypes.data, 4, 1) == 0x7FFFFFFF
    rng = np.random.default_rng(2)
    d = (ctypes.c_int8 * 8)()
    for v in rng.uniform(-1, 1, 200) * 2.0 ** rng.integers(-40, 40, 200):
        ev = math.frexp(abs(v))[1] + int(rng.integers(0, 5))
        for s in (3, 5, 8):
            L.oracle_ozaki_digits(float(v), ev, s, d)
            digs = np.array(d[:s], dtype=np.int64)
            assert np.all(np.abs(digs) <= 127)
            assert np.all(digs * np.sign(v) >= 0)  # truncation keeps the sign
            rec = sum(float(digs[t]) * 2.0 ** (-7 * (t + 1)) for t in range(s))
            assert abs(math.ldexp(v, -ev) - rec) < 2.0 ** (-7 * s)


@pytest.mark.parametrize("s", [3, 4, 5, 6, 7, 8])
def test_ozaki_exact_on_small_integers(s):
    """Entries that fit in the first digit make every slice product exact, so
    the emulated GEMM equals the exact integer GEMM for every slice count."""
    rng = np.random.default_rng(s)
    for ta in "NT":
        for tb in "NT":
            m, n, k = 7, 5, 37
            ra, ca = stored_shape(ta, m, k)
            rb, cb = stored_shape(tb, k, n)
            A = np.asfortranarray(rng.integers(-3, 4, (ra, ca)).astype(np.float64))
            B = np.asfortranarray(rng.integers(-3, 4, (rb, cb)).astype(np.float64))
            C = np.asfortranarray(rng.integers(-3, 4, (m, n)).astype(np.float64))
            exact = 2.0 * (op(A, ta, m, k) @ op(B, tb, k, n)) - 1.0 * C
            oracle.ozaki_gemm(ta, tb, m, n, k, 2.0, A, ra, B, rb, -1.0, C, m, s)
            np.testing.assert_array_equal(C, exact)


def test_ozaki_error_shrinks_with_slices():
    rng = np.random.default_rng(4)
    m, n, k = 40, 30, 300
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    R = np.zeros((m, n), order="F")
    oracle.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m)
    errs = []
    for s in range(3, 9):
        C = np.zeros((m, n), order="F")
        oracle.ozaki_gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, s)
        e = rel_fro(C, R)
        errs.append(e)
        # truncation error: k terms, each off by < 2*2^-7s relative to the maxima
        assert e <= 4 * k * 2.0 ** (-7 * s) * 1.0 / (np.linalg.norm(R) / np.sqrt(m * n)) + 4 * U64
    assert all(b < a for a, b in zip(errs, errs[1:-1]))
    assert errs[-1] < 1e-15


def test_ozaki_nonfinite_rows_and_kchunk():
    L = oracle.lib()
    assert L.oracle_ozaki_kchunk(8) == 16640  # 8*16640*127^2 < 2^31 <= 8*16768*127^2
    assert 8 * 16640 * 127 ** 2 < 2 ** 31 <= 8 * (16640 + 128) * 127 ** 2
    A = np.ones((3, 4), order="F")
    A[1, 2] = np.inf
    B = np.ones((4, 2), order="F")
    C = np.zeros((3, 2), order="F")
    oracle.ozaki_gemm("N", "N", 3, 2, 4, 1.0, A, 3, B, 4, 0.0, C, 3, 4)
    assert np.all(np.isnan(C[1])) and np.all(C[[0, 2]] == 4.0)
    # K split across chunks still reaches FP64-level error
    k = 16640 + 300
    rng = np.random.default_rng(8)
    A = np.asfortranarray(rng.uniform(-1, 1, (2, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, 3)))
    R = np.zeros((2, 3), order="F")
    oracle.gemm("N", "N", 2, 3, k, 1.0, A, 2, B, k, 0.0, R, 2)
    C = np.zeros((2, 3), order="F")
    oracle.ozaki_gemm("N", "N", 2, 3, k, 1.0, A, 2, B, k, 0.0, C, 2, 8)
    assert rel_fro(C, R) < 1e-14


# --------------------------------------------------------------- Ozaki-II
def test_ozaki2_moduli_are_the_greedy_coprime_list():
    mods = [oracle.ozaki2_modulus(j) for j in range(22)]
    assert oracle.ozaki2_modulus(22) == -1 and mods[0] == 256
    assert all(a > b for a, b in zip(mods, mods[1:]))
    for i, a in enumerate(mods):
        assert all(math.gcd(a, b) == 1 for b in mods[:i])
    # greedy: every skipped integer in (mods[-1], 256] shares a factor with a
    # larger modulus
    for c in range(mods[-1] + 1, 257):
        if c not in mods:
            assert any(math.gcd(c, b) > 1 for b in mods if b > c), c


def _oz2_vectors(k, rng):
    """Vectors of k entries spanning the cases of the bit-count rule: uniform,
    heavy-tailed (phi = 4), one spike over tiny entries, every entry at the
    top of its binade (the largest bound sum), subnormal only, mixed zeros."""
    top = np.full(k, np.nextafter(1.0, 0.0)) * rng.choice([-1.0, 1.0], k) * 2.0 ** 7
    spike = rng.uniform(-1e-12, 1e-12, k)
    spike[k // 2] = 3.0
    zeros = rng.uniform(-1, 1, k)
    zeros[::3] = 0.0
    return [rng.uniform(-0.5, 0.5, k),
            (rng.uniform(size=k) - 0.5) * np.exp(4.0 * rng.standard_normal(k)),
            spike, top, rng.uniform(-1, 1, k) * 2.0 ** -1060, zeros]


def _oz2_scaled_vector(x, P):
    e = math.frexp(float(np.max(np.abs(x))))[1] if np.any(x) else 0
    return [oracle.ozaki2_scaled(v, e, P) for v in x]


@pytest.mark.parametrize("k", [1, 37, 300, 1024, 1537])
def test_ozaki2_vector_bits_keep_every_dot_product_inside_the_crt_range(k):
    """Safety, exactly in integers: every vector's scaled integers have
    ||a||^2 <= Mr/2, Mr = M - floor(M/2^10) - 1, so by Cauchy-Schwarz any dot
    product of two vectors' integers is at most Mr/2 < (1 - 2^-10) M/2: the
    reconstruction's X/M stays at least 2^-11 from a half-integer.
    Near-maximality: for the vector with every entry at the top of its
    binade the same integers at P + 1 bits break that norm bound, and no
    vector gets fewer bits than one bit count for k full-size terms with
    L = floor(log2 M), floor((L - 1 - bitlen(k)) / 2), would give."""
    rng = np.random.default_rng(20 + k)
    for nmod in (8, 11, 14, 16, 18, 20, 22):
        M = math.prod(oracle.ozaki2_modulus(j) for j in range(nmod))
        L = M.bit_length() - 1
        Mr = M - M // 1024 - 1
        floor_rule = min((L - 1 - k.bit_length()) // 2, 77)
        for idx, x in enumerate(_oz2_vectors(k, rng)):
            P = oracle.ozaki2_vector_bits(x, nmod)
            assert 10 <= P <= 77, (nmod, idx, P)
            a = _oz2_scaled_vector(x, P)
            assert 2 * sum(v * v for v in a) <= Mr, (nmod, k, idx, P)
            assert P >= floor_rule, (nmod, k, idx, P, floor_rule)
            if idx == 3:  # every entry at the top of its binade
                assert P == oracle.ozaki2_bits(nmod, k)
                if P < 77:
                    a1 = _oz2_scaled_vector(x, P + 1)
                    assert 2 * sum(v * v for v in a1) > Mr, (nmod, k, P)
    assert oracle.ozaki2_bits(7, k) == -1 and oracle.ozaki2_bits(23, k) == -1
    assert oracle.ozaki2_vector_bits(np.ones(4), 7) == -1


def test_ozaki2_vector_bits_do_not_depend_on_the_entry_order():
    """The bound sum is an exact integer, so any summation order (the GPU
    reduces in another one) gives the same bit count; a heavy-tailed vector
    keeps about log2(sqrt(k)) more bits than the all-top vector."""
    rng = np.random.default_rng(24)
    k = 4096
    x = (rng.uniform(size=k) - 0.5) * np.exp(4.0 * rng.standard_normal(k))
    for nmod in (12, 18):
        P = oracle.ozaki2_vector_bits(x, nmod)
        for _ in range(5):
            assert oracle.ozaki2_vector_bits(rng.permutation(x), nmod) == P
        assert oracle.ozaki2_vector_bits(-x, nmod) == P
        assert P >= oracle.ozaki2_bits(nmod, k) + 4


def test_ozaki2_scaled_integers_are_rounded_to_nearest_even():
    from fractions import Fraction
    rng = np.random.default_rng(21)
    xs = list(rng.standard_normal(200) * np.exp(4 * rng.standard_normal(200)))
    xs += [0.0, -0.0, 5e-324, -2.5e-310, 1.5, -2.5, 0.5 + 2.0 ** -52, 2.0 ** 60]
    for x in xs:
        if x == 0.0:
            e = 0
        else:
            e = math.frexp(abs(x))[1] + int(rng.integers(0, 30))  # row max >= |x|
        for P in (24, 53, 60, 77):
            want = round(Fraction(x) * Fraction(2) ** (P - e))  # ties to even
            assert oracle.ozaki2_scaled(x, e, P) == want, (x, e, P)
            assert abs(want) <= 2 ** P


def test_ozaki2_gemm_rounds_tie_products_to_even():
    """k = 1 products of odd integers: 54-bit odd products are exact ties,
    (2^54 - 1) rounds up across a binade; with 16 or more moduli the split
    keeps the integers exact, so the result is Python's correctly rounded
    int -> float conversion of the exact product."""
    a = [1, 3, 5, 7, -3, 9, 2**26 + 1, -(2**25 + 3)]
    b = [(2**53 + 1) // 3, (2**54 - 1) // 3, 2**52 + 1, 2**53 - 1, -(2**52 + 3),
         4503599627370497, 1, 0]
    A = np.asfortranarray(np.array(a, dtype=np.float64).reshape(8, 1))
    B = np.asfortranarray(np.array(b, dtype=np.float64).reshape(1, 8))
    want = np.array([[float(x * y) for y in b] for x in a])
    assert sum(float(x * y) != x * y for x in a for y in b) >= 30  # most need rounding
    for nmod in (16, 19, 22):
        R = np.zeros((8, 8), order="F")
        oracle.ozaki2_gemm("N", "N", 8, 8, 1, 1.0, A, 8, B, 1, 0.0, R, 8, nmod)
        np.testing.assert_array_equal(R, want)


def test_ozaki2_round_is_rne_then_one_ldexp():
    from fractions import Fraction
    rng = np.random.default_rng(22)
    vals = [0, 1, -1, 2 ** 53 + 1, 2 ** 53 + 3, -(2 ** 54 + 2), 2 ** 200 + 2 ** 147,
            2 ** 200 + 2 ** 147 + 1, 3 * 2 ** 147, -(2 ** 254)]
    vals += [int(rng.integers(-2 ** 62, 2 ** 62)) << int(rng.integers(0, 190)) for _ in range(300)]
    vals += [int(rng.integers(-2 ** 62, 2 ** 62)) * (2 ** 64) + int(rng.integers(0, 2 ** 62))
             for _ in range(100)]
    for v in vals:
        for scale in (0, -140, 37):
            # normal range: one correctly rounded conversion
            assert oracle.ozaki2_round(v, scale) == float(Fraction(v) * Fraction(2) ** scale), v
    # results below the normal range: RNE to 53 bits first, then ldexp
    v = (2 ** 53 - 1) * 2 ** 60 + 2 ** 59 + 1
    first = float(Fraction(v))
    assert oracle.ozaki2_round(v, -1130) == math.ldexp(first, -1130)


def test_ozaki2_is_the_correctly_rounded_product_when_inputs_fit():
    """Entries j / 2^20 with |j| < 2^20 become exact integers for every
    modulus count (P - e >= 20), so the result is the exact dot product
    rounded once: math.fsum of the (exact) products."""
    rng = np.random.default_rng(23)
    for nmod in (8, 12, 16, 22):
        for ta in "NT":
            for tb in "NT":
                m, n, k = 6, 5, 301
                ra, ca = stored_shape(ta, m, k)
                rb, cb = stored_shape(tb, k, n)
                A = np.asfortranarray(rng.integers(-2 ** 20 + 1, 2 ** 20, (ra, ca)) / 2.0 ** 20)
                B = np.asfortranarray(rng.integers(-2 ** 20 + 1, 2 ** 20, (rb, cb)) / 2.0 ** 20)
                C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
                oracle.ozaki2_gemm(ta, tb, m, n, k, 1.0, A, ra, B, rb, 0.0, C, m, nmod)
                oa, ob = op(A, ta, m, k), op(B, tb, k, n)
                for i in range(m):
                    for j in range(n):
                        assert C[i, j] == math.fsum(oa[i] * ob[:, j]), (nmod, ta, tb, i, j)


def test_ozaki2_semantics_error_and_chunks():
    rng = np.random.default_rng(24)
    # beta != 1 first (beta == 0 keeps NaN), alpha == 0 reads neither operand
    A = np.ones((3, 4), order="F")
    B = np.ones((4, 2), order="F")
    C = np.full((3, 2), np.nan, order="F")
    oracle.ozaki2_gemm("N", "N", 3, 2, 4, 1.0, A, 3, B, 4, 0.0, C, 3, 16)
    assert np.all(np.isnan(C))
    C = np.full((3, 2), 2.0, order="F")
    oracle.ozaki2_gemm("N", "N", 3, 2, 4, 0.0, A, 3, B, 4, 3.0, C, 3, 16)
    assert np.all(C == 6.0)
    A[1, 2] = np.inf
    C = np.zeros((3, 2), order="F")
    oracle.ozaki2_gemm("N", "N", 3, 2, 4, 1.0, A, 3, B, 4, 0.0, C, 3, 16)
    assert np.all(np.isnan(C[1])) and np.all(C[[0, 2]] == 4.0)
    # error against Dot2 falls with the modulus count on wide exponent ranges
    m, n, k = 20, 16, 400
    A = np.asfortranarray(rng.uniform(-0.5, 0.5, (m, k)) * np.exp(4 * rng.standard_normal((m, k))))
    B = np.asfortranarray(rng.uniform(-0.5, 0.5, (k, n)) * np.exp(4 * rng.standard_normal((k, n))))
    R = np.zeros((m, n), order="F")
    oracle.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m)
    errs = []
    for nmod in (8, 10, 12, 14, 16, 18, 20, 22):
        C = np.zeros((m, n), order="F")
        oracle.ozaki2_gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, nmod)
        errs.append(rel_fro(C, R))
    assert all(b <= a for a, b in zip(errs, errs[1:])), errs
    assert errs[0] > 1e-9 and errs[-1] < 1e-15, errs
    # K beyond one chunk: chunk sums rounded and added in order onto beta*C
    from fractions import Fraction
    k = 65536 + 100
    a = rng.integers(-2 ** 10, 2 ** 10, (1, k)) / 2.0 ** 10
    b = rng.integers(-2 ** 10, 2 ** 10, (k, 1)) / 2.0 ** 10
    A, B = np.asfortranarray(a), np.asfortranarray(b)
    C = np.array([[0.25]], order="F")
    oracle.ozaki2_gemm("N", "N", 1, 1, k, 1.0, A, 1, B, k, -2.0, C, 1, 22)
    s0 = float(sum(Fraction(x) * Fraction(y) for x, y in zip(a[0, :65536], b[:65536, 0])))
    s1 = float(sum(Fraction(x) * Fraction(y) for x, y in zip(a[0, 65536:], b[65536:, 0])))
    assert C[0, 0] == (s0 + (-0.5)) + s1


@pytest.mark.parametrize("ta,tb,dist,param", [("N", "N", 1, 4.0), ("T", "N", 0, 0.0),
                                               ("N", "T", 1, 1.0), ("T", "T", 1, 0.0)])
def test_dot2_pairs_equal_the_generator_sampled_entries(ta, tb, dist, param):
    """Dot2 on explicit rows/columns gives the generator-based sampled
    entries bit for bit when fed the same generator inputs."""
    m, n, k, seed = 37, 29, 515, 8
    alpha, beta = 1.5, -0.5
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    A = oracle.fill(ra, ca, ra, seed, 0, np.float64, dist, param)
    B = oracle.fill(rb, cb, rb, seed, 1, np.float64, dist, param)
    C = oracle.fill(m, n, m, seed, 2, np.float64, dist, param)
    rng = np.random.default_rng(3)
    is_, js = rng.integers(0, m, 200), rng.integers(0, n, 200)
    want, ab_w = oracle.sampled(np.float64, ta, tb, m, n, k, alpha, beta, seed, is_, js,
                                dist, param)
    rows, cols = op(A, ta, m, k), op(B, tb, k, n).T
    got, ab_g = oracle.dot2_pairs(rows, cols, is_, js, alpha, beta * C[is_, js])
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(ab_g, ab_w)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_emulated_pairs_equal_the_whole_matrix_restatements(ta, tb):
    """The sampled-entry restatements (used on products too large to restate
    whole) give the whole-matrix restatements' entries bit for bit: Ozaki-I
    at every slice count, Ozaki-II at several modulus counts, on heavy-tailed
    entries, with alpha/beta, repeated pairs and a non-finite row."""
    rng = np.random.default_rng(41)
    m, n, k = 23, 19, 700
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    A = np.asfortranarray(rng.uniform(-1, 1, (ra, ca)) * np.exp(4 * rng.standard_normal((ra, ca))))
    B = np.asfortranarray(rng.uniform(-1, 1, (rb, cb)) * np.exp(2 * rng.standard_normal((rb, cb))))
    opA = op(A, ta, m, k)
    opA[5, 17] = np.inf  # row 5 of op(A) is non-finite
    C0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    rows, cols = np.ascontiguousarray(op(A, ta, m, k)), np.ascontiguousarray(op(B, tb, k, n).T)
    ri = np.concatenate([rng.integers(0, m, 60), [5, 5, 0]])
    cj = np.concatenate([rng.integers(0, n, 60), [0, 3, 0]])
    for fn, whole, counts in ((oracle.ozaki_pairs, oracle.ozaki_gemm, range(3, 9)),
                              (oracle.ozaki2_pairs, oracle.ozaki2_gemm, (8, 13, 16, 18, 22))):
        for cnt in counts:
            R = C0.copy(order="F")
            whole(ta, tb, m, n, k, 1.25, A, ra, B, rb, -0.5, R, m, cnt)
            got = fn(rows, cols, ri, cj, 1.25, -0.5 * C0[ri, cj], cnt)
            np.testing.assert_array_equal(got, R[ri, cj])
    assert np.all(np.isnan(got[ri == 5]))
    with pytest.raises(ValueError):  # beyond one K chunk
        oracle.ozaki2_pairs(np.ones((1, 65537)), np.ones((1, 65537)), [0], [0], 1.0, 0.0, 16)


# --------------------------------------------------------------- BF16x3
def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(9)
    x = np.concatenate([
        rng.standard_normal(5000).astype(np.float32) * np.float32(2.0) ** rng.integers(-30, 30, 5000).astype(np.float32),
        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 65504.0, 3e38, 1e-40,
                  0.0, -0.0, np.inf, -np.inf], dtype=np.float32)])
    t = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().astype(np.uint16)
    L = oracle.lib()
    mine = np.array([L.oracle_bf16_rne(float(v)) for v in x], dtype=np.uint16)
    np.testing.assert_array_equal(mine, t)


def test_bf16_split_and_bf16x3_gemm():
    L = oracle.lib()
    rng = np.random.default_rng(10)
    hi, lo = ctypes.c_uint16(), ctypes.c_uint16()

    def f(h):
        return np.frombuffer(np.array([h << 16], dtype=np.uint32).tobytes(), dtype=np.float32)[0]
    for v in rng.uniform(-1, 1, 500).astype(np.float32):
        L.oracle_bf16_split(float(v), ctypes.byref(hi), ctypes.byref(lo))
        r = float(v) - float(f(hi.value)) - float(f(lo.value))
        assert abs(r) <= 2.0 ** -17 * abs(float(v)) + 1e-45
    # inputs already in BF16: lo = 0 and the product is exact (Dot2 == bf16x3)
    m, n, k = 9, 11, 40
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(np.float32))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(np.float32))
    Ab = np.vectorize(lambda v: f(L.oracle_bf16_rne(float(v))))(A).astype(np.float32)
    Bb = np.vectorize(lambda v: f(L.oracle_bf16_rne(float(v))))(B).astype(np.float32)
    Ab, Bb = np.asfortranarray(Ab), np.asfortranarray(Bb)
    C1 = np.zeros((m, n), dtype=np.float32, order="F")
    C2 = np.zeros((m, n), dtype=np.float32, order="F")
    oracle.bf16x3_gemm("N", "N", m, n, k, 1.0, Ab, m, Bb, k, 0.0, C1, m)
    oracle.gemm("N", "N", m, n, k, 1.0, Ab, m, Bb, k, 0.0, C2, m)
    np.testing.assert_array_equal(C1, C2)
    # general FP32 inputs: error ~2^-16 relative, well above FP32 rounding
    C3 = np.zeros((m, n), dtype=np.float32, order="F")
    oracle.bf16x3_gemm("T", "N", m, n, k, 1.0, np.asfortranarray(A.T), k, B, k, 0.0, C3, m)
    R = op(A, "N", m, k).astype(np.float64) @ B.astype(np.float64)
    bound = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
    assert max_comp(C3, R, bound) <= 3 * 2.0 ** -16


# --------------------------------------------------------------- complex GEMM
CTRANS = ["N", "T", "C"]


def _cplx(rng, shape, dtype=np.complex128):
    return np.asfortranarray((rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
                             .astype(dtype))


def _cop(X, t):
    """op(X) of a stored complex matrix: 'N', 'T', or 'C' (conjugate transpose)."""
    return X if t in "Nn" else (X.conj().T if t in "Cc" else X.T)


@pytest.mark.parametrize("ta", CTRANS)
@pytest.mark.parametrize("tb", CTRANS)
def test_complex_gemm_of_real_matrices_is_dgemm(ta, tb):
    """With zero imaginary parts (alpha = 1, beta = 0) the real part is
    oracle_dgemm's result bit for bit (the same Dot2 sums, the extra terms
    are exact zeros) and the imaginary part is zero; conjugation changes
    nothing."""
    rng = np.random.default_rng(600)
    m, n, k = 7, 5, 33
    ra, ca = stored_shape(ta, m, k)
    rb, cb = stored_shape(tb, k, n)
    Ar = np.asfortranarray(rng.uniform(-1, 1, (ra, ca)))
    Br = np.asfortranarray(rng.uniform(-1, 1, (rb, cb)))
    Cz = np.zeros((m, n), dtype=np.complex128, order="F")
    oracle.complex_gemm(ta, tb, m, n, k, 1 + 0j, np.asfortranarray(Ar + 0j), ra,
                        np.asfortranarray(Br + 0j), rb, 0j, Cz, m)
    R = np.zeros((m, n), order="F")
    oracle.gemm("N" if ta == "N" else "T", "N" if tb == "N" else "T", m, n, k, 1.0, Ar, ra, Br,
                rb, 0.0, R, m)
    np.testing.assert_array_equal(Cz.real, R)
    assert np.all(Cz.imag == 0.0)


@pytest.mark.parametrize("t", ["C", "c"])
def test_complex_gemm_conjugate_transpose_is_transpose_of_the_conjugate(t):
    """'C' on A gives exactly 'T' on conj(A) as stored (the imaginary plane
    negated); likewise for B; 't' and any other character act as 'T'."""
    rng = np.random.default_rng(601)
    m, n, k = 6, 4, 9
    A, B, C = _cplx(rng, (k, m)), _cplx(rng, (n, k)), _cplx(rng, (m, n))
    for ta, tb, A2, B2, ta2, tb2 in ((t, "T", A, B, "T", "T"), ("T", t, A, B, "T", "T")):
        R1, R2 = C.copy(order="F"), C.copy(order="F")
        oracle.complex_gemm(ta, tb, m, n, k, 1.5 - 0.5j, A, k, B, n, 0.25j, R1, m)
        Ac = np.asfortranarray(A.conj()) if ta in "Cc" else A
        Bc = np.asfortranarray(B.conj()) if tb in "Cc" else B
        oracle.complex_gemm(ta2, tb2, m, n, k, 1.5 - 0.5j, Ac, k, Bc, n, 0.25j, R2, m)
        np.testing.assert_array_equal(R1, R2)
    Bt = np.asfortranarray(B.T)
    for other in ("t", "X"):
        R1, R2 = C.copy(order="F"), C.copy(order="F")
        oracle.complex_gemm(other, "N", m, n, k, 1 + 0j, A, k, Bt, k, 0j, R1, m)
        oracle.complex_gemm("T", "N", m, n, k, 1 + 0j, A, k, Bt, k, 0j, R2, m)
        np.testing.assert_array_equal(R1, R2)


@pytest.mark.parametrize("ta", CTRANS)
@pytest.mark.parametrize("tb", CTRANS)
def test_complex_gemm_against_exact_rationals(ta, tb):
    """Tiny shapes against the exact product (Python Fractions): the product
    is rounded once per part, then alpha P + beta C costs a few roundings."""
    from fractions import Fraction
    rng = np.random.default_rng(602)
    m, n, k = 3, 4, 5
    ra, ca = stored_shape(ta, m, k)
    rb, cb = stored_shape(tb, k, n)
    A, B, C = _cplx(rng, (ra, ca)), _cplx(rng, (rb, cb)), _cplx(rng, (m, n))
    alpha, beta = 0.75 - 1.25j, -0.5 + 0.125j
    R = C.copy(order="F")
    oracle.complex_gemm(ta, tb, m, n, k, alpha, A, ra, B, rb, beta, R, m)
    oa, ob = _cop(A, ta), _cop(B, tb)
    F = Fraction
    for i in range(m):
        for j in range(n):
            pr = sum(F(oa[i, p].real) * F(ob[p, j].real) - F(oa[i, p].imag) * F(ob[p, j].imag)
                     for p in range(k))
            pi = sum(F(oa[i, p].real) * F(ob[p, j].imag) + F(oa[i, p].imag) * F(ob[p, j].real)
                     for p in range(k))
            cr, ci = F(C[i, j].real), F(C[i, j].imag)
            wr = F(beta.real) * cr - F(beta.imag) * ci + F(alpha.real) * pr - F(alpha.imag) * pi
            wi = F(beta.real) * ci + F(beta.imag) * cr + F(alpha.real) * pi + F(alpha.imag) * pr
            scale = abs(alpha) * float(sum(abs(oa[i, p]) * abs(ob[p, j]) for p in range(k))) \
                + abs(beta) * abs(C[i, j])
            assert abs(R[i, j].real - float(wr)) <= 8 * U64 * scale, (i, j)
            assert abs(R[i, j].imag - float(wi)) <= 8 * U64 * scale, (i, j)


def test_complex_gemm_semantics():
    """beta != 1 scales C first, also beta = 0 (NaN stays NaN); alpha = 0 or
    k = 0 leaves beta C; beta = 1 with alpha = 0 leaves C untouched; the
    single-precision form rounds the same Dot2 sums once to float."""
    rng = np.random.default_rng(603)
    m, n, k = 4, 3, 6
    A, B, C = _cplx(rng, (m, k)), _cplx(rng, (k, n)), _cplx(rng, (m, n))
    C[1, 2] = complex(np.nan, 0.0)
    R = C.copy(order="F")
    oracle.complex_gemm("N", "N", m, n, k, 1 + 0j, A, m, B, k, 0j, R, m)
    assert np.isnan(R[1, 2].real) and np.isnan(R[1, 2].imag)
    mask = np.ones((m, n), bool)
    mask[1, 2] = False
    np.testing.assert_allclose(R[mask], (A @ B)[mask], rtol=0, atol=1e-14)
    for alpha, kk, beta in ((0j, k, 2 - 1j), (1 + 1j, 0, 0.5j), (0j, k, 1 + 0j)):
        R = C.copy(order="F")
        oracle.complex_gemm("N", "N", m, n, kk, alpha, A, m, B, k, beta, R, m)
        want = C if beta == 1 else (beta.real * C.real - beta.imag * C.imag) + 1j * (
            beta.real * C.imag + beta.imag * C.real)
        np.testing.assert_array_equal(R, want)
    A32, B32 = A.astype(np.complex64), B.astype(np.complex64)
    C32 = np.asfortranarray(_cplx(rng, (m, n)).astype(np.complex64))
    R32 = C32.copy(order="F")
    oracle.complex_gemm("N", "C", m, n, k, 1 - 2j, A32, m, np.asfortranarray(B32.T), n, 0.5 + 0j,
                        R32, m)
    R64 = C32.astype(np.complex128)
    oracle.complex_gemm("N", "C", m, n, k, 1 - 2j, A32.astype(np.complex128), m,
                        np.asfortranarray(B32.T.astype(np.complex128)), n, 0.5 + 0j,
                        np.asfortranarray(R64), m)
    np.testing.assert_allclose(R32, R64, rtol=0, atol=4e-6 * np.max(np.abs(R64)))


# --------------------------------------------------------------- triangular
TRI_CASES = [(s, u, t, d) for s in "LR" for u in "UL" for t in "NT" for d in "NU"]


def _tri_dense(A, uplo, diag):
    T = np.triu(A) if uplo == "U" else np.tril(A)
    if diag == "U":
        np.fill_diagonal(T, 1.0)
    return T


@pytest.mark.parametrize("side,uplo,trans,diag", TRI_CASES)
def test_triangular_ops_never_read_the_other_triangle_or_a_unit_diagonal(side, uplo, trans,
                                                                          diag):
    """NaN in the unused triangle (and on a unit diagonal) changes nothing:
    the results equal those with zeros there, bit for bit."""
    rng = np.random.default_rng(800)
    m, n = 9, 6
    k = m if side == "L" else n
    A = np.asfortranarray(rng.uniform(-1, 1, (k, k)) / k + np.diag(2 + rng.uniform(0, 1, k)))
    clean = np.asfortranarray(_tri_dense(A, uplo, "N"))
    dirty = clean.copy(order="F")
    other = np.tril(np.ones((k, k)), -1) if uplo == "U" else np.triu(np.ones((k, k)), 1)
    dirty[other == 1] = np.nan
    if diag == "U":
        np.fill_diagonal(dirty, np.nan)
    B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    for solve in (0, 1):
        X1, X2 = B.copy(order="F"), B.copy(order="F")
        oracle.tri(solve, side, uplo, trans, diag, m, n, 1.25, clean, k, X1, m)
        oracle.tri(solve, side, uplo, trans, diag, m, n, 1.25, dirty, k, X2, m)
        np.testing.assert_array_equal(X1, X2)


@pytest.mark.parametrize("side,uplo,trans,diag", TRI_CASES)
def test_triangular_ops_on_exact_integer_systems(side, uplo, trans, diag):
    """Integer triangles with unit or power-of-two diagonals and integer X:
    B = op(A) X (or X op(A)) is exact in FP64, TRMM must return it exactly
    and TRSM must recover X exactly (every substitution step is exact)."""
    rng = np.random.default_rng(801)
    m, n = 8, 5
    k = m if side == "L" else n
    A = np.asfortranarray(rng.integers(-3, 4, (k, k)).astype(np.float64))
    np.fill_diagonal(A, 2.0 ** rng.integers(0, 3, k) * rng.choice([-1.0, 1.0], k))
    T = _tri_dense(A, uplo, diag)
    opT = T.T if trans == "T" else T
    X = np.asfortranarray(rng.integers(-5, 6, (m, n)).astype(np.float64))
    B = np.asfortranarray(opT @ X if side == "L" else X @ opT)
    got = X.copy(order="F")
    oracle.tri(0, side, uplo, trans, diag, m, n, 1.0, A, k, got, m)
    np.testing.assert_array_equal(got, B)
    sol = B.copy(order="F")
    oracle.tri(1, side, uplo, trans, diag, m, n, 1.0, A, k, sol, m)
    np.testing.assert_array_equal(sol, X)


@pytest.mark.parametrize("side,uplo,trans,diag", TRI_CASES[::3])
def test_triangular_solve_against_exact_rationals(side, uplo, trans, diag):
    """A random well-conditioned system solved exactly with Fractions: the
    oracle's X is within a few units in the last place of it."""
    from fractions import Fraction as F
    rng = np.random.default_rng(802)
    m, n = 5, 4
    k = m if side == "L" else n
    A = rng.uniform(-1, 1, (k, k)) / k + np.diag(2 + rng.uniform(0, 1, k))
    A = np.asfortranarray(A)
    T = _tri_dense(A, uplo, diag)
    opT = [[F(v) for v in row] for row in (T.T if trans == "T" else T)]
    B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    sol = B.copy(order="F")
    oracle.tri(1, side, uplo, trans, diag, m, n, 0.75, A, k, sol, m)
    # exact solve by Gaussian elimination on the (triangular) system
    def solve(M, rhs):
        N = len(M)
        aug = [list(M[i]) + [rhs[i]] for i in range(N)]
        for c in range(N):
            p = next(r for r in range(c, N) if aug[r][c] != 0)
            aug[c], aug[p] = aug[p], aug[c]
            for r in range(N):
                if r != c and aug[r][c] != 0:
                    f = aug[r][c] / aug[c][c]
                    aug[r] = [x - f * y for x, y in zip(aug[r], aug[c])]
        return [aug[i][N] / aug[i][i] for i in range(N)]
    if side == "L":
        for j in range(n):
            x = solve(opT, [F(0.75) * F(B[i, j]) for i in range(m)])
            for i in range(m):
                assert abs(sol[i, j] - float(x[i])) <= 16 * U64 * max(1.0, abs(float(x[i]))), (i, j)
    else:
        opTt = [list(col) for col in zip(*opT)]  # x op(A) = b  <=>  op(A)^T x^T = b^T
        for i in range(m):
            x = solve(opTt, [F(0.75) * F(B[i, j]) for j in range(n)])
            for j in range(n):
                assert abs(sol[i, j] - float(x[j])) <= 16 * U64 * max(1.0, abs(float(x[j]))), (i, j)


def test_triangular_alpha_zero_and_single_precision():
    """alpha = 0 writes zeros without reading A or B (NaN in both); the
    float forms equal the double computation rounded to float within the
    float rounding of the sums."""
    rng = np.random.default_rng(803)
    m, n = 6, 4
    A = np.full((m, m), np.nan, order="F")
    B = np.full((m, n), np.nan, order="F")
    for solve in (0, 1):
        X = B.copy(order="F")
        oracle.tri(solve, "L", "U", "N", "N", m, n, 0.0, A, m, X, m)
        assert np.all(X == 0.0)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, m)) / m + np.diag(2 + rng.uniform(0, 1, m)))
    B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    for solve in (0, 1):
        X32 = B.astype(np.float32, order="F")
        oracle.tri(solve, "L", "L", "T", "N", m, n, 1.5, A.astype(np.float32, order="F"), m,
                   X32, m)
        X64 = B.astype(np.float32).astype(np.float64, order="F")
        oracle.tri(solve, "L", "L", "T", "N", m, n, 1.5,
                   A.astype(np.float32).astype(np.float64, order="F"), m, X64, m)
        np.testing.assert_allclose(X32, X64, rtol=4 * U32 * m, atol=0)
