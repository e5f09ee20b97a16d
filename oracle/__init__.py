"""CPU oracle for the SGEMM/DGEMM hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package, and only as the checker or the CPU
baseline.  The shipped libraries never load it.

Two things live here:
  * ``liboracle.so`` (gemm_oracle.c): a restatement of the reference's
    semantics (wjc404/GEMM_AVX2 src/gemm_driver.c:118-197) with Dot2
    accumulation, the counter-based input generator, and bit-exact
    restatements of the emulated forms (Ozaki-I, Ozaki-II, split-BF16x3);
  * ``_ref/DGEMM.so`` / ``_ref/SGEMM.so``: the reference itself, compiled from
    its sources by oracle/Makefile when they are available.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_lib = None
_ref = {}

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def build(quiet=True):
    """Compiles liboracle.so (and _ref/ when the reference sources exist)."""
    r = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if not quiet:
        print(r.stdout)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("GEMMX_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.oracle_counter_u64.restype = ctypes.c_uint64
    L.oracle_counter_u64.argtypes = [ctypes.c_uint64] * 3
    L.oracle_uniform_f64.restype = ctypes.c_double
    L.oracle_uniform_f64.argtypes = [ctypes.c_uint64] * 3
    L.oracle_uniform_f32.restype = ctypes.c_float
    L.oracle_uniform_f32.argtypes = [ctypes.c_uint64] * 3
    for name, t in (("oracle_fill_f64", ctypes.c_void_p), ("oracle_fill_f32", ctypes.c_void_p)):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                      ctypes.c_uint64, ctypes.c_int, ctypes.c_double]
    for name, fl in (("oracle_dgemm", _dp), ("oracle_sgemm", _fp)):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _ip, _ip, _ip, fl, ctypes.c_void_p, _ip,
                      ctypes.c_void_p, _ip, fl, ctypes.c_void_p, _ip]
    L.oracle_dgemm_longdouble.restype = None
    L.oracle_dgemm_longdouble.argtypes = [
        ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_int]
    for name, fl in (("oracle_sampled_f64", ctypes.c_double), ("oracle_sampled_f32", ctypes.c_float)):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      fl, fl, ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_ozaki_exponent.restype = ctypes.c_int
    L.oracle_ozaki_exponent.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.oracle_ozaki_digits.restype = None
    L.oracle_ozaki_digits.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.oracle_ozaki_kchunk.restype = ctypes.c_int
    L.oracle_ozaki_kchunk.argtypes = [ctypes.c_int]
    L.oracle_ozaki_dgemm.restype = None
    L.oracle_ozaki_dgemm.argtypes = [
        ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.oracle_ozaki2_modulus.restype = ctypes.c_int
    L.oracle_ozaki2_modulus.argtypes = [ctypes.c_int]
    L.oracle_ozaki2_bits.restype = ctypes.c_int
    L.oracle_ozaki2_bits.argtypes = [ctypes.c_int, ctypes.c_int]
    L.oracle_ozaki2_vector_bits.restype = ctypes.c_int
    L.oracle_ozaki2_vector_bits.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int]
    L.oracle_ozaki2_scaled.restype = None
    L.oracle_ozaki2_scaled.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.POINTER(ctypes.c_int64)]
    L.oracle_ozaki2_round.restype = ctypes.c_double
    L.oracle_ozaki2_round.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
    L.oracle_ozaki2_dgemm.restype = None
    L.oracle_ozaki2_dgemm.argtypes = [
        ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.oracle_dot2_pairs.restype = None
    L.oracle_dot2_pairs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    for fn in ("oracle_ozaki_pairs", "oracle_ozaki2_pairs"):
        f = getattr(L, fn)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                      ctypes.c_int, ctypes.c_void_p]
    for fn, fl in (("oracle_zgemm", ctypes.c_double), ("oracle_cgemm", ctypes.c_float)):
        f = getattr(L, fn)
        f.restype = None
        f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.POINTER(fl), ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                      ctypes.c_int, ctypes.POINTER(fl), ctypes.c_void_p, ctypes.c_int]
    for fn, fl in (("oracle_trmm", ctypes.c_double), ("oracle_trsm", ctypes.c_double),
                   ("oracle_trmm_f32", ctypes.c_float), ("oracle_trsm_f32", ctypes.c_float)):
        f = getattr(L, fn)
        f.restype = None
        f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_int,
                      ctypes.c_int, fl, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                      ctypes.c_int]
    L.oracle_bf16_rne.restype = ctypes.c_uint16
    L.oracle_bf16_rne.argtypes = [ctypes.c_float]
    L.oracle_bf16_split.restype = None
    L.oracle_bf16_split.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_uint16),
                                    ctypes.POINTER(ctypes.c_uint16)]
    L.oracle_bf16x3_sgemm.restype = None
    L.oracle_bf16x3_sgemm.argtypes = [
        ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
        ctypes.c_void_p, ctypes.c_int]
    _lib = L
    return L


def ref(dtype: str):
    """The reference's dgemm_/sgemm_ from _ref/, or None when not built."""
    name = "DGEMM.so" if dtype == "d" else "SGEMM.so"
    if name in _ref:
        return _ref[name]
    path = os.path.join(REF_DIR, name)
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    fl = _dp if dtype == "d" else _fp
    f = getattr(L, f"{dtype}gemm_")
    f.restype = None
    f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _ip, _ip, _ip, fl, ctypes.c_void_p, _ip,
                  ctypes.c_void_p, _ip, fl, ctypes.c_void_p, _ip]
    _ref[name] = f
    return f


# ---------------------------------------------------------------- helpers
def _ptr(a):
    return a.ctypes.data


def fill(rows, cols, ld, seed, mid, dtype=np.float64, dist=0, param=0.0):
    """Column-major rows x cols matrix (ld x cols buffer, Fortran order) from
    the counter-based generator; padding rows are NaN."""
    buf = np.empty((ld, max(cols, 1)), dtype=dtype, order="F")
    f = lib().oracle_fill_f64 if dtype == np.float64 else lib().oracle_fill_f32
    f(_ptr(buf), rows, cols, ld, seed, mid, dist, param)
    return buf


def _call(fn, dtype, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    I = ctypes.c_int
    fl = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    fn(transa.encode(), transb.encode(), ctypes.byref(I(m)), ctypes.byref(I(n)),
       ctypes.byref(I(k)), ctypes.byref(fl(alpha)), _ptr(a), ctypes.byref(I(lda)), _ptr(b),
       ctypes.byref(I(ldb)), ctypes.byref(fl(beta)), _ptr(c), ctypes.byref(I(ldc)))


def gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Oracle GEMM in place on numpy buffers (dtype picks D or S)."""
    fn = lib().oracle_dgemm if c.dtype == np.float64 else lib().oracle_sgemm
    _call(fn, c.dtype, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)


def ref_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """The reference's own dgemm_/sgemm_ in place; returns False if absent."""
    fn = ref("d" if c.dtype == np.float64 else "s")
    if fn is None:
        return False
    _call(fn, c.dtype, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    return True


def sampled(dtype, transa, transb, m, n, k, alpha, beta, seed, is_, js, dist=0, param=0.0):
    """Expected C[is, js] (and sum|a||b|) for generator inputs, via Dot2."""
    is_ = np.ascontiguousarray(is_, dtype=np.int32)
    js = np.ascontiguousarray(js, dtype=np.int32)
    out = np.empty(len(is_), dtype=np.float64)
    ab = np.empty(len(is_), dtype=np.float64)
    f = lib().oracle_sampled_f64 if dtype == np.float64 else lib().oracle_sampled_f32
    f(transa.encode()[:1], transb.encode()[:1], m, n, k, alpha, beta, seed, dist, param,
      len(is_), _ptr(is_), _ptr(js), _ptr(out), _ptr(ab))
    return out, ab


def ozaki_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, slices):
    lib().oracle_ozaki_dgemm(transa.encode()[:1], transb.encode()[:1], m, n, k, alpha, _ptr(a),
                             lda, _ptr(b), ldb, beta, _ptr(c), ldc, slices)


def ozaki2_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, moduli):
    lib().oracle_ozaki2_dgemm(transa.encode()[:1], transb.encode()[:1], m, n, k, alpha,
                              _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc, moduli)


def _pairs(fn, rows, cols, ri, cj, alpha, c1, count_arg):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    cols = np.ascontiguousarray(cols, dtype=np.float64)
    k = rows.shape[1]
    assert cols.shape[1] == k
    ri = np.ascontiguousarray(ri, dtype=np.int32)
    cj = np.ascontiguousarray(cj, dtype=np.int32)
    assert ri.size == cj.size and ri.max(initial=0) < rows.shape[0] and cj.max(initial=0) < cols.shape[0]
    c1 = np.ascontiguousarray(np.broadcast_to(np.asarray(c1, dtype=np.float64), ri.shape))
    out = np.empty(ri.size, dtype=np.float64)
    rc = fn(_ptr(rows), _ptr(cols), k, ri.size, _ptr(ri), _ptr(cj), alpha, _ptr(c1), count_arg,
            _ptr(out))
    if rc != 0:
        raise ValueError("unsupported arguments for the sampled restatement")
    return out


def dot2_pairs(rows, cols, ri, cj, alpha, c1):
    """Dot2 entries alpha*(row ri[t] . col cj[t]) + c1[t] (as sampled(), for
    explicit operands); returns (values, sum|a||b|)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    cols = np.ascontiguousarray(cols, dtype=np.float64)
    ri = np.ascontiguousarray(ri, dtype=np.int32)
    cj = np.ascontiguousarray(cj, dtype=np.int32)
    assert rows.shape[1] == cols.shape[1] and ri.size == cj.size
    c1 = np.ascontiguousarray(np.broadcast_to(np.asarray(c1, dtype=np.float64), ri.shape))
    out = np.empty(ri.size, dtype=np.float64)
    ab = np.empty(ri.size, dtype=np.float64)
    lib().oracle_dot2_pairs(_ptr(rows), _ptr(cols), rows.shape[1], ri.size, _ptr(ri), _ptr(cj),
                            alpha, _ptr(c1), _ptr(out), _ptr(ab))
    return out, ab


def ozaki_pairs(rows, cols, ri, cj, alpha, c1, slices):
    """Ozaki-I entries alpha*(row ri[t] . col cj[t]) + c1[t]; rows (nr, k) are
    whole rows of op(A), cols (nc, k) whole columns of op(B)."""
    return _pairs(lib().oracle_ozaki_pairs, rows, cols, ri, cj, alpha, c1, slices)


def ozaki2_pairs(rows, cols, ri, cj, alpha, c1, moduli):
    """Ozaki-II entries, as ozaki_pairs."""
    return _pairs(lib().oracle_ozaki2_pairs, rows, cols, ri, cj, alpha, c1, moduli)


def complex_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """zgemm (complex128 arrays) or cgemm (complex64), Fortran-ordered, in place
    on c; alpha and beta Python complex numbers."""
    single = c.dtype == np.complex64
    fl = ctypes.c_float if single else ctypes.c_double
    al, be = (fl * 2)(alpha.real, alpha.imag), (fl * 2)(beta.real, beta.imag)
    fn = lib().oracle_cgemm if single else lib().oracle_zgemm
    fn(transa.encode()[:1], transb.encode()[:1], m, n, k, al, _ptr(a), lda, _ptr(b), ldb, be,
       _ptr(c), ldc)


def tri(solve, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb):
    """TRSM (solve) or TRMM on float64 or float32 Fortran arrays, in place on b."""
    single = b.dtype == np.float32
    name = ("oracle_trsm" if solve else "oracle_trmm") + ("_f32" if single else "")
    ch = [c.encode()[:1] for c in (side, uplo, transa, diag)]
    getattr(lib(), name)(*ch, m, n, alpha, _ptr(a), lda, _ptr(b), ldb)


def ozaki2_modulus(j):
    return int(lib().oracle_ozaki2_modulus(j))


def ozaki2_bits(moduli, k):
    """The smallest per-vector bit count over vectors of k entries."""
    return int(lib().oracle_ozaki2_bits(moduli, k))


def ozaki2_vector_bits(x, moduli):
    """Bit count P of one vector (a 1-D float64 array)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    return int(lib().oracle_ozaki2_vector_bits(_ptr(x), x.size, 1, moduli))


def ozaki2_scaled(x, e, P):
    """rint(x * 2^(P-e)) as a Python int."""
    lo, hi = ctypes.c_uint64(), ctypes.c_int64()
    lib().oracle_ozaki2_scaled(float(x), e, P, ctypes.byref(lo), ctypes.byref(hi))
    return (hi.value << 64) | lo.value


def ozaki2_round(value, scale):
    """A signed integer (|value| < 2^255) rounded to 53 bits and scaled."""
    v = value & ((1 << 256) - 1)
    limbs = (ctypes.c_uint64 * 4)(*[(v >> (64 * t)) & ((1 << 64) - 1) for t in range(4)])
    return float(lib().oracle_ozaki2_round(limbs, scale))


def bf16x3_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    lib().oracle_bf16x3_sgemm(transa.encode()[:1], transb.encode()[:1], m, n, k, alpha,
                              _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc)


def longdouble_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    lib().oracle_dgemm_longdouble(transa.encode()[:1], transb.encode()[:1], m, n, k, alpha,
                                  _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc)
