"""MI355X (gfx950) SGEMM/DGEMM: Python binding of DGEMM.so / SGEMM.so.

The product is the pair of C-ABI shared libraries in ``lib/`` (built by
``python -m gemm_avx2_amd.build``); this module only loads them with ctypes so
that tests, the benchmark and Python callers can reach the same entry points a
Fortran or C program links against.  It mirrors the reference's single entry
point, ``dgemm_``/``sgemm_`` of wjc404/GEMM_AVX2 src/gemm_driver.c:118, with
the same argument order and meaning:

    dgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)

``a``, ``b`` and ``c`` may be numpy arrays (host memory: staged to the GPU and
back), torch tensors on the GPU (used in place), or raw integer addresses.
Like the reference, nothing is returned and C is updated in place; unlike it,
invalid arguments and GPU errors are reported through :func:`last_error`
(``dgemm`` raises :class:`GemmError` when ``check=True``).

There is no CPU fallback: when the libraries are missing every call raises.
"""
from __future__ import annotations

import ctypes
import os

__all__ = [
    "LIB_DIR", "GemmError", "load", "dgemm", "sgemm", "dgemm64", "sgemm64", "zgemm", "cgemm",
    "zgemm3m", "cgemm3m", "dtrmm", "dtrsm", "strmm", "strsm", "dgemm_dev",
    "sgemm_dev", "dgemm_dev_noalloc", "sgemm_dev_noalloc", "dgemm_strided_batched_dev", "sgemm_strided_batched_dev",
    "dgemm_ozaki_dev", "dgemm_ozaki_workspace", "dgemm_ozaki2_dev", "dgemm_ozaki2_workspace",
    "dgemm_ozaki2_bits", "set_phase_timing", "phase_times", "sgemm_bf16x3_dev", "sgemm_bf16x3_workspace",
    "set_dgemm_mode", "set_sgemm_mode", "set_tile_config", "set_split_config",
    "tune_lookup_split", "build_rocm", "set_product_config", "load_tuning", "tune_lookup", "tune_lookup_emulated",
    "tune_lookup_emulated32", "tune_lookup_strided", "last_error", "fill",
    "BatchPlan", "set_batch_config", "batch_config", "MODE_NATIVE", "MODE_OZAKI", "MODE_BF16X3", "MODE_OZAKI2", "MODE_AUTO",
]

LIB_DIR = os.environ.get("GEMMX_LIB_DIR") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "lib")

MODE_NATIVE, MODE_OZAKI, MODE_BF16X3, MODE_OZAKI2, MODE_AUTO = 0, 1, 2, 3, 4
OK, ERR_ARG, ERR_NODEVICE, ERR_ALLOC, ERR_LAUNCH, ERR_UNSUPPORTED = range(6)


class GemmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_libs: dict = {}

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_i64_p = ctypes.POINTER(ctypes.c_int64)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_float_p = ctypes.POINTER(ctypes.c_float)


def _declare(lib, dtype):
    fl = ctypes.c_double if dtype == "d" else ctypes.c_float
    flp = _c_double_p if dtype == "d" else _c_float_p
    p = dtype
    f = getattr(lib, f"{p}gemm_")
    f.restype = None
    f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _c_int_p, _c_int_p, _c_int_p, flp,
                  ctypes.c_void_p, _c_int_p, ctypes.c_void_p, _c_int_p, flp, ctypes.c_void_p,
                  _c_int_p]
    for name in (("dtrmm_", "dtrsm_") if dtype == "d" else ("strmm_", "strsm_")):
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = [ctypes.c_char_p] * 4 + [_c_int_p, _c_int_p, flp, ctypes.c_void_p, _c_int_p,
                                              ctypes.c_void_p, _c_int_p]
    for name in (("zgemm_", "zgemm3m_") if dtype == "d" else ("cgemm_", "cgemm3m_")):
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _c_int_p, _c_int_p, _c_int_p, flp,
                      ctypes.c_void_p, _c_int_p, ctypes.c_void_p, _c_int_p, flp,
                      ctypes.c_void_p, _c_int_p]
    f = getattr(lib, f"{p}gemm_64_")
    f.restype = None
    f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _c_i64_p, _c_i64_p, _c_i64_p, flp,
                  ctypes.c_void_p, _c_i64_p, ctypes.c_void_p, _c_i64_p, flp, ctypes.c_void_p,
                  _c_i64_p]
    for name in (f"{p}gemm_x_dev", f"{p}gemm_x_dev_noalloc"):
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      fl, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, fl,
                      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    f = getattr(lib, f"{p}gemm_x_strided_batched_dev")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, fl,
                  ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_int64, fl, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                  ctypes.c_int, ctypes.c_void_p]
    f = getattr(lib, f"{p}gemm_x_last_error")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.POINTER(ctypes.c_char_p)]
    f = getattr(lib, f"{p}gemm_x_set_tile_config")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int]
    f = getattr(lib, f"{p}gemm_x_set_product_config")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int]
    f = getattr(lib, f"{p}gemm_x_load_tuning")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char_p]
    if dtype == "d":
        lib.dgemm_x_set_mode.restype = ctypes.c_int
        lib.dgemm_x_set_mode.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.dgemm_x_ozaki_workspace.restype = ctypes.c_size_t
        lib.dgemm_x_ozaki_workspace.argtypes = [ctypes.c_int] * 4
        lib.dgemm_x_ozaki_dev.restype = ctypes.c_int
        lib.dgemm_x_ozaki_dev.argtypes = [
            ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
            ctypes.c_size_t, ctypes.c_void_p]
        lib.dgemm_x_ozaki2_workspace.restype = ctypes.c_size_t
        lib.dgemm_x_ozaki2_workspace.argtypes = [ctypes.c_int] * 4
        lib.dgemm_x_ozaki2_bits.restype = ctypes.c_int
        lib.dgemm_x_ozaki2_bits.argtypes = [ctypes.c_int] * 2
        lib.dgemm_x_ozaki2_dev.restype = ctypes.c_int
        lib.dgemm_x_ozaki2_dev.argtypes = lib.dgemm_x_ozaki_dev.argtypes
    else:
        lib.sgemm_x_set_mode.restype = ctypes.c_int
        lib.sgemm_x_set_mode.argtypes = [ctypes.c_int]
        lib.sgemm_x_bf16x3_workspace.restype = ctypes.c_size_t
        lib.sgemm_x_bf16x3_workspace.argtypes = [ctypes.c_int] * 3
        lib.sgemm_x_bf16x3_dev.restype = ctypes.c_int
        lib.sgemm_x_bf16x3_dev.argtypes = [
            ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
            ctypes.c_void_p]


def _declare_batch(lib, dtype):
    fl = ctypes.c_double if dtype == "d" else ctypes.c_float
    f = getattr(lib, f"{dtype}gemm_x_batch_create")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_char_p,
                  ctypes.c_char_p, _c_int_p, _c_int_p, _c_int_p, ctypes.POINTER(fl),
                  ctypes.POINTER(ctypes.c_void_p), _c_int_p, ctypes.POINTER(ctypes.c_void_p),
                  _c_int_p, ctypes.POINTER(fl), ctypes.POINTER(ctypes.c_void_p), _c_int_p]
    f = getattr(lib, f"{dtype}gemm_x_batch_run")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    f = getattr(lib, f"{dtype}gemm_x_batch_flops")
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_void_p]
    f = getattr(lib, f"{dtype}gemm_x_batch_destroy")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p]
    f = getattr(lib, f"{dtype}gemm_x_batch_set_config")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int]
    f = getattr(lib, f"{dtype}gemm_x_batch_config")
    f.restype = ctypes.c_int
    f.argtypes = []


def load(name: str):
    """Loads lib/<name> ('DGEMM.so', 'SGEMM.so', 'libgemmx_gen.so') once."""
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise GemmError(ERR_UNSUPPORTED,
                        f"{path} is not built; run `python -m gemm_avx2_amd.build`")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    if name == "DGEMM.so":
        _declare(lib, "d")
        _declare_batch(lib, "d")
    elif name == "SGEMM.so":
        _declare(lib, "s")
        _declare_batch(lib, "s")
    elif name == "libgemmx_gen.so":
        for fn, t in (("gemmx_fill_f64", ctypes.c_void_p), ("gemmx_fill_f32", ctypes.c_void_p)):
            f = getattr(lib, fn)
            f.restype = ctypes.c_int
            f.argtypes = [t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                          ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    _libs[name] = lib
    return lib


def _addr(x) -> int:
    """Address of a numpy array, torch tensor, ctypes object or raw int."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    if hasattr(x, "ctypes"):
        return int(x.ctypes.data)
    return ctypes.addressof(x)


def _stream(s) -> int:
    if s is None:
        return 0
    if hasattr(s, "cuda_stream"):
        return int(s.cuda_stream)
    return int(s)


def last_error(dtype: str = "d"):
    """(code, message) of the calling thread's last call into DGEMM/SGEMM."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    msg = ctypes.c_char_p()
    code = getattr(lib, f"{dtype}gemm_x_last_error")(ctypes.byref(msg))
    return code, (msg.value or b"").decode()


def _check(dtype, code=None):
    c, msg = last_error(dtype)
    if code is not None and code != OK:
        raise GemmError(code, msg)
    if c != OK:
        raise GemmError(c, msg)


def _fortran(dtype, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check):
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    fl = ctypes.c_double if dtype == "d" else ctypes.c_float
    ta = transa.encode() if isinstance(transa, str) else bytes(transa)
    tb = transb.encode() if isinstance(transb, str) else bytes(transb)
    I = ctypes.c_int
    getattr(lib, f"{dtype}gemm_")(ta, tb, ctypes.byref(I(m)), ctypes.byref(I(n)),
                                   ctypes.byref(I(k)), ctypes.byref(fl(alpha)), _addr(a),
                                   ctypes.byref(I(lda)), _addr(b), ctypes.byref(I(ldb)),
                                   ctypes.byref(fl(beta)), _addr(c), ctypes.byref(I(ldc)))
    if check:
        _check(dtype)


def dgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check=True):
    """The drop-in dgemm_ (src/gemm_driver.c:118): C = alpha*op(A)op(B) + beta*C."""
    _fortran("d", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check)


def sgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check=True):
    """The drop-in sgemm_ (src/gemm_driver.c:118 without -DDOUBLE)."""
    _fortran("s", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check)


def _fortran64(dtype, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check):
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    fl = ctypes.c_double if dtype == "d" else ctypes.c_float
    ta = transa.encode() if isinstance(transa, str) else bytes(transa)
    tb = transb.encode() if isinstance(transb, str) else bytes(transb)
    I = ctypes.c_int64
    getattr(lib, f"{dtype}gemm_64_")(ta, tb, ctypes.byref(I(m)), ctypes.byref(I(n)),
                                      ctypes.byref(I(k)), ctypes.byref(fl(alpha)), _addr(a),
                                      ctypes.byref(I(lda)), _addr(b), ctypes.byref(I(ldb)),
                                      ctypes.byref(fl(beta)), _addr(c), ctypes.byref(I(ldc)))
    if check:
        _check(dtype)


def dgemm64(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check=True):
    """dgemm_64_: dgemm_ with 64-bit integer arguments (ILP64)."""
    _fortran64("d", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check)


def sgemm64(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check=True):
    """sgemm_64_: sgemm_ with 64-bit integer arguments (ILP64)."""
    _fortran64("s", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check)


def _complex(name, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check):
    dtype = "d" if name[0] == "z" else "s"
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    fl = ctypes.c_double if dtype == "d" else ctypes.c_float
    ta = transa.encode() if isinstance(transa, str) else bytes(transa)
    tb = transb.encode() if isinstance(transb, str) else bytes(transb)
    al, be = complex(alpha), complex(beta)
    I = ctypes.c_int
    getattr(lib, name)(ta, tb, ctypes.byref(I(m)), ctypes.byref(I(n)), ctypes.byref(I(k)),
                       (fl * 2)(al.real, al.imag), _addr(a), ctypes.byref(I(lda)), _addr(b),
                       ctypes.byref(I(ldb)), (fl * 2)(be.real, be.imag), _addr(c),
                       ctypes.byref(I(ldc)))
    if check:
        _check(dtype)


def zgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check=True):
    """zgemm_: double complex C = alpha op(A) op(B) + beta C ('C' conjugates),
    four real MFMA products; operands complex128 (host arrays or device
    tensors), sizes in complex elements."""
    _complex("zgemm_", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check)


def zgemm3m(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check=True):
    """zgemm3m_: as zgemm with three real products."""
    _complex("zgemm3m_", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check)


def cgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check=True):
    """cgemm_: single complex (complex64) form of zgemm."""
    _complex("cgemm_", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check)


def cgemm3m(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check=True):
    _complex("cgemm3m_", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, check)


def _tri(name, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check):
    dtype = name[0]
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    fl = ctypes.c_double if dtype == "d" else ctypes.c_float
    ch = [(c.encode() if isinstance(c, str) else bytes(c)) for c in (side, uplo, transa, diag)]
    I = ctypes.c_int
    getattr(lib, name)(*ch, ctypes.byref(I(m)), ctypes.byref(I(n)), ctypes.byref(fl(alpha)),
                       _addr(a), ctypes.byref(I(lda)), _addr(b), ctypes.byref(I(ldb)))
    if check:
        _check(dtype)


def dtrmm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check=True):
    """dtrmm_: B := alpha op(A) B (side 'L') or alpha B op(A), A triangular."""
    _tri("dtrmm_", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check)


def dtrsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check=True):
    """dtrsm_: solves op(A) X = alpha B (side 'L') or X op(A) = alpha B, X over B."""
    _tri("dtrsm_", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check)


def strmm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check=True):
    _tri("strmm_", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check)


def strsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check=True):
    _tri("strsm_", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, check)


def _ch(t):
    return (t.encode() if isinstance(t, str) else bytes(t))[:1]


def _strided(dtype, transa, transb, m, n, k, alpha, a, lda, stride_a, b, ldb, stride_b, beta,
             c, ldc, stride_c, batch, stream):
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    rc = getattr(lib, f"{dtype}gemm_x_strided_batched_dev")(
        _ch(transa), _ch(transb), m, n, k, alpha, _addr(a), lda, stride_a, _addr(b), ldb,
        stride_b, beta, _addr(c), ldc, stride_c, batch, _stream(stream))
    if rc:
        _check(dtype, rc)


def dgemm_strided_batched_dev(transa, transb, m, n, k, alpha, a, lda, stride_a, b, ldb,
                              stride_b, beta, c, ldc, stride_c, batch, stream=None):
    """`batch` device-pointer DGEMMs of one shape, operands `stride_*`
    elements apart, enqueued on `stream`."""
    _strided("d", transa, transb, m, n, k, alpha, a, lda, stride_a, b, ldb, stride_b, beta, c,
             ldc, stride_c, batch, stream)


def sgemm_strided_batched_dev(transa, transb, m, n, k, alpha, a, lda, stride_a, b, ldb,
                              stride_b, beta, c, ldc, stride_c, batch, stream=None):
    _strided("s", transa, transb, m, n, k, alpha, a, lda, stride_a, b, ldb, stride_b, beta, c,
             ldc, stride_c, batch, stream)


def dgemm_dev(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    """Device-pointer DGEMM enqueued on `stream` (no host synchronisation)."""
    lib = load("DGEMM.so")
    rc = lib.dgemm_x_dev(_ch(transa), _ch(transb), m, n, k, alpha, _addr(a), lda, _addr(b),
                         ldb, beta, _addr(c), ldc, _stream(stream))
    if rc:
        _check("d", rc)


def sgemm_dev(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    lib = load("SGEMM.so")
    rc = lib.sgemm_x_dev(_ch(transa), _ch(transb), m, n, k, alpha, _addr(a), lda, _addr(b),
                         ldb, beta, _addr(c), ldc, _stream(stream))
    if rc:
        _check("s", rc)


def dgemm_dev_noalloc(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                      stream=None):
    """dgemm_dev without any device allocation (split-K partials in the
    library's static device scratch; overlapping calls on one GPU race)."""
    lib = load("DGEMM.so")
    rc = lib.dgemm_x_dev_noalloc(_ch(transa), _ch(transb), m, n, k, alpha, _addr(a), lda,
                                 _addr(b), ldb, beta, _addr(c), ldc, _stream(stream))
    if rc:
        _check("d", rc)


def sgemm_dev_noalloc(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                      stream=None):
    lib = load("SGEMM.so")
    rc = lib.sgemm_x_dev_noalloc(_ch(transa), _ch(transb), m, n, k, alpha, _addr(a), lda,
                                 _addr(b), ldb, beta, _addr(c), ldc, _stream(stream))
    if rc:
        _check("s", rc)


def dgemm_ozaki_workspace(m, n, k, slices) -> int:
    return int(load("DGEMM.so").dgemm_x_ozaki_workspace(m, n, k, slices))


def dgemm_ozaki_dev(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, slices,
                    work=None, work_bytes=0, stream=None):
    """Emulated DGEMM (Ozaki-I, INT8 slices) on device pointers."""
    lib = load("DGEMM.so")
    rc = lib.dgemm_x_ozaki_dev(_ch(transa), _ch(transb), m, n, k, alpha, _addr(a), lda,
                               _addr(b), ldb, beta, _addr(c), ldc, slices, _addr(work),
                               work_bytes, _stream(stream))
    if rc:
        _check("d", rc)


def dgemm_ozaki2_workspace(m, n, k, moduli) -> int:
    return int(load("DGEMM.so").dgemm_x_ozaki2_workspace(m, n, k, moduli))


def dgemm_ozaki2_bits(moduli, k) -> int:
    return int(load("DGEMM.so").dgemm_x_ozaki2_bits(moduli, k))


def dgemm_ozaki2_dev(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, moduli,
                     work=None, work_bytes=0, stream=None):
    """Emulated DGEMM (Ozaki-II, CRT over INT8 residues) on device pointers."""
    lib = load("DGEMM.so")
    rc = lib.dgemm_x_ozaki2_dev(_ch(transa), _ch(transb), m, n, k, alpha, _addr(a), lda,
                                _addr(b), ldb, beta, _addr(c), ldc, moduli, _addr(work),
                                work_bytes, _stream(stream))
    if rc:
        _check("d", rc)


def set_phase_timing(on: bool = True):
    """Per-thread phase timing of the emulated DGEMMs (see phase_times)."""
    lib = load("DGEMM.so")
    lib.dgemm_x_set_phase_timing.restype = ctypes.c_int
    lib.dgemm_x_set_phase_timing.argtypes = [ctypes.c_int]
    lib.dgemm_x_set_phase_timing(1 if on else 0)


def phase_times():
    """(split, products, recombination) ms of this thread's last emulated
    DGEMM issued with phase timing on; waits for it."""
    lib = load("DGEMM.so")
    lib.dgemm_x_phase_times.restype = ctypes.c_int
    lib.dgemm_x_phase_times.argtypes = [ctypes.POINTER(ctypes.c_float)]
    out = (ctypes.c_float * 3)()
    rc = lib.dgemm_x_phase_times(out)
    if rc:
        _check("d", rc)
    return tuple(float(v) for v in out)


def sgemm_bf16x3_workspace(m, n, k) -> int:
    return int(load("SGEMM.so").sgemm_x_bf16x3_workspace(m, n, k))


def sgemm_bf16x3_dev(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                     work=None, work_bytes=0, stream=None):
    """Emulated SGEMM (split-BF16x3) on device pointers."""
    lib = load("SGEMM.so")
    rc = lib.sgemm_x_bf16x3_dev(_ch(transa), _ch(transb), m, n, k, alpha, _addr(a), lda,
                                _addr(b), ldb, beta, _addr(c), ldc, _addr(work), work_bytes,
                                _stream(stream))
    if rc:
        _check("s", rc)


class BatchPlan:
    """A grouped batch of independent GEMMs on device pointers, captured once
    as a hipGraph on the current device and replayed by run().  `problems` is
    a sequence of (transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c,
    ldc) with a/b/c device tensors or addresses; each problem has the
    semantics of dgemm_/sgemm_.  No reference counterpart (the reference
    computes one GEMM per call, src/gemm_driver.c:118)."""

    def __init__(self, dtype: str, problems):
        self.dtype = dtype
        self._lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
        fl = ctypes.c_double if dtype == "d" else ctypes.c_float
        cnt = len(problems)
        I = ctypes.c_int * max(cnt, 1)
        F = fl * max(cnt, 1)
        P = ctypes.c_void_p * max(cnt, 1)
        cols = list(zip(*problems)) if cnt else [()] * 13
        ta = b"".join(_ch(t) for t in cols[0]) or b"N"
        tb = b"".join(_ch(t) for t in cols[1]) or b"N"
        m, n, k = (I(*cols[i]) for i in (2, 3, 4))
        alpha, beta = F(*cols[5]), F(*cols[10])
        a, b, c = (P(*[_addr(x) for x in cols[i]]) for i in (6, 8, 11))
        lda, ldb, ldc = (I(*cols[i]) for i in (7, 9, 12))
        self._plan = ctypes.c_void_p()
        rc = getattr(self._lib, f"{dtype}gemm_x_batch_create")(
            ctypes.byref(self._plan), cnt, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        if rc:
            self._plan = None
            _check(dtype, rc)

    @property
    def flops(self) -> float:
        return float(getattr(self._lib, f"{self.dtype}gemm_x_batch_flops")(self._plan))

    def run(self, stream=None):
        if self._plan is None:
            raise GemmError(ERR_ARG, "batch plan destroyed")
        rc = getattr(self._lib, f"{self.dtype}gemm_x_batch_run")(self._plan, _stream(stream))
        if rc:
            _check(self.dtype, rc)

    def destroy(self):
        if self._plan is not None:
            getattr(self._lib, f"{self.dtype}gemm_x_batch_destroy")(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


BATCH_CONFIGS = 4


def batch_config(dtype: str) -> int:
    """The batch kernel configuration a plan of `dtype` created now would use
    (the one set, else GEMMX_BATCH_CFG, else the tuning table, else 0)."""
    return int(getattr(load("DGEMM.so" if dtype == "d" else "SGEMM.so"),
                       f"{dtype}gemm_x_batch_config")())


def set_batch_config(dtype: str, cfg: int):
    """Batch kernel tile configuration for plans created afterwards ('d' or
    's'); -1 returns to GEMMX_BATCH_CFG / the default."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    rc = getattr(lib, f"{dtype}gemm_x_batch_set_config")(cfg)
    if rc:
        _check(dtype, rc)


TILE_CONFIGS = 10


def set_tile_config(dtype: str, cfg: int):
    """Forces native tile configuration `cfg` for every call ('d' or 's');
    -1 returns to the tuning table."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    rc = getattr(lib, f"{dtype}gemm_x_set_tile_config")(cfg)
    if rc:
        _check(dtype, rc)


def set_split_config(dtype: str, split: int):
    """Forces `split` K ranges (split-K, 1 = one pass) for every native call
    ('d' or 's'); -1 returns to the tuning table / heuristic."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    rc = getattr(lib, f"{dtype}gemm_x_set_split_config")(split)
    if rc:
        _check(dtype, rc)


def tune_lookup_split(dtype, transa, transb, m, n, k):
    """(K ranges, partial-tile bytes) of a native call with these arguments
    (1 range = one pass)."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    f = getattr(lib, f"{dtype}gemm_x_tune_lookup_split")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_size_t)]
    b = ctypes.c_size_t()
    return int(f(_ch(transa), _ch(transb), m, n, k, ctypes.byref(b))), int(b.value)


def build_rocm(dtype: str = "d") -> str:
    """"major.minor" of the ROCm release the library was built with (the
    `rocm` key of the tuning tables it accepts)."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    f = getattr(lib, f"{dtype}gemm_x_build_rocm")
    f.restype = ctypes.c_char_p
    return f().decode()


def set_product_config(dtype: str, cfg: int):
    """Product kernel of the emulated forms in the library of `dtype` ('d':
    Ozaki-I and Ozaki-II, 's': BF16x3): 1 direct-to-LDS (default), 0
    register-staged, -1 back to GEMMX_SLICE_CFG / the default."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    rc = getattr(lib, f"{dtype}gemm_x_set_product_config")(cfg)
    if rc:
        _check(dtype, rc)


def load_tuning(path, dtype: str = "d") -> int:
    """Loads a tuning table into the library of `dtype`; returns entries read,
    or -1 when the file is unreadable, of another format version, arch or
    ROCm release, or names no ROCm release."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    return int(getattr(lib, f"{dtype}gemm_x_load_tuning")(
        None if path is None else str(path).encode()))


def tune_lookup(dtype, transa, transb, m, n, k) -> int:
    """Tile configuration a call with these arguments would use."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    f = getattr(lib, f"{dtype}gemm_x_tune_lookup")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return int(f(_ch(transa), _ch(transb), m, n, k))


def tune_lookup_strided(dtype, transa, transb, m, n, k, batch) -> int:
    """Tile configuration a strided batch of `batch` such problems uses."""
    lib = load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    f = getattr(lib, f"{dtype}gemm_x_tune_lookup_strided")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char, ctypes.c_char] + [ctypes.c_int] * 4
    return int(f(_ch(transa), _ch(transb), m, n, k, batch))


def tune_lookup_emulated(transa, transb, m, n, k):
    """(mode, count) dgemm_ uses in MODE_AUTO for these arguments."""
    lib = load("DGEMM.so")
    f = lib.dgemm_x_tune_lookup_emulated
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_int)]
    c = ctypes.c_int()
    mode = int(f(_ch(transa), _ch(transb), m, n, k, ctypes.byref(c)))
    return mode, c.value


def tune_lookup_emulated32(transa, transb, m, n, k) -> int:
    """The mode (MODE_NATIVE or MODE_BF16X3) sgemm_ uses in MODE_AUTO for
    these arguments."""
    lib = load("SGEMM.so")
    f = lib.sgemm_x_tune_lookup_emulated
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return int(f(_ch(transa), _ch(transb), m, n, k))


def set_dgemm_mode(mode: int, slices: int = 8):
    rc = load("DGEMM.so").dgemm_x_set_mode(mode, slices)
    if rc:
        _check("d", rc)


def set_sgemm_mode(mode: int):
    rc = load("SGEMM.so").sgemm_x_set_mode(mode)
    if rc:
        _check("s", rc)


def fill(t, rows, cols, ld, seed, mid, dist=0, param=0.0, stream=None):
    """Fills a device matrix (torch tensor or address) from the counter-based
    generator: element (i, j) draws counter i + j*rows of stream (seed, mid);
    rows..ld-1 are NaN.  dist 0 U[-1,1), 1 (u-.5)exp(param*g), 2 exp(param*g)."""
    lib = load("libgemmx_gen.so")
    dt = getattr(t, "dtype", None)
    is32 = dt is not None and "float32" in str(dt)
    fn = lib.gemmx_fill_f32 if is32 else lib.gemmx_fill_f64
    rc = fn(_addr(t), rows, cols, ld, seed, mid, dist, param, _stream(stream))
    if rc:
        raise GemmError(rc, "gemmx_fill failed")
