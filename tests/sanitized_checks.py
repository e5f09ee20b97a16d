"""Host-side checks run in a child process under AddressSanitizer and UBSan
(started by test_sanitizers.py with the sanitizer runtime preloaded and
GEMMX_LIB_DIR / GEMMX_ORACLE_LIB pointing at the sanitized builds).  They
drive the host C layer through every path that needs no GPU -- argument
checking of every entry point, the tuning-table reader on valid and malformed
tables, the lookups, the grouped-batch plan builder up to its device lookup
-- and the oracle's restatements on small inputs.  Any sanitizer report ends
the process with a non-zero status.  Not a pytest module (no test_ prefix)."""
import ctypes
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import gemm_avx2_amd as gx  # noqa: E402
import oracle  # noqa: E402
from gemm_avx2_amd import autotune as at  # noqa: E402


def expect_error(fn, *args):
    try:
        fn(*args)
    except gx.GemmError as e:
        return e.code
    raise AssertionError(f"{fn.__name__}{args[:5]} did not fail")


def host_layer():
    A = np.ones((4, 4), order="F")
    for d, f, f64 in (("d", gx.dgemm, gx.dgemm64), ("s", gx.sgemm, gx.sgemm64)):
        dt = np.float64 if d == "d" else np.float32
        M = np.ones((4, 4), dtype=dt, order="F")
        # invalid arguments are reported before any device lookup
        assert expect_error(f, "N", "N", -1, 4, 4, 1.0, M, 4, M, 4, 0.0, M, 4) == gx.ERR_ARG
        assert expect_error(f, "N", "N", 4, 4, 4, 1.0, M, 3, M, 4, 0.0, M, 4) == gx.ERR_ARG
        assert expect_error(f, "T", "N", 4, 4, 8, 1.0, M, 4, M, 8, 0.0, M, 2) == gx.ERR_ARG
        assert expect_error(f64, "N", "N", 1 << 40, 4, 4, 1.0, M, 4, M, 4, 0.0, M, 4) == gx.ERR_ARG
        # m = 0 or n = 0 returns at once; valid calls find no device here
        f("N", "N", 0, 4, 4, 1.0, M, 4, M, 4, 0.0, M, 4)
        code = expect_error(f, "N", "N", 4, 4, 4, 1.0, M, 4, M, 4, 0.0, M, 4)
        assert code in (gx.ERR_NODEVICE, gx.ERR_LAUNCH), code
        msg = gx.last_error(d)
        assert isinstance(msg, tuple) or msg is not None
        for tr in ("NN", "NT", "TN", "TT"):
            for shape in ((1, 1, 1), (7, 9, 300), (512, 512, 65536), (100000, 3, 70000),
                          (2 ** 31 - 1, 1, 1), (0, 0, 0)):
                gx.tune_lookup(d, tr[0], tr[1], *shape)
                gx.tune_lookup_split(d, tr[0], tr[1], *shape)
        gx.set_split_config(d, 7)
        gx.tune_lookup_split(d, "N", "N", 100, 100, 5000)
        gx.set_split_config(d, -1)
    gx.tune_lookup_emulated("N", "N", 8192, 8192, 8192)
    for fn in (gx.zgemm, gx.cgemm):
        dt = np.complex128 if fn is gx.zgemm else np.complex64
        Z = np.ones((4, 4), dtype=dt, order="F")
        assert expect_error(fn, "C", "N", 4, 4, 4, 1.0, Z, 2, Z, 4, 0.0, Z, 4) == gx.ERR_ARG
    for fn in (gx.dtrsm, gx.strmm):
        dt = np.float64 if fn is gx.dtrsm else np.float32
        T = np.ones((4, 4), dtype=dt, order="F")
        assert expect_error(fn, "L", "U", "N", "N", 4, 4, 1.0, T, 2, T, 4) == gx.ERR_ARG
    assert expect_error(gx.dgemm_ozaki2_dev, "N", "N", 4, 4, 4, 1.0, 0, 4, 0, 4, 0.0, 0, 4, 7) \
        == gx.ERR_ARG
    # grouped batch: a bad problem is refused; a valid plan reaches the device lookup
    lib = gx.load("DGEMM.so")
    n = 3
    ta, tb = (ctypes.c_char * n)(*b"NTN"), (ctypes.c_char * n)(*b"NNT")
    I3 = ctypes.c_int * n
    D3 = ctypes.c_double * n
    P3 = ctypes.c_void_p * n
    bufs = [np.ones(64 * 64) for _ in range(3)]
    ptrs = P3(*[b.ctypes.data for b in bufs])
    plan = ctypes.c_void_p()
    for m in ((8, 8, 8), (8, -1, 8)):
        rc = lib.dgemm_x_batch_create(ctypes.byref(plan), n, ta, tb, I3(*m), I3(8, 8, 8),
                                      I3(8, 8, 8), D3(1, 1, 1), ptrs, I3(8, 8, 8), ptrs,
                                      I3(8, 8, 8), D3(0, 0, 0), ptrs, I3(8, 8, 8))
        assert rc != 0
        if rc == 0:
            lib.dgemm_x_batch_destroy(plan)


def tables():
    good, bad = dict(), []
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t.txt")
        at.write_table(path, {("f64", "NN", 13, 13, 13): (0, 2), ("f32", "TT", 4, 15, 15): 5,
                              ("emu", "NN", 13, 13, 13): ("ozaki2", 19), ("batch", "f64"): 1},
                       {}, "gfx950")
        for dt in ("d", "s"):
            assert gx.load_tuning(path, dt) == 4
        rocm = gx.build_rocm()
        bad = [
            "format 2\nrocm %s\n" % rocm + "f64 NN 13 13 13 " + "9" * 400 + "\n",
            "format 2\nrocm %s\n" % rocm + "x" * 1000 + "\n" + "f64 NN 13 13 13 1 1 1 1\n",
            "format 2\nrocm " + "7" * 300 + "\n",
            "format 2\nrocm %s\nemu NN 13 13 13 ozaki2 -5\nbatch f64 99999999999\n" % rocm,
            "format 2\nrocm %s\nf64 NN -1 99 13 1 65\nf64 NN 13 13 13 1 0\n" % rocm,
            "\x00\x01\x02 format 2\n", "", "#\n" * 1000,
            "format 2\narch " + "g" * 500 + "\n",
        ]
        for i, text in enumerate(bad):
            p = os.path.join(d, f"bad{i}.txt")
            with open(p, "w") as f:
                f.write(text)
            for dt in ("d", "s"):
                gx.load_tuning(p, dt)
                gx.tune_lookup(dt, "N", "N", 8192, 8192, 8192)
        gx.load_tuning(None, "d")
        gx.load_tuning(None, "s")


def oracle_functions():
    rng = np.random.default_rng(1)
    for ta, tb in (("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")):
        m, n, k = 13, 11, 37
        A = np.asfortranarray(rng.uniform(-1, 1, (40, 40)))
        B = np.asfortranarray(rng.uniform(-1, 1, (40, 40)))
        C = np.asfortranarray(rng.uniform(-1, 1, (40, 40)))
        oracle.gemm(ta, tb, m, n, k, 1.5, A, 40, B, 40, -0.5, C.copy(order="F"), 40)
        oracle.ozaki_gemm(ta, tb, m, n, k, 1.5, A, 40, B, 40, -0.5, C.copy(order="F"), 40, 6)
        oracle.ozaki2_gemm(ta, tb, m, n, k, 1.5, A, 40, B, 40, -0.5, C.copy(order="F"), 40, 17)
        oracle.bf16x3_gemm(ta, tb, m, n, k, 1.5, A.astype(np.float32, order="F"), 40,
                           B.astype(np.float32, order="F"), 40, -0.5,
                           C.astype(np.float32, order="F"), 40)
        oracle.longdouble_gemm(ta, tb, m, n, k, 1.5, A, 40, B, 40, -0.5, C.copy(order="F"), 40)
        Z = (A + 1j * B).astype(np.complex128, order="F")
        oracle.complex_gemm(ta, tb, m, n, k, 1 + 0.5j, Z, 40, Z, 40, 0.5j, Z.copy(order="F"), 40)
    rows, cols = rng.uniform(-1, 1, (5, 300)), rng.uniform(-1, 1, (4, 300))
    ri, cj = np.array([0, 4, 2]), np.array([3, 0, 1])
    oracle.dot2_pairs(rows, cols, ri, cj, 1.0, 0.0)
    oracle.ozaki_pairs(rows, cols, ri, cj, 1.0, 0.0, 8)
    oracle.ozaki2_pairs(rows, cols, ri, cj, 1.0, 0.0, 22)
    oracle.sampled(np.float64, "N", "T", 20, 30, 40, 1.0, 0.5, 3, ri, cj, 1, 4.0)
    T = np.asfortranarray(rng.uniform(-1, 1, (16, 16)) + 16 * np.eye(16))
    for solve in (0, 1):
        for side in "LR":
            oracle.tri(solve, side, "U", "T", "N", 16, 16, 2.0, T, 16, T.copy(order="F"), 16)
    oracle.fill(7, 5, 9, 3, 1, np.float32, 2, 2.0)
    for x in (0.0, 5e-324, -1.5, 2.0 ** 70):
        oracle.ozaki2_scaled(x, 3, 40)
    oracle.ozaki2_round(-(2 ** 200) - 12345, -300)


if __name__ == "__main__":
    host_layer()
    tables()
    oracle_functions()
    print("sanitized checks ok", flush=True)
