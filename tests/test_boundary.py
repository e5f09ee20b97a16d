"""The drop-in boundary on the CPU (no GPU needed): the libraries load, export
exactly what include/gemm_mi355x.h declares (dgemm_/sgemm_ as in
wjc404/GEMM_AVX2 src/gemm_driver.c:118 plus the *_x_* extensions), and the
failure paths leave C untouched: invalid arguments, and no HIP device (this
container has none)."""
import os
import re
import subprocess

import numpy as np
import pytest

import gemm_avx2_amd as gx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gemm_mi355x.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b([a-z_0-9]+)\s*\(", text)) - {"if", "sizeof"}


def _exports(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(gx.LIB_DIR, lib)],
                         capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_exports_match_header():
    decl = _declared()
    d, s = _exports("DGEMM.so"), _exports("SGEMM.so")
    assert "dgemm_" in d and "sgemm_" in s
    assert d == {f for f in decl if f.startswith(("dgemm_", "zgemm", "dtr"))}
    assert s == {f for f in decl if f.startswith(("sgemm_", "cgemm", "str"))}
    assert not (d & s)


def test_libraries_load():
    for name in ("DGEMM.so", "SGEMM.so", "libgemmx_gen.so"):
        assert gx.load(name) is not None


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_ilp64_and_strided_batch_argument_checks(dtype):
    """The ILP64 form rejects dimensions beyond 2^31 - 1 and checks like
    dgemm_; the strided batch rejects a negative count or stride and C
    matrices that overlap.  All before any device work (no GPU needed)."""
    dt = np.float64 if dtype == "d" else np.float32
    f64 = gx.dgemm64 if dtype == "d" else gx.sgemm64
    A = np.ones((4, 4), dtype=dt, order="F")
    C = np.full((4, 4), 7.0, dtype=dt, order="F")
    for args, word in [((2 ** 31, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4), "beyond"),
                       ((4, 4, -2 ** 40, 1.0, A, 4, A, 4, 0.0, C, 4), "beyond"),
                       ((4, 4, 4, 1.0, A, 3, A, 4, 0.0, C, 4), "lda"),
                       ((-1, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4), "negative")]:
        f64("N", "N", *args, check=False)
        code, msg = gx.last_error(dtype)
        assert code == gx.ERR_ARG and word in msg, (code, msg)
        np.testing.assert_array_equal(C, 7.0)
    sb = gx.dgemm_strided_batched_dev if dtype == "d" else gx.sgemm_strided_batched_dev
    fake = 1 << 20  # never dereferenced: every call below fails its checks
    for args, word in [((4, 4, 4, 1.0, fake, 4, 16, fake, 4, 16, 0.0, fake, 4, 16, -1), "batch"),
                       ((4, 4, 4, 1.0, fake, 4, 16, fake, 4, 16, 0.0, fake, 4, 15, 2), "overlap"),
                       ((4, 4, 4, 1.0, fake, 4, -16, fake, 4, 16, 0.0, fake, 4, 16, 2),
                        "negative stride"),
                       ((4, 4, 4, 1.0, fake, 3, 16, fake, 4, 16, 0.0, fake, 4, 16, 2), "lda")]:
        with pytest.raises(gx.GemmError) as ei:
            sb("N", "N", *args)
        assert ei.value.code == gx.ERR_ARG and word in str(ei.value), str(ei.value)


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_invalid_arguments_leave_c_untouched(dtype):
    dt = np.float64 if dtype == "d" else np.float32
    f = gx.dgemm if dtype == "d" else gx.sgemm
    A = np.ones((4, 4), dtype=dt, order="F")
    C = np.full((4, 4), 7.0, dtype=dt, order="F")
    for args, word in [((4, 4, 4, 1.0, A, 3, A, 4, 0.0, C, 4), "lda"),
                       ((4, 4, 4, 1.0, A, 4, A, 3, 0.0, C, 4), "ldb"),
                       ((4, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 3), "ldc"),
                       ((-1, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4), "negative"),
                       ((4, 4, 4, 1.0, None, 4, A, 4, 0.0, C, 4), "NULL")]:
        f("N", "N", *args, check=False)
        code, msg = gx.last_error(dtype)
        assert code == gx.ERR_ARG and word in msg, (code, msg)
        np.testing.assert_array_equal(C, 7.0)
    # ld checks follow the stored shape: transa T reads A as k x m
    At = np.ones((2, 5), dtype=dt, order="F")
    Cc = np.zeros((5, 3), dtype=dt, order="F")
    B = np.ones((2, 3), dtype=dt, order="F")
    f("T", "N", 5, 3, 2, 1.0, At, 1, B, 2, 0.0, Cc, 5, check=False)
    assert gx.last_error(dtype)[0] == gx.ERR_ARG


@pytest.mark.skipif(_has_gpu(), reason="needs a machine without a HIP device")
@pytest.mark.parametrize("dtype", ["d", "s"])
def test_no_device_reports_and_leaves_c_untouched(dtype):
    dt = np.float64 if dtype == "d" else np.float32
    f = gx.dgemm if dtype == "d" else gx.sgemm
    A = np.ones((4, 4), dtype=dt, order="F")
    C = np.full((4, 4), 7.0, dtype=dt, order="F")
    f("N", "N", 4, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4, check=False)
    code, msg = gx.last_error(dtype)
    assert code == gx.ERR_NODEVICE, msg
    np.testing.assert_array_equal(C, 7.0)
    with pytest.raises(gx.GemmError):
        f("N", "N", 4, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4)
    # empty problems are no-ops that succeed without a device, as in the reference
    f("N", "N", 0, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4)
    assert gx.last_error(dtype)[0] == gx.OK


def test_mode_switches():
    assert gx.load("DGEMM.so").dgemm_x_set_mode(gx.MODE_OZAKI, 9) == gx.ERR_ARG
    assert gx.load("DGEMM.so").dgemm_x_set_mode(gx.MODE_OZAKI, 7) == gx.OK
    assert gx.load("DGEMM.so").dgemm_x_set_mode(gx.MODE_NATIVE, 0) == gx.OK
    assert gx.load("SGEMM.so").sgemm_x_set_mode(gx.MODE_OZAKI) == gx.ERR_ARG
    assert gx.load("SGEMM.so").sgemm_x_set_mode(gx.MODE_NATIVE) == gx.OK
    ws = gx.dgemm_ozaki_workspace(8192, 8192, 8192, 8)
    # exponents + 8 INT8 slices of A and B (K padded to 128) + 8 INT32 weight sums
    need = 8 * 2 * 8192 * 8192 + 8 * 4 * 8192 * 8192
    assert need <= ws < need + (1 << 20)
    wb = gx.sgemm_bf16x3_workspace(1000, 700, 333)
    # hi and lo BF16 slices of A and B, K padded to 64 elements
    assert wb >= 2 * 2 * (1000 + 700) * 384


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_product_config_accepts_only_its_two_kernels(dtype):
    lib = gx.load("DGEMM.so" if dtype == "d" else "SGEMM.so")
    f = getattr(lib, f"{dtype}gemm_x_set_product_config")
    for bad in (2, 6, -2):
        assert f(bad) == gx.ERR_ARG
    for good in (0, 1, -1):
        assert f(good) == gx.OK
    with pytest.raises(gx.GemmError):
        gx.set_product_config(dtype, 3)
    gx.set_product_config(dtype, -1)


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_batch_checks_arguments_before_the_device(dtype):
    """Grouped batch (no reference counterpart): per-problem argument checks
    name the problem; a valid batch on a machine without a device reports
    ERR_NODEVICE (on a GPU machine it builds)."""
    bad = [("N", "N", 4, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 4),
           ("T", "N", 4, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 3)]
    with pytest.raises(gx.GemmError) as ei:
        gx.BatchPlan(dtype, bad)
    assert ei.value.code == gx.ERR_ARG and "problem 1" in str(ei.value)
    with pytest.raises(gx.GemmError) as ei:
        gx.BatchPlan(dtype, [("N", "N", -1, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 4)])
    assert ei.value.code == gx.ERR_ARG
    if not _has_gpu():
        with pytest.raises(gx.GemmError) as ei:
            gx.BatchPlan(dtype, bad[:1])
        assert ei.value.code == gx.ERR_NODEVICE


def test_ozaki2_bits_match_the_oracle():
    """The library's smallest Ozaki-II integer width P over vectors of k
    entries (host arithmetic, no device) equals the oracle's for every
    modulus count and length."""
    import oracle
    for nmod in range(6, 25):
        for k in (0, 1, 100, 8191, 8192, 65535, 65536, 65537, 10 ** 6):
            want = oracle.ozaki2_bits(nmod, k)
            assert gx.dgemm_ozaki2_bits(nmod, k) == want, (nmod, k)
    assert gx.dgemm_ozaki2_workspace(64, 64, 64, 7) == 0
    assert gx.dgemm_ozaki2_workspace(64, 64, 64, 20) > 20 * 64 * 128 * 2
