"""The allocation-free device entry points (dgemm_x_dev_noalloc,
sgemm_x_dev_noalloc; no counterpart in the reference) against the existing
device entry points dgemm_x_dev / sgemm_x_dev, through the C ABI: the same
kernels and split-K choice, so the same bits, on two calls in a row (the
second smaller, reusing the static scratch the first left filled), a call
that needs no scratch, and the refusal of a call whose split-K partial tiles
exceed the 64 MiB of static scratch (C untouched).  The small call is also
checked against the oracle within the native tolerance."""
import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_util import U32, U64, abs_bound, from_device, max_comp, to_device

pytestmark = pytest.mark.gpu


def _run(torch, fn, d, m, n, k, A, B, C):
    dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
    fn("N", "N", m, n, k, 1.5, dA, m, dB, k, -0.5, dC, m)
    torch.cuda.synchronize()
    return from_device(dC, C.shape)


@pytest.mark.parametrize("d", ["d", "s"])
def test_noalloc_matches_the_device_entry(cuda, d):
    import torch
    dt = np.float64 if d == "d" else np.float32
    dev = gx.dgemm_dev if d == "d" else gx.sgemm_dev
    noalloc = gx.dgemm_dev_noalloc if d == "d" else gx.sgemm_dev_noalloc
    rng = np.random.default_rng(21)
    # 512 x 512 x 16384 and 200 x 136 x 3000 split K; 2048 x 2048 x 64 does not
    for (m, n, k) in [(512, 512, 16384), (200, 136, 3000), (2048, 2048, 64)]:
        splits, need = gx.tune_lookup_split(d, "N", "N", m, n, k)
        print(f"case {d} {m}x{n}x{k} splits {splits} scratch {need}", flush=True)
        assert need <= 64 << 20 and (splits > 1) == (k > 64)
        A = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dt))
        B = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(dt))
        C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
        want = _run(torch, dev, d, m, n, k, A, B, C)
        got = _run(torch, noalloc, d, m, n, k, A, B, C)
        np.testing.assert_array_equal(got, want)
        if m * n * k <= 200 * 136 * 3000:
            R = C.astype(np.float64, order="F")
            oracle.gemm("N", "N", m, n, k, 1.5, A.astype(np.float64, order="F"), m,
                        B.astype(np.float64, order="F"), k, -0.5, R, m)
            bound = abs_bound(A, B, C, "N", "N", m, n, k, 1.5, -0.5)
            assert max_comp(got, R, bound) <= 64 * k * (U64 if d == "d" else U32)


@pytest.mark.parametrize("d", ["d", "s"])
def test_noalloc_refuses_more_than_its_scratch(cuda, d):
    import torch
    dt = np.float64 if d == "d" else np.float32
    noalloc = gx.dgemm_dev_noalloc if d == "d" else gx.sgemm_dev_noalloc
    m = n = 2048
    k = 4096
    rng = np.random.default_rng(22)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dt))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(dt))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
    try:
        gx.set_tile_config(d, 0)
        gx.set_split_config(d, 64)  # 64 ranges of 256 C tiles of 128 x 128: far past 64 MiB
        assert gx.tune_lookup_split(d, "N", "N", m, n, k)[1] > 64 << 20
        dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
        with pytest.raises(gx.GemmError) as ei:
            noalloc("N", "N", m, n, k, 1.5, dA, m, dB, k, -0.5, dC, m)
        assert ei.value.code == gx.ERR_UNSUPPORTED
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_device(dC, C.shape), C)
    finally:
        gx.set_tile_config(d, -1)
        gx.set_split_config(d, -1)
