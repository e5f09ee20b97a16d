"""Device-allocation failure handling on the GPU, through the C ABI.

GEMMX_FAIL_ALLOC=1 makes every device allocation the library attempts fail
as if the device were out of memory.  The reference allocates nothing on a
device and checks nothing (src/gemm_driver.c:135-142), so this is the
failure path of this build's own additions: the device scratch that host
operands and library-workspace emulation are staged in, the stream-ordered
pool of split-K partial tiles, and the grouped-batch plan.  The checks run in
a child process, so that its library state starts empty (a scratch that has
already grown would not be reallocated) and the switch is read from its
environment:

* calls that need a new allocation return GEMMX_ERR_ALLOC and leave C
  untouched, bit for bit (host C, device C, every transpose pair);
* calls that need none still run (device operands, no split), and a split-K
  call whose partial tiles cannot be allocated runs in one pass with the
  native tolerance against the oracle;
* a batch plan is refused with GEMMX_ERR_ALLOC;
* once the switch is cleared, the same calls succeed and match the oracle.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.gpu
def test_injected_allocation_failures(cuda):
    env = dict(os.environ, GEMMX_FAIL_ALLOC="1", GEMMX_QUIET="1")
    r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__)], env=env, cwd=HERE,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "fail-alloc checks passed" in r.stdout


def _child():
    sys.path.insert(0, os.path.dirname(HERE))
    sys.path.insert(0, HERE)
    import torch
    import gemm_avx2_amd as gx
    import oracle
    from gemm_util import U32, U64, abs_bound, max_comp, rel_fro, stored_shape

    assert os.environ.get("GEMMX_FAIL_ALLOC") == "1"
    rng = np.random.default_rng(5)

    def case(dt, ta, tb, m, n, k):
        ra, ca = stored_shape(ta, m, k)
        rb, cb = stored_shape(tb, k, n)
        A = np.asfortranarray(rng.uniform(-1, 1, (ra, ca)).astype(dt))
        B = np.asfortranarray(rng.uniform(-1, 1, (rb, cb)).astype(dt))
        C = np.asfortranarray(rng.uniform(-1, 1, (m + 3, n)).astype(dt))
        C[m:] = np.nan  # rows past m: never written
        return A, B, C, ra, rb

    def expect(code, fn, *args, **kw):
        try:
            fn(*args, **kw)
        except gx.GemmError as e:
            assert e.code == code, (fn.__name__, e.code, str(e))
            return
        raise AssertionError(f"{fn.__name__} did not fail")

    def check(dt, ta, tb, m, n, k, A, B, C0, G, alpha=1.5, beta=-0.5):
        R = C0.astype(np.float64, order="F")
        oracle.gemm(ta, tb, m, n, k, alpha, A.astype(np.float64, order="F"), A.shape[0],
                    B.astype(np.float64, order="F"), B.shape[0], beta, R, C0.shape[0])
        bound = abs_bound(A, B, C0, ta, tb, m, n, k, alpha, beta)
        u = U64 if dt == np.float64 else U32
        assert max_comp(G[:m], R[:m], bound) <= 64 * k * u
        assert rel_fro(G[:m], R[:m]) <= (1e-14 if dt == np.float64 else 1e-5)
        assert np.all(np.isnan(G[m:]))

    def same_bits(x, y):
        return x.tobytes() == y.tobytes()

    trans = [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")]
    for dt, fn in ((np.float64, gx.dgemm), (np.float32, gx.sgemm)):
        for ta, tb in trans:
            m, n, k = 97, 70, 130
            A, B, C, ra, rb = case(dt, ta, tb, m, n, k)
            # host operands: the device scratch they are staged in cannot grow
            Ch = C.copy(order="F")
            print(f"host {fn.__name__} {ta}{tb}", flush=True)
            expect(gx.ERR_ALLOC, fn, ta, tb, m, n, k, 1.5, A, ra, B, rb, -0.5, Ch, m + 3)
            assert same_bits(Ch, C), "C changed by a failed call"
            # device operands, one pass: nothing to allocate, the call runs
            tdt = torch.float64 if dt == np.float64 else torch.float32
            dA, dB = (torch.from_numpy(X.ravel(order="F")).to("cuda") for X in (A, B))
            dC = torch.from_numpy(C.ravel(order="F")).to("cuda")
            assert dA.dtype == tdt
            print(f"device {fn.__name__} {ta}{tb}", flush=True)
            fn(ta, tb, m, n, k, 1.5, dA, ra, dB, rb, -0.5, dC, m + 3)
            torch.cuda.synchronize()
            check(dt, ta, tb, m, n, k, A, B, C,
                  dC.cpu().numpy().reshape(C.shape, order="F"))

    # split-K on device pointers: the pool cannot give the partial tiles, so
    # the product runs in one pass
    for d, dt, dev in (("d", np.float64, gx.dgemm_dev), ("s", np.float32, gx.sgemm_dev)):
        m, n, k = 64, 48, 6000
        A, B, C, ra, rb = case(dt, "N", "T", m, n, k)
        gx.set_split_config(d, 8)
        try:
            assert gx.tune_lookup_split(d, "N", "T", m, n, k)[0] == 8
            dA, dB = (torch.from_numpy(X.ravel(order="F")).to("cuda") for X in (A, B))
            dC = torch.from_numpy(C.ravel(order="F")).to("cuda")
            print(f"split {d} without partials", flush=True)
            dev("N", "T", m, n, k, 1.5, dA, ra, dB, rb, -0.5, dC, m + 3)
            torch.cuda.synchronize()
            check(dt, "N", "T", m, n, k, A, B, C, dC.cpu().numpy().reshape(C.shape, order="F"))
        finally:
            gx.set_split_config(d, -1)

    # emulated DGEMM with the library's workspace: refused, device C untouched
    m, n, k = 80, 60, 200
    A, B, C, ra, rb = case(np.float64, "N", "N", m, n, k)
    dA, dB = (torch.from_numpy(X.ravel(order="F")).to("cuda") for X in (A, B))
    dC = torch.from_numpy(C.ravel(order="F")).to("cuda")
    print("ozaki2 library workspace", flush=True)
    expect(gx.ERR_ALLOC, gx.dgemm_ozaki2_dev, "N", "N", m, n, k, 1.0, dA, ra, dB, rb, 0.0, dC,
           m + 3, 16)
    expect(gx.ERR_ALLOC, gx.dgemm_ozaki_dev, "N", "N", m, n, k, 1.0, dA, ra, dB, rb, 0.0, dC,
           m + 3, 6)
    torch.cuda.synchronize()
    assert same_bits(dC.cpu().numpy(), C.ravel(order="F"))
    # dgemm_ routed to Ozaki-II: its workspace comes from the same scratch
    gx.set_dgemm_mode(gx.MODE_OZAKI2, 16)
    try:
        expect(gx.ERR_ALLOC, gx.dgemm, "N", "N", m, n, k, 1.0, dA, ra, dB, rb, 0.0, dC, m + 3)
    finally:
        gx.set_dgemm_mode(gx.MODE_NATIVE)
    torch.cuda.synchronize()
    assert same_bits(dC.cpu().numpy(), C.ravel(order="F"))

    # a grouped batch plan: its descriptors cannot be uploaded
    print("batch plan", flush=True)
    expect(gx.ERR_ALLOC, gx.BatchPlan, "d",
           [("N", "N", m, n, k, 1.0, dA, ra, dB, rb, 0.0, dC, m + 3)])

    # the switch cleared: the refused calls now run and match the oracle
    os.environ["GEMMX_FAIL_ALLOC"] = "0"
    A2, B2, C2, ra2, rb2 = case(np.float64, "T", "N", 97, 70, 130)
    Ch = C2.copy(order="F")
    print("host dgemm after recovery", flush=True)
    gx.dgemm("T", "N", 97, 70, 130, 1.5, A2, ra2, B2, rb2, -0.5, Ch, 100)
    check(np.float64, "T", "N", 97, 70, 130, A2, B2, C2, Ch)
    print("batch plan after recovery", flush=True)
    plan = gx.BatchPlan("d", [("N", "N", m, n, k, 1.5, dA, ra, dB, rb, -0.5, dC, m + 3)])
    plan.run()
    torch.cuda.synchronize()
    plan.destroy()
    check(np.float64, "N", "N", m, n, k, A, B, C, dC.cpu().numpy().reshape(C.shape, order="F"))
    print("fail-alloc checks passed", flush=True)


if __name__ == "__main__":
    _child()
