"""Parity at full size through the C ABI.

Emulated and native DGEMM, m = n = k = 8192, NN, ld 8192, alpha 1, beta 0,
A and B entries (u - 0.5) exp(phi g), phi in {0, 1, 4}, seed 3 (the emu8192
cases of tests/golden: the reference's own dgemm_ outputs at 2048 sampled
entries per phi, computed from the same generator on the host).  At the
sampled entries, with the whole rows of op(A) and columns of op(B) read back
from the device (the device generator's exp/log/cos may differ from the host
library's in the last bit, so the exact values are formed from what the GPU
multiplied):
  * Ozaki-II (14..22 moduli) and Ozaki-I (8 slices) equal the oracle's
    sampled-entry restatements bit for bit;
  * native FP64 is within 64 k u (|A||B|) componentwise and 1e-14 relative
    Frobenius of the Dot2 values and of the reference's values, and so is
    Ozaki-II from 17 moduli on (at phi = 4, 16 moduli reach 1.6e-14);
  * the error falls with the modulus count, and 20 moduli are no less
    accurate than native FP64 at any phi, in relative Frobenius and in the
    largest componentwise error (relative to |A||B|) over the 2048 entries;
  * the arithmetic dgemm_ takes in automatic mode at this shape is no less
    accurate than native FP64 at any phi on both measures.

Split-BF16x3 and native SGEMM, 16384^3 NN and 65536 x 256 x 256 TN, alpha 1,
beta 0, entries U[-1,1) and log-normal exp(2g): 512 sampled entries against
Dot2 values of the same (device) inputs,
  native:  componentwise <= 64 k 2^-24 (|A||B|), relative Frobenius <= 1e-5;
  BF16x3:  componentwise <= 3 2^-16 + 64 k 2^-24 (the split keeps 16 of the
           24 significand bits of each operand), relative Frobenius <= 1e-4.
"""
import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_util import U32, U64, load_fixtures

pytestmark = pytest.mark.gpu
EMU = [c for c in load_fixtures() if c["group"] == "emu8192"]
MODULI = (14, 15, 16, 17, 18, 19, 20, 21, 22)
SLICES = 8


def _errors(got, exact, ab):
    err = np.abs(got - exact)
    return {"rel_fro": float(np.linalg.norm(err) / np.linalg.norm(exact)),
            "max_comp": float(np.max(err / (ab + 1e-300)))}


@pytest.fixture(scope="module")
def dgemm_buffers(cuda):
    import torch
    n = 8192
    ws = max(gx.dgemm_ozaki2_workspace(n, n, n, max(MODULI)),
             gx.dgemm_ozaki_workspace(n, n, n, SLICES))
    bufs = {"A": torch.empty(n * n, dtype=torch.float64, device="cuda"),
            "B": torch.empty(n * n, dtype=torch.float64, device="cuda"),
            "C": torch.empty(n * n, dtype=torch.float64, device="cuda"),
            "ws": torch.empty(ws, dtype=torch.uint8, device="cuda")}
    yield bufs
    del bufs
    torch.cuda.empty_cache()


@pytest.mark.parametrize("case", EMU, ids=lambda c: f"phi{c['param']:g}")
def test_dgemm_8192_emulated_and_native(dgemm_buffers, case):
    import torch
    m, n, k = case["m"], case["n"], case["k"]
    assert (m, n, k, case["lda"], case["ldb"], case["ldc"]) == (8192,) * 6
    assert case["transa"] == case["transb"] == "N" and case["alpha"] == 1 and case["beta"] == 0
    seed, phi = case["seed"], case["param"]
    print(f"case dgemm 8192^3 NN phi={phi:g} seed={seed}", flush=True)
    A, B, C, ws = (dgemm_buffers[x] for x in ("A", "B", "C", "ws"))
    gx.fill(A, m, k, m, seed, 0, 1, phi)
    gx.fill(B, k, n, k, seed, 1, 1, phi)
    torch.cuda.synchronize()

    is_, js = case["is"], case["js"]
    ui, ri = np.unique(is_, return_inverse=True)
    uj, cj = np.unique(js, return_inverse=True)
    rows = A.view(k, m)[:, torch.from_numpy(ui).cuda()].T.contiguous().cpu().numpy()
    cols = B.view(n, k)[torch.from_numpy(uj).cuda()].contiguous().cpu().numpy()
    exact, ab = oracle.dot2_pairs(rows, cols, ri, cj, 1.0, 0.0)
    flat = torch.from_numpy(is_.astype(np.int64) + js.astype(np.int64) * m).cuda()
    tol = 64 * k * U64

    def sample(fn):
        gx.fill(C, m, n, m, seed, 2)  # beta = 0: C is multiplied by 0, not read as data
        fn()
        torch.cuda.synchronize()
        return C[flat].cpu().numpy()

    errs = {}
    # native FP64 MFMA: the reference's tolerance against Dot2 and its own values
    nat = sample(lambda: gx.dgemm_dev("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m))
    errs["native"] = _errors(nat, exact, ab)
    ref = _errors(nat, case["values"], ab)
    assert errs["native"]["max_comp"] <= tol and errs["native"]["rel_fro"] <= 1e-14
    assert ref["max_comp"] <= tol and ref["rel_fro"] <= 1e-14
    # Ozaki-II: bit for bit the restatement; FP64 tolerance from 17 moduli
    results = {}
    for nm in MODULI:
        got = sample(lambda nm=nm: gx.dgemm_ozaki2_dev("N", "N", m, n, k, 1.0, A, m, B, k, 0.0,
                                                        C, m, nm, ws, ws.numel()))
        want = oracle.ozaki2_pairs(rows, cols, ri, cj, 1.0, 0.0, nm)
        np.testing.assert_array_equal(got, want, err_msg=f"ozaki2 n{nm} phi {phi:g}")
        errs[f"ozaki2_n{nm}"] = _errors(got, exact, ab)
        results[nm] = got
        if nm >= 17:
            e_ref = _errors(got, case["values"], ab)
            assert errs[f"ozaki2_n{nm}"]["max_comp"] <= tol and e_ref["max_comp"] <= tol, nm
            assert errs[f"ozaki2_n{nm}"]["rel_fro"] <= 1e-14 and e_ref["rel_fro"] <= 1e-14, nm
    # Ozaki-I: bit for bit the restatement
    got = sample(lambda: gx.dgemm_ozaki_dev("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m,
                                            SLICES, ws, ws.numel()))
    np.testing.assert_array_equal(got, oracle.ozaki_pairs(rows, cols, ri, cj, 1.0, 0.0, SLICES))
    errs[f"ozaki_s{SLICES}"] = _errors(got, exact, ab)
    for name, e in errs.items():
        print(f"  phi={phi:g} {name:10s} rel_fro={e['rel_fro']:.3e} max_comp={e['max_comp']:.3e}",
              flush=True)

    for a, b in zip(MODULI, MODULI[1:]):
        for meas in ("rel_fro", "max_comp"):
            assert errs[f"ozaki2_n{b}"][meas] <= errs[f"ozaki2_n{a}"][meas], (a, b, meas)
    for meas in ("rel_fro", "max_comp"):
        assert errs["ozaki2_n20"][meas] <= errs["native"][meas], meas
    # dgemm_ in automatic mode (device operands through the Fortran ABI)
    mode, count = gx.tune_lookup_emulated("N", "N", m, n, k)
    try:
        gx.set_dgemm_mode(gx.MODE_AUTO)
        auto = sample(lambda: gx.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m))
    finally:
        gx.set_dgemm_mode(gx.MODE_NATIVE)
    if mode == gx.MODE_OZAKI2:
        np.testing.assert_array_equal(auto, results[count] if count in results else
                                      oracle.ozaki2_pairs(rows, cols, ri, cj, 1.0, 0.0, count))
    else:
        assert mode == gx.MODE_NATIVE
        np.testing.assert_array_equal(auto, nat)
    e_auto = _errors(auto, exact, ab)
    print(f"  phi={phi:g} auto mode {mode} count {count}: rel_fro={e_auto['rel_fro']:.3e} "
          f"max_comp={e_auto['max_comp']:.3e}", flush=True)
    assert e_auto["rel_fro"] <= errs["native"]["rel_fro"]
    assert e_auto["max_comp"] <= errs["native"]["max_comp"]


BF16_SHAPES = [("N", "N", 16384, 16384, 16384), ("T", "N", 65536, 256, 256)]


@pytest.mark.parametrize("dist,param", [(0, 0.0), (2, 2.0)], ids=["uniform", "lognormal2"])
@pytest.mark.parametrize("ta,tb,m,n,k", BF16_SHAPES,
                         ids=lambda v: str(v) if not isinstance(v, str) else v)
def test_sgemm_bf16x3_and_native_full_size(cuda, ta, tb, m, n, k, dist, param):
    import torch
    print(f"case sgemm {ta}{tb} {m}x{n}x{k} dist={dist} param={param}", flush=True)
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    A = torch.empty(ra * ca, dtype=torch.float32, device="cuda")
    B = torch.empty(rb * cb, dtype=torch.float32, device="cuda")
    C = torch.empty(m * n, dtype=torch.float32, device="cuda")
    ws = torch.empty(gx.sgemm_bf16x3_workspace(m, n, k), dtype=torch.uint8, device="cuda")
    gx.fill(A, ra, ca, ra, 5, 0, dist, param)
    gx.fill(B, rb, cb, rb, 5, 1, dist, param)
    torch.cuda.synchronize()
    rng = np.random.default_rng(17)
    is_, js = rng.integers(0, m, 512), rng.integers(0, n, 512)
    ui, ri = np.unique(is_, return_inverse=True)
    uj, cj = np.unique(js, return_inverse=True)
    tui, tuj = torch.from_numpy(ui).cuda(), torch.from_numpy(uj).cuda()
    # op(A) row i: A(i, :) (N, column-major m x k) or A(:, i) (T, stored k x m)
    rows = (A.view(k, m)[:, tui].T if ta == "N" else A.view(m, k)[tui]).double()
    cols = (B.view(n, k)[tuj] if tb == "N" else B.view(k, n)[:, tuj].T).double()
    exact, ab = oracle.dot2_pairs(rows.contiguous().cpu().numpy(), cols.contiguous().cpu().numpy(),
                                  ri, cj, 1.0, 0.0)
    flat = torch.from_numpy(is_.astype(np.int64) + js.astype(np.int64) * m).cuda()
    out = {}
    for form in ("native", "bf16x3"):
        gx.fill(C, m, n, m, 5, 2)
        if form == "native":
            gx.sgemm_dev(ta, tb, m, n, k, 1.0, A, ra, B, rb, 0.0, C, m)
        else:
            gx.sgemm_bf16x3_dev(ta, tb, m, n, k, 1.0, A, ra, B, rb, 0.0, C, m, ws, ws.numel())
        torch.cuda.synchronize()
        out[form] = _errors(C[flat].double().cpu().numpy(), exact, ab)
        print(f"  {form:7s} rel_fro={out[form]['rel_fro']:.3e} "
              f"max_comp={out[form]['max_comp']:.3e}", flush=True)
    assert out["native"]["max_comp"] <= 64 * k * U32 and out["native"]["rel_fro"] <= 1e-5
    assert out["bf16x3"]["max_comp"] <= 3 * 2.0 ** -16 + 64 * k * U32
    assert out["bf16x3"]["rel_fro"] <= 1e-4
    del A, B, C, ws
    torch.cuda.empty_cache()
