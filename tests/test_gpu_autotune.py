"""The autotuner on the GPU: it writes a table the libraries load, the choice
for each entry is the one it measured, every entry's configuration gives
results within the native tolerance against the oracle (oracle_dgemm /
oracle_sgemm), its split-K factor is the one the libraries run, and a
resumed run measures nothing again."""
import numpy as np
import pytest

import gemm_avx2_amd as gx
import oracle
from gemm_avx2_amd import autotune as at
from gemm_util import U32, U64, abs_bound, from_device, max_comp, stored_shape, to_device

pytestmark = pytest.mark.gpu


def test_autotune_table_is_used_and_correct(cuda, tmp_path, capsys):
    import torch
    out = str(tmp_path / "tune.txt")
    shapes = "256,256,256;640,200,300"
    assert at.main(["--out", out, "--shapes", shapes, "--trans", "NN,TN", "--reps", "2",
                    "--batch", "off"]) == 0
    entries, notes, arch, rocm = at.read_table(out)
    assert len(entries) == 8 and arch.startswith("gfx") and rocm == gx.build_rocm()
    capsys.readouterr()
    # resumed: nothing left to measure, same entries
    assert at.main(["--out", out, "--shapes", shapes, "--trans", "NN,TN", "--resume",
                    "--batch", "off"]) == 0
    assert "[1/" not in capsys.readouterr().out
    assert at.read_table(out)[0] == entries
    rng = np.random.default_rng(5)
    try:
        for d in ("d", "s"):
            assert gx.load_tuning(out, d) == 8
        for (dt, tr, mb, nb, kb), (cfg, split) in sorted(entries.items()):
            d = "d" if dt == "f64" else "s"
            npdt = np.float64 if d == "d" else np.float32
            m, n, k = (int(v) for v in notes[(dt, tr, mb, nb, kb)].split()[0].split("x"))
            assert gx.tune_lookup(d, tr[0], tr[1], m, n, k) == cfg
            # the measured split: K ranges of whole 16-deep slices
            kc = -(-(-(-k // split)) // 16) * 16
            want = -(-k // kc) if cfg in at.SPLIT_CONFIGS and split > 1 else 1
            assert gx.tune_lookup_split(d, tr[0], tr[1], m, n, k)[0] == want
            ra, ca = stored_shape(tr[0], m, k)
            rb, cb = stored_shape(tr[1], k, n)
            A = np.asfortranarray(rng.uniform(-1, 1, (ra, ca)).astype(npdt))
            B = np.asfortranarray(rng.uniform(-1, 1, (rb, cb)).astype(npdt))
            C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(npdt))
            R = C.copy(order="F")
            oracle.gemm(tr[0], tr[1], m, n, k, 1.25, A, ra, B, rb, 0.5, R, m)
            dA, dB, dC = to_device(A, torch), to_device(B, torch), to_device(C, torch)
            (gx.dgemm_dev if d == "d" else gx.sgemm_dev)(tr[0], tr[1], m, n, k, 1.25, dA, ra,
                                                           dB, rb, 0.5, dC, m)
            torch.cuda.synchronize()
            G = from_device(dC, C.shape)
            tol = 64 * k * (U64 if d == "d" else U32)
            assert max_comp(G, R, abs_bound(A, B, C, tr[0], tr[1], m, n, k, 1.25, 0.5)) <= tol
    finally:
        import os
        for d in ("d", "s"):
            gx.load_tuning(at.DEFAULT_OUT if os.path.exists(at.DEFAULT_OUT) else None, d)


def test_autotune_emulated_choice_drives_auto_mode(cuda, tmp_path):
    """The autotuner's DGEMM arithmetic choice for a 1024^3 bucket is what
    dgemm_ uses in the automatic mode: bit-identical to the oracle's Ozaki-II
    restatement when it chose Ozaki-II, within the native tolerance
    otherwise."""
    import os
    out = str(tmp_path / "tune.txt")
    assert at.main(["--out", out, "--shapes", "1024,1024,1024", "--trans", "NN",
                    "--dtypes", "d", "--reps", "2"]) == 0
    entries, notes, _, _ = at.read_table(out)
    scheme, cnt = entries[("emu", "NN", 10, 10, 10)]
    assert scheme in ("native", "ozaki2")
    rng = np.random.default_rng(12)
    m = n = k = 1024
    A = np.asfortranarray((rng.uniform(0, 1, (m, k)) - 0.5) * np.exp(rng.standard_normal((m, k))))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    try:
        assert gx.load_tuning(out, "d") == len(entries)
        assert gx.tune_lookup_emulated("N", "N", m, n, k) == (
            gx.MODE_OZAKI2 if scheme == "ozaki2" else gx.MODE_NATIVE, cnt)
        gx.set_dgemm_mode(gx.MODE_AUTO, 0)
        G = C.copy(order="F")
        gx.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, G, m)
        R = C.copy(order="F")
        if scheme == "ozaki2":
            oracle.ozaki2_gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m, cnt)
            np.testing.assert_array_equal(G, R)
        else:
            oracle.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m)
            bound = abs_bound(A, B, C, "N", "N", m, n, k, 1.0, 0.0)
            assert max_comp(G, R, bound) <= 64 * k * U64
    finally:
        gx.set_dgemm_mode(gx.MODE_NATIVE)
        gx.load_tuning(at.DEFAULT_OUT if os.path.exists(at.DEFAULT_OUT) else None, "d")


@pytest.mark.parametrize("d", ["d", "s"])
def test_autotune_batch_measurement(cuda, d):
    """The grouped-batch measurement times every kernel configuration on a
    mixed batch (here 96 problems up to 256), checks each configuration's C
    against configuration 0's and returns the fastest agreeing one; the
    forced configuration is released afterwards."""
    import torch
    before = gx.batch_config(d)
    times, ok, best = at.measure_batch(torch, gx, d, 2, mix=(96, 16, 256))
    assert len(times) == at.BATCH_CONFIGS and all(ok) and 0 <= best < at.BATCH_CONFIGS
    assert times[best] == min(times)
    assert gx.batch_config(d) == before  # no forced value left behind
