"""The autotuner on the GPU: it writes a table the libraries load, the choice
for each entry is the one it measured, every entry's configuration gives
results within the native tolerance against the oracle (oracle_dgemm /
oracle_sgemm), its split-K factor is the one the libraries run, and a
resumed run measures nothing again."""
import os

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
                    "--batch", "off", "--strided", "off"]) == 0
    entries, notes, arch, rocm = at.read_table(out)
    assert len(entries) == 8 and arch.startswith("gfx") and rocm == gx.build_rocm()
    capsys.readouterr()
    # resumed: nothing left to measure, same entries
    assert at.main(["--out", out, "--shapes", shapes, "--trans", "NN,TN", "--resume",
                    "--batch", "off", "--strided", "off"]) == 0
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
            bk = at.TILE_BK[d][cfg]
            kc = -(-(-(-k // split)) // bk) * bk
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
    """The autotuner's arithmetic choices for a 1024^3 bucket (DGEMM: native,
    Ozaki-II or Ozaki-I; SGEMM: native or BF16x3) are what the libraries read
    back, and dgemm_ in the automatic mode follows its entry: bit-identical to
    the oracle's Ozaki-II or Ozaki-I restatement when it chose one, within the
    native tolerance otherwise."""
    import os
    out = str(tmp_path / "tune.txt")
    assert at.main(["--out", out, "--shapes", "1024,1024,1024", "--trans", "NN",
                    "--dtypes", "d,s", "--reps", "2", "--strided", "off"]) == 0
    entries, notes, _, _ = at.read_table(out)
    scheme, cnt = entries[("emu", "NN", 10, 10, 10)]
    assert scheme in ("native", "ozaki2", "ozaki")
    # the SGEMM choice is recorded too, and the library reads it back
    s32 = entries[("emu32", "NN", 10, 10, 10)]
    assert s32 in ("native", "bf16x3")
    rng = np.random.default_rng(12)
    m = n = k = 1024
    A = np.asfortranarray((rng.uniform(0, 1, (m, k)) - 0.5) * np.exp(rng.standard_normal((m, k))))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    try:
        assert gx.load_tuning(out, "d") == len(entries)
        assert gx.tune_lookup_emulated("N", "N", m, n, k) == (
            {"ozaki2": gx.MODE_OZAKI2, "ozaki": gx.MODE_OZAKI}.get(scheme, gx.MODE_NATIVE), cnt)
        assert gx.load_tuning(out, "s") == len(entries)
        assert gx.tune_lookup_emulated32("N", "N", m, n, k) == (
            gx.MODE_BF16X3 if s32 == "bf16x3" else gx.MODE_NATIVE)
        gx.set_dgemm_mode(gx.MODE_AUTO, 0)
        G = C.copy(order="F")
        gx.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, G, m)
        R = C.copy(order="F")
        if scheme == "ozaki2":
            oracle.ozaki2_gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m, cnt)
            np.testing.assert_array_equal(G, R)
        elif scheme == "ozaki":
            oracle.ozaki_gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m, cnt)
            np.testing.assert_array_equal(G, R)
        else:
            oracle.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, R, m)
            bound = abs_bound(A, B, C, "N", "N", m, n, k, 1.0, 0.0)
            assert max_comp(G, R, bound) <= 64 * k * U64
    finally:
        gx.set_dgemm_mode(gx.MODE_NATIVE)
        for d in ("d", "s"):
            gx.load_tuning(at.DEFAULT_OUT if os.path.exists(at.DEFAULT_OUT) else None, d)


def test_arithmetic_choice_ignores_stale_c(cuda):
    """The arithmetic measurement does not depend on what C's memory held
    before: beta = 0 still multiplies C (dgemm_'s rule), so it clears C.
    With NaN left in the allocator's free blocks, native SGEMM's error stays
    finite and BF16x3 (16 of an input's 24 bits) does not match it at 1024^3,
    where native's own rounding is about a tenth of BF16x3's."""
    import re
    import torch
    n = 1024
    junk = [torch.full((n * n,), float("nan"), dtype=torch.float32, device="cuda")
            for _ in range(4)]
    del junk  # A, B and C below take these blocks
    scheme, note = at.measure_emulated32(torch, gx, "NN", n, n, n, 1)
    assert scheme == "native" and "does not match its error" in note, note
    r_fro, r_comp = (float(v) for v in re.search(r"error x(\S+) / x(\S+) of native",
                                                  note).groups())
    assert np.isfinite(r_fro) and r_fro > 1.0 and np.isfinite(r_comp), note


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
    assert best == at.batch_pick(times, ok)
    assert gx.batch_config(d) == before  # no forced value left behind


def test_table_entries_drive_both_auto_modes(cuda, tmp_path):
    """Table entries written by the autotuner choose the arithmetic of both
    automatic modes: an `emu ... ozaki 6` entry makes dgemm_ run Ozaki-I with
    six slices (bit-identical to dgemm_x_ozaki_dev on the same inputs), an
    `emu32 ... bf16x3` entry makes sgemm_ run split-BF16x3 (bit-identical to
    sgemm_x_bf16x3_dev), and `native` entries the native kernels (within the
    native tolerance of the oracle)."""
    import os
    import torch
    path = str(tmp_path / "t.txt")
    at.write_table(path, {("emu", "NN", 10, 10, 10): ("ozaki", 6),
                          ("emu32", "NN", 10, 10, 10): "bf16x3",
                          ("emu", "TN", 10, 10, 10): ("native", 0),
                          ("emu32", "TN", 10, 10, 10): "native"}, {}, None)
    m = n = k = 1024
    rng = np.random.default_rng(21)
    try:
        for d in ("d", "s"):  # each library holds its own table
            assert gx.load_tuning(path, d) == 4
        gx.set_dgemm_mode(gx.MODE_AUTO, 0)
        gx.set_sgemm_mode(gx.MODE_AUTO)
        for dt, fn in ((np.float64, gx.dgemm), (np.float32, gx.sgemm)):
            for ta in ("N", "T"):
                print(f"case {fn.__name__} auto {ta}N {m}^3", flush=True)
                A = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dt))
                B = np.asfortranarray(rng.uniform(-1, 1, (k, n)).astype(dt))
                C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
                G = C.copy(order="F")
                fn(ta, "N", m, n, k, 1.5, A, m, B, k, -0.5, G, m)
                if ta == "N":
                    dA, dB, dC = (torch.from_numpy(X.ravel(order="F")).cuda() for X in (A, B, C))
                    if dt == np.float64:
                        gx.dgemm_ozaki_dev("N", "N", m, n, k, 1.5, dA, m, dB, k, -0.5, dC, m, 6)
                    else:
                        ws = torch.empty(gx.sgemm_bf16x3_workspace(m, n, k), dtype=torch.uint8,
                                         device="cuda")
                        gx.sgemm_bf16x3_dev("N", "N", m, n, k, 1.5, dA, m, dB, k, -0.5, dC, m,
                                            ws, ws.numel())
                    torch.cuda.synchronize()
                    np.testing.assert_array_equal(G, dC.cpu().numpy().reshape((m, n), order="F"))
                else:
                    R = C.astype(np.float64, order="F")
                    oracle.gemm("T", "N", m, n, k, 1.5, A.astype(np.float64, order="F"), m,
                                B.astype(np.float64, order="F"), k, -0.5, R, m)
                    bound = abs_bound(A, B, C, "T", "N", m, n, k, 1.5, -0.5)
                    u = U64 if dt == np.float64 else U32
                    assert max_comp(G, R, bound) <= 64 * k * u
    finally:
        gx.set_dgemm_mode(gx.MODE_NATIVE)
        gx.set_sgemm_mode(gx.MODE_NATIVE)
        for d in ("d", "s"):
            gx.load_tuning(at.DEFAULT_OUT if os.path.exists(at.DEFAULT_OUT) else None, d)


@pytest.mark.parametrize("dtype", ["d", "s"])
def test_strided_measurement_picks_an_agreeing_configuration(cuda, dtype, tmp_path):
    """The strided-batch measurement times every tile configuration on one
    batch, checks each against configuration 0's result and records the
    fastest agreeing one, which the library then uses for such batches."""
    times, ok, best, count = at.measure_strided(__import__("torch"), gx, dtype, "TN", 128, 128,
                                                128, 2, count=512)
    assert count == 512 and sorted(times) == list(range(gx.TILE_CONFIGS))
    assert all(ok.values()), ok
    assert best == min(times, key=times.get)
    out = str(tmp_path / "tune.txt")
    key = ("strided", "f64" if dtype == "d" else "f32", "TN", 7, 7, 7)
    at.write_table(out, {key: best}, {}, None)
    try:
        assert gx.load_tuning(out, dtype) == 1
        assert gx.tune_lookup_strided(dtype, "T", "N", 128, 128, 128, 512) == best
    finally:
        gx.load_tuning(os.path.join(gx.LIB_DIR, "gemm_mi355x_tune.txt"), dtype)
