"""The autotuner on the GPU: it writes a table the libraries load, the choice
for each entry is the one it measured, every entry's configuration gives
results within the native tolerance against the oracle (oracle_dgemm /
oracle_sgemm), and a resumed run measures nothing again."""
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
    assert at.main(["--out", out, "--shapes", shapes, "--trans", "NN,TN", "--reps", "2"]) == 0
    entries, notes, arch = at.read_table(out)
    assert len(entries) == 8 and arch.startswith("gfx")
    capsys.readouterr()
    # resumed: nothing left to measure, same entries
    assert at.main(["--out", out, "--shapes", shapes, "--trans", "NN,TN", "--resume"]) == 0
    assert "[1/" not in capsys.readouterr().out
    assert at.read_table(out)[0] == entries
    rng = np.random.default_rng(5)
    try:
        for d in ("d", "s"):
            assert gx.load_tuning(out, d) == 8
        for (dt, tr, mb, nb, kb), cfg in sorted(entries.items()):
            d = "d" if dt == "f64" else "s"
            npdt = np.float64 if d == "d" else np.float32
            m, n, k = (int(v) for v in notes[(dt, tr, mb, nb, kb)].split()[0].split("x"))
            assert gx.tune_lookup(d, tr[0], tr[1], m, n, k) == cfg
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
