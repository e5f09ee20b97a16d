"""The benchmark's host side on the CPU: it refuses more than one GPU
before touching a device, and its grouped-batch workload (the batch mix of
the measurements) is the fixed, seeded problem set the profiles and DESIGN
numbers refer to."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
    assert not any(line.startswith("{") for line in r.stdout.splitlines())


def test_batch_mix_is_fixed():
    shapes, offs = bench.batch_shapes()
    assert len(shapes) == bench.BATCH_COUNT == 4096
    assert shapes == bench.batch_shapes()[0]  # seeded: the same problems every time
    flops = sum(2 * m * n * k for _, _, m, n, k, *_ in shapes)
    assert abs(flops / 1e9 - 110.232) < 1e-3
    for ta, tb, m, n, k, alpha, beta, (oa, ob, oc) in shapes:
        assert ta in "NT" and tb in "NT"
        assert bench.BATCH_LO <= min(m, n, k) and max(m, n, k) <= bench.BATCH_HI
        assert 0.5 <= alpha <= 2.0 and -1.0 <= beta <= 1.0
        assert oa % 2 == ob % 2 == oc % 2 == 0  # 16-byte aligned starts for FP64
    # the three operand buffers hold every problem back to back
    assert offs == [sum((x + 1) // 2 * 2 for x in col) for col in
                    zip(*[(m * k, k * n, m * n) for _, _, m, n, k, *_ in shapes])]
