"""The host C layer (gemmx_*.c) and the oracle built with AddressSanitizer
and UBSan (host code only; the gfx950 objects are linked as built), driven
on the CPU through argument checking, the tuning-table reader (valid and
malformed tables), the lookups, the batch plan builder and the oracle's
restatements (tests/sanitized_checks.py, in a child process with the
sanitizer runtime preloaded).  Any report fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.slow
def test_host_layer_and_oracle_under_asan_and_ubsan(tmp_path):
    from gemm_avx2_amd import build as b
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    try:
        b.build_sanitized(str(tmp_path / "lib"))
    except RuntimeError as e:
        pytest.skip(f"no HIP objects to link against: {e}")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitized",
                        f"SAN_DIR={tmp_path / 'oracle'}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ)
    env.update({"LD_PRELOAD": asan, "GEMMX_LIB_DIR": str(tmp_path / "lib"),
                "GEMMX_ORACLE_LIB": str(tmp_path / "oracle" / "liboracle.so"),
                "GEMMX_QUIET": "1", "HIP_VISIBLE_DEVICES": "",
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:exitcode=66",
                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1:exitcode=67",
                "OMP_NUM_THREADS": "4"})
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "sanitized_checks.py")],
                       env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "sanitized checks ok" in r.stdout, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
