"""The edge-tile load index math (gemm_avx2_amd/csrc/gemmx_index.h, used by
the tile kernels' clamped loads) checked on the host: tests/clamp_index_check.cpp
built with the host C++ compiler against that header and run.  See the
checker's header for what it asserts (valid chunks keep their place; reads
past an operand's minimal extent only in the last line's aligned 16-byte
block)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "gemm_avx2_amd", "csrc")


def test_clamped_load_indices(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "clamp_index_check")
    subprocess.run([cxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I", CSRC,
                    os.path.join(HERE, "clamp_index_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout
