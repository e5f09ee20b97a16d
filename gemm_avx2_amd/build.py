"""Builds the gfx950 libraries of this package in-tree.

Outputs (gemm_avx2_amd/lib/):
  DGEMM.so         dgemm_, its ILP64 form dgemm_64_, the dgemm_x_* extensions, the
                   double complex zgemm_ / zgemm3m_ and dtrmm_ / dtrsm_
  SGEMM.so         sgemm_, sgemm_64_, the sgemm_x_* extensions, cgemm_ / cgemm3m_,
                   strmm_ / strsm_
  libgemmx_gen.so  device-side input generator used by bench.py and the tests
  gemm_mi355x_tune.txt  the shipped gfx950 tuning table (tuning/), if present

Each HIP source is compiled once with hipcc for gfx950; the host C layer is
compiled with gcc; the two BLAS libraries are linked from the same objects and
differ only in their export list (the reference builds one source tree twice,
-DDOUBLE selecting dgemm_, wjc404/GEMM_AVX2 Makefile:14-18).

    python -m gemm_avx2_amd.build [-j N] [--verbose]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIP_SOURCES = ["gemm_native.hip"] + [f"gemm_native_{t}_{tr}.hip" for t in ("f64", "f32")
                                      for tr in ("nn", "nt", "tn", "tt")] + [
    "gemm_split.hip", "gemm_products.hip", "gemm_emulated.hip", "gemm_batch.hip",
    "gemm_complex.hip", "gemm_tri.hip"]
C_SOURCES = ["gemmx_api.c", "gemmx_tune.c", "gemmx_batch.c", "gemmx_stage.c"]
GEN_SOURCES = ["gemmx_gen.hip"]
# tuning table measured on MI355X by gemm_avx2_amd.autotune (--set full)
SHIPPED_TABLE = os.path.join(HERE, "tuning", "gemm_mi355x_tune_gfx950.txt")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-fvisibility=hidden", "-ffp-contract=fast",
    # Packed FP32 arithmetic is kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
CC_FLAGS = [
    "-O2", "-fPIC", "-std=gnu11", "-fvisibility=hidden", "-Wall", "-Wextra",
    "-Wno-unused-parameter", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
]
EXPORTS = {
    "DGEMM.so": ["dgemm_", "dgemm_64_", "dgemm_x_*", "zgemm_", "zgemm3m_", "dtrmm_", "dtrsm_"],
    "SGEMM.so": ["sgemm_", "sgemm_64_", "sgemm_x_*", "cgemm_", "cgemm3m_", "strmm_", "strsm_"],
}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... ({r.returncode})")
    return r


def _hipcc():
    p = os.path.join(ROCM, "bin", "hipcc")
    return p if os.path.exists(p) else shutil.which("hipcc")


def build(jobs: int = 8, verbose: bool = False, lib_dir: str = LIB, build_dir: str = BUILD,
          extra_hip_flags=()) -> dict:
    """Compiles every source and links the libraries; returns their paths."""
    LIB, BUILD = lib_dir, build_dir
    hipcc = _hipcc()
    if not hipcc:
        raise RuntimeError("hipcc not found (ROCm is required to build)")
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)

    def obj(src):
        return os.path.join(BUILD, os.path.splitext(src)[0] + ".o")

    steps = []
    for src in HIP_SOURCES + GEN_SOURCES:
        steps.append([hipcc, *HIPCC_FLAGS, *extra_hip_flags, "-c", os.path.join(CSRC, src),
                      "-o", obj(src)])
    for src in C_SOURCES:
        steps.append(["gcc", *CC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj(src)])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), steps))

    outputs = {}
    objs = [obj(s) for s in HIP_SOURCES + C_SOURCES]
    for lib, syms in EXPORTS.items():
        vs = os.path.join(BUILD, lib + ".map")
        with open(vs, "w") as f:
            f.write("{\n  global:\n" + "".join(f"    {s};\n" for s in syms) + "  local: *;\n};\n")
        out = os.path.join(LIB, lib)
        tmp = out + ".tmp"
        _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
              f"-Wl,--version-script={vs}", "-Wl,-Bsymbolic", "-Wl,-z,defs",
              f"-L{ROCM}/lib", "-lamdhip64", "-lpthread", "-ldl"], verbose)
        os.replace(tmp, out)
        outputs[lib] = out
    gen = os.path.join(LIB, "libgemmx_gen.so")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", gen + ".tmp",
          *[obj(s) for s in GEN_SOURCES], f"-L{ROCM}/lib", "-lamdhip64"], verbose)
    os.replace(gen + ".tmp", gen)
    outputs["libgemmx_gen.so"] = gen
    # the shipped tuning table goes beside the libraries, where they look first
    # (a table written there by the autotuner is replaced on every build)
    if os.path.exists(SHIPPED_TABLE):
        table = os.path.join(LIB, "gemm_mi355x_tune.txt")
        shutil.copyfile(SHIPPED_TABLE, table)
        outputs["gemm_mi355x_tune.txt"] = table
    return outputs


SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fno-omit-frame-pointer", "-g"]


def build_sanitized(out_dir: str, build_dir: str = BUILD, verbose: bool = False) -> dict:
    """The host C layer compiled with AddressSanitizer and UBSan (gcc, host
    code only) and linked with the gfx950 objects of the last build() into
    out_dir/{DGEMM,SGEMM}.so, for the CPU-side sanitizer tests.  The HIP
    objects are used as built (GPU code is never sanitized here)."""
    hip_objs = [os.path.join(build_dir, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES]
    missing = [o for o in hip_objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError(f"build() first: {missing[0]} is missing")
    os.makedirs(out_dir, exist_ok=True)
    objs = []
    for src in C_SOURCES:
        o = os.path.join(out_dir, os.path.splitext(src)[0] + ".san.o")
        _run(["gcc", *CC_FLAGS, *SAN_FLAGS, "-c", os.path.join(CSRC, src), "-o", o], verbose)
        objs.append(o)
    outputs = {}
    for lib, syms in EXPORTS.items():
        vs = os.path.join(out_dir, lib + ".map")
        with open(vs, "w") as f:
            f.write("{\n  global:\n" + "".join(f"    {s};\n" for s in syms) + "  local: *;\n};\n")
        out = os.path.join(out_dir, lib)
        _run(["g++", "-shared", "-fPIC", *SAN_FLAGS, "-o", out, *hip_objs, *objs,
              f"-Wl,--version-script={vs}", f"-L{ROCM}/lib", "-lamdhip64", "-lpthread", "-ldl"],
             verbose)
        outputs[lib] = out
    if os.path.exists(SHIPPED_TABLE):
        shutil.copyfile(SHIPPED_TABLE, os.path.join(out_dir, "gemm_mi355x_tune.txt"))
    return outputs


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("-j", "--jobs", type=int, default=8)
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args(argv)
    for name, path in build(args.jobs, args.verbose).items():
        print(f"{name}: {path}")


if __name__ == "__main__":
    main()
