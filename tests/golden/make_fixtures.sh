#!/bin/bash
# Regenerates the golden vectors in this directory from the reference's own
# sgemm_/dgemm_, compiled from its sources by oracle/Makefile into oracle/_ref/.
# Needs the reference sources (REF, default /root/reference); run on the CPU.
#   tests/golden/make_fixtures.sh        small, block-edge and 1024^3 cases
#   tests/golden/make_fixtures.sh big    also the 8192^3 cases (~2 min, 8 cores)
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
make -C "$ROOT/oracle" oracle ref >/dev/null
gcc -O2 -fopenmp -o "$ROOT/oracle/_ref/gen_fixtures" "$HERE/gen_fixtures.c" \
    -L"$ROOT/oracle" -loracle -Wl,-rpath,"$ROOT/oracle" -ldl -lm
export OMP_NUM_THREADS="${OMP_NUM_THREADS:-$(nproc)}" OMP_PROC_BIND=close
"$ROOT/oracle/_ref/gen_fixtures" "$ROOT/oracle/_ref/DGEMM.so" \
    "$ROOT/oracle/_ref/SGEMM.so" "$HERE" "$([ "${1:-}" = big ] && echo 1 || echo 0)"
sha256sum "$HERE/values.f64" "$HERE/index.i32" "$HERE/manifest.json" | sed "s|$HERE/||" > "$HERE/SHA256SUMS"
