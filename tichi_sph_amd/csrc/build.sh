#!/bin/bash
# Builds libtichi_sph.so (gfx950) in-tree.  Usage: build.sh [outdir]
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="${1:-$here/..}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Wall -Wno-unused-variable -Wno-unused-function
       -Xclang -target-feature -Xclang -packed-fp32-ops)
srcs=(tsph_api tsph_grid tsph_step tsph_ref tsph_adjoint)
objs=()
pids=()
for src in "${srcs[@]}"; do
  rm -f "$out/_$src.o" "$out/_$src.log"
  ( "$HIPCC" "${FLAGS[@]}" -c "$here/$src.hip" -o "$out/_$src.o" > "$out/_$src.log" 2>&1 ) &
  pids+=($!)
  objs+=("$out/_$src.o")
done
fail=0
for i in "${!srcs[@]}"; do
  wait "${pids[$i]}" || fail=1
  grep -v "packed-fp32-ops' is not a recognized feature" "$out/_${srcs[$i]}.log" || true
  rm -f "$out/_${srcs[$i]}.log"
done
if [ "$fail" != 0 ]; then
  rm -f "${objs[@]}"
  echo "build.sh: compilation failed" >&2
  exit 1
fi
"$HIPCC" --offload-arch=gfx950 -shared -fPIC "${objs[@]}" -o "$out/libtichi_sph.so.tmp"
mv "$out/libtichi_sph.so.tmp" "$out/libtichi_sph.so"
rm -f "${objs[@]}"
