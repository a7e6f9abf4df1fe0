#!/bin/bash
# Builds libtichi_sph.so (gfx950) in-tree.  Usage: build.sh [outdir]
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="${1:-$here/..}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Wall -Wno-unused-variable -Wno-unused-function
       -Xclang -target-feature -Xclang -packed-fp32-ops)
objs=()
for src in tsph_api tsph_grid tsph_step tsph_ref; do
  "$HIPCC" "${FLAGS[@]}" -c "$here/$src.hip" -o "$out/_$src.o" 2>&1 | grep -v "packed-fp32-ops' is not a recognized feature" || true
  test -f "$out/_$src.o"
  objs+=("$out/_$src.o")
done
"$HIPCC" --offload-arch=gfx950 -shared -fPIC "${objs[@]}" -o "$out/libtichi_sph.so.tmp"
mv "$out/libtichi_sph.so.tmp" "$out/libtichi_sph.so"
rm -f "${objs[@]}"
