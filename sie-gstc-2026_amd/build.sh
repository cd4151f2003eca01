#!/bin/bash
# Builds libsie_amd.so (HIP kernels for gfx950 + the C ABI) in place.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/sie_mom_amd/libsie_amd.so"
SRC="$HERE/csrc"
JOBS=${MAX_JOBS:-8}
FLAGS=(-O3 -std=c++17 -fPIC --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops
       -munsafe-fp-atomics -Wno-unused-command-line-argument -I"$HERE/../include")
TMP="$HERE/_build"
mkdir -p "$TMP"
pids=()
for f in "$SRC"/*.hip "$SRC"/*.cpp; do
  b=$(basename "$f")
  hipcc "${FLAGS[@]}" -c -x hip "$f" -o "$TMP/$b.o" &
  pids+=($!)
  if (( ${#pids[@]} >= JOBS )); then wait "${pids[0]}"; pids=("${pids[@]:1}"); fi
done
for p in "${pids[@]}"; do wait "$p"; done
hipcc -shared --offload-arch=gfx950 -o "$OUT.tmp" "$TMP"/*.o
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
