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
# the C++ interface layer is host code over the C ABI: a plain C++ compiler
# (its public headers carry the reference's host types, which are not meant to
# meet the HIP compiler's own vector types)
for f in "$SRC"/host/*.cpp; do
  b=$(basename "$f")
  g++ -std=c++17 -O2 -fPIC -D__HIP_PLATFORM_AMD__ -I"$HERE/../include" -I/opt/rocm/include -c "$f" -o "$TMP/host_$b.o" &
  pids+=($!)
  if (( ${#pids[@]} >= JOBS )); then wait "${pids[0]}"; pids=("${pids[@]:1}"); fi
done
for p in "${pids[@]}"; do wait "$p"; done
hipcc -shared --offload-arch=gfx950 -o "$OUT.tmp" "$TMP"/*.o
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
# the C++ interface check program used by tests/test_cpp_interface.py (a plain
# C++17 caller of include/sie_mom/*.hpp, linked against the library)
CHK="$HERE/../tests/cpp/sie_cpp_check"
g++ -std=c++17 -O2 -D__HIP_PLATFORM_AMD__ -I"$HERE/../include" -I/opt/rocm/include "$CHK.cpp" \
    -L"$HERE/sie_mom_amd" -lsie_amd -L/opt/rocm/lib -lamdhip64 \
    -Wl,-rpath,'$ORIGIN/../../sie-gstc-2026_amd/sie_mom_amd' -Wl,-rpath,/opt/rocm/lib -o "$CHK.tmp"
mv "$CHK.tmp" "$CHK"
echo "built $CHK"
