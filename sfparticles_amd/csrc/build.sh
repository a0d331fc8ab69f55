#!/bin/bash
# Build libsfp_hip.so for gfx950 in-tree: each translation unit is compiled in
# parallel, then linked.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="${1:-$HERE/../libsfp_hip.so}"
B="$HERE/_build"
mkdir -p "$B"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -ffp-contract=off -fno-fast-math
       -Xclang -target-feature -Xclang -packed-fp32-ops -Wall -Wno-unused-function
       -I"$HERE" -I"$HERE/../../include")
pids=()
for src in sfp_lib sfp_step sfp_mlp; do
  /opt/rocm/bin/hipcc "${FLAGS[@]}" -c -o "$B/$src.o" "$HERE/$src.hip" &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
/opt/rocm/bin/hipcc --offload-arch=gfx950 -shared -fPIC -o "$OUT.tmp" "$B/sfp_lib.o" "$B/sfp_step.o" "$B/sfp_mlp.o"
mv "$OUT.tmp" "$OUT"
