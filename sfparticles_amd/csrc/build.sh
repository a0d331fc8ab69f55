#!/bin/bash
# Build libsfp_hip.so for gfx950 in-tree (next to this script's package).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="${1:-$HERE/../libsfp_hip.so}"
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -ffp-contract=off -fno-fast-math \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-function \
  -I"$HERE" -I"$HERE/../../include" \
  -o "$OUT.tmp" "$HERE/sfp_lib.hip"
mv "$OUT.tmp" "$OUT"
