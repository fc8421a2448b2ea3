#!/bin/bash
# Builds libssegpu.so for gfx950 in-tree.  -ffp-contract=off keeps every FP64
# expression as written (bit-exact with the CPU restatement); packed FP32 ops
# are disabled.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="${HERE}/../libssegpu.so"
hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -ffp-contract=off -fno-fast-math \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wno-unused-command-line-argument \
  -o "${OUT}.tmp" "${HERE}/capi.hip"
mv "${OUT}.tmp" "${OUT}"
