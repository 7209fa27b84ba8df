#!/bin/bash
# Builds libcmmdft_hip.so for gfx950 in this directory's parent package.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${HERE}/../libcmmdft_hip.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"${HIPCC}" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -I"${HERE}" -I"${HERE}/../../include" \
  -o "${OUT}.tmp" "${HERE}/cmm_api.hip"
mv "${OUT}.tmp" "${OUT}"
echo "built ${OUT}"
