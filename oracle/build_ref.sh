#!/bin/bash
# Builds oracle/_ref/libref_framework.so: the reference's framework kernels
# (compiled from their sources under the reference tree, not copied) behind the
# launch harness oracle/ref_harness.hip.  Test infrastructure only: the GPU
# parity tests load it, when present, to compare the product path with the
# reference's own kernels.  Exits 0 without building if the tree is absent.
set -euo pipefail
REF="${VBG_REFERENCE_ROOT:-/root/reference}"
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
if [ ! -f "$REF/ROCm/framework/kernel.h" ]; then
  echo "build_ref: reference tree not found at $REF; skipping"
  exit 0
fi
mkdir -p "$HERE/_ref"
/opt/rocm/bin/hipcc -O3 -std=c++17 -fPIC -shared --offload-arch=gfx950 \
  -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-result \
  -I "$REF/ROCm/framework" "$HERE/ref_harness.hip" -o "$HERE/_ref/libref_framework.so"
echo "build_ref: built $HERE/_ref/libref_framework.so"
