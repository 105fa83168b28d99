#!/bin/bash
# Builds the gfx950 FEM library in-tree: plask_amd/libplaskfem.so
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/../libplaskfem.so"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wno-unused-result -Wno-unused-value \
    -o "$OUT.tmp" "$HERE/pf_capi.hip"
mv "$OUT.tmp" "$OUT"
