#!/bin/bash
# Builds the gfx950 FEM library in-tree: plask_amd/libplaskfem.so
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/../libplaskfem.so"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wno-unused-result -Wno-unused-value \
    -o "$OUT.tmp" "$HERE/pf_capi.hip" "$HERE/warmnet.hip"
mv "$OUT.tmp" "$OUT"
# host solver cores + pybind11 module (links the C ABI library by rpath)
PYINC=$(python3 -c "import sysconfig; print(sysconfig.get_paths()['include'])")
PBINC=$(python3 -c "import pybind11; print(pybind11.get_include())")
EXT=$(python3 -c "import sysconfig; print(sysconfig.get_config_var('EXT_SUFFIX'))")
g++ -O2 -std=c++17 -fPIC -shared -fvisibility=hidden -I"$PYINC" -I"$PBINC" \
    -o "$HERE/../_core$EXT.tmp" "$HERE/host/fem_solvers.cpp" "$HERE/host/pymodule.cpp" \
    -L"$HERE/.." -lplaskfem -Wl,-rpath,'$ORIGIN'
mv "$HERE/../_core$EXT.tmp" "$HERE/../_core$EXT"
