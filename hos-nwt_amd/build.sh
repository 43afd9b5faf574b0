#!/bin/bash
# Builds libhosnwt.so (gfx950) in-tree.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/lib"
hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -o "$here/lib/libhosnwt.so" "$here/csrc/hosnwt.hip" -L/opt/rocm/lib -lrocblas
