#!/bin/bash
# Builds libhosnwt.so (gfx950) in-tree.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/lib"
hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -o "$here/lib/libhosnwt.so" "$here/csrc/hosnwt.hip" -L/opt/rocm/lib -lrocblas
# Fortran iso_c_binding module and driver (AMD flang)
if command -v amdflang >/dev/null 2>&1; then
  mkdir -p "$here/lib/mod"
  amdflang -O2 -module-dir "$here/lib/mod" -c "$here/fortran/hosnwt_gpu.f90" -o "$here/lib/hosnwt_gpu.o"
  amdflang -O2 -module-dir "$here/lib/mod" "$here/fortran/hos_nwt_gpu.f90" "$here/lib/hosnwt_gpu.o" \
    -o "$here/lib/hos_nwt_gpu" -L"$here/lib" -lhosnwt -Wl,-rpath,'$ORIGIN'
fi
