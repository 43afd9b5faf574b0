#!/bin/bash
# Builds libhosnwt.so (gfx950) in-tree.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/lib"
flags="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Xclang -target-feature -Xclang -packed-fp32-ops"
# the two translation units compile in parallel
hipcc $flags -c "$here/csrc/hosnwt.hip" -o "$here/lib/hosnwt.o" &
p1=$!
hipcc $flags -c "$here/csrc/mlp.hip" -o "$here/lib/mlp.o" &
p2=$!
wait $p1
wait $p2
hipcc --offload-arch=gfx950 -shared -fPIC -o "$here/lib/libhosnwt.so" "$here/lib/hosnwt.o" "$here/lib/mlp.o"
# Fortran iso_c_binding module and driver (AMD flang)
if command -v amdflang >/dev/null 2>&1; then
  mkdir -p "$here/lib/mod"
  amdflang -O2 -module-dir "$here/lib/mod" -c "$here/fortran/hosnwt_gpu.f90" -o "$here/lib/hosnwt_gpu.o"
  amdflang -O2 -module-dir "$here/lib/mod" "$here/fortran/hos_nwt_gpu.f90" "$here/lib/hosnwt_gpu.o" \
    -o "$here/lib/hos_nwt_gpu" -L"$here/lib" -lhosnwt -Wl,-rpath,'$ORIGIN'
fi
