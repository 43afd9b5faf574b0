#!/bin/bash
# Copies the reference's benchmark inputs and recorded outputs (data files of
# Benchmark/<case>/ and Benchmark/<case>/Ref_data/) into this folder as fixtures.
# Run in the development container, where the reference tree is present.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
src=/root/reference/Benchmark
for c in Irreg_2D Irreg_3D Regular_3D Sloshing; do
  mkdir -p "$here/$c"
  for f in "$src/$c"/input_HOS-NWT.dat "$src/$c"/prob.inp "$src/$c"/*.cfg "$src/$c"/*.dat "$src/$c"/*.txt "$src/$c"/Ref_data/*.dat; do
    [ -f "$f" ] && cp "$f" "$here/$c/"
  done
done
chmod -R u+w "$here"
