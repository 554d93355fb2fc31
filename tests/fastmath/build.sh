#!/bin/sh
# Builds the exhaustive sqrt/reciprocal check (a GPU test library) with the
# library's numeric flags into tests/fastmath/_build/.
set -e
cd "$(dirname "$0")"
mkdir -p _build
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared -ffp-contract=off \
  -fhip-fp32-correctly-rounded-divide-sqrt -fno-gpu-flush-denormals-to-zero -fno-fast-math \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  fastmath_check.hip -o _build/libfastmath_check.tmp.so
mv _build/libfastmath_check.tmp.so _build/libfastmath_check.so
