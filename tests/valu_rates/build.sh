#!/bin/sh
# Builds the VALU issue-rate measurement program into tests/valu_rates/_build/.
set -e
cd "$(dirname "$0")"
mkdir -p _build
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O3 -std=c++17 \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  valu_rates.hip -o _build/valu_rates.tmp
mv _build/valu_rates.tmp _build/valu_rates
