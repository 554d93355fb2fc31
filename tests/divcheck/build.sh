#!/bin/sh
# Builds the exhaustive division checker (host only) into tests/divcheck/_build/.
set -e
cd "$(dirname "$0")"
mkdir -p _build
g++ -std=c++17 -O3 -march=native -ffp-contract=off -fno-fast-math -fopenmp divcheck.cpp -o _build/divcheck.tmp
mv _build/divcheck.tmp _build/divcheck
