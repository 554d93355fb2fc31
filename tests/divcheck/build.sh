#!/bin/sh
# Builds the exhaustive division checker (host only) into tests/divcheck/_build/.
# Each run compiles to a private temporary file and renames it into place, so
# test processes that build concurrently never run a half-written binary.
set -e
cd "$(dirname "$0")"
mkdir -p _build
tmp="_build/divcheck.$$.tmp"
g++ -std=c++17 -O3 -march=native -ffp-contract=off -fno-fast-math -fopenmp divcheck.cpp -o "$tmp"
mv -f "$tmp" _build/divcheck
