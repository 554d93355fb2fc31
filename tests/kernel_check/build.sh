#!/bin/sh
# Host build of the kernel headers for the CPU kernel-check tests.  Each run
# compiles to a private temporary file and renames it into place, so test
# processes that build concurrently never load a half-written library.
set -e
cd "$(dirname "$0")"
mkdir -p _build
tmp="_build/libkernel_check.$$.tmp.so"
g++ -std=c++17 -O2 -fPIC -shared -ffp-contract=off -fno-fast-math -Wall -Wno-unused-function -Wno-unknown-pragmas \
  kernel_check.cpp -o "$tmp"
mv -f "$tmp" _build/libkernel_check.so
