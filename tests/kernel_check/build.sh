#!/bin/sh
# Host build of the kernel headers for the CPU kernel-check tests.  Each run
# compiles to a private temporary file and renames it into place, so test
# processes that build concurrently never load a half-written library.
# SANITIZE=1 builds _build/libkernel_check_san.so instead, with
# AddressSanitizer and UBSan (load it with libasan preloaded).
set -e
cd "$(dirname "$0")"
mkdir -p _build
if [ "${SANITIZE:-0}" = 1 ]; then
  name=libkernel_check_san.so
  flags="-O1 -g -fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer"
else
  name=libkernel_check.so
  flags="-O2"
fi
tmp="_build/$name.$$.tmp.so"
g++ -std=c++17 $flags -fPIC -shared -ffp-contract=off -fno-fast-math -Wall -Wno-unused-function -Wno-unknown-pragmas \
  kernel_check.cpp -o "$tmp"
mv -f "$tmp" "_build/$name"
