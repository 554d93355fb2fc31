#!/bin/sh
# Host build of the kernel headers for the CPU kernel-check tests.
set -e
cd "$(dirname "$0")"
mkdir -p _build
g++ -std=c++17 -O2 -fPIC -shared -ffp-contract=off -fno-fast-math -Wall -Wno-unused-function -Wno-unknown-pragmas \
  kernel_check.cpp -o _build/libkernel_check.so
