#!/bin/bash
# Compiles tests/shim/shim_test.cpp (host C++) against libise_amd.so into $1.
set -e
ROOT="$(cd "$(dirname "$0")/../.." && pwd)"
PKG="$ROOT/incidence-system-enumeration_amd"
g++ -std=c++17 -O2 -Wall -Wextra -D__HIP_PLATFORM_AMD__ -I "$ROOT/include" -I "$PKG/shim" -I /opt/rocm/include \
    "$ROOT/tests/shim/shim_test.cpp" -o "$1" -L "$PKG" -lise_amd -Wl,-rpath,"$PKG" \
    -L /opt/rocm/lib -Wl,-rpath-link,/opt/rocm/lib
