#!/bin/bash
# Compiles tests/shim/cd_tools_test.cpp with incidence-system-enumeration_amd/shim/cd_tools.cpp
# (host C++) against libise_amd.so into $1.
set -e
ROOT="$(cd "$(dirname "$0")/../.." && pwd)"
PKG="$ROOT/incidence-system-enumeration_amd"
g++ -std=c++17 -O2 -Wall -Wextra -I "$ROOT/include" -I "$PKG/shim" \
    "$ROOT/tests/shim/cd_tools_test.cpp" "$PKG/shim/cd_tools.cpp" -o "$1" -L "$PKG" -lise_amd -Wl,-rpath,"$PKG" \
    -L /opt/rocm/lib -Wl,-rpath-link,/opt/rocm/lib
