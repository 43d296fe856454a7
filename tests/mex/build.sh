#!/bin/bash
# Compiles the DD3_GPU gateway against the stand-in MATLAB API into tests/mex/_build.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
mkdir -p "$HERE/_build"
g++ -std=c++17 -O1 -Wall -I"$ROOT/include" -I"$HERE" \
  "$ROOT/ctrecon_amd/mex/DD3_GPU_mex.cpp" "$HERE/mex_driver.cpp" \
  -L"$ROOT/ctrecon_amd" -lctrecon_amd -Wl,-rpath,'$ORIGIN/../../../ctrecon_amd' \
  -ldl -o "$HERE/_build/mex_driver"
