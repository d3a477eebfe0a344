#!/usr/bin/env bash
# Builds libavbd_hip.so (the HIP/gfx950 AVBD step and its C ABI) in this directory.
# FP flags are part of the parity contract: no FP contraction, no fast math, correctly
# rounded fp32 division/sqrt (hipcc's default), denormals kept; packed-fp32 instructions
# are disabled.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
    -ffp-contract=off -fno-fast-math -fhip-fp32-correctly-rounded-divide-sqrt \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wall -Wno-unused-function -Wno-unused-result \
    -o "$HERE/libavbd_hip.so" "$HERE/csrc/engine.hip"

# C++ facade test program (links the library; runs only where a GPU is present)
mkdir -p "$HERE/bin"
g++ -std=c++17 -O2 -ffp-contract=off -I"$HERE/../include" -I/opt/rocm/include \
    -o "$HERE/bin/facade_test" "$HERE/../tests/cpp/facade_test.cpp" \
    -L"$HERE" -lavbd_hip -Wl,-rpath,'$ORIGIN/..'
