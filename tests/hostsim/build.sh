#!/bin/sh
# Host build of the product's lane engine for the CPU test suite (no device code runs).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O2 -std=c++17 -fPIC -shared -ffp-contract=off \
    -I"$ROOT/photic_amd/csrc" -o "$HERE/libph_hostsim.so" "$HERE/ph_hostsim.hip"
