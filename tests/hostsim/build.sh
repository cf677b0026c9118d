#!/bin/sh
# Host builds of the product's lane engine for the CPU test suite (no device code runs):
# libph_hostsim.so with the host C library's exp/log/pow (as the oracle uses), and
# libph_hostsim_cr.so with the correctly rounded ones the GPU kernels use.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O2 -std=c++17 -fPIC -shared -ffp-contract=off \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -DPH_HOST_LIBM -I"$ROOT/photic_amd/csrc" -o "$HERE/libph_hostsim.so" "$HERE/ph_hostsim.hip"
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O2 -std=c++17 -fPIC -shared -ffp-contract=off \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -I"$ROOT/photic_amd/csrc" -o "$HERE/libph_hostsim_cr.so" "$HERE/ph_hostsim.hip"
