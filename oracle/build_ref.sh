#!/bin/bash
# Builds the reference-backed checker oracle/_ref/chkn_ref from the reference's own
# src/chicken sources, which stay where they lie (nothing is copied).  The sources are
# header-style CUDA C++ whose __host__/__device__ qualifiers HIP defines natively, so
# hipcc compiles them as host code without any stand-in header.  x86-64 host code without
# -mfma has no fused multiply-add, as in the reference's own gcc -O2 CPU build.
# Test infrastructure only; a no-op when the reference tree is absent.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${CHKN_REFERENCE_DIR:-/root/reference}/src/chicken"
if [ ! -f "$REF/block.cu" ]; then
    echo "reference sources not found at $REF; oracle/_ref not built"
    exit 0
fi
mkdir -p "$HERE/_ref"
# Two gas-model flavours, as the reference's makefile builds them (GAS_MODEL).
for flavour in IDEAL_AIR:chkn_ref AB_REACTING_GAS:chkn_ref_ab; do
    /opt/rocm/bin/hipcc -x hip --offload-host-only -std=c++17 -O2 -fopenmp -D"${flavour%%:*}" \
        -ffp-contract=off -Wno-deprecated-literal-operator -w \
        -I"$REF" "$HERE/ref_harness.cpp" -o "$HERE/_ref/${flavour##*:}" -lz
    echo "built $HERE/_ref/${flavour##*:}"
done
