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
# The gas-model flavours (GAS_MODEL), each with double and float numbers (FLOAT_NUMBERS),
# as the reference's makefile builds them.
for flavour in "IDEAL_AIR:chkn_ref" "AB_REACTING_GAS:chkn_ref_ab" \
               "IDEAL_AIR -DFLOAT_NUMBERS:chkn_ref_f32" "AB_REACTING_GAS -DFLOAT_NUMBERS:chkn_ref_ab_f32"; do
    # shellcheck disable=SC2086
    /opt/rocm/bin/hipcc -x hip --offload-host-only -std=c++17 -O2 -fopenmp -D${flavour%%:*} \
        -ffp-contract=off -Wno-deprecated-literal-operator -w \
        -I"$REF" "$HERE/ref_harness.cpp" -o "$HERE/_ref/${flavour##*:}" -lz
    echo "built $HERE/_ref/${flavour##*:}"
done
