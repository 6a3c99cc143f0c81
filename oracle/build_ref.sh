#!/usr/bin/env bash
# Compiles DMD's own translation units (read in place under the reference
# tree) plus oracle/ref_shim.cpp into oracle/_ref/libdmdref.so.
# TEST INFRASTRUCTURE ONLY.  Needs only the C++ standard library:
# src/force/coulomb_pme.cpp (fftw3.h) and src/sim/simulation_engine.cpp
# (fftw3.h, hdf5.h) are not buildable in this image and are left out.
set -euo pipefail
REF=${DMD_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
    echo "build_ref: no reference tree at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
SRCS=(
  core/cell.cpp core/system_data.cpp core/error.cpp
  force/force_component.cpp force/force_engine.cpp force/lennard_jones.cpp
  force/coulomb_direct.cpp force/coulomb_exclusion.cpp force/harmonic_bond.cpp
  force/harmonic_angle.cpp force/periodic_dihedral.cpp force/harmonic_improper.cpp
  integrate/integrator.cpp integrate/constraints.cpp
  thermostat/thermostat.cpp thermostat/berendsen.cpp thermostat/nose_hoover.cpp
  thermostat/andersen.cpp
  barostat/barostat.cpp barostat/berendsen_barostat.cpp barostat/andersen_barostat.cpp
  sim/checkpoint.cpp
)
ARGS=()
for s in "${SRCS[@]}"; do ARGS+=("$REF/src/$s"); done
g++ -std=c++20 -O2 -fPIC -ffp-contract=off -fno-fast-math -shared \
    -I"$REF/src" -o "$OUT/libdmdref.so.tmp" "${ARGS[@]}" "$HERE/ref_shim.cpp"
mv "$OUT/libdmdref.so.tmp" "$OUT/libdmdref.so"
echo "build_ref: $OUT/libdmdref.so"

# tests/adapter_check.cpp: include/dmdx_dmd.hpp inside DMD's own classes,
# linked against DMD's translation units and the HIP library
LIB="$HERE/../dmd_amd/libdmdx.so"
H5LIB="$HERE/../dmd_amd/libdmdx_h5md.so"
HDF5_ROOT=${DMDX_HDF5_ROOT:-/opt/conda}
if [ -f "$LIB" ] && [ -f "$H5LIB" ]; then
  ADP=(
    core/cell.cpp core/system_data.cpp core/error.cpp
    force/force_component.cpp force/force_engine.cpp force/lennard_jones.cpp
    force/coulomb_direct.cpp force/coulomb_exclusion.cpp force/harmonic_bond.cpp
    force/harmonic_angle.cpp force/periodic_dihedral.cpp force/harmonic_improper.cpp
    integrate/integrator.cpp sim/checkpoint.cpp
  )
  AARGS=()
  for s in "${ADP[@]}"; do AARGS+=("$REF/src/$s"); done
  g++ -std=c++20 -O2 -ffp-contract=off -fno-fast-math -I"$REF/src" -I"$HERE/../include" \
      -I"$HDF5_ROOT/include" \
      -o "$OUT/adapter_check.tmp" "$HERE/../tests/adapter_check.cpp" "${AARGS[@]}" \
      -L"$HERE/../dmd_amd" -ldmdx -ldmdx_h5md "$HDF5_ROOT/lib/libhdf5.so" \
      -Wl,-rpath,"\$ORIGIN/../../dmd_amd:/usr/lib/x86_64-linux-gnu:$HDF5_ROOT/lib"
  mv "$OUT/adapter_check.tmp" "$OUT/adapter_check"
  echo "build_ref: $OUT/adapter_check"
  # the same program with its host code (DMD's units, the adapters) under
  # AddressSanitizer + UndefinedBehaviorSanitizer, for `--no-gpu`; g++ only,
  # no device code is built here
  g++ -std=c++20 -O1 -g -fno-omit-frame-pointer -ffp-contract=off -fno-fast-math \
      -fsanitize=address,undefined -fno-sanitize-recover=all \
      -I"$REF/src" -I"$HERE/../include" -I"$HDF5_ROOT/include" \
      -o "$OUT/adapter_check_san.tmp" "$HERE/../tests/adapter_check.cpp" "${AARGS[@]}" \
      -L"$HERE/../dmd_amd" -ldmdx -ldmdx_h5md "$HDF5_ROOT/lib/libhdf5.so" \
      -Wl,-rpath,"\$ORIGIN/../../dmd_amd:/usr/lib/x86_64-linux-gnu:$HDF5_ROOT/lib"
  mv "$OUT/adapter_check_san.tmp" "$OUT/adapter_check_san"
  echo "build_ref: $OUT/adapter_check_san"
else
  echo "build_ref: adapter_check skipped (libdmdx.so or libdmdx_h5md.so not built)" >&2
fi
