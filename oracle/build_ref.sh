#!/usr/bin/env bash
# Builds oracle/_ref/libsofia_ref.so from the reference Fortran sources where they lie
# (read-only, never copied) plus oracle/ref_driver.f90.  Test infrastructure only: the
# library produces golden vectors in this development container.
#
# Built: global, mod_minimg, mod_force_nonbond, mod_force_correction, mod_force_dihed,
#        mod_force_bond, mod_force_angle, mod_pbc, mod_force_impro_amber,
#        mod_force_dihed_amber  (each compiles unmodified with amdflang).
# Not buildable here (would need stand-ins for files the image lacks, so they are not built):
#   mod_neighbor.f90  -- `use ifport` (Intel Fortran portability module)
#   mod_shake.f90     -- Intel's `call abort('message')` extension
#   mod_ensemble.f90, mod_langevin.f90 -- ifport / MKL VSL (`include 'mkl_vsl.f90'`)
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
src="${SOFIA_REFERENCE:-/root/reference}/Source"
out="$here/_ref"
if [ ! -d "$src" ]; then
  echo "reference sources not present ($src); skipping oracle/_ref" >&2
  exit 0
fi
FC="${FC:-amdflang}"
command -v "$FC" >/dev/null || FC=/opt/rocm/bin/amdflang
mkdir -p "$out/obj"
FFLAGS="-O2 -fPIC -fopenmp -ffp-contract=off -w"
cd "$out/obj"
for m in global mod_minimg mod_force_nonbond mod_force_correction mod_force_dihed \
         mod_force_bond mod_force_angle mod_pbc mod_force_impro_amber mod_force_dihed_amber; do
  "$FC" $FFLAGS -c "$src/$m.f90" -o "$m.o"
done
"$FC" $FFLAGS -c "$here/ref_driver.f90" -o ref_driver.o
"$FC" -shared -fopenmp -o "$out/libsofia_ref.so" ref_driver.o global.o mod_minimg.o \
  mod_force_nonbond.o mod_force_correction.o mod_force_dihed.o mod_force_bond.o \
  mod_force_angle.o mod_pbc.o mod_force_impro_amber.o mod_force_dihed_amber.o
echo "built $out/libsofia_ref.so"
