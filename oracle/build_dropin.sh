#!/bin/bash
# Builds, from the reference sources where they lie, the reference's complete
# program twice for the drop-in check (outputs only into oracle/_ref/dropin/):
#
#   ref_program     all 14 reference files, unmodified except the block size,
#                   root size and base level in Parameters.f90 (nx=ny=16,
#                   L0=64, minLevels=1): the reference's own driver, three root
#                   blocks, Tf steps, Tecplot output.
#   dropin_program  the same driver and set-up code linked with the product's
#                   Fortran shim (amr-lbm-openmp-2d_amd/fortran/amrlbm_shim.f90)
#                   and libamrlbm.so in place of the hot-path files; the five
#                   per-block routines the shim replaces are made weak in
#                   LBM_Subroutines.o with objcopy -W.
#
# Both print the reference's Output table; a GPU test checks that the drop-in
# prints the same table and writes the same Tecplot files as the reference.
set -euo pipefail
REF=${AMRLBM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
OUT=$HERE/_ref/dropin
FC=${FC:-amdflang}
FFLAGS=${FFLAGS:-"-O3 -fopenmp"}
LIBDIR=$ROOT/amr-lbm-openmp-2d_amd/lib

[ -f "$REF/LBM_Subroutines.f90" ] || { echo "build_dropin.sh: reference sources not found at $REF" >&2; exit 1; }
[ -f "$LIBDIR/libamrlbm.so" ] || { echo "build_dropin.sh: build libamrlbm.so first" >&2; exit 1; }

rm -rf "$OUT"
mkdir -p "$OUT/src" "$OUT/ref" "$OUT/drop"
tr -d '\r' < "$REF/Parameters.f90" | sed -E \
  -e "s/^([[:space:]]*INTEGER,PARAMETER :: nx) = .*/\1 = 16/" \
  -e "s/^([[:space:]]*INTEGER,PARAMETER :: ny) = .*/\1 = 16/" \
  -e "s/^([[:space:]]*INTEGER,PARAMETER :: L0) *= .*/\1 = 64/" \
  -e "s/^([[:space:]]*INTEGER,PARAMETER :: minLevels) = .*/\1 = 1/" > "$OUT/src/Parameters.f90"
grep -q "nx = 16" "$OUT/src/Parameters.f90"
grep -q "L0 = 64" "$OUT/src/Parameters.f90"

# the complete reference program
cd "$OUT/ref"
$FC $FFLAGS -J . -c "$REF/D2Q9_Mod.f90" -o D2Q9_Mod.o
$FC $FFLAGS -J . -c "$OUT/src/Parameters.f90" -o Parameters.o
objs="D2Q9_Mod.o Parameters.o"
for f in LBM_Subroutines Interpolations BC_Cell Neighbors Prolong Restrict Refine_Derefine \
         Fill_Ghost_Cells_C Fill_Ghost_Cells_C2 Fill_Ghost_Cells_hg TecPlot_ASCII AMR_Rising_2D_OMP; do
  $FC $FFLAGS -J . -c "$REF/$f.f90" -o $f.o
  objs="$objs $f.o"
done
$FC $FFLAGS $objs -o "$OUT/ref_program"

# the reference's driver and set-up code on the GPU library
cd "$OUT/drop"
$FC $FFLAGS -J . -c "$REF/D2Q9_Mod.f90" -o D2Q9_Mod.o
$FC $FFLAGS -J . -c "$OUT/src/Parameters.f90" -o Parameters.o
$FC $FFLAGS -J . -c "$REF/LBM_Subroutines.f90" -o LBM_Subroutines.o
objcopy -W collision_h_g_amr_ -W perfect_shift_amr_ -W streaming_laxwendroff_amr_ \
        -W macroscopic_values_h_amr_ -W macroscopic_values_g_amr_ LBM_Subroutines.o
objs="D2Q9_Mod.o Parameters.o LBM_Subroutines.o"
for f in Neighbors TecPlot_ASCII AMR_Rising_2D_OMP; do
  $FC $FFLAGS -J . -c "$REF/$f.f90" -o $f.o
  objs="$objs $f.o"
done
$FC $FFLAGS -J . -c "$ROOT/amr-lbm-openmp-2d_amd/fortran/amrlbm_shim.f90" -o amrlbm_shim.o
$FC $FFLAGS amrlbm_shim.o $objs -L"$LIBDIR" -lamrlbm \
    -Wl,-rpath,'$ORIGIN/../../../amr-lbm-openmp-2d_amd/lib' -o "$OUT/dropin_program"

cd "$HERE"
rm -rf "$OUT/src" "$OUT/ref" "$OUT/drop"
echo "built $OUT/ref_program $OUT/dropin_program"
