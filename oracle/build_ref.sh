#!/bin/bash
# Builds the reference's Fortran hot path, from its sources where they lie, into
# oracle/_ref/<config>/ref_harness (test infrastructure only: the checker).
#
# Only two edits are made, both on copies written under oracle/_ref/<config>/src
# (git-ignored and deleted after the build):
#   Parameters.f90           nx, ny, L0 and minLevels (compile-time PARAMETERs,
#                            Parameters.f90:12-17) per workload;
#   AMR_Rising_2D_OMP.f90    the driver-side routines the harness calls, taken
#                            as unmodified line ranges, with the bubble radius
#                            line (AMR_Rising_2D_OMP.f90:154) set per workload.
# Every other reference file is compiled straight from the reference tree.
#
# usage: oracle/build_ref.sh [config ...]     (default: tiny cfg1 cfg2)
#   config   L0    minLevels  bubble radius  (nx = ny = 16, 2 roots)
#   tiny     64    1          10             -> 64 x 128 finest-equivalent, 2 levels
#   cfg1     128   1          16             -> 128 x 256, 3 levels
#   cfg2     1024  3          64             -> 1024 x 2048, 4 levels (3..6)
set -euo pipefail
REF=${AMRLBM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-amdflang}
FFLAGS=${FFLAGS:-"-O3 -fopenmp"}

if [ ! -f "$REF/LBM_Subroutines.f90" ]; then
  echo "build_ref.sh: reference sources not found at $REF" >&2
  exit 1
fi

configs=("$@")
[ ${#configs[@]} -eq 0 ] && configs=(tiny cfg1 cfg2)

for cfg in "${configs[@]}"; do
  case $cfg in
    tiny) L0=64;   MINL=1; RAD=10.d0 ;;
    cfg1) L0=128;  MINL=1; RAD=16.d0 ;;
    cfg2) L0=1024; MINL=3; RAD=64.d0 ;;
    *) echo "unknown config $cfg" >&2; exit 1 ;;
  esac
  D=$OUT/$cfg
  rm -rf "$D"
  mkdir -p "$D/src" "$D/mod" "$D/obj"
  tr -d '\r' < "$REF/Parameters.f90" | sed -E \
    -e "s/^([[:space:]]*INTEGER,PARAMETER :: nx) = .*/\1 = 16/" \
    -e "s/^([[:space:]]*INTEGER,PARAMETER :: ny) = .*/\1 = 16/" \
    -e "s/^([[:space:]]*INTEGER,PARAMETER :: L0) *= .*/\1 = $L0/" \
    -e "s/^([[:space:]]*INTEGER,PARAMETER :: minLevels) = .*/\1 = $MINL/" \
    > "$D/src/Parameters.f90"
  grep -q "nx = 16" "$D/src/Parameters.f90"
  grep -q "L0 = $L0" "$D/src/Parameters.f90"
  grep -q "minLevels = $MINL" "$D/src/Parameters.f90"
  # driver-side routines: Initialize_Base_Block (79-115), Initialize_AMR_Rising_Bubble
  # and Set_Rising_Bubble_Parameters (142-263), Phase_Field_LBM_h_g_AMR and
  # Determine_Min_Max_Levels (381-452), Output (456-489), Initial_Mesh,
  # Add_Child_Blocks, Array_of_Leaf_Blocks (493-610)
  tr -d '\r' < "$REF/AMR_Rising_2D_OMP.f90" | sed -n '79,115p;142,263p;381,452p;456,489p;493,610p' \
    | sed -E "s/^([[:space:]]*R  = )L0\/10\.d0/\1$RAD/" > "$D/src/amr_driver_routines.f90"
  grep -q "R  = $RAD" "$D/src/amr_driver_routines.f90"

  cd "$D/obj"
  MODF="-J $D/mod"
  $FC $FFLAGS $MODF -c "$REF/D2Q9_Mod.f90" -o D2Q9_Mod.o
  $FC $FFLAGS $MODF -c "$D/src/Parameters.f90" -o Parameters.o
  objs="D2Q9_Mod.o Parameters.o"
  for f in LBM_Subroutines Interpolations BC_Cell Neighbors Prolong Restrict Refine_Derefine \
           Fill_Ghost_Cells_C Fill_Ghost_Cells_C2 Fill_Ghost_Cells_hg TecPlot_ASCII; do
    $FC $FFLAGS $MODF -c "$REF/$f.f90" -o $f.o
    objs="$objs $f.o"
  done
  $FC $FFLAGS $MODF -c "$D/src/amr_driver_routines.f90" -o amr_driver_routines.o
  $FC $FFLAGS $MODF -c "$HERE/ref_harness.f90" -o ref_harness.o
  $FC $FFLAGS $objs amr_driver_routines.o ref_harness.o -o "$D/ref_harness"
  cd "$HERE"
  rm -rf "$D/src" "$D/obj" "$D/mod"
  echo "built $D/ref_harness"
done
