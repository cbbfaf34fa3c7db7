#!/bin/bash
# Builds the reference CaMa-Flood Fortran (read in place from /root/reference/src,
# unmodified) plus the test drivers oracle/ref_driver.F90 and (serial64 only)
# oracle/ref_forcing.F90 and oracle/ref_diag.F90 into oracle/_ref/.
# oracle/_ref/ is test infrastructure: it pins the C oracle (golden dumps) and is
# timed as the CPU baseline. Nothing in the product links or loads it.
#
# Variants (each in its own module/object directory):
#   serial64 : FP64 (no -DSinglePrec_CMF), -O2 -ffp-contract=off, no OpenMP.
#              Parity build: no FMA contraction, deterministic.
#   omp64    : FP64, -O3 -fopenmp -DNoAtom_CMF. Timing build (OpenMP CPU baseline).
#   omp32    : reference default single-precision mode (-DSinglePrec_CMF),
#              -O3 -fopenmp -DNoAtom_CMF. CPU counterpart of the mixed path.
#   serial32 : single-precision mode, -O2 -ffp-contract=off, no OpenMP: the
#              fixtures of the single-precision drop-in (JPRB = 4 host).
#
# OpenMP builds need a build copy of two files: flang rejects the loop index INUM
# listed in THREADPRIVATE (cmf_calc_outflw_mod.F90:394, cmf_ctrl_tracer_mod.F90:542).
# Loop indices of OpenMP loops are predetermined private, so dropping "INUM," from
# those two lists changes nothing semantically. The copies live only in
# oracle/_ref/<variant>/fix/ (git-ignored).
set -euo pipefail
REF=${CMF_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
VARIANTS=${*:-serial64 omp64 omp32}

SRCS="parkind1 yos_cmf_input yos_cmf_time yos_cmf_map yos_cmf_prog yos_cmf_diag cmf_utils_mod
cmf_calc_outflw_mod cmf_calc_pthout_mod cmf_calc_fldstg_mod cmf_calc_stonxt_mod cmf_calc_diag_mod cmf_opt_outflw_mod
cmf_ctrl_mpi_mod cmf_ctrl_damout_mod cmf_ctrl_tracer_mod cmf_ctrl_levee_mod cmf_ctrl_forcing_mod
cmf_ctrl_boundary_mod cmf_ctrl_upsinf_mod cmf_ctrl_output_mod cmf_ctrl_restart_mod cmf_ctrl_physics_mod
cmf_ctrl_time_mod cmf_ctrl_maps_mod cmf_ctrl_vars_mod cmf_ctrl_nmlist_mod
cmf_drv_control_mod cmf_drv_advance_mod"

if [ ! -d "$REF/src" ]; then
  echo "build_ref: reference sources not found at $REF/src; skipping" >&2
  exit 0
fi

for V in $VARIANTS; do
  case $V in
    serial64) FFLAGS="-O2 -cpp -ffp-contract=off"; OMP="" ;;
    omp64)    FFLAGS="-O3 -cpp -fopenmp -DNoAtom_CMF"; OMP=1 ;;
    omp32)    FFLAGS="-O3 -cpp -fopenmp -DNoAtom_CMF -DSinglePrec_CMF"; OMP=1 ;;
    serial32) FFLAGS="-O2 -cpp -ffp-contract=off -DSinglePrec_CMF"; OMP="" ;;
    *) echo "unknown variant $V" >&2; exit 2 ;;
  esac
  D=$OUT/$V
  mkdir -p "$D/obj" "$D/fix"
  OBJS=""
  for s in $SRCS; do
    f=$REF/src/$s.F90
    if [ -n "$OMP" ] && { [ $s = cmf_calc_outflw_mod ] || [ $s = cmf_ctrl_tracer_mod ]; }; then
      sed -e '/THREADPRIVATE/s/INUM, *//' "$f" > "$D/fix/$s.F90"
      f=$D/fix/$s.F90
    fi
    "$FC" $FFLAGS -module-dir "$D/obj" -c "$f" -o "$D/obj/$s.o"
    OBJS="$OBJS $D/obj/$s.o"
  done
  "$FC" $FFLAGS -module-dir "$D/obj" -I "$D/obj" -c "$HERE/ref_driver.F90" -o "$D/obj/ref_driver.o"
  "$FC" $FFLAGS -o "$D/cmf_ref" "$D/obj/ref_driver.o" $OBJS
  echo "build_ref: built $D/cmf_ref"
  if [ $V = serial64 ]; then
    # the reference's CMF_FORCING_PUT (runoff interpolation) on a given input matrix
    "$FC" $FFLAGS -module-dir "$D/obj" -I "$D/obj" -c "$HERE/ref_forcing.F90" -o "$D/obj/ref_forcing.o"
    "$FC" $FFLAGS -o "$D/cmf_ref_forcing" "$D/obj/ref_forcing.o" $OBJS
    echo "build_ref: built $D/cmf_ref_forcing"
    # the reference's output-period diagnostics (CMF_DIAG_*_OUTPUT)
    "$FC" $FFLAGS -module-dir "$D/obj" -I "$D/obj" -c "$HERE/ref_diag.F90" -o "$D/obj/ref_diag.o"
    "$FC" $FFLAGS -o "$D/cmf_ref_diag" "$D/obj/ref_diag.o" $OBJS
    echo "build_ref: built $D/cmf_ref_diag"
  fi
done
