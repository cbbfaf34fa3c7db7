#!/bin/bash
# Links the unmodified reference Fortran host (read in place from
# /root/reference/src) with this package's CMF_CTRL_PHYSICS_MOD replacement and
# libcmf_hip.so: the drop-in configuration a maintainer would build.
# Outputs (git-ignored), in cama-flood_v4_amd/fortran/_build/:
#   cmf_dropin          the reference's CMF_CALC_DIAG_MOD; main program: the test
#                       driver oracle/ref_driver.F90 (writes the state dumps the
#                       tests compare with the reference's own)
#   cmf_dropin_devdiag  this package's device-side cmf_calc_diag_mod.F90; main
#                       program: this package's cmf_dropin_main.F90
#   cmf_dropin32        as cmf_dropin, with the reference in its default
#                       single-precision mode (-DSinglePrec_CMF, JPRB = 4)
set -euo pipefail
REF=${CMF_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
PKG=$(dirname "$HERE")
ROOT=$(dirname "$PKG")
OUT=$HERE/_build
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
FFLAGS="-O2 -cpp -ffp-contract=off"
if [ ! -d "$REF/src" ]; then echo "build_dropin: no reference sources at $REF/src; skipping" >&2; exit 0; fi
if [ ! -f "$PKG/cmf/libcmf_hip.so" ]; then echo "build_dropin: build libcmf_hip.so first" >&2; exit 1; fi
SRCS="parkind1 yos_cmf_input yos_cmf_time yos_cmf_map yos_cmf_prog yos_cmf_diag cmf_utils_mod
cmf_calc_outflw_mod cmf_calc_pthout_mod cmf_calc_fldstg_mod cmf_calc_stonxt_mod DIAG cmf_opt_outflw_mod
cmf_ctrl_mpi_mod cmf_ctrl_damout_mod cmf_ctrl_tracer_mod cmf_ctrl_levee_mod cmf_ctrl_forcing_mod
cmf_ctrl_boundary_mod cmf_ctrl_upsinf_mod cmf_ctrl_output_mod cmf_ctrl_restart_mod PHYSICS DEVDIAG
cmf_ctrl_time_mod cmf_ctrl_maps_mod cmf_ctrl_vars_mod cmf_ctrl_nmlist_mod
cmf_drv_control_mod cmf_drv_advance_mod"

# build_variant NAME DEVICE_DIAG MAIN [EXTRA_FFLAGS]: DEVICE_DIAG=1 links this
# package's device-side CMF_CALC_DIAG_MOD (compiled after the physics module it
# uses) in place of the reference's; MAIN is the main program's source
build_variant() {
  local name=$1 devdiag=$2 main=$3 extra=${4:-} obj=$OUT/obj_$1 objs="" f o
  local FFLAGS="$FFLAGS $extra"
  mkdir -p "$obj"
  for s in $SRCS; do
    case $s in
      PHYSICS) f=$HERE/cmf_ctrl_physics_mod.F90; o=$obj/cmf_ctrl_physics_mod_hip.o ;;
      DIAG)    [ "$devdiag" = 1 ] && continue; f=$REF/src/cmf_calc_diag_mod.F90; o=$obj/cmf_calc_diag_mod.o ;;
      DEVDIAG) [ "$devdiag" = 1 ] || continue; f=$HERE/cmf_calc_diag_mod.F90; o=$obj/cmf_calc_diag_mod_hip.o ;;
      *)       f=$REF/src/$s.F90; o=$obj/$s.o ;;
    esac
    "$FC" $FFLAGS -module-dir "$obj" -c "$f" -o "$o"
    objs="$objs $o"
  done
  "$FC" $FFLAGS -module-dir "$obj" -I "$obj" -c "$main" -o "$obj/main.o"
  "$FC" $FFLAGS -o "$OUT/$name" "$obj/main.o" $objs \
    -L"$PKG/cmf" -lcmf_hip -Wl,-rpath,'$ORIGIN/../../cmf' -Wl,-rpath,/opt/rocm/lib
  echo "build_dropin: built $OUT/$name"
}
mkdir -p "$OUT"
build_variant cmf_dropin 0 "$ROOT/oracle/ref_driver.F90"
build_variant cmf_dropin_devdiag 1 "$HERE/cmf_dropin_main.F90"
build_variant cmf_dropin32 0 "$ROOT/oracle/ref_driver.F90" -DSinglePrec_CMF
