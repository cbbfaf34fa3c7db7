#!/bin/bash
# Links the unmodified reference Fortran host (read in place from
# /root/reference/src) with this package's CMF_CTRL_PHYSICS_MOD replacement and
# libcmf_hip.so: the drop-in configuration a maintainer would build. The driver
# program is oracle/ref_driver.F90 (the same one that writes the golden dumps), so
# the run can be compared with the reference's own dumps.
# Output: cama-flood_v4_amd/fortran/_build/cmf_dropin (git-ignored).
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
mkdir -p "$OUT/obj"
SRCS="parkind1 yos_cmf_input yos_cmf_time yos_cmf_map yos_cmf_prog yos_cmf_diag cmf_utils_mod
cmf_calc_outflw_mod cmf_calc_pthout_mod cmf_calc_fldstg_mod cmf_calc_stonxt_mod cmf_calc_diag_mod cmf_opt_outflw_mod
cmf_ctrl_mpi_mod cmf_ctrl_damout_mod cmf_ctrl_tracer_mod cmf_ctrl_levee_mod cmf_ctrl_forcing_mod
cmf_ctrl_boundary_mod cmf_ctrl_upsinf_mod cmf_ctrl_output_mod cmf_ctrl_restart_mod PHYSICS
cmf_ctrl_time_mod cmf_ctrl_maps_mod cmf_ctrl_vars_mod cmf_ctrl_nmlist_mod
cmf_drv_control_mod cmf_drv_advance_mod"
OBJS=""
for s in $SRCS; do
  if [ $s = PHYSICS ]; then
    f=$HERE/cmf_ctrl_physics_mod.F90; o=$OUT/obj/cmf_ctrl_physics_mod_hip.o
  else
    f=$REF/src/$s.F90; o=$OUT/obj/$s.o
  fi
  "$FC" $FFLAGS -module-dir "$OUT/obj" -c "$f" -o "$o"
  OBJS="$OBJS $o"
done
"$FC" $FFLAGS -module-dir "$OUT/obj" -I "$OUT/obj" -c "$ROOT/oracle/ref_driver.F90" -o "$OUT/obj/ref_driver.o"
"$FC" $FFLAGS -o "$OUT/cmf_dropin" "$OUT/obj/ref_driver.o" $OBJS \
  -L"$PKG/cmf" -lcmf_hip -Wl,-rpath,'$ORIGIN/../../cmf' -Wl,-rpath,/opt/rocm/lib
echo "build_dropin: built $OUT/cmf_dropin"
