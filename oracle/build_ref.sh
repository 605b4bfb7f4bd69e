#!/usr/bin/env bash
# Builds the reference Geoflow-SPH Fortran sources (read where they lie under
# /root/reference) into oracle/_ref/gfsph_ref, an FP64 executable driven by
# oracle/ref_harness.f90.  Test infrastructure only: nothing in the product
# path links or runs it.
#
# Precision recipe (mirrors what the reference's own kinds imply):
#   -fdefault-real-8   promotes REAL / kind(0.0) (irk) to FP64; const(:) stays
#                      REAL(4) because it is declared real(ink)
#   -fdefault-double-8 keeps DOUBLE PRECISION at 8 bytes
#   -ffp-contract=off  no FMA contraction, so the neighbour predicate rounds
#                      exactly as written
#
# Compatibility edits.  flang rejects 8 lines of the reference; they are edited
# on the fly with sed into a scratch file that is deleted right after it is
# compiled (only objects and the executable remain under oracle/_ref):
#   6 SPH_SW_2019.f90:395,536  nested implied-do parentheses (window output)
#   6 SPH_SW_2019.f90:402,543  same
#   6 SPH_SW_2019.f90:9258     "if (if_frictional)" on an INTEGER -> ".ne.0"
#   7 SPH_MAIN_2019.f90:3264,3279 "if (if_density_comput)" INTEGER -> ".ne.0"
#   5 SPH_Main_Vars_2019.f90:175  "integer (irk) iffix" -> integer (ink); with
#                                 irk promoted to 8 the semi-implicit solver's
#                                 INTEGER(4) dummy no longer matches (off path)
# The values tested are 0/1, so the logical edits keep the semantics.
set -euo pipefail

HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC="${GFSPH_REFERENCE_SRC:-/root/reference/Codes Geoflow_SPH}"
OUT="$HERE/_ref"
BLD="$OUT/build"
FC="${FC:-flang}"
FFLAGS="-O2 -fdefault-real-8 -fdefault-double-8 -ffp-contract=off"

if [ ! -d "$SRC" ]; then
  echo "build_ref.sh: reference sources not found at $SRC; skipping" >&2
  exit 0
fi
command -v "$FC" >/dev/null 2>&1 || { echo "build_ref.sh: $FC not found; skipping" >&2; exit 0; }

mkdir -p "$BLD"
cd "$BLD"

compile_asis() {   # $1 = file stem under $SRC, $2 = object name
  "$FC" -c $FFLAGS -module-dir "$BLD" "$SRC/$1.f90" -o "$BLD/$2.o"
}

compile_asis "1 SPH_FEM_variable_types_2019" r1
compile_asis "3 SPH_FEM_Driver_time_vars_2019" r3
compile_asis "2 SPH_FEM_topo_2019" r2
compile_asis "4 SPH_SW_Vars_2019" r4
sed -e '175s/integer (irk), allocatable:: iffix/integer (ink), allocatable:: iffix/' \
    "$SRC/5 SPH_Main_Vars_2019.f90" > "$BLD/r5_edit.f90"
"$FC" -c $FFLAGS -module-dir "$BLD" "$BLD/r5_edit.f90" -o "$BLD/r5.o"
rm -f "$BLD/r5_edit.f90"
compile_asis "5a My_SPH_Tools_2019" r5a
compile_asis "5b GFL_Global_Vars_2019" r5b
compile_asis "5c SPH_GFL_SW_Interactions_2019" r5c
compile_asis "5d SPH_GCsolver_2019" r5d
compile_asis "5e SPH_Quicksort_2019" r5e
compile_asis "5f SPH_Vulnerability" r5f

sed -e '395s/((\s*h_cw(icw),vn_cw(icw),q_cw(icw)),icw/(h_cw(icw),vn_cw(icw),q_cw(icw),icw/' \
    -e '536s/((\s*h_cw(icw),vn_cw(icw),q_cw(icw)),icw/(h_cw(icw),vn_cw(icw),q_cw(icw),icw/' \
    -e '402s/(\s*(time, h_cw(icw),vn_cw(icw),q_cw(icw)),icw/(time, h_cw(icw),vn_cw(icw),q_cw(icw),icw/' \
    -e '543s/(\s*(time, h_cw(icw),vn_cw(icw),q_cw(icw)),icw/(time, h_cw(icw),vn_cw(icw),q_cw(icw),icw/' \
    -e '9258s/if (if_frictional)/if (if_frictional.ne.0)/' \
    "$SRC/6 SPH_SW_2019.f90" > "$BLD/r6_edit.f90"
"$FC" -c $FFLAGS -module-dir "$BLD" "$BLD/r6_edit.f90" -o "$BLD/r6.o"
rm -f "$BLD/r6_edit.f90"

sed -e '3264s/if (if_density_comput)/if (if_density_comput.ne.0)/' \
    -e '3279s/if (if_density_comput)/if (if_density_comput.ne.0)/' \
    "$SRC/7 SPH_MAIN_2019.f90" > "$BLD/r7_edit.f90"
"$FC" -c $FFLAGS -module-dir "$BLD" "$BLD/r7_edit.f90" -o "$BLD/r7.o"
rm -f "$BLD/r7_edit.f90"

compile_asis "7b GFL_MAIN_2019" r7b

"$FC" -c $FFLAGS -cpp -module-dir "$BLD" "$HERE/ref_harness.f90" -o "$BLD/harness.o"
"$FC" $FFLAGS -o "$OUT/gfsph_ref" \
  "$BLD"/r1.o "$BLD"/r3.o "$BLD"/r2.o "$BLD"/r4.o "$BLD"/r5.o "$BLD"/r5a.o \
  "$BLD"/r5b.o "$BLD"/r5c.o "$BLD"/r5d.o "$BLD"/r5e.o "$BLD"/r5f.o \
  "$BLD"/r6.o "$BLD"/r7.o "$BLD"/r7b.o "$BLD"/harness.o
echo "build_ref.sh: built $OUT/gfsph_ref"

# The same harness driving the GPU step through the product's ISO_C_BINDING
# layer (SPH_MAIN_GFSPH), linked against the in-tree libgfsph.so.
PKG="$HERE/../deslizamientos-y-flujos_amd"
if [ -f "$PKG/gfsph/libgfsph.so" ]; then
  "$FC" -c $FFLAGS -module-dir "$BLD" "$PKG/fortran/gfsph_c.f90" -o "$BLD/gfsph_c.o"
  "$FC" -c $FFLAGS -module-dir "$BLD" "$PKG/fortran/sph_main_gfsph.f90" -o "$BLD/sph_main_gfsph.o"
  "$FC" -c $FFLAGS -cpp -DGFSPH_GPU -module-dir "$BLD" "$HERE/ref_harness.f90" -o "$BLD/harness_gpu.o"
  "$FC" $FFLAGS -o "$OUT/gfsph_ref_gpu" \
    "$BLD"/r1.o "$BLD"/r3.o "$BLD"/r2.o "$BLD"/r4.o "$BLD"/r5.o "$BLD"/r5a.o \
    "$BLD"/r5b.o "$BLD"/r5c.o "$BLD"/r5d.o "$BLD"/r5e.o "$BLD"/r5f.o \
    "$BLD"/r6.o "$BLD"/r7.o "$BLD"/r7b.o "$BLD"/gfsph_c.o "$BLD"/sph_main_gfsph.o "$BLD"/harness_gpu.o \
    -L"$PKG/gfsph" -lgfsph -Wl,-rpath,'$ORIGIN/../../deslizamientos-y-flujos_amd/gfsph'
  echo "build_ref.sh: built $OUT/gfsph_ref_gpu"
fi
