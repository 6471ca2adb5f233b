#!/usr/bin/env bash
# Builds the reference NHWAVE serial (NSPCG) path from its own sources, read in
# place under the reference tree, into oracle/_ref/ (git-ignored).  Nothing from
# the reference is copied into tracked paths: the preprocessed .f90 files and
# objects are build outputs that live only in oracle/_ref/.
#
# Flags only, no source edits:
#   -DDOUBLE_PRECISION -DFROUDE_CAP   the reference's own FP64 feature set
#                                     (src/Makefile.mills.mpif90:10,24)
#   no -DPARALLEL                     serial NSPCG path (src/pressure.F:81-145)
#   no -DINTEL                        avoids `use IFPORT` (src/initialize.F:119-121)
#   -D'rand(x)=...error stop...'      LLVM flang has no GNU rand(); it is read
#                                     only in random-phase set-up of spectral
#                                     wavemakers (src/initialize.F:826-828,941).
#                                     Each call site becomes `0.0; error stop
#                                     '<message>'; TIME=0.0...` (three statements
#                                     on the line; the last is never reached), so
#                                     a run that reaches one stops with a message
#                                     instead of using a made-up phase.
#   -Dmyid=Zero                       the serial build reads `myid` outside any
#                                     PARALLEL guard (src/nhwave.F:6395-6397,
#                                     src/waves.F:382,642) but only declares it
#                                     under PARALLEL (src/mod_global.F:48-49);
#                                     rank 0 is the only rank, so the
#                                     GLOBAL parameter Zero stands for it.
#   oracle/ref_stubs.c                `phi_3d_exch_` (unguarded call in les_dyn,
#                                     src/nhwave.F:8077) sits on a LES path the
#                                     hot-path inputs never execute; the stub
#                                     stops the run with a message.  The link
#                                     fails on any other undefined symbol.
#   -mllvm -inline-threshold=-10000   keeps fluxes/eval_duvw/update_mask/preview
#                                     as real calls from the driver so that they
#                                     can be substituted at link time.
#
# Outputs:
#   oracle/_ref/nhwave_ref        the unmodified serial reference
#   oracle/_ref/nhwave_ref_dump   same, with `preview` replaced by
#                                 oracle/ref_dump.f90 (raw FP64 state dumps)
#   oracle/_ref/nhwave_ref_solves the unmodified reference with NSPCG's driver
#                                 wrapped by oracle/ref_solve_log.f90 (the
#                                 true residual of every Poisson solve)
#   oracle/_ref/obj/*.o, *.mod    reference objects (checker only)
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
SELF="$HERE/$(basename "$0")"
SRC="${NHWAVE_REF_SRC:-/root/reference/src}"
OUT="$HERE/_ref"
FC="${FC:-/opt/rocm/lib/llvm/bin/amdflang}"
FFLAGS="-O2 -ffp-contract=off"
CPPDEFS=(-DDOUBLE_PRECISION -DFROUDE_CAP -Dmyid=Zero
         "-Drand(x)=0.0; error stop 'nhwave_ref: rand() unavailable (spectral wavemaker)'; TIME=0.0")
if [ ! -d "$SRC" ]; then
  echo "build_ref: reference sources not found at $SRC; skipping" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
cd "$OUT/obj"
MODS="mod_global mod_util nhwave initialize two_layer_slide fluid_slide deformable_slide source_terms waves pressure immersed"
for m in $MODS; do
  if [ ! -f "$m.o" ] || [ "$SRC/$m.F" -nt "$m.o" ] || [ "$SELF" -nt "$m.o" ]; then
    cpp -P -traditional "${CPPDEFS[@]}" "$SRC/$m.F" > "$m.f90"
    extra=""
    if [ "$m" = nhwave ]; then extra="-mllvm -inline-threshold=-10000"; fi
    "$FC" $FFLAGS $extra -c "$m.f90" -o "$m.o"
  fi
done
if [ ! -f nspcg.o ] || [ "$SRC/nspcg.f" -nt nspcg.o ]; then
  "$FC" $FFLAGS -c "$SRC/nspcg.f" -o nspcg.o
fi
gcc -O2 -c "$HERE/ref_stubs.c" -o ref_stubs.o
OBJS=""
for m in $MODS; do OBJS="$OBJS $m.o"; done
"$FC" $FFLAGS -o "$OUT/nhwave_ref" $OBJS nspcg.o ref_stubs.o 2> link.log || { cat link.log; exit 1; }
# dump variant: `preview` becomes a weak symbol in nhwave.o and the strong one
# from ref_dump.f90 wins at link time.
objcopy --weaken-symbol=preview_ nhwave.o nhwave_weakprev.o
"$FC" $FFLAGS -c "$HERE/ref_dump.f90" -o ref_dump.o
DOBJS=$(echo "$OBJS" | sed 's/ nhwave.o / nhwave_weakprev.o /')
"$FC" $FFLAGS -o "$OUT/nhwave_ref_dump" $DOBJS ref_dump.o nspcg.o ref_stubs.o 2>> link.log || { cat link.log; exit 1; }
# solve-log variant: NSPCG's driver renamed in a copy of its object file, the
# wrapper of ref_solve_log.f90 logging each solve's true residual
objcopy --redefine-sym nspcg_=nspcg_orig_ nspcg.o nspcg_orig.o
"$FC" $FFLAGS -c "$HERE/ref_solve_log.f90" -o ref_solve_log.o
"$FC" $FFLAGS -o "$OUT/nhwave_ref_solves" $OBJS ref_solve_log.o nspcg_orig.o ref_stubs.o 2>> link.log || { cat link.log; exit 1; }
echo "build_ref: built $OUT/nhwave_ref, $OUT/nhwave_ref_dump and $OUT/nhwave_ref_solves"
