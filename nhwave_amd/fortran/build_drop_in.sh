#!/usr/bin/env bash
# Links the reference NHWAVE driver, unmodified, against libnhwave_amd: the
# drop-in form of the GPU hot path.
#
# 1. Compiles the reference's serial sources, read in place (NHWAVE_REF_SRC,
#    default /root/reference/src), into nhwave_amd/fortran/_build/ with the
#    flags of its FP64 hot-path build (-DDOUBLE_PRECISION -DFROUDE_CAP; serial;
#    -Dmyid=Zero because the serial build reads `myid` outside PARALLEL
#    guards; LLVM flang has no GNU rand, so each rand() call site, reached
#    only by spectral wavemakers, becomes an `error stop` with a message; the
#    driver object with inlining off so that its calls stay real calls).
#    nhw_serial_stubs.c supplies the one routine the serial build never
#    defines (phi_3D_exch, LES path), as a stop with a message; the link
#    fails on any other undefined symbol.
# 2. Weakens the definitions the shim replaces: estimate_dt_, update_mask_,
#    update_wind_, update_vars_, fluxes_, eval_duvw_, update_wave_bc_,
#    sponge_damping_, eval_turb_ (nhwave.o),
#    source_terms_ (source_terms.o), poisson_solver_ (pressure.o).
# 3. Compiles nhw_shim.F90 against the reference's own GLOBAL module and links
#    everything with libnhwave_amd.so into nhwave_amd/fortran/nhwave_gpu.
#
# The result reads the same input.txt / depth.txt / eta0.txt / uvw0.txt and
# writes the same outputs as the reference executable.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
PKG="$(cd "$HERE/.." && pwd)"
SRC="${NHWAVE_REF_SRC:-/root/reference/src}"
OUT="$HERE/_build"
FC="${FC:-/opt/rocm/lib/llvm/bin/amdflang}"
FFLAGS="-O2 -ffp-contract=off"
CPPDEFS=(-DDOUBLE_PRECISION -DFROUDE_CAP -Dmyid=Zero
         "-Drand(x)=0.0; error stop 'nhwave_ref: rand() unavailable (spectral wavemaker)'; TIME=0.0")
SELF="$HERE/$(basename "$0")"
if [ ! -d "$SRC" ]; then
  echo "build_drop_in: reference sources not found at $SRC; skipping" >&2
  exit 0
fi
if [ ! -f "$PKG/libnhwave_amd.so" ]; then
  echo "build_drop_in: $PKG/libnhwave_amd.so missing; build the library first" >&2
  exit 1
fi
mkdir -p "$OUT"
cd "$OUT"
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
objcopy --weaken-symbol=estimate_dt_ --weaken-symbol=update_mask_ --weaken-symbol=update_wind_ \
        --weaken-symbol=update_vars_ --weaken-symbol=fluxes_ --weaken-symbol=eval_duvw_ \
        --weaken-symbol=update_wave_bc_ --weaken-symbol=sponge_damping_ --weaken-symbol=eval_turb_ \
        nhwave.o nhwave_weak.o
objcopy --weaken-symbol=source_terms_ source_terms.o source_terms_weak.o
objcopy --weaken-symbol=poisson_solver_ pressure.o pressure_weak.o
"$FC" $FFLAGS -cpp -DDOUBLE_PRECISION -DFROUDE_CAP -c "$HERE/nhw_shim.F90" -o nhw_shim.o
gcc -O2 -c "$HERE/nhw_serial_stubs.c" -o nhw_serial_stubs.o
OBJS="mod_global.o mod_util.o nhwave_weak.o initialize.o two_layer_slide.o fluid_slide.o deformable_slide.o source_terms_weak.o waves.o pressure_weak.o immersed.o nspcg.o nhw_serial_stubs.o"
"$FC" $FFLAGS -o "$HERE/nhwave_gpu" nhw_shim.o $OBJS -L"$PKG" -lnhwave_amd \
  -Wl,-rpath,'$ORIGIN/..' -Wl,-rpath-link,/opt/rocm/lib 2> link.log \
  || { cat link.log; exit 1; }
echo "build_drop_in: built $HERE/nhwave_gpu"
