#!/bin/bash
# Builds the reference's own per-function sources (read from /root/reference,
# never copied) plus our ref_driver.cpp into oracle/_ref/libref_<variant>.so.
# Test infrastructure only: the libraries check the oracle restatement
# function by function.  Variants differ only in which of the reference's own
# Config.h / Constants.h is first on the include path:
#   mhd      DRAGON/Config.h (MHD), TESTMODE runtime switches
#   hydro    Examples/1D/Sod/Config.h (hydro), TESTMODE
#   mhdprod  DRAGON/Config.h (MHD), production switches (CT beta-gated, MC, HLLD)
#   mhdg14   Examples/MHDRotor/{Config.h,Constants.h} (gamma = 1.4), TESTMODE
# Unsplit.cpp and Splitting.cpp are not built: they include <format>, which
# this image's C++ library does not provide.  CTU.cpp is built into the hydro
# variant for correctState and computeCTUFlux_X/Y/Z, which need neither; its
# ctu_sweep_* functions call Unsplit.cpp and are dropped at link time
# (-ffunction-sections, --gc-sections: nothing references them).
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
CXX=${REFCXX:-/opt/rocm/lib/llvm/bin/clang++}
[ -d "$REF/DRAGON" ] || { echo "reference tree not found at $REF; skipping"; exit 0; }
mkdir -p "$OUT"
INC=""
for d in DRAGON DRAGON/FluidElement DRAGON/Hydro DRAGON/Hydro/ExtendedArray DRAGON/Hydro/Reconstruction \
         DRAGON/Hydro/Riemann DRAGON/Hydro/Godunov DRAGON/Hydro/CFL DRAGON/MHD DRAGON/Boundary \
         DRAGON/Refinement DRAGON/main DRAGONWING DRAGONWING/Memory_Management DRAGONWING/Multithreading; do
  INC="$INC -I$REF/$d"
done
SRCS="DRAGON/FluidElement/FluidElement.cpp DRAGON/FluidElement/Arithmetic.cpp
      DRAGON/Hydro/Reconstruction/Limiters.cpp DRAGON/Hydro/Reconstruction/MUSCL.cpp
      DRAGON/Hydro/Riemann/HLL.cpp DRAGON/Hydro/Riemann/HLLD.cpp DRAGON/Hydro/Riemann/Exact.cpp
      DRAGON/Hydro/Riemann/Flux.cpp DRAGON/Hydro/Riemann/Roe.cpp DRAGON/Hydro/CFL/CFL.cpp
      DRAGON/MHD/Electric.cpp DRAGON/MHD/B_fields.cpp"
FLAGS="-std=c++23 -O2 -fPIC -ffp-contract=off -fno-fast-math -ffunction-sections -fdata-sections -fvisibility=hidden -w"
build() {
  local name=$1; shift
  local pre="$1"; shift
  local extra="$1"; shift
  local defs="$*"
  local objs=""
  local tmp=$OUT/obj_$name
  mkdir -p "$tmp"
  for s in $SRCS $extra; do
    local o=$tmp/$(basename "$s" .cpp).o
    $CXX $FLAGS $defs $pre $INC -c "$REF/$s" -o "$o"
    objs="$objs $o"
  done
  $CXX $FLAGS $defs $pre $INC -c "$HERE/ref_driver.cpp" -o "$tmp/ref_driver.o"
  $CXX -shared -Wl,--gc-sections -Wl,--no-undefined -o "$OUT/libref_$name.so" $objs "$tmp/ref_driver.o"
  rm -rf "$tmp"
  echo "built $OUT/libref_$name.so"
}
build mhd "" "" -DTESTMODE
build hydro "-I$REF/Examples/1D/Sod" "DRAGON/Hydro/Godunov/CTU.cpp" -DTESTMODE
build mhdprod "" ""
build mhdg14 "-I$REF/Examples/MHDRotor" "" -DTESTMODE
