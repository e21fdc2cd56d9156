#!/usr/bin/env bash
# Compile the reference's own BFM operator library (Fortran, unmodified, read in
# place under /root/reference/src) plus the test-only harness ref_harness.F90
# into oracle/_ref/ref_harness.  Test infrastructure only.
#
# Second product: oracle/_ref/bsacl_module_driver, the reference's own Fortran
# binding of BsaCL (BsaCL/BsaCL_Fortran/*.F90, unmodified) plus the test-only
# bsacl_module_driver.F90, linked against this repository's libbsahip.so
# (built first by bsalib_amd/build.py) in place of the reference BsaCL runtime.
#
# Third product: oracle/_ref/bsa_app, the reference's standalone program
# app/bsa.F90 (unmodified) compiled against this repository's module BsaLib.
#
# Units compiled (dependency order): CONSTANTS, Timer, BsaIO, Settings,
# StructureData, WindType, utility, BsaLibData(+Impl), IO/Settings/Structure/
# Wind implementations, BsaLib_Functions(+Impl).  The reference's own build
# system is not used.  functionsImpl.F90 references LAPACK dsyev_ only on the
# Mesher-POD path, which the harness never calls; the image has no LAPACK, so
# that one symbol is left unresolved at link time (no stand-in is provided).
set -euo pipefail
REF=${BSA_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$(cd "$HERE/.." && pwd)/_ref
SRC=$REF/src
FC=${FC:-amdflang}
FFLAGS="-cpp -O2 -fopenmp -fPIC -I$SRC/CONSTANTS"
[ -d "$SRC" ] || { echo "reference sources not found at $SRC" >&2; exit 3; }
rm -rf "$OUT/obj"
mkdir -p "$OUT/obj"
cd "$OUT/obj"
UNITS="CONSTANTS/CONSTANTS.F90 timing/Timer.F90 IO/BsaIO.F90 settings/Settings.F90
structure/StructureData.F90 wind/WindType.F90 utils/utility.F90 bsa/data/BsaLibData.F90
bsa/data/BsaLibDataImpl.F90 IO/BsaIOImpl.F90 settings/SettingsImpl.F90
structure/StructureImpl.F90 wind/WindSetImpl.F90 wind/WindPSDImpl.F90
bsa/functions/functions.F90 bsa/functions/functionsImpl.F90"
OBJS=""
for u in $UNITS; do
  o=$(basename "$u" .F90).o
  $FC $FFLAGS -c "$SRC/$u" -o "$o"
  OBJS="$OBJS $o"
done
$FC $FFLAGS -c "$HERE/ref_harness.F90" -o ref_harness.o
$FC -fopenmp -o "$OUT/ref_harness" ref_harness.o $OBJS -Wl,--unresolved-symbols=ignore-all
echo "built $OUT/ref_harness"

CLF=$REF/BsaCL/BsaCL_Fortran
LIBDIR=$(cd "$OUT/../.." && pwd)/bsalib_amd/_lib
if [ -f "$LIBDIR/libbsahip.so" ]; then
  mkdir -p "$OUT/obj/bsacl"
  cd "$OUT/obj/bsacl"
  CLFLAGS="-cpp -O2 -I$REF/BsaCL/include"
  CLOBJS="bsacl_c.o bsacl.o bsacl_impl.o"
  # dependency order: the submodule bsacl_impl uses module BsaCL_c
  $FC $CLFLAGS -c "$CLF/bsacl_c.F90" -o bsacl_c.o
  $FC $CLFLAGS -c "$CLF/bsacl.F90" -o bsacl.o
  $FC $CLFLAGS -c "$CLF/bsacl_impl.F90" -o bsacl_impl.o
  $FC $CLFLAGS -c "$HERE/bsacl_module_driver.F90" -o bsacl_module_driver.o
  $FC -o "$OUT/bsacl_module_driver" bsacl_module_driver.o $CLOBJS -L"$LIBDIR" -lbsahip \
      -Wl,-rpath,'$ORIGIN/../../bsalib_amd/_lib'
  echo "built $OUT/bsacl_module_driver"
else
  echo "libbsahip.so not built yet: skipping bsacl_module_driver" >&2
fi

# Third product: oracle/_ref/bsa_app, the reference's standalone program
# (app/bsa.F90, unmodified) compiled against this repository's module BsaLib
# (bsalib_amd/_lib/fortran) and linked to libbsalibf.a + libbsahip.so.  It reads
# bsa.bsadata / bsa.extdata with the reference's own reader.
FMOD=$(cd "$OUT/../.." && pwd)/bsalib_amd/_lib/fortran
if [ -f "$FMOD/libbsalibf.a" ] && [ -f "$LIBDIR/libbsahip.so" ]; then
  mkdir -p "$OUT/obj/app"
  cd "$OUT/obj/app"
  $FC -cpp -O2 -I"$FMOD" -c "$REF/app/bsa.F90" -o bsa.o
  $FC -o "$OUT/bsa_app" bsa.o "$FMOD/libbsalibf.a" -L"$LIBDIR" -lbsahip \
      -Wl,-rpath,'$ORIGIN/../../bsalib_amd/_lib'
  echo "built $OUT/bsa_app"
else
  echo "Fortran front not built yet: skipping bsa_app" >&2
fi
