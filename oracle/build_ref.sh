#!/bin/sh
# Builds oracle/_ref/libo3dref_{rm,wl,exp,v2}.so from the reference's own kernel
# headers, read in place from the reference checkout (default /root/reference).
# Test infrastructure only.  The core is a compile-time #define in CoreFunc.h
# (lines 35-39), so each core is compiled from a temporary copy of CoreFunc.h,
# made outside this repository, in which only that #define is changed.  The
# reference's own build system is not used; nothing but the .so files is written
# into the repository.
set -eu
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${O3D_REFERENCE:-/root/reference}/src
OUT="$HERE/_ref"
CXX=${CXX:-g++}
if [ ! -f "$REF/Kernels.h" ] || [ ! -f "$REF/CoreFunc.h" ] || [ ! -f "$REF/MathHelper.h" ]; then
  echo "build_ref.sh: reference headers not found under $REF; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/o3dref.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
for pair in RM:rm WL:wl EXPONENTIAL:exp V2:v2; do
  macro=${pair%%:*}; name=${pair##*:}
  mkdir -p "$TMP/$name"
  # Kernels.h includes "CoreFunc.h" from its own directory, so it sits beside the
  # core-selected copy; MathHelper.h is found in the reference tree through -I.
  cp "$REF/Kernels.h" "$TMP/$name/Kernels.h"
  sed -e 's|^#define USE_WL_KERNEL|//#define USE_WL_KERNEL|' \
      -e "s|^//#define USE_${macro}_KERNEL|#define USE_${macro}_KERNEL|" \
      "$REF/CoreFunc.h" > "$TMP/$name/CoreFunc.h"
  n=$(grep -c '^#define USE_.*_KERNEL' "$TMP/$name/CoreFunc.h")
  if [ "$n" != 1 ] || ! grep -q "^#define USE_${macro}_KERNEL" "$TMP/$name/CoreFunc.h"; then
    echo "build_ref.sh: could not select core $macro" >&2; exit 1
  fi
  # strict IEEE scalar semantics: no FMA contraction, no fast-math
  $CXX -std=c++17 -O2 -ffp-contract=off -fno-fast-math -fPIC -shared -fopenmp \
       -I"$TMP/$name" -I"$REF" "$HERE/ref_harness.cpp" -o "$OUT/libo3dref_$name.so"
done
# the reference's VTK XML writer (VtkXmlWriter.h, vendored tinyxml2 and cppcodec),
# driven through Points::write_vtk's and Surfaces::write_vtk's call sequences by ref_vtk_harness.cpp
LIB=${O3D_REFERENCE:-/root/reference}/lib
if [ -f "$REF/VtkXmlWriter.h" ] && [ -f "$LIB/tinyxml2/tinyxml2.cpp" ]; then
  $CXX -std=c++17 -O2 -fPIC -shared -I"$REF" -I"$LIB" "$HERE/ref_vtk_harness.cpp" "$LIB/tinyxml2/tinyxml2.cpp" \
       -o "$OUT/libo3dref_vtk.so"
fi
echo "build_ref.sh: built $(ls "$OUT" | tr '\n' ' ')"
