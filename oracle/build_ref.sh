#!/bin/bash
# Test infrastructure only.  Builds the reference's own C++ for the Galerkin singular integrals
# (DIRECTFN) and the face-face cubature, compiled from the sources where they lie, into oracle/_ref/,
# each behind a small C entry point (oracle/refshim/).  Needs the reference checkout (default
# /root/reference); the MEX gateways (which need mex.h) are not built.  Usage: oracle/build_ref.sh [REF]
set -euo pipefail
REF=${1:-/root/reference}
SRC=$REF/src/src_vie/src_operators/mexcubatures
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
[ -d "$SRC" ] || { echo "build_ref.sh: $SRC not found" >&2; exit 1; }
mkdir -p "$OUT"
CXX=${CXX:-g++}
FLAGS="-O2 -shared -fPIC -w"
$CXX $FLAGS -I"$SRC" "$HERE/refshim/directfn_st_shim.cpp" "$SRC/directfn_ws_st_const.cpp" -o "$OUT/libvsref_st.so"
$CXX $FLAGS -I"$SRC" "$HERE/refshim/directfn_ea_shim.cpp" "$SRC/directfn_ws_ea_const.cpp" -o "$OUT/libvsref_ea.so"
$CXX $FLAGS -I"$SRC" "$HERE/refshim/directfn_va_shim.cpp" "$SRC/directfn_ws_va_const.cpp" -o "$OUT/libvsref_va.so"
$CXX $FLAGS "$HERE/refshim/cubature_nop_shim.cpp" "$SRC/cubature_nop_const.cpp" -o "$OUT/libvsref_cub.so"
echo "built $OUT"
