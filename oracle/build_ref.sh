#!/bin/bash
# Compiles the reference's host-only preprocessing and sparse-curl sources, as
# they lie under the reference tree, together with oracle/ref_driver.cpp into
# oracle/_ref/libfame_ref_host.so.  Test infrastructure only; the output stays
# out of git.  Does nothing when the reference tree is absent (GPU box).
set -e
REF="${FAME_REFERENCE_DIR:-/root/reference}"
HERE="$(cd "$(dirname "$0")" && pwd)"
[ -d "$REF/FAME_Preprocessing" ] || { echo "build_ref: reference tree not present, skipping"; exit 0; }
mkdir -p "$HERE/_ref"
SRCS=( "$REF"/FAME_Preprocessing/*.cpp "$REF"/FAME_Tools/*.cpp "$REF/FAME_Main/FAME_Matrix_Curl.cpp" "$REF/FAME_Main/FAME_Malloc_mtx_C.cpp" )
g++ -O2 -fPIC -shared -std=gnu++11 -w -I"$REF/include" -o "$HERE/_ref/libfame_ref_host.so.tmp" "$HERE/ref_driver.cpp" "${SRCS[@]}"
mv "$HERE/_ref/libfame_ref_host.so.tmp" "$HERE/_ref/libfame_ref_host.so"
echo "build_ref: built $HERE/_ref/libfame_ref_host.so"
