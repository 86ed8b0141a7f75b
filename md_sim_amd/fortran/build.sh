#!/bin/bash
# Builds the Fortran interface module and programs against libmdforce.so (amdflang).
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
LIBDIR=$(cd "$HERE/.." && pwd)
FC=${FC:-amdflang}
OUT="$HERE/_fbuild"
mkdir -p "$OUT"
$FC -O2 -module-dir "$OUT" -c "$HERE/mdforce_c.f90" -o "$OUT/mdforce_c.o"
for prog in mdf_force_once; do
    $FC -O2 -module-dir "$OUT" -I"$OUT" -c "$HERE/$prog.f90" -o "$OUT/$prog.o"
    $FC -o "$HERE/$prog" "$OUT/$prog.o" "$OUT/mdforce_c.o" -L"$LIBDIR" -lmdforce -Wl,-rpath,"\$ORIGIN/.."
done
