#!/bin/bash
# Builds the golden-vector harness from the reference's own Fortran sources, in the
# development container only.  Outputs go to oracle/_ref/ (never committed).
#   slab: src/forcem.f90      (open z, slab correction, wall tail correction)
#   bulk: src/forcem_back.f90 (3-D periodic pair loops, no slab/wall terms)
# No fast-math and no FMA contraction, so the vectors are the plain FP64 path.
set -euo pipefail
REF=${MD_SIM_REF_SRC:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
MPI_INC=${MPI_INC:-/opt/conda/include}
MPI_LIB=${MPI_LIB:-/opt/conda/lib}
FC=${FC:-amdflang}
FLAGS="-O2 -ffp-contract=off -w -I$MPI_INC"
if [ ! -d "$REF" ]; then echo "reference sources not found at $REF" >&2; exit 3; fi
for variant in slab bulk; do
    if [ $variant = slab ]; then FORCE=forcem.f90; else FORCE=forcem_back.f90; fi
    D="$OUT/$variant"
    mkdir -p "$D"
    objs=()
    for f in global.f90 comm.f90 rndmod.f90 sysdef.f90 config.f90 nnlist.f90 measure.f90 $FORCE move.f90; do
        o="$D/${f%.f90}.o"
        $FC $FLAGS -module-dir "$D" -c "$REF/$f" -o "$o"
        objs+=("$o")
    done
    $FC $FLAGS -module-dir "$D" -I"$D" -c "$HERE/ref_harness.f90" -o "$D/ref_harness.o"
    $FC -o "$OUT/ref_$variant" "$D/ref_harness.o" "${objs[@]}" \
        -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
done
echo "built $OUT/ref_slab $OUT/ref_bulk"
