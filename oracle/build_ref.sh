#!/bin/bash
# Builds the golden-vector harness from the reference's own Fortran sources, in the
# development container only.  Outputs go to oracle/_ref/ (never committed).
#   slab: src/forcem.f90      (open z, slab correction, wall tail correction)
#   bulk: src/forcem_back.f90 (3-D periodic pair loops, no slab/wall terms)
# No fast-math and no FMA contraction, so the vectors are the plain FP64 path; the _fast
# timing builds at the end use the reference's own flags.
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
# NVE drivers: the reference's time loop with its own force module (md_nve_slab,
# md_nve_bulk), and the same loop with the engine's drop-in module force
# (md_sim_amd/fortran/force_mdf.f90) in place of forcem.f90 (md_nve_gpu)
for variant in slab bulk; do
    D="$OUT/$variant"
    $FC $FLAGS -module-dir "$D" -I"$D" -c "$HERE/md_nve.f90" -o "$D/md_nve.o"
    objs=()
    if [ $variant = slab ]; then FORCE=forcem.f90; else FORCE=forcem_back.f90; fi
    for f in global.f90 comm.f90 rndmod.f90 sysdef.f90 config.f90 nnlist.f90 measure.f90 $FORCE move.f90; do
        objs+=("$D/${f%.f90}.o")
    done
    $FC -o "$OUT/md_nve_$variant" "$D/md_nve.o" "${objs[@]}" -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
done
LIBDIR=$(cd "$HERE/../md_sim_amd" && pwd)
FDIR="$HERE/../md_sim_amd/fortran"
D="$OUT/dropin"
mkdir -p "$D"
objs=()
for f in global.f90 comm.f90 rndmod.f90 sysdef.f90 config.f90 nnlist.f90 measure.f90; do
    o="$D/${f%.f90}.o"
    $FC $FLAGS -module-dir "$D" -c "$REF/$f" -o "$o"
    objs+=("$o")
done
$FC $FLAGS -module-dir "$D" -c "$FDIR/mdforce_c.f90" -o "$D/mdforce_c.o"
$FC $FLAGS -module-dir "$D" -I"$D" -c "$FDIR/force_mdf.f90" -o "$D/force_mdf.o"
$FC $FLAGS -module-dir "$D" -I"$D" -c "$REF/move.f90" -o "$D/move.o"
$FC $FLAGS -module-dir "$D" -I"$D" -c "$HERE/md_nve.f90" -o "$D/md_nve.o"
objs+=("$D/mdforce_c.o" "$D/force_mdf.o" "$D/move.o")
# the drop-in module on the CPU oracle (defect-free FP64 trajectory for the drift test)
gcc -O2 -fPIC -shared -ffp-contract=off -o "$OUT/libmdforce_oracle.so" "$HERE/oracle.c" "$HERE/orc_as_mdf.c" -lm
$FC -o "$OUT/md_nve_oracle" "$D/md_nve.o" "${objs[@]}" -L"$OUT" -lmdforce_oracle -Wl,-rpath,'$ORIGIN' \
    -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
# the drop-in module on the engine
if [ -f "$LIBDIR/libmdforce.so" ]; then
    $FC -o "$OUT/md_nve_gpu" "$D/md_nve.o" "${objs[@]}" -L"$LIBDIR" -lmdforce -Wl,-rpath,'$ORIGIN/../../md_sim_amd' \
        -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
fi
# timing builds for the CPU baseline: the same harness with the reference's own compiler
# flags (src/Makefile:2, -O2 -ffast-math), ref_slab_fast / ref_bulk_fast
FAST="-O2 -ffast-math -w -I$MPI_INC"
for variant in slab bulk; do
    if [ $variant = slab ]; then FORCE=forcem.f90; else FORCE=forcem_back.f90; fi
    D="$OUT/${variant}_fast"
    mkdir -p "$D"
    objs=()
    for f in global.f90 comm.f90 rndmod.f90 sysdef.f90 config.f90 nnlist.f90 measure.f90 $FORCE move.f90; do
        o="$D/${f%.f90}.o"
        $FC $FAST -module-dir "$D" -c "$REF/$f" -o "$o"
        objs+=("$o")
    done
    $FC $FAST -module-dir "$D" -I"$D" -c "$HERE/ref_harness.f90" -o "$D/ref_harness.o"
    $FC -o "$OUT/ref_${variant}_fast" "$D/ref_harness.o" "${objs[@]}" \
        -L"$MPI_LIB" -lmpifort -lmpi -Wl,-rpath,"$MPI_LIB"
done
echo "built $OUT/ref_slab $OUT/ref_bulk $OUT/ref_slab_fast $OUT/ref_bulk_fast $OUT/md_nve_slab $OUT/md_nve_bulk $OUT/md_nve_oracle $OUT/md_nve_gpu"
