#!/bin/sh
# Builds the CPU restatement (oracle/awp_oracle.c) as oracle/lib/libawp_oracle.so
# holding the FP32 (*_f32), FP64 (*_f64) and mixed-storage (*_mix) variants.  Test infrastructure
# only.  -ffp-contract=off and no -march keep the FP32 operation order of the
# reference Fortran (no fused multiply-add, no reassociation).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/lib"
mkdir -p "$OUT"
CFLAGS="-O2 -fPIC -fopenmp -ffp-contract=off -fno-fast-math -std=c11 -Wall -Wno-unused-function"
gcc $CFLAGS -c "$HERE/awp_oracle.c" -o "$OUT/awp_oracle_f32.o"
gcc $CFLAGS -DREAL=double -DREAL_IS_DOUBLE -c "$HERE/awp_oracle.c" -o "$OUT/awp_oracle_f64.o"
gcc $CFLAGS -DMIXED -c "$HERE/awp_oracle.c" -o "$OUT/awp_oracle_mix.o"
gcc -shared -fopenmp "$OUT/awp_oracle_f32.o" "$OUT/awp_oracle_f64.o" "$OUT/awp_oracle_mix.o" -lm -o "$OUT/libawp_oracle.so.tmp"
mv "$OUT/libawp_oracle.so.tmp" "$OUT/libawp_oracle.so"
rm -f "$OUT/awp_oracle_f32.o" "$OUT/awp_oracle_f64.o" "$OUT/awp_oracle_mix.o"
