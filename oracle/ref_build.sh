#!/bin/bash
# Builds the reference's CPU path (CpuCrowdSim, nav builder, exit choice, the metrics' CPU reference; the GL-side
# classes of crowd_sim.cpp and metrics.cpp link but are never called) from its own sources under
# /root/reference, plus ref_driver.cpp, into oracle/_ref/ref_driver. Test infrastructure only: the driver
# writes the golden vectors tests/golden/make_golden.py packs. Needs /root/reference (absent on GPU boxes).
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
mkdir -p "$OUT"
FLAGS="-std=c++17 -O2 -fopenmp -ffp-contract=off -fno-fast-math -I$REF/include -I$REF/external/glm -I$REF/external/glad/include"
OBJS=""
for f in cpu_crowd_sim sdf venue nav_field exit_choice shader assets metrics crowd_sim gpu_caps; do
    if [ ! -f "$OUT/$f.o" ] || [ "$REF/src/$f.cpp" -nt "$OUT/$f.o" ]; then
        g++ $FLAGS -c "$REF/src/$f.cpp" -o "$OUT/$f.o"
    fi
    OBJS="$OBJS $OUT/$f.o"
done
if [ ! -f "$OUT/glad.o" ]; then gcc -O2 -I"$REF/external/glad/include" -c "$REF/external/glad/src/glad.c" -o "$OUT/glad.o"; fi
g++ $FLAGS "$HERE/ref_driver.cpp" $OBJS "$OUT/glad.o" -o "$OUT/ref_driver" -ldl
echo "built $OUT/ref_driver"
