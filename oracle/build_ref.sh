#!/bin/sh
# Builds the reference's CPU path (Multicore specification + templates) from
# its own sources where they lie, into oracle/_ref/ (git-ignored):
#   _ref/DrivenCavityCPU     the reference driver (DrivenCavity.cpp, -DENABLE_GPU=0), as-shipped flags (-O0)
#   _ref/gip_ref_harness     oracle/ref_harness.cpp over the same sources, -O2 (no FMA contraction on x86-64)
#   _ref/gip_ref_harness_O0  the same at the as-shipped flags (no -O, compile_cpu.sh:3-15)
#   _ref/gip_ref_harness_O3  the same at -O3 -march=x86-64-v3 (AVX2 + FMA: the optimised CPU baseline;
#                            any x86-64 host of the last decade runs it)
# Toolchain: the MPICH wrapper in /opt/conda with g++ as its compiler.
# Two flags beyond the reference's compile_cpu.sh: -include cstring
# (GIP_VectorMulticore.cpp uses memcpy without <cstring>) and a static
# libstdc++ (conda's runtime is older than g++'s).
# Exits 0 without building when the reference sources are absent.
set -e
REF=${GIP_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
REPO=$(cd "$HERE/.." && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/Algorithms/GIP_DNS.h" ]; then
    echo "build_ref: reference sources not found under $REF; skipped"
    exit 0
fi
if [ ! -x /opt/conda/bin/mpic++ ]; then
    echo "build_ref: /opt/conda/bin/mpic++ not found; skipped"
    exit 0
fi
export PATH=/opt/conda/bin:$PATH MPICH_CXX=g++
mkdir -p "$OUT/o0" "$OUT/o2" "$OUT/o3"
SRCS="src/BasicInterfaces/GIP_Topo.cpp src/AbstractInterfaces/GIP_Vector.cpp
src/Specifications/Multicore/GIP_VectorMulticore.cpp src/AbstractInterfaces/GIP_Matrix.cpp
src/Specifications/Multicore/GIP_MatrixMulticore.cpp src/AbstractInterfaces/GIP_Timer.cpp
src/Specifications/Multicore/GIP_TimerMulticore.cpp src/BasicInterfaces/GIP_Parameters.cpp
src/AbstractInterfaces/GIP_Arch.cpp src/Specifications/Multicore/GIP_ArchCPU.cpp
src/Specifications/Multicore/GIP_NonLinearMulticore.cpp src/AbstractInterfaces/GIP_NonLinear.cpp"
LINK="-fopenmp -static-libstdc++ -static-libgcc -L/opt/conda/lib -lmpi -Wl,-rpath,/opt/conda/lib"
O0=""
O2=""
O3=""
for s in $SRCS; do
    b=$(basename "$s" .cpp)
    mpic++ -include cstring -c "$REF/$s" -fopenmp -o "$OUT/o0/$b.o"
    mpic++ -include cstring -O2 -c "$REF/$s" -fopenmp -o "$OUT/o2/$b.o"
    mpic++ -include cstring -O3 -march=x86-64-v3 -c "$REF/$s" -fopenmp -o "$OUT/o3/$b.o"
    O0="$O0 $OUT/o0/$b.o"
    O2="$O2 $OUT/o2/$b.o"
    O3="$O3 $OUT/o3/$b.o"
done
mpic++ -include cstring -c "$REF/DrivenCavity.cpp" -fopenmp -DENABLE_GPU=0 -o "$OUT/o0/DrivenCavity.o"
g++ "$OUT/o0/DrivenCavity.o" $O0 -o "$OUT/DrivenCavityCPU" $LINK
mpic++ -include cstring -O2 -c "$HERE/ref_harness.cpp" -fopenmp -I"$REF" -I"$REPO/include" -o "$OUT/o2/ref_harness.o"
g++ "$OUT/o2/ref_harness.o" $O2 -o "$OUT/gip_ref_harness" $LINK -L"$REPO/gip_amd" -lgip_cavity -Wl,-rpath,'$ORIGIN/../../gip_amd'
mpic++ -include cstring -c "$HERE/ref_harness.cpp" -fopenmp -I"$REF" -I"$REPO/include" -o "$OUT/o0/ref_harness.o"
g++ "$OUT/o0/ref_harness.o" $O0 -o "$OUT/gip_ref_harness_O0" $LINK -L"$REPO/gip_amd" -lgip_cavity -Wl,-rpath,'$ORIGIN/../../gip_amd'
mpic++ -include cstring -O3 -march=x86-64-v3 -c "$HERE/ref_harness.cpp" -fopenmp -I"$REF" -I"$REPO/include" -o "$OUT/o3/ref_harness.o"
g++ "$OUT/o3/ref_harness.o" $O3 -o "$OUT/gip_ref_harness_O3" $LINK -L"$REPO/gip_amd" -lgip_cavity -Wl,-rpath,'$ORIGIN/../../gip_amd'
# Drop-in check: the reference's unmodified templates and abstract/basic
# interfaces instantiated with this repository's HIP specification, built by
# the maintainer's drop-in recipe (its marked blocks run verbatim in a
# temporary copy of the reference tree).
if [ -f "$REPO/gip_amd/libgip_hip.so" ]; then
    python3 "$HERE/dropin_build.py" "$REF" "$OUT/DrivenCavityHIP_reftemplates"
fi
rm -rf "$OUT/o0" "$OUT/o2" "$OUT/o3"
echo "build_ref: built $OUT/DrivenCavityCPU $OUT/gip_ref_harness"
