#!/bin/bash
# Builds the reference's own C sources, where they lie under the reference
# tree, into oracle/_ref/ -- TEST INFRASTRUCTURE ONLY (parity checker and the
# "reference" CPU baseline).  Nothing from the reference tree is copied into
# the repository: the only derived file, the FP64 variant of
# swe_code/main/include/swe_config.h (its RUN_32BIT line, :75, commented out,
# as the reference's own FP64 builds had it), lives in a temporary directory
# that is removed when the script ends.
#
# Outputs (git-ignored, shipped with the tree):
#   _ref/libswe_ref_<layout>.so  reference sources + ref_harness.c, strict
#                                FP64 (-O2 -ffp-contract=off), per layout
#   _ref/libswe_ref32_<layout>.so  the same with the reference's header as
#                                shipped (RUN_32BIT: fType = float)
#   _ref/swe_ref_<layout>.exe    the reference main() built like its best CPU
#                                configuration (-O3 -fopenmp -DRHS_INNER_SIMD)
# Layouts: default, cfdl, sfdl, cfdl_sfdl (compile_default.sh:3-6 matrix);
# SIMD_LENGTH=4 as in arch/gnu/config.swe.
#
# Build-flag notes (no source edits): -DUSE_GCC is not passed (it stacks
# "omp simd" on "omp parallel for", rk4_rbffd_swe.c:96-98, which gcc 11
# rejects) and <time.h> is force-included (profiling.c:31 uses
# CLOCK_MONOTONIC without including it).
set -euo pipefail
REF=${SWE_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
OUT="$HERE/_ref"
S="$REF/swe_code"
if [ ! -f "$S/main/rk4_rbffd_swe.c" ]; then
    echo "build_ref: reference tree not found at $REF; skipping" >&2
    exit 0
fi
SIMD=${SWE_REF_SIMD:-4}
ARCH=${SWE_REF_ARCH:--march=x86-64-v3}
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
sed 's|^#define RUN_32BIT|// RUN_32BIT off: FP64 build|' "$S/main/include/swe_config.h" > "$TMP/swe_config.h"
mkdir -p "$OUT"
INC="-I$TMP -I$S/main/include -I$S/io/include -I$S/rcm/include -I$S/layout/include -I$S/mpi/include -I$S/ocl/include"
INC32="-I$S/main/include -I$S/io/include -I$S/rcm/include -I$S/layout/include -I$S/mpi/include -I$S/ocl/include"
LIBSRC="$S/main/rk4_rbffd_swe.c $S/main/runtime_params.c $S/main/profiling.c $S/io/input.c $S/rcm/rcm.c $S/rcm/reorder_nodes.c $S/layout/layout.c $S/layout/matrix_transformations.c $S/mpi/init_patches.c $S/mpi/halos.c"
COMMON="-std=gnu99 -include time.h -DUSE_BINIO -DSIMD_LENGTH=$SIMD -DPADDED_NUM=$SIMD -w"
for layout in default cfdl sfdl cfdl_sfdl; do
    case $layout in
        default)   LF="" ;;
        cfdl)      LF="-DUSE_CFDL" ;;
        sfdl)      LF="-DUSE_SFDL" ;;
        cfdl_sfdl) LF="-DUSE_CFDL -DUSE_SFDL" ;;
    esac
    gcc $COMMON $LF -O2 -ffp-contract=off -fopenmp -fPIC -shared $INC \
        "$HERE/ref_harness.c" $LIBSRC -lm -o "$OUT/libswe_ref_$layout.so"
    gcc $COMMON $LF -O2 -ffp-contract=off -fopenmp -fPIC -shared $INC32 \
        "$HERE/ref_harness.c" $LIBSRC -lm -o "$OUT/libswe_ref32_$layout.so"
    gcc $COMMON $LF -O3 $ARCH -fopenmp -DRHS_INNER_SIMD $INC \
        "$S/main/main.c" $LIBSRC -lm -o "$OUT/swe_ref_$layout.exe"
done
echo "build_ref: built reference libraries and executables into $OUT"
