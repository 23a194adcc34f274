#!/bin/bash
# Builds the reference's own driver (swe_code/main/main.c) with the HIP
# back-end dropped in -- TEST INFRASTRUCTURE ONLY: it proves the C ABI of
# libswe_hip.so (include/swe_hip.h) against the reference's main(), its host
# modules (reader, RCM, layouts, patches), verify_output and profiling.
# The driver is patched in a temporary directory by dropin_main.py (which
# holds the inserted lines); nothing from the reference tree is copied into
# the repository.  Needs rbf-fd-swe_amd/libswe_hip.so built first.
#
# Outputs (git-ignored, shipped with the tree like the rest of _ref/):
#   _ref/swe_ref_hip_<layout>.exe    FP64 (RUN_32BIT commented out, as build_ref.sh)
#   _ref/swe_ref32_hip_<layout>.exe  the reference's header as shipped (FP32)
set -euo pipefail
REF=${SWE_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
ROOT=$(cd "$HERE/.." && pwd)
OUT="$HERE/_ref"
S="$REF/swe_code"
if [ ! -f "$S/main/main.c" ]; then
    echo "build_dropin: reference tree not found at $REF; skipping" >&2
    exit 0
fi
if [ ! -f "$ROOT/rbf-fd-swe_amd/libswe_hip.so" ]; then
    echo "build_dropin: rbf-fd-swe_amd/libswe_hip.so not built" >&2
    exit 1
fi
SIMD=${SWE_REF_SIMD:-4}
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
mkdir -p "$TMP/fp64" "$OUT"
sed 's|^#define RUN_32BIT|// RUN_32BIT off: FP64 build|' "$S/main/include/swe_config.h" > "$TMP/fp64/swe_config.h"
python3 "$HERE/dropin_main.py" "$S/main/main.c" "$TMP/main_hip.c"
INC="-I$S/main/include -I$S/io/include -I$S/rcm/include -I$S/layout/include -I$S/mpi/include -I$S/ocl/include -I$ROOT/include"
LIBSRC="$S/main/rk4_rbffd_swe.c $S/main/runtime_params.c $S/main/profiling.c $S/io/input.c $S/rcm/rcm.c $S/rcm/reorder_nodes.c $S/layout/layout.c $S/layout/matrix_transformations.c $S/mpi/init_patches.c $S/mpi/halos.c"
COMMON="-std=gnu99 -include time.h -DUSE_BINIO -DSIMD_LENGTH=$SIMD -DPADDED_NUM=$SIMD -w -O2 -fopenmp"
LINK="-L$ROOT/rbf-fd-swe_amd -lswe_hip -Wl,-rpath,\$ORIGIN/../../rbf-fd-swe_amd -lm"
for layout in default cfdl sfdl cfdl_sfdl; do
    case $layout in
        default)   LF="" ;;
        cfdl)      LF="-DUSE_CFDL" ;;
        sfdl)      LF="-DUSE_SFDL" ;;
        cfdl_sfdl) LF="-DUSE_CFDL -DUSE_SFDL" ;;
    esac
    gcc $COMMON $LF -I"$TMP/fp64" $INC "$TMP/main_hip.c" $LIBSRC $LINK -o "$OUT/swe_ref_hip_$layout.exe"
    gcc $COMMON $LF $INC "$TMP/main_hip.c" $LIBSRC $LINK -o "$OUT/swe_ref32_hip_$layout.exe"
done
echo "build_dropin: built the reference driver with the HIP back-end into $OUT"
