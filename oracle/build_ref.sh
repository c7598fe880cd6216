#!/usr/bin/env bash
# Build the reference's own CPU programs from their sources in place under
# /root/reference/src, with the flags of src/Makefile (g++ -O1), into
# oracle/_ref/.  Test infrastructure only: the binaries are checkers for the
# oracle, never part of the product.  Nothing is copied into the repository:
# variants with other compile-time bin counts are produced by streaming the
# source through sed into g++.
#
# Outputs (all git-ignored):
#   _ref/tri32      3pcf_C_version_triangle_parametrization    (THETA_NBINS 32, as shipped)
#   _ref/tri25      same with THETA_NBINS 25 (the sandbox histograms' setting)
#   _ref/trij32     3pcf_C_version_triangle_parametrization_JBJ
#   _ref/plain      3pcf_C_version                              (64 bins, as shipped)
#   _ref/plain16    3pcf_C_version with DEFAULT_NBINS 16
#   _ref/two        2pcf_C_version                              (700 bins on [0,70), as shipped)
set -euo pipefail
SRC=${REF_SRC:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$SRC" ]; then
    echo "reference sources not present at $SRC; nothing to build" >&2
    exit 0
fi
mkdir -p "$OUT"
CXX=${CXX:-g++}
FLAGS="-O1 -w"
$CXX $FLAGS -o "$OUT/tri32" "$SRC/3pcf_C_version_triangle_parametrization.cc" -lm
sed 's/^#define THETA_NBINS 32/#define THETA_NBINS 25/' "$SRC/3pcf_C_version_triangle_parametrization.cc" \
    | $CXX $FLAGS -x c++ -o "$OUT/tri25" - -lm
$CXX $FLAGS -o "$OUT/trij32" "$SRC/3pcf_C_version_triangle_parametrization_JBJ.cc" -lm
$CXX $FLAGS -o "$OUT/plain" "$SRC/3pcf_C_version.cc" -lm
sed 's/^#define DEFAULT_NBINS 64/#define DEFAULT_NBINS 16/' "$SRC/3pcf_C_version.cc" \
    | $CXX $FLAGS -x c++ -o "$OUT/plain16" - -lm
$CXX $FLAGS -o "$OUT/two" "$SRC/2pcf_C_version.cc" -lm
echo "built: $(ls "$OUT" | tr '\n' ' ')"
