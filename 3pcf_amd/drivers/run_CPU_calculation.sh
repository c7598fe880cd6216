#!/usr/bin/env bash
# Drop-in for examples/run_CPU_calculation.csh (and run_GPU_calculation.csh):
#   run_CPU_calculation.sh [ngals_in_thousands=1] [which_part=all|0|1|2|3]
# Runs DDD / DDR / DRR / RRR with the file orders of
# examples/run_CPU_calculation.csh:65-88 and its output names
# {TERM}_{tag}_{ngals}k.dat.  Inputs, tag, extra flags and the program can be
# overridden: INPUT0, INPUT1, TAG, GLOBAL_PARAMS, EXECUTABLE.
set -euo pipefail
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
BIN_DIR=${BIN_DIR:-$HERE/../bin}
EXECUTABLE=${EXECUTABLE:-$BIN_DIR/3pcf_C_version_triangle_parametrization}
ngals=${1:-1}
which_part=${2:-all}
input0=${INPUT0:-../sample_data/MICE_5degx5deg_${ngals}k_randomized.dat}
input1=${INPUT1:-../sample_data/flat_MICE_5degx5deg_${ngals}k.dat}
tag=${TAG:-evenbinning_CPU_5degx5deg}
read -r -a params <<< "${GLOBAL_PARAMS:-}"

term() {  # term name, three inputs
    echo "#####################"
    time "$EXECUTABLE" "$2" "$3" "$4" "${params[@]}" -o "$1_${tag}_${ngals}k.dat"
}
case "$which_part" in
    all) term DDD "$input0" "$input0" "$input0"; term DDR "$input0" "$input0" "$input1"
         term DRR "$input0" "$input1" "$input1"; term RRR "$input1" "$input1" "$input1"
         echo "#####################" ;;
    0) term DDD "$input0" "$input0" "$input0" ;;
    1) term DDR "$input0" "$input0" "$input1" ;;
    2) term DRR "$input0" "$input1" "$input1" ;;
    3) term RRR "$input1" "$input1" "$input1" ;;
    *) echo "which_part must be all, 0, 1, 2 or 3" >&2; exit 2 ;;
esac
