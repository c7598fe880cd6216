#!/usr/bin/env bash
# Drop-in for examples/run_CPU_calculation_2pcf.csh:
#   run_CPU_calculation_2pcf.sh [ngals_in_thousands=1] [which_part=all|0|1|2]
# Runs DD / DR / RR with the file orders of
# examples/run_CPU_calculation_2pcf.csh:36-53 and its output names
# {TERM}_{tag}_{ngals}k.dat.  Inputs, tag, extra flags and the program can be
# overridden: INPUT0, INPUT1, TAG, GLOBAL_PARAMS, EXECUTABLE.
set -euo pipefail
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
BIN_DIR=${BIN_DIR:-$HERE/../bin}
EXECUTABLE=${EXECUTABLE:-$BIN_DIR/2pcf_C_version}
ngals=${1:-1}
which_part=${2:-all}
input0=${INPUT0:-../sample_data/MICE_5degx5deg_${ngals}k_randomized.dat}
input1=${INPUT1:-../sample_data/flat_MICE_5degx5deg_${ngals}k.dat}
tag=${TAG:-2pcf_TESTING_evenbinning_CPU_5degx5deg_LRG_binning}
read -r -a params <<< "${GLOBAL_PARAMS:-}"

term() {  # term name, two inputs
    echo "#####################"
    time "$EXECUTABLE" "$2" "$3" "${params[@]}" -o "$1_${tag}_${ngals}k.dat"
}
case "$which_part" in
    all) term DD "$input0" "$input0"; term DR "$input0" "$input1"; term RR "$input1" "$input1" ;;
    0) term DD "$input0" "$input0" ;;
    1) term DR "$input0" "$input1" ;;
    2) term RR "$input1" "$input1" ;;
    *) echo "which_part must be all, 0, 1 or 2" >&2; exit 2 ;;
esac
