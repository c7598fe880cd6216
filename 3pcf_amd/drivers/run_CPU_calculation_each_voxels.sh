#!/usr/bin/env bash
# Drop-in for examples/run_CPU_calculation_each_voxels.csh (and
# run_GPU_calculation_each_voxels.csh):
#   run_CPU_calculation_each_voxels.sh [ngals=1] [which_part=0] [vox_div=2] [index=000000000]
# One term of one index-chunk voxel (-X vox_div -x index), output
# {TERM}_voxel{div}_{index}_{tag}_{ngals}k.dat
# (examples/run_CPU_calculation_each_voxels.csh:60-67).  Overrides: INPUT0,
# INPUT1, TAG, GLOBAL_PARAMS, EXECUTABLE.
set -euo pipefail
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
BIN_DIR=${BIN_DIR:-$HERE/../bin}
EXECUTABLE=${EXECUTABLE:-$BIN_DIR/3pcf_C_version_triangle_parametrization}
ngals=${1:-1}
which_part=${2:-0}
vox_div=${3:-2}
index=${4:-000000000}
input0=${INPUT0:-../sample_data/MICE_${ngals}k_randomized.dat}
input1=${INPUT1:-../sample_data/flat_MICE_${ngals}k.dat}
tag=${TAG:-MICE_full_sky_CPU_LRG_binning_SLAC}
read -r -a params <<< "${GLOBAL_PARAMS:-}"
echo "$index"
run() { time "$EXECUTABLE" "$2" "$3" "$4" "${params[@]}" -o "$1_voxel${vox_div}_${index}_${tag}_${ngals}k.dat" -X "$vox_div" -x "$index"; }
case "$which_part" in
    0) run DDD "$input0" "$input0" "$input0" ;;
    1) run DDR "$input0" "$input0" "$input1" ;;
    2) run DRR "$input0" "$input1" "$input1" ;;
    3) run RRR "$input1" "$input1" "$input1" ;;
    *) echo "which_part must be 0, 1, 2 or 3" >&2; exit 2 ;;
esac
