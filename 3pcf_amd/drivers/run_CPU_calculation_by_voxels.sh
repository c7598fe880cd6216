#!/usr/bin/env bash
# Drop-in for examples/run_CPU_calculation_by_voxels.csh:
#   run_CPU_calculation_by_voxels.sh [ngals=1] [which_part=all|0..3]
# Loops over every vox_div^3 index-chunk voxel (VOX_DIV, default 2), one
# program run per voxel (examples/run_CPU_calculation_by_voxels.csh:61-85).
# BATCH=1: one program run per term instead, counting every voxel of that
# term in one batched device launch (--all-voxels); same files, same bytes.
set -euo pipefail
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
BIN_DIR=${BIN_DIR:-$HERE/../bin}
EXECUTABLE=${EXECUTABLE:-$BIN_DIR/3pcf_C_version_triangle_parametrization}
ngals=${1:-1}
which_part=${2:-all}
vox_div=${VOX_DIV:-2}
input0=${INPUT0:-../sample_data/MICE_5degx5deg_${ngals}k_randomized.dat}
input1=${INPUT1:-../sample_data/flat_MICE_5degx5deg_${ngals}k.dat}
tag=${TAG:-evenbinning_CPU_5degx5deg_LRG_binning_XSEDE}
read -r -a params <<< "${GLOBAL_PARAMS:-}"
echo "#####################"
if [[ "${BATCH:-0}" == 1 ]]; then
  batch() { time "$EXECUTABLE" "$2" "$3" "$4" "${params[@]}" -o "$1_voxel${vox_div}_{index}_${tag}_${ngals}k.dat" -X "$vox_div" --all-voxels; }
  case "$which_part" in
    0) batch DDD "$input0" "$input0" "$input0" ;;
    1) batch DDR "$input0" "$input0" "$input1" ;;
    2) batch DRR "$input0" "$input1" "$input1" ;;
    3) batch RRR "$input1" "$input1" "$input1" ;;
    all) batch DDD "$input0" "$input0" "$input0"; batch DDR "$input0" "$input0" "$input1"
         batch DRR "$input0" "$input1" "$input1"; batch RRR "$input1" "$input1" "$input1" ;;
  esac
  exit 0
fi
for ((i = 0; i < vox_div; i++)); do
  for ((j = 0; j < vox_div; j++)); do
    for ((k = 0; k < vox_div; k++)); do
      index=$(printf "%03d%03d%03d" "$i" "$j" "$k")
      echo "$index"
      run() { time "$EXECUTABLE" "$2" "$3" "$4" "${params[@]}" -o "$1_voxel${vox_div}_${index}_${tag}_${ngals}k.dat" -X "$vox_div" -x "$index"; }
      case "$which_part" in
        0) run DDD "$input0" "$input0" "$input0" ;;
        1) run DDR "$input0" "$input0" "$input1" ;;
        2) run DRR "$input0" "$input1" "$input1" ;;
        3) run RRR "$input1" "$input1" "$input1" ;;
        all) run DDD "$input0" "$input0" "$input0"; run DDR "$input0" "$input0" "$input1"
             run DRR "$input0" "$input1" "$input1"; run RRR "$input1" "$input1" "$input1" ;;
      esac
    done
  done
done
