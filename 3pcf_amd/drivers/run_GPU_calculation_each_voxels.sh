#!/usr/bin/env bash
# Drop-in for examples/run_GPU_calculation_each_voxels.csh.
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
ngals=${1:-1}
INPUT0=${INPUT0:-../sample_data/MICE_5degx5deg_${ngals}k_randomized.dat} \
INPUT1=${INPUT1:-../sample_data/flat_MICE_5degx5deg_${ngals}k.dat} \
TAG=${TAG:-evenbinning_GPU_5degx5deg_LRG_binning_TACC} exec "$HERE/run_CPU_calculation_each_voxels.sh" "$@"
