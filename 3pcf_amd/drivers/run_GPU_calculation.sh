#!/usr/bin/env bash
# Drop-in for examples/run_GPU_calculation.csh: the same term loop as
# run_CPU_calculation.sh with the reference's GPU tag.
HERE=$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)
TAG=${TAG:-evenbinning_GPU_naive_LRG_binning} exec "$HERE/run_CPU_calculation.sh" "$@"
