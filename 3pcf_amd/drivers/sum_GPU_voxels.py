#!/usr/bin/env python3
"""Drop-in for examples/sum_GPU_voxels.py: sums single-row (theta only)
histogram files; same bytes as the reference (examples/sum_GPU_voxels.py:
42-54).  See sum_voxels.py for the shared parsing."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from sum_voxels import sum_files  # noqa: E402

if __name__ == "__main__":
    sys.stdout.write(sum_files(sys.argv[1:], gpu_format=True))
