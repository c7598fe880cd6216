#!/usr/bin/env python3
"""Sum per-voxel histogram files into one (drop-in for examples/sum_voxels.py).

    python3 sum_voxels.py FILE [FILE ...] > summed.dat

Same input parsing and output bytes as the reference reducer
(examples/sum_voxels.py:18-56): the body values of all files are summed
elementwise; the three catalog sizes, bin counts and (lo, hi, width) lines are
taken from the last file; the body is written as n0*n1 lines of n2 "%d "
values; one extra newline ends the output (the reference's print).

The reference sums in float64, exact below 2**53; here the sum is exact for
any count (Python integers), which gives the same bytes wherever float64 was
exact.  With --gpu-format the body is the single row of the reference's GPU
program output (examples/sum_GPU_voxels.py:42-54).
"""
from __future__ import annotations

import sys


def read_hist(path: str):
    with open(path) as f:
        vals = f.read().split()
    if len(vals) < 15:
        raise ValueError(f"{path}: not a histogram file (header has 15 tokens)")
    ngals = [int(v) for v in vals[0:3]]
    nbins = [int(v) for v in vals[3:6]]
    ranges = [[float(v) for v in vals[6 + 3 * i:9 + 3 * i]] for i in range(3)]
    body = [int(float(v)) if not v.lstrip("-").isdigit() else int(v) for v in vals[15:]]
    return ngals, nbins, ranges, body


def sum_files(paths: list[str], gpu_format: bool = False) -> str:
    total = None
    ngals = nbins = ranges = None
    for p in paths:
        ngals, nbins, ranges, body = read_hist(p)
        if total is None:
            total = [0] * len(body)
        if len(body) != len(total):
            raise ValueError(f"{p}: {len(body)} values, expected {len(total)}")
        total = [a + b for a, b in zip(total, body)]
    if total is None:
        raise ValueError("no input files")
    out = []
    for n in ngals:
        out.append("%d\n" % n)
    out.append("%d %d %d\n" % (nbins[0], nbins[1], nbins[2]))
    for h in ranges:
        out.append("%-6.3f %-6.3f %-6.3f\n" % (h[0], h[1], h[2]))
    if gpu_format:
        if len(total) != nbins[2]:
            raise ValueError("GPU-format files hold one row of n2 values")
        out.append("".join("%d " % v for v in total))
        out.append("\n")
    else:
        n0, n1, n2 = nbins
        if len(total) != n0 * n1 * n2:
            raise ValueError(f"body has {len(total)} values, header says {n0}x{n1}x{n2}")
        for r in range(n0 * n1):
            out.append("".join("%d " % v for v in total[r * n2:(r + 1) * n2]))
            out.append("\n")
    return "".join(out) + "\n"


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    gpu = False
    if argv and argv[0] == "--gpu-format":
        gpu = True
        argv = argv[1:]
    sys.stdout.write(sum_files(argv, gpu))
    return 0


if __name__ == "__main__":
    sys.exit(main())
