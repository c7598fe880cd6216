#!/usr/bin/env python3
"""Benchmark of the brute-force triplet-counting path on one MI355X.

Headline (default): 50k clustered data + 50k uniform randoms, all four terms
(DDD, DDR, DRR, RRR) of sorted (r1,r2,r3) histograms, FAST path (FP32 pair
stage with FP64 recheck, integer triplet kernel), histograms exact.

  python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one tpcf_count() of the whole four-term job with the catalogs
already resident in device memory.  The JSON line reports triplets/s (unique
triplets classified per second of the timed steps), the roofline of the
dominant kernel (the triplet kernel, timed with HIP events recorded by the
library on its own stream), and a CPU baseline: the oracle on the host cores
over a bounded sample of the same catalogs.
"""
from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")

METRIC = "triplets/s at 50k+50k DDD..RRR; batched voxel jobs/s + p99 vs plain, exact"
# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz vector lane-ops, and the
# nominal work of one triplet in the pair-bin form (6 VALU lane-ops).
VALU_PEAK = 256 * 64 * 2.4e9
OPS_PER_TRIPLET = 6.0


def headline_catalogs(nd: int, nr: int):
    d = synth.neyman_scott(nd, nclust=200, sigma=2.0, size=200.0, seed=0)
    r = synth.uniform_cube(nr, size=200.0, seed=1)
    return d, r


def headline_spec(rel: float = 1e-5):
    # 20 log bins per side on [1, 350) Mpc/h: every pair of the 200 Mpc/h cube
    # (diagonal 346.4) is in range, so every triplet is counted.
    return tp.r123_spec(20, 1.0, 350.0, "log", rel)


def measured_traffic(args):
    """HBM bytes per launch of the triplet kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs of this workload (scaled by the copy-kernel
    calibration), or None for other workloads."""
    path = os.path.join(ROOT, "profiles", "r123_triplet_kernel_traffic.json")
    if not (args.nd == 50000 and args.nr == 50000 and args.mode == "fast" and os.path.exists(path)):
        return None
    t = json.load(open(path))
    return (t["fetch_size_kib"] * t["fetch_factor"] + t["write_size_kib"] * t["write_factor"]) * 1024.0


def cpu_baseline(d, r, seconds_hint: float = 20.0):
    """Oracle (clean64, all host threads) on the first m data + m randoms."""
    from oracle import oracle as orc

    threads = max(1, min(16, os.cpu_count() or 1))
    m = 1500
    segs = [d[:m], r[:m]]
    t0 = time.perf_counter()
    orc.clean_count(segs, tp.FOUR_TERMS, orc.R123, [orc.axis(20, 1.0, 350.0, "log")], nthreads=threads)
    dt = time.perf_counter() - t0
    trip = math.comb(2 * m, 3)
    return {
        "value": trip / dt,
        "unit": "triplets/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {m} data + {m} randoms of the same catalogs, four terms ({trip} triplets), "
                  f"clean64 oracle, {dt:.1f} s",
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nd", type=int, default=50000)
    ap.add_argument("--nr", type=int, default=50000)
    ap.add_argument("--mode", choices=sorted(tp.MODES), default="fast")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (calibrates FETCH_SIZE/WRITE_SIZE "
                         "counters against known bytes) and exit")
    args = ap.parse_args(argv)
    if args.copy_probe_mib:
        import torch

        src = torch.ones(args.copy_probe_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": src.numel() * 4, "copy_probe_bytes_written": dst.numel() * 4}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the path shards by splitting work items over devices, see DESIGN.md)", file=sys.stderr)
        return 2

    d, r = headline_catalogs(args.nd, args.nr)
    spec = headline_spec()
    mode = tp.MODES[args.mode]
    with tp.Context(0) as ctx:
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        job = tp.four_term_job(cd, cr)
        for _ in range(args.warmup):
            ctx.count(job, spec, mode)
        stats = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out, st = ctx.count(job, spec, mode)
            stats.append(st)
        dt = time.perf_counter() - t0
        cd.close()
        cr.close()

    triplets = stats[0].triplets
    counted = int(out.sum())
    value = triplets * args.steps / dt
    tri_s = sum(s.triplet_ms for s in stats) / len(stats) / 1e3
    achieved = OPS_PER_TRIPLET * triplets / tri_s if tri_s > 0 else 0.0
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 pair bins, u32/u64 counts (pairs fp32 + fp64 recheck)",
        "data": "synthetic",
        "config": {
            "workload": f"{args.nd // 1000}k clustered data + {args.nr // 1000}k uniform randoms, all four terms "
                        "DDD/DDR/DRR/RRR, sorted (r1,r2,r3), 20 log bins per side on [1,350) Mpc/h",
            "geometry": "Neyman-Scott data (200 clusters, sigma 2 Mpc/h) and uniform randoms in a 200 Mpc/h cube, "
                        "seeds 0/1",
            "n_data": args.nd,
            "n_randoms": args.nr,
            "mode": args.mode,
            "triplets_per_step": triplets,
            "triplets_in_range": counted,
            "pairs_rechecked_fraction": (stats[0].pairs_rechecked / stats[0].pairs) if stats[0].pairs else 0.0,
            "pair_stage_ms": stats[0].pair_ms,
            "triplet_kernel_ms": stats[0].triplet_ms,
            # nominal LDS traffic of the pair-bin form: 8 B per triplet against
            # 128 B/clk/CU x 256 CUs x 2.4 GHz
            "lds_roofline_frac": (8.0 * triplets / tri_s) / (128 * 256 * 2.4e9) if tri_s > 0 else 0.0,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": VALU_PEAK / 1e12,
            "unit": "TOP/s (VALU lane-ops, 6 per triplet)",
            "frac": achieved / VALU_PEAK,
            "traffic": measured_traffic(args),
        },
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(d, r)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
