#!/usr/bin/env python3
"""Benchmark of the brute-force triplet-counting path on one MI355X.

Headline (default): 50k clustered data + 50k uniform randoms, all four terms
(DDD, DDR, DRR, RRR) of sorted (r1,r2,r3) histograms, FAST path (FP32 pair
stage with FP64 recheck, integer triplet kernel), histograms exact.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--budget-s S]

One step = one tpcf_count() of the whole four-term job with the catalogs
already resident in device memory (tens of seconds).  At most one warmup
step runs (nothing is compiled at run time), and timed steps stop early when
one more would end past the --budget-s wall budget: "steps" is the number
timed, "steps_requested" the number asked for.  Each step writes a progress
line to stderr.  The JSON line reports triplets/s (unique triplets
classified per second of the timed steps), the roofline of the dominant
kernel (the triplet kernel, timed with HIP events recorded by the library on
its own stream), the batched half of the metric (64 voxels of 150k + 150k in
one prepared launch against the FP64 plain kernel one voxel at a time, every
histogram compared), and CPU baselines timed on bounded samples of the same
catalogs: the reference C path (one thread, -O1) and the clean64 oracle on
the host's threads.
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
# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz vector lane-ops.  The
# (r1,r2,r3) triplet kernel counts the VALU lane-ops of its own cost model
# (stats.mask_ops: per visited row the y ballots, the x weights, and per
# (y, z) bin pair of a y bin present in the row two ANDs, two popcounts and
# the multiply-adds; per accumulator field its wave sum), so `achieved` is
# the work the kernel visits, not a per-triplet constant.
VALU_PEAK = 256 * 64 * 2.4e9


# stats.triplet_form: which (r1,r2,r3) LDS triplet kernel counted
FORMS = {0: "run-merging", 1: "bit-mask", 2: "split"}


def valu_roofline(ops: float, seconds: float) -> dict:
    """Roofline entry of the (r1,r2,r3) bit-mask triplet kernel: its counted
    VALU lane-ops over its HIP-event time, against the VALU peak."""
    achieved = ops / seconds if seconds > 0 else 0.0
    return {
        "bound": "valu",
        "achieved": achieved / 1e12,
        "peak": VALU_PEAK / 1e12,
        "unit": "TOP/s (VALU lane-ops of the bit-mask triplet kernel's cost model, counted by the kernel; "
                "matrix-core passes: 4 per operand, per 16x16 block and bin pair 1 + 3 + NX, 3.5 per accumulator)",
        "frac": achieved / VALU_PEAK,
        "traffic": None,
    }


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


def cpu_baseline(d, r, threads_max: int = 16):
    """CPU rates on bounded samples of the headline catalogs, timed on this
    host in this run.

    value: the reference C path -- the oracle's refcompat32 restatement of
    the reference's triangle loop (src/3pcf_C_version_triangle_parametrization.cc:477-533,
    FP32 distances, acosf, the reference's bins S 50 [2,12), Q 16 [0.9,4.1),
    theta/pi 32 [0,1)) built at the reference's -O1 and run on one thread
    over DDD/DDR/DRR/RRR of the first m data and m randoms.  Also reported:
    the clean64 oracle (FP64, the headline's (r1,r2,r3) spec) on all host
    threads up to 16."""
    from oracle import oracle as orc

    # reference C path, one thread
    m = 420
    cats = {"D": np.asarray(d[:m]), "R": np.asarray(r[:m])}
    ref_axes = [(50, 2.0, 12.0), (16, 0.9, 4.1), (32, 0.0, 1.0)]
    iters = 0
    t0 = time.perf_counter()
    for names, which in (("DDD", 0), ("DDR", 1), ("DRR", 1), ("RRR", 0)):
        _, nc = orc.ref_tri([cats[c] for c in names], which, 0, ref_axes, nthreads=1, o1=True)
        iters += nc
    ref_dt = time.perf_counter() - t0
    ref_trip = math.comb(2 * m, 3)
    sys.stderr.write(f"cpu baseline: reference C path {ref_trip / ref_dt:.3e} triplets/s ({ref_dt:.1f} s)\n")
    sys.stderr.flush()

    # clean64 oracle, all host threads
    threads = max(1, min(threads_max, os.cpu_count() or 1))
    mc = 1500
    t0 = time.perf_counter()
    orc.clean_count([d[:mc], r[:mc]], tp.FOUR_TERMS, orc.R123, [orc.axis(20, 1.0, 350.0, "log")], nthreads=threads)
    dt = time.perf_counter() - t0
    trip = math.comb(2 * mc, 3)
    sys.stderr.write(f"cpu baseline: clean64 port {trip / dt:.3e} triplets/s on {threads} threads ({dt:.1f} s)\n")
    sys.stderr.flush()
    return {
        "value": ref_trip / ref_dt,
        "unit": "triplets/s",
        "cores": 1,
        "kind": "port",
        "sample": f"reference C path: refcompat32 restatement of the reference's triangle loop (FP32, acosf, its "
                  f"S/Q/theta bins), gcc -O1, one thread, DDD/DDR/DRR/RRR of the first {m} data + {m} randoms "
                  f"of the headline catalogs ({ref_trip} unique triplets, {iters} reference loop iterations, "
                  f"{ref_dt:.1f} s)",
        "reference_loop_iterations_per_s": iters / ref_dt,
        "clean64_port": {
            "value": trip / dt,
            "unit": "triplets/s",
            "cores": threads,
            "kind": "port",
            "sample": f"clean64 oracle (FP64, the headline's 20 log bins on [1,350)), first {mc} data + {mc} "
                      f"randoms, four terms ({trip} triplets), {dt:.1f} s",
        },
    }


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return 0.0
    k = min(len(xs) - 1, max(0, int(math.ceil(q / 100.0 * len(xs))) - 1))
    return xs[k]


def bench_batch(args) -> int:
    """64 angular voxels (8x8 cells of a 20x20 deg wedge, r in [500,800]
    Mpc/h) of 150k clustered data + 150k randoms with lognormal occupancy
    500..8000; all four terms of every voxel, 20 log bins on [1,50) Mpc/h;
    FAST path as one prepared batch (autotuned shape, HIP graph) against the
    FP64 plain kernel running the voxel jobs one at a time.  Also: the culled
    form (CELLS) prepared and by count_batch, and the same voxels in the
    (s,q,theta) parametrization as a prepared batch against plain FP64."""
    vox = synth.angular_voxels(8, 150000, seed=0)
    sys.stderr.write("batch: 64 voxels of 150k + 150k\n")
    sys.stderr.flush()
    spec = tp.r123_spec(20, 1.0, 50.0, "log", 1e-5)
    reps = max(1, args.steps)
    with tp.Context(0) as ctx:
        cats = [(ctx.catalog(d), ctx.catalog(r)) for d, r in vox]
        jobs = [tp.four_term_job(cd, cr) for cd, cr in cats]
        t0 = time.perf_counter()
        b = ctx.batch(jobs, [spec] * len(jobs), tp.MODE_FAST, autotune=True, graph=True)
        prep_s = time.perf_counter() - t0
        for _ in range(args.warmup):
            b.run()
        lat = []
        tri_ms = []
        for _ in range(reps):
            t1 = time.perf_counter()
            outs, st = b.run()
            lat.append(time.perf_counter() - t1)
            tri_ms.append(st.device_ms)
        fast = [o.copy() for o in outs]
        info = b.info
        b.close()
        # The culled form (Morton-ordered voxels, block triples out of reach
        # skipped) as a prepared batch (culled once, graph-replayed) and as
        # one count_batch call, against the same call unculled.
        specs = [spec] * len(jobs)
        bc = ctx.batch(jobs, specs, tp.MODE_CELLS, autotune=True, graph=True)
        bc.run()
        clat = []
        for _ in range(reps):
            t1 = time.perf_counter()
            couts_p, _ = bc.run()
            clat.append(time.perf_counter() - t1)
        cells_prepared_exact = all(np.array_equal(a, b_) for a, b_ in zip(couts_p, fast))
        bc.close()
        ctx.count_batch(jobs, specs, tp.MODE_CELLS)
        t1 = time.perf_counter()
        couts, cst = ctx.count_batch(jobs, specs, tp.MODE_CELLS)
        cells_s = time.perf_counter() - t1
        t1 = time.perf_counter()
        uouts, ust = ctx.count_batch(jobs, specs, tp.MODE_FAST)
        uncull_s = time.perf_counter() - t1
        cells_exact = all(np.array_equal(a, b_) for a, b_ in zip(couts, fast)) and \
            all(np.array_equal(a, b_) for a, b_ in zip(uouts, fast))
        # plain FP64, one voxel job per launch
        plain_lat = []
        exact = True
        for j, job in enumerate(jobs):
            t1 = time.perf_counter()
            out, _ = ctx.count(job, spec, tp.MODE_PLAIN)
            plain_lat.append(time.perf_counter() - t1)
            exact &= bool(np.array_equal(out, fast[j]))
        sys.stderr.write("batch: r123 prepared, culled and plain done\n")
        sys.stderr.flush()
        # The same voxels in the triangle parametrization (what the
        # reference's voxel drivers run): s [2,20) 10 log, q [1,3) 8,
        # theta/pi 16, as one prepared graph-replayed batch against the FP64
        # plain kernel one voxel job at a time.
        sspec = tp.binspec(tp.SQT, [tp.axis(10, 2.0, 20.0, "log"), tp.axis(8, 1.0, 3.0, "lin"),
                                    tp.axis(16, 0.0, 1.0, "lin")])
        bs = ctx.batch(jobs, [sspec] * len(jobs), tp.MODE_FAST, autotune=False, graph=True)
        bs.run()
        slat = []
        for _ in range(reps):
            t1 = time.perf_counter()
            souts, sst = bs.run()
            slat.append(time.perf_counter() - t1)
        sfast = [o.copy() for o in souts]
        bs.close()
        splain_s = 0.0
        sexact = True
        for j, job in enumerate(jobs):
            t1 = time.perf_counter()
            out, _ = ctx.count(job, sspec, tp.MODE_PLAIN)
            splain_s += time.perf_counter() - t1
            sexact &= bool(np.array_equal(out, sfast[j]))
            if j % 16 == 15:
                sys.stderr.write(f"batch: sqt plain {j + 1}/{len(jobs)} voxels, {splain_s:.1f} s\n")
                sys.stderr.flush()
    nv = len(jobs)
    mean = sum(lat) / len(lat)
    fast_rate = nv / mean
    plain_rate = nv / sum(plain_lat)
    triplets = st.triplets
    dev_s = sum(tri_ms) / len(tri_ms) / 1e3
    summed = np.sum(np.stack(fast), axis=0)
    line = {
        "metric": "batched voxel jobs/s (64 voxels x 4 terms) vs plain FP64 one job at a time, exact",
        "value": fast_rate,
        "unit": "voxel-jobs/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": mean * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 pair bins, u32/u64 counts (pairs fp32 + fp64 recheck)",
        "data": "synthetic",
        "config": {
            "workload": "64 angular voxels (8x8 of a 20x20 deg wedge, r 500-800 Mpc/h), 150k clustered data + "
                        "150k randoms, lognormal occupancy 501-7998 per voxel, four terms per voxel, 20 log bins "
                        "on [1,50) Mpc/h",
            "triplets_per_step": triplets,
            "triplets_in_range": int(summed.sum()),
            "batch_latency_p50_ms": pct(lat, 50) * 1e3,
            "batch_latency_p99_ms": pct(lat, 99) * 1e3,
            "prepare_s": prep_s,
            "launch": info.splitlines()[0] if info else "",
            "plain_fp64_voxel_jobs_per_s": plain_rate,
            "plain_job_latency_p50_ms": pct(plain_lat, 50) * 1e3,
            "plain_job_latency_p99_ms": pct(plain_lat, 99) * 1e3,
            "speedup_vs_plain": fast_rate / plain_rate,
            "histograms_identical_to_plain": exact,
            "count_batch_fast_ms": uncull_s * 1e3,
            "count_batch_fast_triplet_form": FORMS.get(ust.triplet_form, "none"),
            "count_batch_fast_tile_bins_mean": ust.tile_bins_mean,
            "count_batch_fast_tile_fill": ust.tile_fill,
            "count_batch_cells_ms": cells_s * 1e3,
            "count_batch_cells_triplet_kernel_ms": cst.triplet_ms,
            "cells_voxel_jobs_per_s": nv / cells_s,
            "cells_histograms_identical": cells_exact,
            "cells_prepared_voxel_jobs_per_s": nv / (sum(clat) / len(clat)),
            "cells_prepared_latency_p50_ms": pct(clat, 50) * 1e3,
            "cells_prepared_latency_p99_ms": pct(clat, 99) * 1e3,
            "cells_prepared_identical": cells_prepared_exact,
            "sqt_spec": "s [2,20) 10 log, q [1,3) 8 lin, theta/pi [0,1) 16 lin",
            "sqt_prepared_voxel_jobs_per_s": nv / (sum(slat) / len(slat)),
            "sqt_prepared_latency_p50_ms": pct(slat, 50) * 1e3,
            "sqt_prepared_latency_p99_ms": pct(slat, 99) * 1e3,
            "sqt_triplets_rechecked": sst.pairs_rechecked,
            "sqt_plain_fp64_voxel_jobs_per_s": nv / splain_s,
            "sqt_speedup_vs_plain": (nv / (sum(slat) / len(slat))) / (nv / splain_s),
            "sqt_histograms_identical_to_plain": sexact,
        },
        # No roofline: the kernel skips the tiles, rows and C blocks with no
        # pair in range, so a per-triplet figure over all the job's triplets
        # would not describe the work done.
    }
    print(json.dumps(line))
    sys.stderr.write(info)
    return 0


def request_spec(g):
    """A bin spec uniform over {r123-lin, r123-log, sqt-lin, sqt-log} with
    8..32 bins per axis."""
    kind = int(g.integers(0, 4))
    nb = [int(g.integers(8, 33)) for _ in range(3)]
    if kind == 0:
        return tp.r123_spec(nb[0], 0.0, 150.0, "lin"), "r123-lin"
    if kind == 1:
        return tp.r123_spec(nb[0], 1.0, 150.0, "log"), "r123-log"
    scale = "lin" if kind == 2 else "log"
    s_axis = tp.axis(nb[0], 2.0, 20.0, scale)
    return tp.binspec(tp.SQT, [s_axis, tp.axis(nb[1], 1.0, 3.0, "lin"), tp.axis(nb[2], 0.0, 1.0, "lin")]), \
        "sqt-" + scale


def request_stream(n: int, n_max: int, seed: int = 0):
    """Poisson arrivals (mean 20 ms apart); catalog size log-uniform in
    [1000, n_max]; a uniform random non-empty subset of the four terms; a bin
    spec from request_spec()."""
    g = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    reqs = []
    for k in range(n):
        t += g.exponential(0.020)
        size = int(round(math.exp(g.uniform(math.log(1000), math.log(n_max)))))
        mask = int(g.integers(1, 16))
        terms = [tp.FOUR_TERMS[b] for b in range(4) if mask >> b & 1]
        spec, kind = request_spec(g)
        reqs.append((t, size, terms, spec, 1000 + k, kind))
    return reqs


def bench_serve(args) -> int:
    reqs = request_stream(args.requests, args.n_max)
    with tp.Context(0) as ctx:
        jobs = []
        keep = []
        for (_, size, terms, spec, seed, _) in reqs:
            d = synth.neyman_scott(size, nclust=40, sigma=2.0, size=100.0, seed=seed)
            r = synth.uniform_cube(size, 100.0, seed=seed + 7)
            cd, cr = ctx.catalog(d), ctx.catalog(r)
            keep += [cd, cr]
            jobs.append(tp.make_job([(cd, 0, size), (cr, 0, size)], terms))
        # warm the workers and kernels
        t, _ = ctx.submit(jobs[0], reqs[0][3])
        ctx.wait(t)
        sys.stderr.write(f"serve: {len(jobs)} catalogs resident, submitting\n")
        sys.stderr.flush()
        start = time.perf_counter()
        tickets = []
        for (arr, _, _, spec, _, _), job in zip(reqs, jobs):
            now = time.perf_counter() - start
            if arr > now:
                time.sleep(arr - now)
            tickets.append(ctx.submit(job, spec))
        lat = [0.0] * len(tickets)
        outs = [out for _, out in tickets]
        pending = set(range(len(tickets)))
        while pending:
            for k in sorted(pending):
                done, st = ctx.poll(tickets[k][0])
                if done:
                    pending.discard(k)
                    lat[k] = st.wall_ms / 1e3
                    sys.stderr.write(f"serve: request {k} done, latency {st.wall_ms:.0f} ms, "
                                     f"{len(pending)} pending\n")
                    sys.stderr.flush()
            if pending:
                time.sleep(0.01)
        span = time.perf_counter() - start
        tot_trip = sum(ctx.footprint_triplets(job, spec) for (_, _, _, spec, _, _), job in zip(reqs, jobs))
        # Plain FP64 one job at a time, on a bounded sample: jobs in arrival
        # order whose plain time (estimated from the rate measured so far)
        # fits the remaining budget; each also runs alone on the FAST path,
        # and both results are compared with the job's result in the mix.
        budget = args.plain_budget
        rate = 1.0e11  # plain triplets/s before the first measurement
        used = 0.0
        clock = 0.0
        plain_lat, plain_t, solo_t = [], [], []
        sample_trip = 0
        exact = True
        for (arr, _, _, spec, _, _), job, o in zip(reqs, jobs, outs):
            trip = ctx.footprint_triplets(job, spec)
            if used + trip / rate > budget:
                continue
            t1 = time.perf_counter()
            out, st = ctx.count(job, spec, tp.MODE_PLAIN)
            dt = time.perf_counter() - t1
            t1 = time.perf_counter()
            solo, _ = ctx.count(job, spec, tp.MODE_FAST)
            solo_t.append(time.perf_counter() - t1)
            used += dt
            rate = max(rate * 0.5, st.triplets / dt) if st.triplets else rate
            sys.stderr.write(f"serve: plain sample job {len(plain_t)} ({st.triplets} triplets) {dt:.1f} s\n")
            sys.stderr.flush()
            sample_trip += st.triplets
            clock = max(clock, arr) + dt
            plain_lat.append(clock - arr)
            plain_t.append(dt)
            exact &= bool(np.array_equal(out, o)) and bool(np.array_equal(solo, o))
    n = len(reqs)
    line = {
        "metric": "serving throughput (jobs/s) with p50/p99 latency vs plain FP64 one job at a time, exact",
        "value": n / span,
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": span * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 pair bins, u32/u64 counts (pairs fp32 + fp64 recheck)",
        "data": "synthetic",
        "config": {
            "workload": f"{n} requests, Poisson arrivals (mean 50/s), N log-uniform in [1000,{args.n_max}] data + "
                        "as many randoms, random non-empty term subset, spec uniform over r123-lin/r123-log/"
                        "sqt-lin/sqt-log with 8-32 bins per axis",
            "spec_mix": {k: sum(1 for q in reqs if q[5] == k) for k in ("r123-lin", "r123-log", "sqt-lin", "sqt-log")},
            "triplets": tot_trip,
            "latency_p50_ms": pct(lat, 50) * 1e3,
            "latency_p99_ms": pct(lat, 99) * 1e3,
            "plain_sample": f"{len(plain_t)} of {n} jobs (arrival order, those whose FP64 plain time fits a "
                            f"{budget:.0f} s budget), {sample_trip} triplets",
            "plain_fp64_jobs_per_s_on_sample": len(plain_t) / max(sum(plain_t), 1e-9),
            "fast_alone_jobs_per_s_on_sample": len(solo_t) / max(sum(solo_t), 1e-9),
            "speedup_alone_on_sample": sum(plain_t) / max(sum(solo_t), 1e-9),
            "plain_latency_p50_ms_on_sample": pct(plain_lat, 50) * 1e3,
            "plain_latency_p99_ms_on_sample": pct(plain_lat, 99) * 1e3,
            "sample_identical_plain_alone_and_in_mix": exact,
        },
    }
    print(json.dumps(line))
    return 0


def bench_sqt(args) -> int:
    """Four-term (s,q,theta) histograms: 10k clustered data + 10k uniform
    randoms in a 200 Mpc/h cube, s in [2,20) 10 log bins, q in [1,3) 8 bins,
    theta/pi in [0,1) 16 bins; FAST (FP32 candidate arithmetic with FP64
    recheck) against FAST64 and the FP64 plain kernel, all exact."""
    n = args.sqt_n
    d = synth.neyman_scott(n, nclust=200, sigma=2.0, size=200.0, seed=0)
    r = synth.uniform_cube(n, 200.0, seed=1)
    spec = tp.binspec(tp.SQT, [tp.axis(10, 2.0, 20.0, "log"), tp.axis(8, 1.0, 3.0, "lin"),
                               tp.axis(16, 0.0, 1.0, "lin")])
    reps = max(1, args.steps)
    with tp.Context(0) as ctx:
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        job = tp.four_term_job(cd, cr)
        for _ in range(args.warmup):
            ctx.count(job, spec, tp.MODE_FAST)
        stats = []
        t0 = time.perf_counter()
        for _ in range(reps):
            out, st = ctx.count(job, spec, tp.MODE_FAST)
            stats.append(st)
        dt = time.perf_counter() - t0
        ctx.count(job, spec, tp.MODE_CELLS)
        t1 = time.perf_counter()
        outc, stc = ctx.count(job, spec, tp.MODE_CELLS)
        dtc = time.perf_counter() - t1
        out64 = plain = None
        dt64 = dtp = None
        if not args.no_plain:
            t1 = time.perf_counter()
            out64, _ = ctx.count(job, spec, tp.MODE_FAST64)
            dt64 = time.perf_counter() - t1
            t1 = time.perf_counter()
            plain, _ = ctx.count(job, spec, tp.MODE_PLAIN)
            dtp = time.perf_counter() - t1
    triplets = stats[0].triplets
    tri_s = sum(x.triplet_ms for x in stats) / len(stats) / 1e3
    value = triplets * reps / dt
    entries = int(out.sum())
    line = {
        "metric": "(s,q,theta) four-term triplets/s, FAST vs plain FP64, exact",
        "value": value,
        "unit": "triplets/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / reps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 pair codes, f32 entries with f64 recheck, u32/u64 counts",
        "data": "synthetic",
        "config": {
            "workload": f"{n // 1000}k Neyman-Scott data (200 clusters, sigma 2 Mpc/h) + {n // 1000}k uniform "
                        "randoms in a 200 Mpc/h cube, DDD/DDR/DRR/RRR, s [2,20) 10 log, q [1,3) 8 lin, "
                        "theta/pi [0,1) 16 lin",
            "triplets_per_step": triplets,
            "histogram_entries": entries,
            "triangles_rechecked_fp64": stats[0].pairs_rechecked,
            "triplet_kernel_ms": stats[0].triplet_ms,
            "pair_stage_ms": stats[0].pair_ms,
            "fast64_s": dt64,
            "plain_fp64_s": dtp,
            "plain_fp64_triplets_per_s": triplets / dtp if dtp else None,
            "speedup_vs_plain": dtp / (dt / reps) if dtp else None,
            "fast_equals_fast64": bool(np.array_equal(out, out64)) if out64 is not None else None,
            "fast_equals_plain": bool(np.array_equal(out, plain)) if plain is not None else None,
            "cells_ms_per_step": dtc * 1e3,
            "cells_triplet_kernel_ms": stc.triplet_ms,
            "cells_equals_fast": bool(np.array_equal(out, outc)),
        },
        # No roofline: the bit-mask filter settles most triplets 64 at a time
        # without entry arithmetic, so the nominal per-triplet figure does not
        # describe the work done (the profiles give the measured counters).
    }
    print(json.dumps(line))
    return 0


PAIR_OPS = 8.0  # FP64 lane-ops of one exact squared distance: 3 sub, 3 mul, 2 add


def bench_pairs(args) -> int:
    """Pair counts (DD, DR, RR) on the reference's binning (700 linear bins
    on [0, 70) Mpc/h): clustered data + uniform randoms in a 400 Mpc/h cube,
    FAST (FP32 bin guess, FP64 decision) against the plain FP64 kernel, and
    the oracle on the host cores over a sample, all exact."""
    n = args.pairs_n
    d = synth.neyman_scott(n, nclust=2000, sigma=3.0, size=400.0, seed=0)
    r = synth.uniform_cube(n, 400.0, seed=1)
    ax = tp.axis(*tp.TWO_PCF_AXIS)
    reps = max(1, args.steps)
    with tp.Context(0) as ctx:
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        job = tp.pair_job(cd, cr)
        for _ in range(args.warmup):
            ctx.count_pairs(job, ax)
        stats = []
        t0 = time.perf_counter()
        for _ in range(reps):
            out, st = ctx.count_pairs(job, ax)
            stats.append(st)
        dt = time.perf_counter() - t0
        # The culled form (Morton order + box skipping), same histograms.
        ctx.count_pairs(job, ax, tp.MODE_CELLS)
        cstats = []
        t1 = time.perf_counter()
        for _ in range(reps):
            cells, stc = ctx.count_pairs(job, ax, tp.MODE_CELLS)
            cstats.append(stc)
        dtc = time.perf_counter() - t1
        plain = dtp = None
        if not args.no_plain:
            t1 = time.perf_counter()
            plain, stp = ctx.count_pairs(job, ax, tp.MODE_PLAIN)
            dtp = time.perf_counter() - t1
    pairs = stats[0].pairs
    kern_s = sum(x.pair_ms for x in stats) / len(stats) / 1e3
    achieved = PAIR_OPS * pairs / kern_s
    line = {
        "metric": "2PCF DD/DR/RR pairs/s, FAST vs plain FP64, exact",
        "value": pairs * reps / dt,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / reps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 distances, f32 bin guess, u32/u64 counts",
        "data": "synthetic",
        "config": {
            "workload": f"{n // 1000}k Neyman-Scott data (2000 clusters, sigma 3 Mpc/h) + {n // 1000}k uniform "
                        "randoms in a 400 Mpc/h cube, DD/DR/RR, 700 linear bins on [0,70) Mpc/h",
            "pairs_per_step": pairs,
            "pairs_in_range": int(out.sum()),
            "pair_kernel_ms": stats[0].pair_ms,
            "plain_fp64_s": dtp,
            "plain_kernel_ms": stp.pair_ms if dtp else None,
            "speedup_vs_plain_kernel": stp.pair_ms / stats[0].pair_ms if dtp else None,
            "fast_equals_plain": bool(np.array_equal(out, plain)) if plain is not None else None,
            "cells_ms_per_step": dtc * 1e3 / reps,
            "cells_kernel_ms": sum(x.pair_ms for x in cstats) / len(cstats),
            "cells_pairs_per_s": pairs * reps / dtc,
            "cells_equals_fast": bool(np.array_equal(out, cells)),
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": VALU_PEAK / 1e12,
            "unit": "TOP/s (FP64 VALU lane-ops, 8 per pair for the squared distance; binning and the LDS "
                    "atomic are on top)",
            "frac": achieved / VALU_PEAK,
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as orc

        threads = max(1, min(16, os.cpu_count() or 1))
        m = min(n, 40000)
        t1 = time.perf_counter()
        orc.clean_pairs([d[:m], r[:m]], tp.PAIR_TERMS, orc.axis(*tp.TWO_PCF_AXIS), nthreads=threads)
        dc = time.perf_counter() - t1
        cpairs = m * (m - 1) + m * m
        line["cpu_baseline"] = {
            "value": cpairs / dc,
            "unit": "pairs/s",
            "cores": threads,
            "kind": "port",
            "sample": f"first {m} data + {m} randoms, DD/DR/RR ({cpairs} pairs), clean64 oracle, {dc:.1f} s",
        }
    print(json.dumps(line))
    return 0


def bench_survey(args) -> int:
    """Culled counting at survey scale (TPCF_MODE_CELLS): 100k clustered data +
    100k uniform randoms in a 1000 Mpc/h cube, all four terms, (r1,r2,r3) with
    20 log bins on [1,50) Mpc/h and (s,q,theta) with s [2,20) 10 log, q [1,3)
    8, theta/pi 16.  The rate is unique triplets of the job per second; the
    triplets out of reach are settled by box tests, so the entries counted
    per second are reported beside it."""
    n = args.survey_n
    d = synth.neyman_scott(n, nclust=max(1, n // 100), sigma=3.0, size=1000.0, seed=0)
    r = synth.uniform_cube(n, 1000.0, seed=1)
    specs = [("r123", tp.r123_spec(20, 1.0, 50.0, "log")),
             ("sqt", tp.binspec(tp.SQT, [tp.axis(10, 2.0, 20.0, "log"), tp.axis(8, 1.0, 3.0, "lin"),
                                         tp.axis(16, 0.0, 1.0, "lin")]))]
    reps = max(1, args.steps)
    res = {}
    with tp.Context(0) as ctx:
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        job = tp.four_term_job(cd, cr)
        for name, spec in specs:
            for _ in range(args.warmup):
                ctx.count(job, spec, tp.MODE_CELLS)
            sts = []
            t0 = time.perf_counter()
            for _ in range(reps):
                out, st = ctx.count(job, spec, tp.MODE_CELLS)
                sts.append(st)
            dt = (time.perf_counter() - t0) / reps
            res[name] = (dt, sts[0], int(out.sum()))
    trip = res["r123"][1].triplets
    total_s = sum(v[0] for v in res.values())
    line = {
        "metric": "culled survey counting: unique triplets of the job per second, exact",
        "value": 2 * trip / total_s,
        "unit": "triplets/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": total_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 pair bins / f64 pair codes, u32/u64 counts",
        "data": "synthetic",
        "config": {
            "workload": f"{n // 1000}k Neyman-Scott data ({max(1, n // 100)} clusters, sigma 3 Mpc/h) + {n // 1000}k "
                        "uniform randoms in a 1000 Mpc/h cube, DDD/DDR/DRR/RRR, (r1,r2,r3) 20 log [1,50) and "
                        "(s,q,theta) s [2,20) 10 log, q [1,3) 8, theta/pi 16; TPCF_MODE_CELLS",
            "triplets_per_job": trip,
        },
    }
    for name, (dt, st, entries) in res.items():
        line["config"][f"{name}_ms"] = dt * 1e3
        line["config"][f"{name}_triplet_kernel_ms"] = st.triplet_ms
        line["config"][f"{name}_pair_stage_ms"] = st.pair_ms
        line["config"][f"{name}_entries"] = entries
        line["config"][f"{name}_entries_per_s"] = entries / dt
    st = res["r123"][1]
    line["config"]["r123_kernel_valu_lane_ops"] = st.mask_ops
    line["config"]["r123_triplet_form"] = FORMS.get(st.triplet_form, "none")
    line["config"]["r123_tile_bins_mean"] = st.tile_bins_mean
    line["config"]["r123_tile_fill"] = st.tile_fill
    # the (r1,r2,r3) job's triplet kernel: the work it visits, counted by it
    # (the bit-mask kernel only)
    if st.triplet_form == 1:
        line["roofline"] = valu_roofline(st.mask_ops, st.triplet_ms / 1e3)
    print(json.dumps(line))
    return 0


def bench_precision(args) -> int:
    """Precision form: 20k clustered data + 20k uniform randoms in a
    10x10 deg wedge (r 500-800 Mpc/h), 20 log bins on [1,50) Mpc/h, recheck
    window 1e-5; FAST (FP32 pair stage + FP64 recheck) against FAST64 (FP64
    pair stage), histograms identical, recheck fraction reported."""
    n = args.precision_n
    d = synth.wedge(n, 10.0, 10.0, 500.0, 800.0, seed=0, clustered=True, nclust=200)
    r = synth.wedge(n, 10.0, 10.0, 500.0, 800.0, seed=1)
    spec = tp.r123_spec(20, 1.0, 50.0, "log", 1e-5)
    reps = max(1, args.steps)
    with tp.Context(0) as ctx:
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        job = tp.four_term_job(cd, cr)
        for _ in range(args.warmup):
            ctx.count(job, spec, tp.MODE_FAST)
        stats = []
        t0 = time.perf_counter()
        for _ in range(reps):
            out, st = ctx.count(job, spec, tp.MODE_FAST)
            stats.append(st)
        dt = time.perf_counter() - t0
        out64, st64 = ctx.count(job, spec, tp.MODE_FAST64)
        ctx.count(job, spec, tp.MODE_CELLS)
        t1 = time.perf_counter()
        outc, stc = ctx.count(job, spec, tp.MODE_CELLS)
        dtc = time.perf_counter() - t1
    triplets = stats[0].triplets
    tri_s = sum(x.triplet_ms for x in stats) / len(stats) / 1e3
    line = {
        "metric": "precision form triplets/s (FP32 pair stage + FP64 recheck), exact vs FP64",
        "value": triplets * reps / dt,
        "unit": "triplets/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / reps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 pair bins, u32/u64 counts (pairs fp32 + fp64 recheck)",
        "data": "synthetic",
        "config": {
            "workload": f"{n // 1000}k clustered data + {n // 1000}k uniform randoms in a 10x10 deg wedge, "
                        "r 500-800 Mpc/h, four terms, 20 log bins on [1,50) Mpc/h, recheck window 1e-5",
            "triplets_per_step": triplets,
            "triplets_in_range": int(out.sum()),
            "pairs": stats[0].pairs,
            "pairs_rechecked": stats[0].pairs_rechecked,
            "pairs_rechecked_fraction": stats[0].pairs_rechecked / stats[0].pairs if stats[0].pairs else 0.0,
            "pair_stage_ms": stats[0].pair_ms,
            "pair_stage_fp64_ms": st64.pair_ms,
            "triplet_kernel_ms": stats[0].triplet_ms,
            "fast_equals_fast64": bool(np.array_equal(out, out64)),
            "cells_ms_per_step": dtc * 1e3,
            "cells_triplet_kernel_ms": stc.triplet_ms,
            "cells_equals_fast": bool(np.array_equal(out, outc)),
            "kernel_valu_lane_ops": stats[0].mask_ops,
            "triplet_form": FORMS.get(stats[0].triplet_form, "none"),
            "tile_bins_mean": stats[0].tile_bins_mean,
            "tile_fill": stats[0].tile_fill,
        },
    }
    # The bit-mask kernel counts the work it visits (rows and bin pairs of
    # the block triples it does not skip), so its roofline holds here too;
    # the run-merging kernel settles skipped tiles without counting, so a
    # launch of it carries no roofline.
    if stats[0].triplet_form == 1:
        line["roofline"] = valu_roofline(sum(x.mask_ops for x in stats) / len(stats), tri_s)
    print(json.dumps(line))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nd", type=int, default=50000)
    ap.add_argument("--nr", type=int, default=50000)
    ap.add_argument("--mode", choices=sorted(tp.MODES), default="fast")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="headline: leave out the 64-voxel batch figures")
    ap.add_argument("--budget-s", type=float, default=150.0,
                    help="headline: no step is started that would end past this many seconds of timed steps")
    ap.add_argument("--workload", choices=["headline", "batch", "serve", "sqt", "precision", "pairs", "survey"],
                    default="headline",
                    help="headline: 50k+50k four terms (default); batch: 64 angular voxels of 150k+150k in one "
                         "prepared, graph-replayed persistent launch vs one FP64 plain job at a time; serve: a "
                         "Poisson request stream through the serving queue vs plain one-at-a-time; sqt: 10k+10k "
                         "four-term (s,q,theta) FAST vs plain FP64; precision: 20k+20k wedge, FP32 pair stage "
                         "with FP64 recheck vs FP64 pair stage")
    ap.add_argument("--no-plain", action="store_true",
                    help="sqt: skip the FAST64 and FP64 plain comparison runs (for profiler runs)")
    ap.add_argument("--sqt-n", type=int, default=10000, help="sqt: data (and randoms) catalog size")
    ap.add_argument("--precision-n", type=int, default=20000, help="precision: data (and randoms) catalog size")
    ap.add_argument("--pairs-n", type=int, default=500000, help="pairs: data (and randoms) catalog size")
    ap.add_argument("--survey-n", type=int, default=100000, help="survey: data (and randoms) catalog size")
    ap.add_argument("--requests", type=int, default=48, help="serve: number of requests")
    ap.add_argument("--n-max", type=int, default=50000, help="serve: largest catalog size of a request")
    ap.add_argument("--plain-budget", type=float, default=60.0,
                    help="serve: seconds of plain FP64 time for the sampled comparison")
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
              "(the work queue would be split over devices with one histogram reduction at the end; "
              "only the single-device path is built and measured)", file=sys.stderr)
        return 2
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "sqt":
        return bench_sqt(args)
    if args.workload == "pairs":
        return bench_pairs(args)
    if args.workload == "survey":
        return bench_survey(args)
    if args.workload == "precision":
        return bench_precision(args)

    return bench_headline(args)


def voxel_batch_figures(reps: int = 10) -> dict:
    """The batched half of the metric: 64 angular voxels (8x8 cells of a
    20x20 deg wedge, r 500-800 Mpc/h) of 150k clustered data + 150k randoms,
    lognormal occupancy 500..8000, all four terms of every voxel, 20 log bins
    on [1,50) Mpc/h.  FAST as one prepared batch (one persistent launch,
    autotuned shape, HIP graph replay) against the FP64 plain kernel running
    the voxel jobs one at a time; every voxel's histograms compared."""
    vox = synth.angular_voxels(8, 150000, seed=0)
    spec = tp.r123_spec(20, 1.0, 50.0, "log", 1e-5)
    with tp.Context(0) as ctx:
        cats = [(ctx.catalog(dv), ctx.catalog(rv)) for dv, rv in vox]
        jobs = [tp.four_term_job(cd, cr) for cd, cr in cats]
        t0 = time.perf_counter()
        b = ctx.batch(jobs, [spec] * len(jobs), tp.MODE_FAST, autotune=True, graph=True)
        prep_s = time.perf_counter() - t0
        b.run()
        lat, dev = [], []
        for _ in range(reps):
            t1 = time.perf_counter()
            outs, st = b.run()
            lat.append(time.perf_counter() - t1)
            dev.append(st.device_ms)
        fast = [o.copy() for o in outs]
        launch = b.info.splitlines()[0] if b.info else ""
        b.close()
        sys.stderr.write(f"batch: prepared FAST replay {sum(lat) / len(lat) * 1e3:.1f} ms for {len(jobs)} voxels\n")
        sys.stderr.flush()
        plain_lat = []
        exact = True
        for j, job in enumerate(jobs):
            t1 = time.perf_counter()
            out, _ = ctx.count(job, spec, tp.MODE_PLAIN)
            plain_lat.append(time.perf_counter() - t1)
            exact &= bool(np.array_equal(out, fast[j]))
            if j % 16 == 15:
                sys.stderr.write(f"batch: plain FP64 {j + 1}/{len(jobs)} voxels, {sum(plain_lat):.1f} s\n")
                sys.stderr.flush()
    nv = len(jobs)
    mean = sum(lat) / len(lat)
    plain_rate = nv / sum(plain_lat)
    return {
        "workload": "64 angular voxels (8x8 of a 20x20 deg wedge, r 500-800 Mpc/h), 150k clustered data + 150k "
                    "randoms, lognormal occupancy, four terms per voxel, 20 log bins on [1,50) Mpc/h",
        "voxel_jobs_per_s": nv / mean,
        "replays": reps,
        "replay_p50_ms": pct(lat, 50) * 1e3,
        "replay_p99_ms": pct(lat, 99) * 1e3,
        "device_ms_per_replay": sum(dev) / len(dev),
        "triplets_per_replay": st.triplets,
        "prepare_s": prep_s,
        "launch": launch,
        "plain_fp64_voxel_jobs_per_s": plain_rate,
        "plain_job_p50_ms": pct(plain_lat, 50) * 1e3,
        "plain_job_p99_ms": pct(plain_lat, 99) * 1e3,
        "speedup_vs_plain": (nv / mean) / plain_rate,
        "histograms_identical_to_plain": exact,
    }


def bench_headline(args) -> int:
    """50k + 50k, four terms, (r1,r2,r3): warmup (at most one step: nothing
    is compiled at run time), then timed steps until --steps are done or the
    --budget-s wall budget would be passed by one more step (each step takes
    tens of seconds), then the batched voxel figures and the CPU baselines."""
    d, r = headline_catalogs(args.nd, args.nr)
    spec = headline_spec()
    mode = tp.MODES[args.mode]
    warmup = min(args.warmup, 1)
    with tp.Context(0) as ctx:
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        job = tp.four_term_job(cd, cr)
        for w in range(warmup):
            t1 = time.perf_counter()
            ctx.count(job, spec, mode)
            sys.stderr.write(f"headline: warmup step {w + 1}/{warmup} {time.perf_counter() - t1:.2f} s\n")
            sys.stderr.flush()
        stats, times = [], []
        t0 = time.perf_counter()
        for k in range(args.steps):
            if times and (time.perf_counter() - t0) + max(times) > args.budget_s:
                break
            t1 = time.perf_counter()
            out, st = ctx.count(job, spec, mode)
            times.append(time.perf_counter() - t1)
            stats.append(st)
            sys.stderr.write(f"headline: step {k + 1}/{args.steps} {times[-1]:.2f} s "
                             f"(triplet kernel {st.triplet_ms / 1e3:.2f} s)\n")
            sys.stderr.flush()
        dt = time.perf_counter() - t0
        cd.close()
        cr.close()
    steps = len(stats)
    triplets = stats[0].triplets
    counted = int(out.sum())
    value = triplets * steps / dt
    tri_s = sum(s.triplet_ms for s in stats) / steps / 1e3
    ops = sum(s.mask_ops for s in stats) / steps
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "triplets/s",
        "n_gpus": 1,
        "steps": steps,
        "steps_requested": args.steps,
        "warmup": warmup,
        "warmup_requested": args.warmup,
        "ms_per_step": dt * 1e3 / steps,
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
            "step_s": times,
            "wall_budget_s": args.budget_s,
            "pairs_rechecked_fraction": (stats[0].pairs_rechecked / stats[0].pairs) if stats[0].pairs else 0.0,
            "pair_stage_ms": stats[0].pair_ms,
            "triplet_kernel_ms": tri_s * 1e3,
            "kernel_valu_lane_ops": ops,
            "valu_lane_ops_per_triplet": ops / triplets if triplets else 0.0,
            "triplet_form": FORMS.get(stats[0].triplet_form, "none"),
            "tile_bins_mean": stats[0].tile_bins_mean,
            "tile_fill": stats[0].tile_fill,
            "parallelism": "single GPU",
        },
        "roofline": dict(valu_roofline(ops, tri_s), traffic=measured_traffic(args)),
    }
    if not args.no_batch:
        line["batch"] = voxel_batch_figures(reps=10)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(d, r)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
