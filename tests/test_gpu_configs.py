"""GPU parity at the stated workload of every benchmark configuration.

cfg1  plain DDD: 1000 points uniform in a 100 Mpc/h cube (seed 0), sorted
      (r1,r2,r3), 16 linear bins on [0,40): every mode equals the oracle.
cfg2  four-term (s,q,theta): 10k Neyman-Scott data (200 clusters, sigma 2)
      + 10k uniform randoms in a 200 Mpc/h cube; s [2,20) 10 log, q [1,3) 8,
      theta/pi [0,1) 16: FAST equals the FP64 plain kernel, FAST64 and CELLS,
      and the four terms add up to DDD of the union catalog.
cfg3  precision form: 20k clustered + 20k uniform in a 10x10 deg wedge,
      r 500-800 Mpc/h, 20 log bins on [1,50), recheck window 1e-5: FAST
      equals FAST64, the triplet total is the closed form, the recheck
      fraction is reported, and a 2k + 2k sub-sample equals the oracle.
cfg4  batching form: 64 angular voxels of 150k + 150k (lognormal
      occupancy), four terms each, one prepared graph-replayed launch: every
      voxel equals count_batch FAST64; the smallest voxels equal PLAIN and
      the oracle; the voxel sums are the sums of the voxel histograms.
cfg5  serving form: a seeded request stream with catalog sizes log-uniform
      in [1k, 50k], random term subsets and bin specs from r123-lin/log and
      sqt-lin/log with 8-32 bins per axis: each result is the same in the mix
      and alone, and the smallest jobs equal PLAIN.

Reference loops restated by the oracle and the plain kernels:
src/3pcf_C_version.cc:363-432 and
src/3pcf_C_version_triangle_parametrization.cc:477-533.
"""
import importlib
import math
import time

import numpy as np
import pytest

from oracle import oracle as orc

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")

pytestmark = pytest.mark.gpu


def c3(n):
    return math.comb(n, 3)


def four_term_total(nd, nr):
    return c3(nd) + math.comb(nd, 2) * nr + nd * math.comb(nr, 2) + c3(nr)


def run(ctx, segs, terms, spec, mode):
    cats = [ctx.catalog(s) for s in segs]
    job = tp.make_job([(c, 0, c.n) for c in cats], terms)
    out, st = ctx.count(job, spec, mode)
    for c in cats:
        c.close()
    return out, st


# ---- cfg1 -------------------------------------------------------------------

@pytest.mark.parametrize("mode", [tp.MODE_FAST, tp.MODE_FAST64, tp.MODE_PLAIN, tp.MODE_CELLS])
def test_cfg1_plain_ddd_1000(ctx, mode):
    print(f"case cfg1 N=1000 DDD 16 lin [0,40) mode={mode}", flush=True)
    d = synth.uniform_cube(1000, 100.0, seed=0)
    got, st = run(ctx, [d], [(0, 0, 0)], tp.r123_spec(16, 0.0, 40.0, "lin"), mode)
    want = orc.clean_count([d], [(0, 0, 0)], orc.R123, [orc.axis(16, 0.0, 40.0, "lin")])
    assert st.triplets == c3(1000) == 166167000
    assert np.array_equal(got, want)
    assert int(got.sum()) > 0


# ---- cfg2 -------------------------------------------------------------------

def cfg2_catalogs():
    d = synth.neyman_scott(10000, nclust=200, sigma=2.0, size=200.0, seed=0)
    r = synth.uniform_cube(10000, 200.0, seed=1)
    return d, r


def cfg2_spec():
    return tp.binspec(tp.SQT, [tp.axis(10, 2.0, 20.0, "log"), tp.axis(8, 1.0, 3.0, "lin"),
                               tp.axis(16, 0.0, 1.0, "lin")])


def test_cfg2_sqt_10k_fast_equals_plain(ctx):
    print("case cfg2 10k+10k sqt FAST/FAST64/CELLS vs PLAIN", flush=True)
    d, r = cfg2_catalogs()
    spec = cfg2_spec()
    t0 = time.perf_counter()
    plain, st = run(ctx, [d, r], tp.FOUR_TERMS, spec, tp.MODE_PLAIN)
    print(f"  plain FP64 {time.perf_counter() - t0:.1f} s", flush=True)
    assert st.triplets == four_term_total(10000, 10000)
    assert int(plain.sum()) > 0
    for mode in (tp.MODE_FAST, tp.MODE_FAST64, tp.MODE_CELLS):
        got, sf = run(ctx, [d, r], tp.FOUR_TERMS, spec, mode)
        assert np.array_equal(got, plain), mode
        if mode == tp.MODE_FAST:
            print(f"  FAST {sf.device_ms:.1f} ms, {sf.pairs_rechecked} triangles rechecked in FP64", flush=True)


def test_cfg2_union_identity(ctx):
    """DDD + DDR + DRR + RRR of (D, R) == DDD of the union catalog: every
    triplet of D u R has exactly one of the four compositions."""
    print("case cfg2 union identity", flush=True)
    d, r = cfg2_catalogs()
    spec = cfg2_spec()
    four, _ = run(ctx, [d, r], tp.FOUR_TERMS, spec, tp.MODE_FAST)
    union, st = run(ctx, [np.concatenate([d, r])], [(0, 0, 0)], spec, tp.MODE_FAST)
    assert st.triplets == c3(20000)
    assert np.array_equal(four.sum(axis=0), union[0])


# ---- cfg3 -------------------------------------------------------------------

def cfg3_catalogs(n):
    d = synth.wedge(n, 10.0, 10.0, 500.0, 800.0, seed=0, clustered=True, nclust=200, sigma=2.0)
    r = synth.wedge(n, 10.0, 10.0, 500.0, 800.0, seed=1)
    return d, r


def test_cfg3_wedge_20k_fast_equals_fast64(ctx):
    print("case cfg3 20k+20k wedge FAST vs FAST64", flush=True)
    d, r = cfg3_catalogs(20000)
    spec = tp.r123_spec(20, 1.0, 50.0, "log", 1e-5)
    fast, st = run(ctx, [d, r], tp.FOUR_TERMS, spec, tp.MODE_FAST)
    f64, _ = run(ctx, [d, r], tp.FOUR_TERMS, spec, tp.MODE_FAST64)
    cells, _ = run(ctx, [d, r], tp.FOUR_TERMS, spec, tp.MODE_CELLS)
    assert st.triplets == four_term_total(20000, 20000) == 10665866680000
    assert np.array_equal(fast, f64)
    assert np.array_equal(cells, f64)
    frac = st.pairs_rechecked / st.pairs
    print(f"  recheck fraction {frac:.3e} ({st.pairs_rechecked} of {st.pairs} pairs), "
          f"{int(fast.sum())} triplets in range", flush=True)
    assert st.pairs == math.comb(40000, 2)
    assert 0 < st.pairs_rechecked < 1e-3 * st.pairs
    assert int(fast.sum()) > 0


def test_cfg3_wedge_subsample_equals_oracle(ctx):
    """A seeded 2k + 2k sub-sample of the cfg3 catalogs against the oracle."""
    print("case cfg3 2k+2k sub-sample vs oracle", flush=True)
    d, r = cfg3_catalogs(20000)
    g = np.random.Generator(np.random.PCG64(5))
    ds = d[np.sort(g.choice(len(d), 2000, replace=False))]
    rs = r[np.sort(g.choice(len(r), 2000, replace=False))]
    spec = tp.r123_spec(20, 1.0, 50.0, "log", 1e-5)
    want = orc.clean_count([ds, rs], tp.FOUR_TERMS, orc.R123, [orc.axis(20, 1.0, 50.0, "log")])
    assert int(want.sum()) > 0
    for mode in (tp.MODE_FAST, tp.MODE_FAST64, tp.MODE_PLAIN, tp.MODE_CELLS):
        got, _ = run(ctx, [ds, rs], tp.FOUR_TERMS, spec, mode)
        assert np.array_equal(got, want), mode


# ---- cfg4 -------------------------------------------------------------------

def test_cfg4_64_voxels_150k(ctx):
    print("case cfg4 64 voxels 150k+150k prepared graph batch", flush=True)
    vox = synth.angular_voxels(8, 150000, seed=0)
    sizes = [len(dv) for dv, _ in vox]
    assert len(vox) == 64 and sum(sizes) == 150000
    assert min(sizes) >= 500 and max(sizes) <= 8000
    spec = tp.r123_spec(20, 1.0, 50.0, "log", 1e-5)
    cats = [(ctx.catalog(dv), ctx.catalog(rv)) for dv, rv in vox]
    jobs = [tp.four_term_job(cd, cr) for cd, cr in cats]
    b = ctx.batch(jobs, [spec] * len(jobs), tp.MODE_FAST, autotune=True, graph=True)
    outs, st = b.run()
    first = [o.copy() for o in outs]
    outs, _ = b.run()  # a replay after the first
    assert all(np.array_equal(a, c) for a, c in zip(first, outs))
    b.close()
    assert st.triplets == sum(four_term_total(n, n) for n in sizes)
    ref, _ = ctx.count_batch(jobs, [spec] * len(jobs), tp.MODE_FAST64)
    for v in range(64):
        assert np.array_equal(first[v], ref[v]), v
    # the smallest voxels: the FP64 plain kernel and the oracle
    for v in np.argsort(sizes, kind="stable")[:3]:
        plain, _ = ctx.count(jobs[v], spec, tp.MODE_PLAIN)
        assert np.array_equal(plain, first[v]), v
        dv, rv = vox[v]
        want = orc.clean_count([dv, rv], tp.FOUR_TERMS, orc.R123, [orc.axis(20, 1.0, 50.0, "log")])
        assert np.array_equal(want, first[v]), v
    total = np.sum(np.stack(first), axis=0)
    assert int(total.sum()) > 0
    for cd, cr in cats:
        cd.close()
        cr.close()


# ---- cfg5 -------------------------------------------------------------------

def request_spec(g):
    kind = int(g.integers(0, 4))
    nb = [int(g.integers(8, 33)) for _ in range(3)]
    if kind == 0:
        return tp.r123_spec(nb[0], 0.0, 150.0, "lin")
    if kind == 1:
        return tp.r123_spec(nb[0], 1.0, 150.0, "log")
    scale = "lin" if kind == 2 else "log"
    return tp.binspec(tp.SQT, [tp.axis(nb[0], 2.0, 20.0, scale), tp.axis(nb[1], 1.0, 3.0, "lin"),
                               tp.axis(nb[2], 0.0, 1.0, "lin")])


def test_cfg5_serving_stream_alone_equals_mix(ctx):
    """Six seeded requests, catalog sizes log-uniform in [1k, 50k] (the
    catalogs fill a 400 Mpc/h box, so a 50k job takes seconds), submitted
    together: every result in the mix equals the job run alone, and the two
    smallest equal the FP64 plain kernel."""
    print("case cfg5 serving stream", flush=True)
    g = np.random.Generator(np.random.PCG64(11))
    reqs = []
    for k in range(6):
        n = int(round(math.exp(g.uniform(math.log(1000), math.log(50000)))))
        mask = int(g.integers(1, 16))
        terms = [tp.FOUR_TERMS[b] for b in range(4) if mask >> b & 1]
        reqs.append((n, terms, request_spec(g), 500 + k))
    sizes = [q[0] for q in reqs]
    print(f"  sizes {sizes}", flush=True)
    assert min(sizes) >= 1000 and max(sizes) <= 50000
    keep, jobs = [], []
    for n, terms, spec, seed in reqs:
        d = synth.neyman_scott(n, nclust=40, sigma=2.0, size=400.0, seed=seed)
        r = synth.uniform_cube(n, 400.0, seed=seed + 7)
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        keep += [cd, cr]
        jobs.append(tp.make_job([(cd, 0, n), (cr, 0, n)], terms))
    t0 = time.perf_counter()
    tickets = [ctx.submit(job, q[2]) for job, q in zip(jobs, reqs)]
    for (t, _) in tickets:
        ctx.wait(t)
    print(f"  mix {time.perf_counter() - t0:.1f} s", flush=True)
    for (_, mixed), job, q in zip(tickets, jobs, reqs):
        alone, _ = ctx.count(job, q[2], tp.MODE_FAST)
        assert np.array_equal(mixed, alone)
    for k in np.argsort(sizes, kind="stable")[:2]:
        plain, _ = ctx.count(jobs[k], reqs[k][2], tp.MODE_PLAIN)
        assert np.array_equal(plain, tickets[k][1]), k
    for c in keep:
        c.close()
