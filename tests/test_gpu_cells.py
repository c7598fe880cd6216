"""The culled triplet form (TPCF_MODE_CELLS): segments staged in Morton
order, block triples skipped when a box gap rules out every counted triplet.
Histograms must equal the clean64 oracle and the unculled FAST path bit for
bit, on catalogs much larger than the last bin edge (where most block
triples are skipped) and through the batched and serving entry points.
"""
import importlib

import numpy as np
import pytest

from oracle import oracle as orc

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")

pytestmark = pytest.mark.gpu


def job_of(ctx, segs, terms):
    cats = [ctx.catalog(s) for s in segs]
    return tp.make_job([(c, 0, c.n) for c in cats], terms), cats


def test_r123_culled_large_box_matches_oracle(ctx):
    print("case r123 cells large box", flush=True)
    d = synth.neyman_scott(1500, nclust=40, sigma=3.0, size=300.0, seed=31)
    r = synth.uniform_cube(1400, 300.0, seed=32)
    for n, lo, hi, scale in [(20, 1.0, 30.0, "log"), (16, 0.0, 25.0, "lin"), (40, 0.0, 30.0, "lin")]:
        job, cats = job_of(ctx, [d, r], tp.FOUR_TERMS)
        spec = tp.r123_spec(n, lo, hi, scale)
        got, st = ctx.count(job, spec, tp.MODE_CELLS)
        want = orc.clean_count([d, r], tp.FOUR_TERMS, orc.R123, [orc.axis(n, lo, hi, scale)])
        assert np.array_equal(got, want), (n, lo, hi, scale)
        assert st.triplets == sum(int(x) for x in [orc_c3(1500), orc_c2(1500) * 1400, 1500 * orc_c2(1400),
                                                   orc_c3(1400)])


def orc_c2(n):
    return n * (n - 1) // 2


def orc_c3(n):
    return n * (n - 1) * (n - 2) // 6


@pytest.mark.parametrize("param", [tp.SQT, tp.JBJ])
def test_sqt_and_jbj_culled_large_box_matches_oracle(ctx, param):
    print(f"case sqt/jbj cells param={param}", flush=True)
    d = synth.neyman_scott(1300, nclust=30, sigma=2.0, size=200.0, seed=33)
    r = synth.uniform_cube(1200, 200.0, seed=34)
    if param == tp.SQT:
        axes = [tp.axis(8, 2.0, 10.0, "log"), tp.axis(6, 1.0, 2.5), tp.axis(10, 0.0, 1.0)]
        oaxes = [orc.axis(8, 2.0, 10.0, "log"), orc.axis(6, 1.0, 2.5), orc.axis(10, 0.0, 1.0)]
    else:
        axes = [tp.axis(8, 2.0, 10.0, "log"), tp.axis(5, 0.4, 1.0), tp.axis(6, 0.0, 1.2)]
        oaxes = [orc.axis(8, 2.0, 10.0, "log"), orc.axis(5, 0.4, 1.0), orc.axis(6, 0.0, 1.2)]
    spec = tp.binspec(param, axes)
    job, cats = job_of(ctx, [d, r], tp.FOUR_TERMS)
    got, _ = ctx.count(job, spec, tp.MODE_CELLS)
    fast, _ = ctx.count(job, spec, tp.MODE_FAST)
    assert np.array_equal(got, fast)
    want = orc.clean_count([d, r], tp.FOUR_TERMS, orc.SQT if param == tp.SQT else orc.JBJ, oaxes)
    assert np.array_equal(got, want)


def test_culled_equals_fast_at_scale(ctx):
    """20k + 20k in a 400 Mpc/h cube (beyond the oracle's reach): every
    parametrization's culled histograms equal the unculled FAST ones."""
    print("case cells at scale 20k+20k", flush=True)
    d = synth.neyman_scott(20000, nclust=300, sigma=3.0, size=400.0, seed=35)
    r = synth.uniform_cube(20000, 400.0, seed=36)
    job, cats = job_of(ctx, [d, r], tp.FOUR_TERMS)
    for spec in [tp.r123_spec(20, 1.0, 40.0, "log"),
                 tp.binspec(tp.SQT, [tp.axis(10, 2.0, 20.0, "log"), tp.axis(8, 1.0, 3.0), tp.axis(16, 0.0, 1.0)])]:
        fast, sf = ctx.count(job, spec, tp.MODE_FAST)
        cells, sc = ctx.count(job, spec, tp.MODE_CELLS)
        assert np.array_equal(cells, fast)
        assert int(cells.sum()) > 0
        print(f"  fast {sf.triplet_ms:.1f} ms, cells {sc.triplet_ms:.1f} ms", flush=True)


def test_culled_batch_and_serving_equal_fast(ctx):
    print("case cells batch + serving", flush=True)
    rng = np.random.default_rng(37)
    jobs, specs, keep = [], [], []
    for k in range(6):
        d = synth.neyman_scott(int(rng.integers(400, 1500)), nclust=20, sigma=2.0, size=150.0, seed=40 + k)
        r = synth.uniform_cube(int(rng.integers(400, 1500)), 150.0, seed=50 + k)
        job, cats = job_of(ctx, [d, r], tp.FOUR_TERMS)
        keep.append(cats)
        jobs.append(job)
        specs.append(tp.r123_spec(12, 0.0, 20.0) if k % 2 else
                     tp.binspec(tp.SQT, [tp.axis(6, 2.0, 8.0), tp.axis(4, 1.0, 2.0), tp.axis(8, 0.0, 1.0)]))
    outs, _ = ctx.count_batch(jobs, specs, tp.MODE_CELLS)
    for job, spec, out in zip(jobs, specs, outs):
        fast, _ = ctx.count(job, spec, tp.MODE_FAST)
        assert np.array_equal(out, fast)
    tickets = [ctx.submit(job, spec, tp.MODE_CELLS) for job, spec in zip(jobs, specs)]
    for (t, out), want in zip(tickets, outs):
        ctx.wait(t)
        assert np.array_equal(out, want)


def test_prepared_batch_cells_replays_equal_fast(ctx):
    """A prepared CELLS batch (staged and culled once, graph-replayed)
    gives the unculled histograms on every replay."""
    print("case prepared cells batch", flush=True)
    jobs, want = [], []
    for k in range(5):
        d = synth.neyman_scott(900 + 150 * k, nclust=20, sigma=2.0, size=200.0, seed=80 + k)
        r = synth.uniform_cube(800 + 100 * k, 200.0, seed=90 + k)
        job, cats = job_of(ctx, [d, r], tp.FOUR_TERMS)
        jobs.append(job)
        want.append(ctx.count(job, tp.r123_spec(16, 1.0, 30.0, "log"), tp.MODE_FAST)[0])
    specs = [tp.r123_spec(16, 1.0, 30.0, "log")] * len(jobs)
    for graph in (True, False):
        b = ctx.batch(jobs, specs, tp.MODE_CELLS, autotune=False, graph=graph)
        for _ in range(3):
            outs, _ = b.run()
            for o, w in zip(outs, want):
                assert np.array_equal(o, w)
        b.close()
