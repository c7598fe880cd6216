"""GPU parity of the sorted-side (r1,r2,r3) path against the clean64 oracle.

Every comparison is bit-exact (integer counts).  Inputs are seeded; sizes
keep the oracle to seconds.
"""
import importlib

import numpy as np
import pytest

from oracle import oracle as orc

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")

pytestmark = pytest.mark.gpu
MODES = [tp.MODE_FAST, tp.MODE_FAST64, tp.MODE_PLAIN, tp.MODE_CELLS]


def oracle_r123(segs, terms, n, lo, hi, scale="lin"):
    return orc.clean_count(segs, terms, orc.R123, [orc.axis(n, lo, hi, scale)])


def run(ctx, segs, terms, n, lo, hi, scale="lin", mode=tp.MODE_FAST, rel=1e-5):
    cats = [ctx.catalog(s) for s in segs]
    job = tp.make_job([(c, 0, c.n) for c in cats], terms)
    out, st = ctx.count(job, tp.r123_spec(n, lo, hi, scale, rel), mode)
    return out, st


@pytest.mark.parametrize("mode", MODES)
def test_ddd_uniform_cube(ctx, mode):
    print(f"case ddd_uniform_cube mode={mode} N=300", flush=True)
    d = synth.uniform_cube(300, 100.0, seed=0)
    got, st = run(ctx, [d], [(0, 0, 0)], 16, 0.0, 40.0, mode=mode)
    want = oracle_r123([d], [(0, 0, 0)], 16, 0.0, 40.0)
    assert np.array_equal(got, want)
    assert st.triplets == 300 * 299 * 298 // 6


@pytest.mark.parametrize("mode", MODES)
def test_four_terms_log_bins(ctx, mode):
    print(f"case four_terms_log mode={mode}", flush=True)
    d = synth.neyman_scott(250, nclust=20, sigma=2.0, size=60.0, seed=1)
    r = synth.uniform_cube(230, 60.0, seed=2)
    got, _ = run(ctx, [d, r], tp.FOUR_TERMS, 20, 1.0, 50.0, "log", mode=mode)
    want = oracle_r123([d, r], tp.FOUR_TERMS, 20, 1.0, 50.0, "log")
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", MODES)
def test_long_runs_of_equal_cells(ctx, mode):
    """Tight clusters with every pair in range: in Morton order a lane's cell
    stays the same over many consecutive columns of a full row, so the
    kernel adds long runs (and short ones where clusters meet)."""
    print(f"case long runs mode={mode}", flush=True)
    d = synth.neyman_scott(600, nclust=6, sigma=0.3, size=20.0, seed=21)
    r = synth.uniform_cube(520, 20.0, seed=22)
    got, st = run(ctx, [d, r], tp.FOUR_TERMS, 12, 0.05, 40.0, "log", mode=mode)
    want = oracle_r123([d, r], tp.FOUR_TERMS, 12, 0.05, 40.0, "log")
    assert np.array_equal(got, want)
    assert int(got.sum()) > 0.999 * st.triplets  # nearly every triplet in range: full rows


def test_headline_geometry_fast_equals_plain(ctx):
    """The 50k + 50k workload's geometry and spec at 3k + 3k (clustered data
    and randoms in a 200 Mpc/h cube, 20 log bins on [1,350): every pair in
    range, full rows and long runs): FAST, FAST64 and CELLS equal the FP64
    brute-force kernel, which the small cases pin to the oracle."""
    print("case headline geometry 3k+3k", flush=True)
    d = synth.neyman_scott(3000, nclust=200, sigma=2.0, size=200.0, seed=0)
    r = synth.uniform_cube(3000, 200.0, seed=1)
    plain, st = run(ctx, [d, r], tp.FOUR_TERMS, 20, 1.0, 350.0, "log", mode=tp.MODE_PLAIN)
    assert int(plain.sum()) > 0.999 * st.triplets
    for mode in (tp.MODE_FAST, tp.MODE_FAST64, tp.MODE_CELLS):
        got, _ = run(ctx, [d, r], tp.FOUR_TERMS, 20, 1.0, 350.0, "log", mode=mode)
        assert np.array_equal(got, plain), mode


@pytest.mark.parametrize("mode", MODES)
def test_far_apart_groups_skip_empty_tiles(ctx, mode):
    """Two groups of points farther apart than the last edge: every tile
    between them has no pair in range, so the kernel skips those items and C
    blocks; the histograms still equal the oracle."""
    print(f"case far-apart groups mode={mode}", flush=True)
    g1 = synth.uniform_cube(300, 15.0, seed=31)
    g2 = synth.uniform_cube(260, 15.0, seed=32) + np.array([100.0, 0.0, 0.0])
    d = np.vstack([g1[:170], g2[:150]])
    r = np.vstack([g1[170:], g2[150:]])
    got, _ = run(ctx, [d, r], tp.FOUR_TERMS, 10, 0.5, 30.0, "log", mode=mode)
    want = oracle_r123([d, r], tp.FOUR_TERMS, 10, 0.5, 30.0, "log")
    assert np.array_equal(got, want)


def test_dense_all_in_range_totals(ctx):
    """Bins cover every pair: each term's total equals its closed form."""
    print("case dense_totals", flush=True)
    d = synth.uniform_cube(200, 10.0, seed=3)
    r = synth.uniform_cube(150, 10.0, seed=4)
    got, st = run(ctx, [d, r], tp.FOUR_TERMS, 8, 0.0, 20.0)
    nd, nr = 200, 150
    c2 = lambda n: n * (n - 1) // 2
    c3 = lambda n: n * (n - 1) * (n - 2) // 6
    assert [int(x) for x in got.sum(axis=1)] == [c3(nd), c2(nd) * nr, nd * c2(nr), c3(nr)]
    union, _ = run(ctx, [np.vstack([d, r])], [(0, 0, 0)], 8, 0.0, 20.0)
    assert np.array_equal(union[0], got.sum(axis=0))


@pytest.mark.parametrize("sizes", [(1, 1), (2, 1), (3, 0), (63, 64), (64, 65), (129, 5)])
def test_ragged_sizes(ctx, sizes):
    print(f"case ragged {sizes}", flush=True)
    nd, nr = sizes
    d = synth.uniform_cube(nd, 30.0, seed=5)
    r = synth.uniform_cube(nr, 30.0, seed=6)
    terms = [t for t in tp.FOUR_TERMS]
    got, _ = run(ctx, [d, r], terms, 10, 0.0, 52.0)
    want = oracle_r123([d, r], terms, 10, 0.0, 52.0)
    assert np.array_equal(got, want)


def test_lattice_exact_edges(ctx):
    """Integer lattice: many distances fall exactly on integer edges."""
    print("case lattice", flush=True)
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(5), indexing="ij"), -1).reshape(-1, 3)
    pts = g.astype(np.float64) * 1.0
    for mode in MODES:
        got, st = run(ctx, [pts], [(0, 0, 0)], 12, 0.0, 12.0, mode=mode)
        want = oracle_r123([pts], [(0, 0, 0)], 12, 0.0, 12.0)
        assert np.array_equal(got, want), mode
        if mode == tp.MODE_FAST:
            assert st.pairs_rechecked > 0


def test_duplicate_points(ctx):
    print("case duplicates", flush=True)
    d = synth.uniform_cube(80, 5.0, seed=7)
    d = np.vstack([d, d[:20]])
    for mode in MODES:
        got, _ = run(ctx, [d], [(0, 0, 0)], 6, 0.0, 9.0, mode=mode)
        assert np.array_equal(got, oracle_r123([d], [(0, 0, 0)], 6, 0.0, 9.0)), mode


def test_far_from_origin_wedge(ctx):
    """Coordinates ~650 Mpc/h from the origin: FP32 alone would misbin."""
    print("case wedge", flush=True)
    d = synth.wedge(300, 2.0, 2.0, 600.0, 640.0, seed=8)
    r = synth.wedge(300, 2.0, 2.0, 600.0, 640.0, seed=9)
    got, st = run(ctx, [d, r], tp.FOUR_TERMS, 20, 1.0, 50.0, "log")
    want = oracle_r123([d, r], tp.FOUR_TERMS, 20, 1.0, 50.0, "log")
    assert np.array_equal(got, want)
    assert 0 <= st.pairs_rechecked < st.pairs


def test_wide_bins(ctx):
    """More than 32 bins per side uses the global-atomic kernel."""
    print("case wide", flush=True)
    d = synth.uniform_cube(150, 20.0, seed=10)
    for mode in MODES:
        got, _ = run(ctx, [d], [(0, 0, 0)], 64, 0.0, float(np.sqrt(3 * 24 * 24)), mode=mode)
        want = oracle_r123([d], [(0, 0, 0)], 64, 0.0, float(np.sqrt(3 * 24 * 24)))
        assert np.array_equal(got, want), mode


def test_three_segments_and_subranges(ctx):
    print("case three segments", flush=True)
    a = synth.uniform_cube(90, 20.0, seed=11)
    b = synth.uniform_cube(100, 20.0, seed=12)
    cats = [ctx.catalog(a), ctx.catalog(b)]
    # segments: a[10:70], b[0:100], a[70:90]  -> terms over three labels
    job = tp.make_job([(cats[0], 10, 70), (cats[1], 0, 100), (cats[0], 70, 90)],
                      [(0, 1, 2), (0, 0, 2), (1, 2, 2), (0, 0, 0)])
    spec = tp.r123_spec(10, 0.0, 35.0)
    got, _ = ctx.count(job, spec)
    want = orc.clean_count([a[10:70], b, a[70:90]], [(0, 1, 2), (0, 0, 2), (1, 2, 2), (0, 0, 0)], orc.R123,
                           [orc.axis(10, 0.0, 35.0)])
    assert np.array_equal(got, want)


def test_batch_equals_single(ctx):
    print("case batch", flush=True)
    jobs, specs, wants = [], [], []
    keep = []
    for v in range(6):
        nd, nr = 40 + 37 * v, 50 + 23 * v
        d = synth.uniform_cube(nd, 40.0, seed=100 + v)
        r = synth.uniform_cube(nr, 40.0, seed=200 + v)
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        keep += [cd, cr]
        jobs.append(tp.four_term_job(cd, cr))
        n = 8 + 3 * v
        specs.append(tp.r123_spec(n, 1.0, 60.0, "log" if v % 2 else "lin"))
        wants.append(oracle_r123([d, r], tp.FOUR_TERMS, n, 1.0, 60.0, "log" if v % 2 else "lin"))
    outs, st = ctx.count_batch(jobs, specs)
    for o, w in zip(outs, wants):
        assert np.array_equal(o, w)


def test_plain_flush_path_many_flushes(ctx, monkeypatch):
    """The plain kernel flushes its 32-bit LDS counters to the uint64
    histogram between items; with the limit lowered to one triplet every
    workgroup flushes before each of its items after the first (6000 points:
    4465 items of one term over at most 1024 workgroups).  Same histograms as
    with the default limit and as FAST."""
    print("case plain flush path", flush=True)
    d = synth.neyman_scott(6000, nclust=50, sigma=3.0, size=100.0, seed=31)
    r = synth.uniform_cube(1500, 100.0, seed=32)
    terms = [(0, 0, 0), (0, 0, 1)]
    base, _ = run(ctx, [d, r], terms, 16, 0.0, 40.0, mode=tp.MODE_PLAIN)
    fast, _ = run(ctx, [d, r], terms, 16, 0.0, 40.0, mode=tp.MODE_FAST)
    monkeypatch.setenv("TPCF_PLAIN_FLUSH_LIMIT", "1")
    flushed, _ = run(ctx, [d, r], terms, 16, 0.0, 40.0, mode=tp.MODE_PLAIN)
    assert int(base.sum()) > 0
    assert np.array_equal(flushed, base)
    assert np.array_equal(fast, base)
