"""GPU: tpcf_count_noalloc (no device allocation in the call; the workspace
is the library's fixed 64 MiB scratch) against tpcf_count on the same jobs,
bit-exact, on consecutive calls whose second input is smaller (the scratch
keeps the last call's contents, so what must start at zero is cleared by
every call), and the size limit refused before anything runs."""
import importlib

import numpy as np
import pytest

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")

pytestmark = pytest.mark.gpu


def four_term(ctx, nd, nr, size, seed):
    d = synth.neyman_scott(nd, nclust=30, sigma=2.0, size=size, seed=seed)
    r = synth.uniform_cube(nr, size, seed=seed + 1)
    cd, cr = ctx.catalog(d), ctx.catalog(r)
    return tp.four_term_job(cd, cr), (cd, cr)


def test_noalloc_equals_count_two_calls_second_smaller(ctx):
    print("case noalloc r123: 3000+3000 then 900+700", flush=True)
    spec = tp.r123_spec(20, 1.0, 120.0, "log")
    big, keep1 = four_term(ctx, 3000, 3000, 120.0, seed=3)
    small, keep2 = four_term(ctx, 900, 700, 60.0, seed=5)
    assert ctx.footprint(big, spec) < 63 << 20
    for job in (big, small):
        got, st = ctx.count_noalloc(job, spec, tp.MODE_FAST)
        want, sw = ctx.count(job, spec, tp.MODE_FAST)
        assert st.triplets == sw.triplets
        assert int(want.sum()) > 0
        assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", [tp.MODE_FAST, tp.MODE_FAST64, tp.MODE_PLAIN, tp.MODE_CELLS])
def test_noalloc_every_mode_sqt_then_r123(ctx, mode):
    print(f"case noalloc modes mode={mode}", flush=True)
    sqt = tp.binspec(tp.SQT, [tp.axis(10, 2.0, 20.0, "log"), tp.axis(8, 1.0, 3.0, "lin"),
                              tp.axis(16, 0.0, 1.0, "lin")])
    r123 = tp.r123_spec(16, 0.0, 40.0, "lin")
    j1, keep1 = four_term(ctx, 1500, 1400, 80.0, seed=7)
    j2, keep2 = four_term(ctx, 600, 500, 50.0, seed=9)
    for job, spec in ((j1, sqt), (j2, r123), (j2, sqt)):
        got, _ = ctx.count_noalloc(job, spec, mode)
        want, _ = ctx.count(job, spec, mode)
        assert np.array_equal(got, want)


def test_noalloc_refuses_jobs_beyond_the_scratch(ctx):
    print("case noalloc size limit", flush=True)
    job, keep = four_term(ctx, 20000, 20000, 200.0, seed=11)
    spec = tp.r123_spec(20, 1.0, 50.0, "log")
    assert ctx.footprint(job, spec) > 64 << 20
    with pytest.raises(tp.TpcfError) as e:
        ctx.count_noalloc(job, spec, tp.MODE_FAST)
    assert e.value.code == -3 and "fixed scratch" in str(e.value)
    # still usable afterwards
    small, keep2 = four_term(ctx, 400, 300, 40.0, seed=13)
    got, _ = ctx.count_noalloc(small, spec, tp.MODE_FAST)
    want, _ = ctx.count(small, spec, tp.MODE_FAST)
    assert np.array_equal(got, want)
