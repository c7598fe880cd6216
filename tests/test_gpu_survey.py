"""GPU: a survey split into spatial cells (tpcf_survey_*) counts exactly the
survey's histograms.

The cell jobs (each cell's own points against copies of its 13 higher-index
neighbour cells, cells of side >= the spec's reach) partition the counted
triangles, so their sum equals the histogram of the whole survey: checked
against the oracle on a small survey, against the unsplit CELLS and FAST
runs at 20k + 20k, per cell (the per-cell histograms add up to the total),
for (r1,r2,r3) and (s,q,theta) specs and every mode.  Reference:
examples/test_predivision_of_distances.py:11-35, :123-205 (cells of side
maxsep and neighbour cells), replacing the index-chunk voxels of
src/3pcf_C_version_triangle_parametrization.cc:52-80.
"""
import importlib

import numpy as np
import pytest

from oracle import oracle as orc

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")
pytestmark = pytest.mark.gpu


def full(ctx, d, r, spec, mode):
    cd, cr = ctx.catalog(d), ctx.catalog(r)
    out, _ = ctx.count(tp.four_term_job(cd, cr), spec, mode)
    cd.close()
    cr.close()
    return out


@pytest.mark.parametrize("mode", [tp.MODE_FAST, tp.MODE_FAST64, tp.MODE_PLAIN, tp.MODE_CELLS])
def test_survey_cells_equal_oracle(ctx, mode):
    print(f"case survey small r123 mode={mode}", flush=True)
    d = synth.neyman_scott(1500, nclust=30, sigma=2.0, size=100.0, seed=61)
    r = synth.uniform_cube(1500, 100.0, seed=62)
    spec = tp.r123_spec(16, 1.0, 25.0, "log")
    assert tp.spec_reach(spec) == 25.0
    sv = ctx.survey(d, r, 25.0)
    assert sv.ncells == sv.dims[0] * sv.dims[1] * sv.dims[2] and sv.ncells >= 27
    got, per_cell, _ = sv.count(spec, mode, per_cell=True)
    want = orc.clean_count([d, r], tp.FOUR_TERMS, orc.R123, [orc.axis(16, 1.0, 25.0, "log")])
    assert int(want.sum()) > 0
    assert np.array_equal(got, want)
    assert np.array_equal(per_cell.sum(axis=0), got)
    assert np.count_nonzero(per_cell.sum(axis=(1, 2))) > 1  # triangles owned by several cells
    sv.close()


def test_survey_sqt_equals_unsplit(ctx):
    print("case survey small sqt", flush=True)
    d = synth.neyman_scott(2000, nclust=30, sigma=2.0, size=100.0, seed=63)
    r = synth.uniform_cube(2000, 100.0, seed=64)
    spec = tp.binspec(tp.SQT, [tp.axis(4, 2.0, 6.0, "log"), tp.axis(4, 1.0, 3.0, "lin"), tp.axis(8, 0.0, 1.0, "lin")])
    reach = tp.spec_reach(spec)
    assert 24.0 < reach < 24.001
    sv = ctx.survey(d, r, reach)
    for mode in (tp.MODE_FAST, tp.MODE_PLAIN):
        got, _, _ = sv.count(spec, mode)
        want = full(ctx, d, r, spec, tp.MODE_FAST)
        assert int(want.sum()) > 0
        assert np.array_equal(got, want), mode
    sv.close()


def test_survey_20k_equals_unsplit(ctx):
    print("case survey 20k+20k in a 400 Mpc/h cube", flush=True)
    d = synth.neyman_scott(20000, nclust=200, sigma=2.0, size=400.0, seed=65)
    r = synth.uniform_cube(20000, 400.0, seed=66)
    spec = tp.r123_spec(20, 1.0, 50.0, "log")
    sv = ctx.survey(d, r, 50.0)
    print(f"  {sv.ncells} cells {sv.dims}, {sv.njobs} jobs", flush=True)
    assert sv.ncells >= 343
    got, per_cell, st = sv.count(spec, tp.MODE_FAST, per_cell=True)
    want = full(ctx, d, r, spec, tp.MODE_CELLS)
    assert np.array_equal(got, want)
    assert np.array_equal(per_cell.sum(axis=0), got)
    got_c, _, _ = sv.count(spec, tp.MODE_CELLS)
    assert np.array_equal(got_c, want)
    sv.close()


def test_survey_rejects_specs_beyond_its_cells(ctx):
    print("case survey reach checks", flush=True)
    d = synth.uniform_cube(300, 100.0, seed=67)
    r = synth.uniform_cube(300, 100.0, seed=68)
    sv = ctx.survey(d, r, 20.0)
    with pytest.raises(tp.TpcfError) as e:
        sv.count(tp.r123_spec(10, 1.0, 30.0, "log"))
    assert e.value.code == -1 and "reach" in str(e.value)
    jbj = tp.binspec(tp.JBJ, [tp.axis(5, 1.0, 5.0), tp.axis(5, 0.0, 1.0), tp.axis(5, 0.0, 1.0)])
    with pytest.raises(tp.TpcfError):
        tp.spec_reach(jbj)
    sv.close()
    # an empty random catalog: only DDD has triangles
    sv = ctx.survey(d, np.zeros((0, 3)), 20.0)
    got, _, _ = sv.count(tp.r123_spec(8, 1.0, 20.0, "log"))
    want = orc.clean_count([d], [(0, 0, 0)], orc.R123, [orc.axis(8, 1.0, 20.0, "log")])
    assert np.array_equal(got[0], want[0]) and not got[1:].any()
    sv.close()
