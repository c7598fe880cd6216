"""GPU: the (r1,r2,r3) LDS triplet forms -- bit-mask (work per bin pair of a
block triple), run-merging (work per triplet) and split (items whose (A,B)
tile holds at most 5 bins by the bit-mask kernel, the rest by the
run-merging kernel) -- each forced through TPCF_R123_FORM, on geometries
where each would be chosen: all equal the CPU oracle bit for bit, the stats
name the form that counted, and the automatic choice follows the tiles'
mean bins m and fill f (bit-mask when m <= 7.0 sqrt(f), else split when
m <= 8, else run-merging)."""
import importlib
import math
import os

import numpy as np
import pytest

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")

pytestmark = pytest.mark.gpu


def count_with_form(form, job_fn, spec, mode):
    old = os.environ.get("TPCF_R123_FORM")
    if form is None:
        os.environ.pop("TPCF_R123_FORM", None)
    else:
        os.environ["TPCF_R123_FORM"] = form
    try:
        with tp.Context(0) as ctx:
            job, keep = job_fn(ctx)
            out, st = ctx.count(job, spec, mode)
    finally:
        if old is None:
            os.environ.pop("TPCF_R123_FORM", None)
        else:
            os.environ["TPCF_R123_FORM"] = old
    return out, st


def cube_job(nd, nr, size, seed):
    d = synth.neyman_scott(nd, nclust=20, sigma=2.0, size=size, seed=seed)
    r = synth.uniform_cube(nr, size, seed=seed + 1)

    def make(ctx):
        cd, cr = ctx.catalog(d), ctx.catalog(r)
        return tp.four_term_job(cd, cr), (cd, cr)

    return make, [d, r]


CASES = [
    # bins cover the whole box: few bins per tile (bit-mask territory)
    ("wide bins", 1000, 800, 100.0, 3, tp.r123_spec(16, 1.0, 180.0, "log")),
    # short bins in a large box: many bins per near tile (run-merging territory)
    ("short bins", 1200, 1000, 300.0, 5, tp.r123_spec(20, 1.0, 50.0, "log")),
    # linear bins, n = 32 (the largest LDS histogram)
    ("32 lin", 1000, 900, 120.0, 7, tp.r123_spec(32, 0.0, 200.0, "lin")),
]


@pytest.mark.parametrize("name,nd,nr,size,seed,spec", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("mode", [tp.MODE_FAST, tp.MODE_CELLS])
def test_both_forms_equal_the_oracle(orc, name, nd, nr, size, seed, spec, mode):
    print(f"case forms {name} mode={mode}", flush=True)
    make, cats = cube_job(nd, nr, size, seed)
    want = orc.clean_count(cats, tp.FOUR_TERMS, orc.R123, [orc.axis(spec.ax[0].n, spec.ax[0].lo, spec.ax[0].hi,
                                                                         "log" if spec.ax[0].scale else "lin")])
    for form, code in (("runs", 0), ("bits", 1), ("split", 2)):
        got, st = count_with_form(form, make, spec, mode)
        assert st.triplet_form == code
        assert np.array_equal(got, want), (name, form)
        if code == 1:
            assert st.mask_ops > 0
    got, st = count_with_form(None, make, spec, mode)
    assert st.triplet_form in (0, 1, 2) and st.tile_bins_mean > 0 and 0 < st.tile_fill <= 1
    print(f"  auto: form {st.triplet_form}, mean bins {st.tile_bins_mean:.3f}, fill {st.tile_fill:.3f}", flush=True)
    m, f = st.tile_bins_mean, st.tile_fill
    assert st.triplet_form == (1 if m <= 7.0 * math.sqrt(f) else (2 if m <= 8.0 else 0))
    assert np.array_equal(got, want)


VALU_CHILD = r"""
import importlib, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")
d = synth.neyman_scott(1000, nclust=20, sigma=2.0, size=100.0, seed=3)
r = synth.uniform_cube(800, 100.0, seed=4)
with tp.Context(0) as ctx:
    cd, cr = ctx.catalog(d), ctx.catalog(r)
    out, st = ctx.count(tp.four_term_job(cd, cr), tp.r123_spec(16, 1.0, 180.0, "log"), tp.MODE_FAST)
np.save(sys.argv[2], out)
print("form", st.triplet_form, "ops", st.mask_ops, flush=True)
"""


def test_valu_bit_mask_passes_equal_the_oracle(orc, tmp_path):
    """The bit-mask passes on the VALU (TPCF_R123_MFMA=0, read once per
    process, so in a child process) equal the oracle bit for bit."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    dest = str(tmp_path / "valu.npy")
    env = dict(os.environ, TPCF_R123_MFMA="0", TPCF_R123_FORM="bits")
    res = subprocess.run([sys.executable, "-c", VALU_CHILD, root, dest], env=env, capture_output=True, text=True,
                         timeout=240)
    print(res.stdout, res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    form, ops = res.stdout.split()[1], int(res.stdout.split()[3])
    assert form == "1" and ops > 0
    d = synth.neyman_scott(1000, nclust=20, sigma=2.0, size=100.0, seed=3)
    r = synth.uniform_cube(800, 100.0, seed=4)
    want = orc.clean_count([d, r], tp.FOUR_TERMS, orc.R123, [orc.axis(16, 1.0, 180.0, "log")])
    assert np.array_equal(np.load(dest), want)
