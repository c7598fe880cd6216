"""GPU parity of the pair counts (DD / DR / RR of the two-point function)
against the clean64 oracle (oracle/tpcf_oracle.c: orc_clean_pairs).

Every comparison is bit-exact (integer counts).  Inputs are seeded; sizes
keep the oracle to seconds.  The reference program is src/2pcf_C_version.cc
(700 FP32 bins on [0, 70), self pairs in bin 0); its own outputs pin the
oracle's FP32 restatement in tests/test_oracle_pairs.py, and the clean64
definition here differs from it only as documented (FP64 distances, i < j
pairs, uint64 counters).
"""
import importlib
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

tp = importlib.import_module("3pcf_amd")
synth = importlib.import_module("3pcf_amd.synth")

pytestmark = pytest.mark.gpu
MODES = [tp.MODE_FAST, tp.MODE_FAST64, tp.MODE_PLAIN, tp.MODE_CELLS]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INP = os.path.join(ROOT, "tests", "golden", "inputs")
TWO = os.path.join(ROOT, "3pcf_amd", "bin", "2pcf_C_version")


def c2(n):
    return n * (n - 1) // 2


def run(ctx, segs, terms, ax, mode=tp.MODE_FAST):
    cats = [ctx.catalog(s) for s in segs]
    job = tp.make_job([(c, 0, c.n) for c in cats], terms)
    return ctx.count_pairs(job, tp.axis(*ax), mode)


def want(segs, terms, ax):
    return orc.clean_pairs(segs, terms, orc.axis(*ax))


@pytest.mark.parametrize("mode", MODES)
def test_dd_dr_rr_reference_bins(ctx, mode):
    """The reference's binning (700 bins on [0, 70)) on clustered data and
    uniform randoms whose sizes are not multiples of any block."""
    print(f"case dd_dr_rr_700 mode={mode}", flush=True)
    d = synth.neyman_scott(3001, nclust=40, sigma=3.0, size=150.0, seed=11)
    r = synth.uniform_cube(2555, 150.0, seed=12)
    ax = tp.TWO_PCF_AXIS
    got, st = run(ctx, [d, r], tp.PAIR_TERMS, ax, mode)
    assert np.array_equal(got, want([d, r], tp.PAIR_TERMS, ax))
    assert st.pairs == c2(3001) + 3001 * 2555 + c2(2555)


@pytest.mark.parametrize("mode", MODES)
def test_log_axis(ctx, mode):
    print(f"case log_axis mode={mode}", flush=True)
    d = synth.neyman_scott(2100, nclust=25, sigma=2.0, size=80.0, seed=13)
    r = synth.uniform_cube(1900, 80.0, seed=14)
    ax = (20, 1.0, 50.0, "log")
    got, _ = run(ctx, [d, r], tp.PAIR_TERMS, ax, mode)
    assert np.array_equal(got, want([d, r], tp.PAIR_TERMS, ax))


@pytest.mark.parametrize("ax", [(4096, 0.0, 100.0, "lin"), (1, 0.0, 1000.0, "lin"), (4096, 0.01, 200.0, "log"),
                                (3, 5.0, 6.0, "lin")])
def test_bin_count_extremes(ctx, ax):
    print(f"case bins {ax}", flush=True)
    d = synth.neyman_scott(1500, nclust=15, sigma=4.0, size=120.0, seed=15)
    r = synth.uniform_cube(1300, 120.0, seed=16)
    for mode in (tp.MODE_FAST, tp.MODE_PLAIN, tp.MODE_CELLS):
        got, _ = run(ctx, [d, r], tp.PAIR_TERMS, ax, mode)
        assert np.array_equal(got, want([d, r], tp.PAIR_TERMS, ax)), mode


def test_totals_when_bins_cover_every_pair(ctx):
    print("case totals", flush=True)
    d = synth.uniform_cube(2049, 10.0, seed=17)
    r = synth.uniform_cube(1025, 10.0, seed=18)
    got, _ = run(ctx, [d, r], tp.PAIR_TERMS, (50, 0.0, 20.0, "lin"))
    assert [int(x) for x in got.sum(axis=1)] == [c2(2049), 2049 * 1025, c2(1025)]


def test_lattice_distances_on_edges(ctx):
    """Integer lattice: many distances (1, 2, 3, sqrt 2 ...) fall exactly on
    edges of a half-unit axis; exact FP64 thresholds decide them."""
    print("case lattice", flush=True)
    g = np.array([(x, y, z) for x in range(9) for y in range(9) for z in range(9)], dtype=np.float64)
    g2 = g + 0.5
    for ax in [(24, 0.0, 12.0, "lin"), (12, 1.0, 16.0, "log"), (8, 1.0, 5.0, "lin")]:
        for mode in MODES:
            got, _ = run(ctx, [g, g2], tp.PAIR_TERMS, ax, mode)
            assert np.array_equal(got, want([g, g2], tp.PAIR_TERMS, ax)), (ax, mode)


def test_coincident_points_and_far_offsets(ctx):
    """Duplicated points (distance 0: bin 0 when lo = 0, dropped when lo > 0)
    and a survey 700 Mpc/h from the origin, where the FP32 bin guess is
    furthest off and the FP64 comparison decides."""
    print("case duplicates_offsets", flush=True)
    base = synth.neyman_scott(700, nclust=10, sigma=1.0, size=30.0, seed=19)
    d = np.concatenate([base, base[:300]]) + 700.0
    r = synth.uniform_cube(800, 30.0, seed=20) + 700.0
    for ax in [(64, 0.0, 16.0, "lin"), (64, 0.5, 16.0, "log")]:
        for mode in (tp.MODE_FAST, tp.MODE_PLAIN, tp.MODE_CELLS):
            got, _ = run(ctx, [d, r], tp.PAIR_TERMS, ax, mode)
            assert np.array_equal(got, want([d, r], tp.PAIR_TERMS, ax)), (ax, mode)
    got, _ = run(ctx, [d], [(0, 0)], (64, 0.0, 16.0, "lin"))
    assert int(got[0, 0]) >= 300


def test_empty_and_tiny_segments(ctx):
    print("case empty_tiny", flush=True)
    one = np.array([[1.0, 2.0, 3.0]])
    two = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    for segs, terms, total in [([one], [(0, 0)], 0), ([two], [(0, 0)], 1), ([one, two], [(0, 1)], 2),
                               ([two, one], [(0, 0), (0, 1), (1, 1)], 3)]:
        got, st = run(ctx, segs, terms, (10, 0.0, 10.0, "lin"))
        assert int(got.sum()) == total
        assert np.array_equal(got, want(segs, terms, (10, 0.0, 10.0, "lin")))
    two_d = ctx.catalog(two)
    job = tp.make_job([(two_d, 1, 1), (two_d, 0, 2)], [(0, 0), (0, 1)])
    got, st = ctx.count_pairs(job, tp.axis(10, 0.0, 10.0))
    assert int(got.sum()) == 0 and st.pairs == 0


def test_batch_of_voxel_pairs_matches_single_calls(ctx):
    """A voxel run in one launch: every chunk pair of one catalog (p < q
    products, p == q unique pairs), jobs with different axes; the chunk
    pairs add up to the full catalog's unique pairs."""
    print("case batch_voxels", flush=True)
    d = synth.neyman_scott(4000, nclust=30, sigma=3.0, size=100.0, seed=21)
    cat = ctx.catalog(d)
    div = 3
    per = 4000 // div
    bounds = [(p * per, (p + 1) * per if p < div - 1 else 4000) for p in range(div)]
    jobs, axes = [], []
    for p in range(div):
        for q in range(p, div):
            if p == q:
                jobs.append(tp.make_job([(cat, *bounds[p])], [(0, 0)]))
            else:
                jobs.append(tp.make_job([(cat, *bounds[p]), (cat, *bounds[q])], [(0, 1)]))
            axes.append(tp.axis(700, 0.0, 70.0))
    jobs.append(tp.make_job([(cat, 0, 4000)], [(0, 0)]))
    axes.append(tp.axis(30, 0.5, 60.0, "log"))
    outs, st = ctx.count_pairs_batch(jobs, axes)
    culled, _ = ctx.count_pairs_batch(jobs, axes, tp.MODE_CELLS)
    for job, ax, out, cul in zip(jobs, axes, outs, culled):
        single, _ = ctx.count_pairs(job, ax)
        assert np.array_equal(out, single)
        assert np.array_equal(cul, single)
    full = want([d], [(0, 0)], (700, 0.0, 70.0, "lin"))[0]
    assert np.array_equal(sum(o[0] for o in outs[:-1]), full)
    assert np.array_equal(outs[-1], want([d], [(0, 0)], (30, 0.5, 60.0, "log")))
    assert st.pairs == c2(4000) * 2


def test_large_catalog_properties(ctx):
    """300k points (4.5e10 pairs, beyond the oracle's reach): FAST equals
    PLAIN, totals equal the closed form, and the DD counts are additive over
    a split of the catalog (DD(A+B) = DD(A) + DD(B) + DR(A, B))."""
    print("case large_properties N=300000", flush=True)
    d = synth.neyman_scott(300000, nclust=2000, sigma=3.0, size=500.0, seed=22)
    ax = (700, 0.0, 70.0, "lin")
    fast, st = run(ctx, [d], [(0, 0)], ax, tp.MODE_FAST)
    plain, _ = run(ctx, [d], [(0, 0)], ax, tp.MODE_PLAIN)
    assert np.array_equal(fast, plain)
    a, b = d[:123457], d[123457:]
    parts, _ = run(ctx, [a, b], [(0, 0), (0, 1), (1, 1)], ax)
    assert np.array_equal(parts.sum(axis=0), fast[0])
    cover, _ = run(ctx, [d], [(0, 0)], (8, 0.0, 1000.0, "lin"))
    assert int(cover.sum()) == c2(300000)
    # the culled form on the same catalog: short and long last edges
    for cax in [ax, (40, 0.5, 12.0, "log"), (8, 0.0, 1000.0, "lin")]:
        want_fast, _ = run(ctx, [d], [(0, 0)], cax, tp.MODE_FAST)
        got_cells, _ = run(ctx, [d], [(0, 0)], cax, tp.MODE_CELLS)
        assert np.array_equal(got_cells, want_fast), cax


def test_bad_arguments_raise(ctx):
    d = ctx.catalog(synth.uniform_cube(10, 1.0, seed=0))
    job = tp.make_job([(d, 0, 10)], [(0, 0)])
    for ax in [tp.axis(4097, 0.0, 1.0), tp.axis(0, 0.0, 1.0), tp.axis(5, -1.0, 1.0), tp.axis(5, 0.0, 1.0, "log"),
               tp.axis(5, 0.0, 1e101)]:
        with pytest.raises(tp.TpcfError):
            ctx.count_pairs(job, ax)
    bad = tp.make_job([(d, 0, 10)], [(0, 1)])
    with pytest.raises(tp.TpcfError):
        ctx.count_pairs(bad, tp.axis(5, 0.0, 1.0))
    with pytest.raises(tp.TpcfError):
        ctx.count_pairs(job, tp.axis(5, 0.0, 1.0), mode=7)


# ---- the 2pcf_C_version drop-in ---------------------------------------------


def run_two(tmp_path, args, prog):
    out = tmp_path / "out.dat"
    cmd = [prog, *args, "-o", str(out)]
    subprocess.run(cmd, cwd=INP, check=True, stdout=subprocess.DEVNULL, timeout=120)
    return out.read_bytes()


CLI_CASES = [
    ["d200.dat", "d200.dat"],
    ["d200.dat", "r180.dat"],
    ["r180.dat", "r180.dat"],
    ["MICE_5degx5deg_1k_randomized.dat", "flat_MICE_5degx5deg_1k.dat"],
    ["d200.dat", "./d200.dat", "-X", "2", "-x", "000000000"],
    ["d200.dat", "d200.dat", "-X", "2", "-x", "000001000"],
    ["d200.dat", "d200.dat", "-X", "2", "-x", "001000000"],
    ["d200.dat", "r180.dat", "-X", "3", "-x", "002001000"],
    ["d200.dat", "r180.dat", "--bins", "25", "--lo", "1", "--hi", "60", "--log"],
    ["d200.dat", "r180.dat", "--mode", "cells"],
    ["d200.dat", "d200.dat", "--mode", "cells", "-X", "2", "-x", "000001000"],
]


@pytest.mark.parametrize("args", CLI_CASES, ids=lambda a: "_".join(a).replace(".dat", ""))
def test_cli_matches_oracle_cli(tmp_path, args):
    print(f"case cli {args}", flush=True)
    orc.build()
    mine = run_two(tmp_path, args, TWO)
    oargs = []
    it = iter(args)
    for a in it:
        if a == "--bins":
            n = next(it)
        elif a == "--lo":
            lo = next(it)
        elif a == "--hi":
            hi = next(it)
        elif a == "--log":
            oargs += ["--axis", f"0,{n},{lo},{hi},log"]
        elif a == "--mode":
            next(it)
        else:
            oargs.append(a)
    out = tmp_path / "o.dat"
    subprocess.run([orc.CLI, "clean-two", "--threads", "8", *oargs, "-o", str(out)], cwd=INP, check=True,
                   stdout=subprocess.DEVNULL)
    assert mine == out.read_bytes()


def test_cli_voxel_chunks_add_up_to_full_run(tmp_path):
    print("case cli_voxel_sum", flush=True)
    full = run_two(tmp_path, ["d200.dat", "d200.dat"], TWO).decode().split("\n")
    total = np.zeros(700, dtype=np.uint64)
    for p in range(3):
        for q in range(3):
            body = run_two(tmp_path, ["d200.dat", "d200.dat", "-X", "3", "-x", f"{p:03d}{q:03d}000"], TWO)
            total += np.array([int(v) for v in body.decode().split("\n")[4:704]], dtype=np.uint64)
    assert [int(v) for v in full[4:704]] == [int(v) for v in total]


def test_driver_runs_dd_dr_rr(tmp_path):
    """run_CPU_calculation_2pcf.sh writes DD/DR/RR files equal to direct runs."""
    print("case driver_2pcf", flush=True)
    drv = os.path.join(ROOT, "3pcf_amd", "drivers", "run_CPU_calculation_2pcf.sh")
    env = dict(os.environ, INPUT0=os.path.join(INP, "d200.dat"), INPUT1=os.path.join(INP, "r180.dat"), TAG="t")
    subprocess.run(["bash", drv, "1", "all"], cwd=tmp_path, env=env, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, timeout=300)
    for term, files in (("DD", ["d200.dat", "d200.dat"]), ("DR", ["d200.dat", "r180.dat"]),
                        ("RR", ["r180.dat", "r180.dat"])):
        direct = run_two(tmp_path, files, TWO)
        assert (tmp_path / f"{term}_t_1k.dat").read_bytes() == direct, term
