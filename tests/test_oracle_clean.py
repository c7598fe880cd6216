"""Pin the clean64 oracle (the definition the GPU kernels must match).

No reference implementation of the FP64 semantics exists, so the oracle is
pinned by what the mathematics fixes:
  * edge tables and squared thresholds against independent numpy checks;
  * closed-form triplet totals when every side is in range;
  * the union identity DDD+DDR+DRR+RRR(D,R) == DDD(D u R);
  * an independent numpy brute force (sorted sides, searchsorted on edges,
    cosine rule with its own table) on small inputs;
  * hand-computed exact triangles (3-4-5, equilateral, collinear);
  * invariance under catalog permutation and the voxel partition.
"""
import itertools
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
R123, SQT, JBJ = orc.R123, orc.SQT, orc.JBJ
FOUR = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1)]


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def ax(*a):
    return orc.axis(*a)


# ------------------------------------------------------------------ tables

@pytest.mark.parametrize("spec", [(16, 0.0, 40.0, "lin"), (20, 1.0, 50.0, "log"), (7, 0.3, 0.9, "lin"),
                                  (255, 2.0, 300.0, "log")])
def test_edges_closed_form(spec):
    n, lo, hi, sc = spec
    e = orc.edges(ax(*spec))
    assert e[0] == lo and e[-1] == hi and np.all(np.diff(e) > 0)
    if sc == "lin":
        assert np.allclose(np.diff(e), (hi - lo) / n, rtol=1e-12, atol=1e-12)
    else:
        assert np.allclose(np.diff(np.log(e)), math.log(hi / lo) / n, rtol=1e-10)


def test_sq_thresholds_are_exact_sqrt_boundaries():
    e = np.concatenate([orc.edges(ax(20, 1.0, 50.0, "log")), orc.edges(ax(16, 0.0, 40.0, "lin")),
                        [1e-300, 3.0, 4.0, 0.1, 7.5e6]])
    e = np.unique(e)
    T = orc.sq_thresholds(e)
    for em, t in zip(e, T):
        if em <= 0:
            assert t == 0.0
            continue
        assert math.sqrt(t) >= em
        assert math.sqrt(np.nextafter(t, 0.0)) < em


# ------------------------------------------------------------ closed forms

def c2(n):
    return n * (n - 1) // 2


def c3(n):
    return n * (n - 1) * (n - 2) // 6


def test_closed_form_totals_all_in_range():
    d = rng(1).random((37, 3)) * 10
    r = rng(2).random((29, 3)) * 10
    h = orc.clean_count([d, r], FOUR, R123, [ax(8, 0.0, 20.0)])
    assert [int(x) for x in h.sum(axis=1)] == [c3(37), c2(37) * 29, 37 * c2(29), c3(29)]


@pytest.mark.parametrize("param,axes", [
    (R123, [ax(12, 1.0, 15.0, "log")]),
    (SQT, [ax(10, 1.0, 12.0, "lin"), ax(8, 1.0, 4.0, "lin"), ax(6, 0.0, 1.0, "lin")]),
    (JBJ, [ax(10, 1.0, 12.0, "lin"), ax(8, 0.0, 1.0, "lin"), ax(6, 0.0, 1.5, "lin")]),
])
def test_union_identity(param, axes):
    d = rng(3).random((41, 3)) * 9
    r = rng(4).random((33, 3)) * 9
    four = orc.clean_count([d, r], FOUR, param, axes)
    union = orc.clean_count([np.vstack([d, r])], [(0, 0, 0)], param, axes)
    assert np.array_equal(four.sum(axis=0), union[0])


def test_permutation_invariance():
    d = rng(5).random((45, 3)) * 9
    axes = [ax(10, 1.0, 12.0), ax(8, 1.0, 4.0), ax(6, 0.0, 1.0)]
    a = orc.clean_count([d], [(0, 0, 0)], SQT, axes)
    b = orc.clean_count([d[rng(6).permutation(len(d))]], [(0, 0, 0)], SQT, axes)
    assert np.array_equal(a, b)


# ------------------------------------------------------ numpy brute force

def np_edges(n, lo, hi, sc):
    # The edge tables are defined with the C library's exp/log (one host
    # table shared by every device); math.* is that library, numpy's
    # vectorized exp/cos may differ in the last bit.
    if sc == "log":
        a, b = math.log(lo), math.log(hi)
        e = np.array([math.exp(a + (b - a) * m / n) for m in range(n + 1)])
    else:
        e = np.array([lo + (hi - lo) * m / n for m in range(n + 1)])
    e[0], e[-1] = lo, hi
    return e


def np_bin(v, e):
    b = np.searchsorted(e, v, side="right") - 1
    return np.where((v >= e[0]) & (v < e[-1]), b, -1)


def brute(pts_list, terms, param, specs):
    """Independent restatement: explicit triplet enumeration in numpy."""
    e = [np_edges(*s) for s in specs]
    if param == R123:
        cells = specs[0][0] ** 3
    else:
        cells = specs[0][0] * specs[1][0] * specs[2][0]
    out = np.zeros((len(terms), cells), dtype=np.uint64)
    for t, (a, b, c) in enumerate(terms):
        A, B, C = pts_list[a], pts_list[b], pts_list[c]
        trips = []
        for i in range(len(A)):
            for j in range((i + 1) if a == b else 0, len(B)):
                for k in range((j + 1) if b == c else 0, len(C)):
                    trips.append((A[i], B[j], C[k]))
        if not trips:
            continue
        P = np.array(trips)
        def sd(u, v):
            d = P[:, u] - P[:, v]
            return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        S = np.sort(np.stack([sd(0, 1), sd(0, 2), sd(1, 2)], 1), axis=1)
        L = np.sqrt(S)
        n0 = specs[0][0]
        if param == R123:
            bins = np.stack([np_bin(L[:, m], e[0]) for m in range(3)], 1)
            ok = (bins >= 0).all(1)
            idx = (bins[ok, 0] * n0 + bins[ok, 1]) * n0 + bins[ok, 2]
            np.add.at(out[t], idx, 1)
            continue
        nz = (S > 0).all(1)
        L, S = L[nz], S[nz]
        nq, nt = specs[1][0], specs[2][0]
        if param == JBJ:
            bs, bq, bt = np_bin(L[:, 0], e[0]), np_bin(L[:, 0] / L[:, 1], e[1]), np_bin((L[:, 2] - L[:, 1]) / L[:, 0], e[2])
            ok = (bs >= 0) & (bq >= 0) & (bt >= 0)
            np.add.at(out[t], (bs[ok] * nq + bq[ok]) * nt + bt[ok], 1)
            continue
        cos_edges = np.array([math.cos(math.pi * t) for t in e[2]])  # descending
        for (si, a1, a2, opp) in ((0, 0, 1, 2), (1, 1, 2, 0), (0, 0, 2, 1)):
            bs = np_bin(L[:, si], e[0])
            bq = np_bin(L[:, a2] / L[:, a1], e[1])
            cs = ((S[:, a1] + S[:, a2]) - S[:, opp]) / ((2.0 * L[:, a1]) * L[:, a2])
            # bin m iff C[m+1] < c <= C[m]  <=>  -C[m] <= -c < -C[m+1]
            bt = np.searchsorted(-cos_edges, -cs, side="right") - 1
            bt = np.where((cs <= cos_edges[0]) & (cs > cos_edges[-1]), bt, -1)
            ok = (bs >= 0) & (bq >= 0) & (bt >= 0)
            np.add.at(out[t], (bs[ok] * nq + bq[ok]) * nt + bt[ok], 1)
    return out


@pytest.mark.parametrize("param,specs", [
    (R123, [(9, 1.0, 14.0, "log")]),
    (R123, [(6, 0.0, 9.0, "lin")]),
    (SQT, [(8, 1.0, 10.0, "lin"), (6, 1.0, 3.5, "lin"), (8, 0.0, 1.0, "lin")]),
    (SQT, [(5, 2.0, 10.0, "log"), (4, 1.0, 3.0, "log"), (5, 0.1, 0.9, "lin")]),
    (JBJ, [(8, 1.0, 10.0, "lin"), (5, 0.0, 1.0, "lin"), (6, 0.0, 1.2, "lin")]),
])
def test_numpy_brute_force(param, specs):
    d = rng(7).random((19, 3)) * 8
    r = rng(8).random((17, 3)) * 8
    want = brute([d, r], FOUR, param, specs)
    got = orc.clean_count([d, r], FOUR, param, [ax(*s) for s in specs])
    assert np.array_equal(got, want)


def test_lattice_brute_force_exact_edges():
    """Integer coordinates: squared sides are exact integers and many
    distances equal edges exactly (e.g. 1, 2, 3 with unit-width bins)."""
    g = np.stack(np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij"), -1).reshape(-1, 3)
    pts = g.astype(np.float64)
    for param, specs in ((R123, [(5, 0.0, 5.0, "lin")]),
                         (SQT, [(4, 1.0, 3.0, "lin"), (4, 1.0, 3.0, "lin"), (4, 0.0, 1.0, "lin")])):
        want = brute([pts], [(0, 0, 0)], param, specs)
        got = orc.clean_count([pts], [(0, 0, 0)], param, [ax(*s) for s in specs])
        assert np.array_equal(got, want)


# ------------------------------------------------------ exact triangles

def sqt_single(tri, specs):
    return orc.clean_count([np.array(tri, dtype=np.float64)], [(0, 0, 0)], SQT, [ax(*s) for s in specs])[0]


def cell(specs, i, j, k):
    return (i * specs[1][0] + j) * specs[2][0] + k


def test_right_triangle_345():
    # sides 3,4,5; s bins of width 1 on [1,7), q bins of 0.5 on [1,3),
    # theta/pi bins of 0.25 on [0,1).
    specs = [(6, 1.0, 7.0, "lin"), (4, 1.0, 3.0, "lin"), (4, 0.0, 1.0, "lin")]
    h = sqt_single([[0, 0, 0], [3, 0, 0], [0, 4, 0]], specs)
    want = np.zeros_like(h)
    # (s=3, q=4/3, angle(3,4)=pi/2)  -> s bin 2, q bin 0, theta bin 2
    want[cell(specs, 2, 0, 2)] += 1
    # (s=4, q=5/4, angle(4,5)=atan(3/4)/pi=0.2048) -> s bin 3, q bin 0, theta bin 0
    want[cell(specs, 3, 0, 0)] += 1
    # (s=3, q=5/3, angle(3,5)=atan(4/3)/pi=0.2952) -> s bin 2, q bin 1, theta bin 1
    want[cell(specs, 2, 1, 1)] += 1
    assert np.array_equal(h, want)


def test_equilateral_and_collinear():
    specs = [(4, 0.5, 4.5, "lin"), (3, 1.0, 2.5, "lin"), (3, 0.0, 1.0, "lin")]
    s3 = math.sqrt(3.0)
    h = sqt_single([[0, 0, 0], [2, 0, 0], [1, s3, 0]], specs)
    # sides 2, 2, 2 up to rounding: every entry has theta/pi ~ 1/3
    assert int(h.sum()) == 3
    # collinear 0,1,2: E0 angle pi (theta/pi = 1, out of range), E1 and E2
    # angle 0 (bin 0), q = 2 (bin 2 of [1,1.5,2,2.5)).
    h = sqt_single([[0, 0, 0], [1, 0, 0], [2, 0, 0]], specs)
    want = np.zeros_like(h)
    want[cell(specs, 0, 2, 0)] += 2   # s = 1 -> bin 0 of [0.5,1.5)
    assert np.array_equal(h, want)


def test_coincident_points_drop_triangle():
    specs = [(4, 0.0, 4.0, "lin"), (3, 0.0, 3.0, "lin"), (3, 0.0, 1.0, "lin")]
    assert int(sqt_single([[0, 0, 0], [0, 0, 0], [1, 0, 0]], specs).sum()) == 0


# ------------------------------------------------------ voxel partition

def test_voxel_partition_sums_to_full(tmp_path):
    orc.build()
    pts = rng(9).random((90, 3)) * 20
    f = tmp_path / "c.dat"
    with open(f, "w") as fh:
        for i, p in enumerate(pts):
            fh.write(f"{i} 0 0 0 {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
    def run(extra):
        out = tmp_path / "o.dat"
        subprocess.run([orc.CLI, "clean-r123", "--axis", "0,8,0,30", *extra, str(f), str(f), str(f), "-o", str(out)],
                       check=True, stdout=subprocess.DEVNULL)
        vals = out.read_text().split()
        return np.array(vals[15:], dtype=np.int64)
    full = run([])
    tot = np.zeros_like(full)
    for a, b, c in itertools.product(range(3), repeat=3):
        tot += run(["-X", "3", "-x", f"{a:03d}{b:03d}{c:03d}"])
    assert np.array_equal(tot, full)


def _read_tri(path):
    tok = open(path).read().split()
    nb = [int(t) for t in tok[3:6]]
    return nb, np.array(tok[15:], dtype=np.int64).reshape(nb)


def test_clean64_matches_reference_debbie_ddd():
    """clean64 (FP64, unique triplets, edge tables, third angle from its own
    cosine) against a histogram the reference itself produced: the Debbie
    LRG DDD file (examples/sandbox/DDD_Debbie_LRG_CPU_10degx10deg_LRG_binning_SLAC_1k.dat,
    a full run of src/3pcf_C_version_triangle_parametrization.cc with
    THETA_NBINS 25 on sample_data/MICE_LRGs_10degx10deg.dat).  The reference
    computes in FP32 with theta2 = 1 - theta0 - theta1 (:192-236); after its
    zero-side drop it counts the same triplets.  So the two histograms hold
    the same number of entries and differ only where an FP32 value and the
    FP64 value fall on two sides of an edge: measured, 2 entries of 21569 (one
    across a theta edge, one across a q edge).  A misread entry definition
    (another vertex's sides, q or theta off by a bin) moves thousands."""
    gold_dir = os.path.join(HERE, "golden")
    # columns 5-7 of "id ra dec z x y z" (src/3pcf_C_version_triangle_parametrization.cc:366-394)
    cat = np.loadtxt(os.path.join(gold_dir, "inputs", "MICE_LRGs_10degx10deg.dat"), usecols=(4, 5, 6))
    assert cat.shape == (1316, 3)
    nb, gold = _read_tri(os.path.join(gold_dir, "sandbox", "DDD_Debbie_LRG_CPU_10degx10deg_LRG_binning_SLAC_1k.dat"))
    assert nb == [50, 16, 25]

    def clean(q_lo, q_hi, t_lo, t_hi):
        h = orc.clean_count([cat], [(0, 0, 0)], orc.SQT,
                            [orc.axis(50, 2.0, 12.0), orc.axis(16, q_lo, q_hi), orc.axis(25, t_lo, t_hi)])
        return h[0].reshape(nb).astype(np.int64)

    h = clean(0.9, 4.1, 0.0, 1.0)
    assert int(h.sum()) == int(gold.sum()) == 21569
    assert int(np.abs(h - gold).sum()) <= 8  # at most four entries across an edge
    # The check is sensitive: a q axis off by one bin, or theta bins counted
    # from the other end, move thousands of entries.
    assert int(np.abs(clean(1.1, 4.3, 0.0, 1.0) - gold).sum()) > 1000
    assert int(np.abs(clean(0.9, 4.1, 0.0, 1.0)[:, :, ::-1] - gold).sum()) > 1000
