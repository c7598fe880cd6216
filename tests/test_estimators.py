"""Estimators (zeta, reduced amplitude, 2PCF xi): the C++ library functions
and the tpcf_estimate program against the oracle's numpy restatement of
examples/plot_MICE_3pcf.py and examples/plot_MICE_2pcf.py
(oracle/estimators.py).  Host arithmetic only; no GPU.

The restatement is pinned by hand-worked cells with exact answers, the
invariant that proportional term histograms give zeta = 0 (and xi = 0), the
closed forms of the normalizations (the unique triplet and pair counts), the
NaN/inf handling of the scripts, and the abscissae of the reference's own
2PCF table (examples/master_2pcf_reference.dat, written by
plot_MICE_2pcf.py from 700 bins on [0, 70)).
"""
import importlib
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import estimators as est

tp = importlib.import_module("3pcf_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SANDBOX = os.path.join(ROOT, "tests", "golden", "sandbox")
DEBBIE = os.path.join(SANDBOX, "DDD_Debbie_LRG_CPU_10degx10deg_LRG_binning_SLAC_1k.dat")
TABLE = os.path.join(SANDBOX, "master_2pcf_reference.dat")
PROG = os.path.join(ROOT, "3pcf_amd", "bin", "tpcf_estimate")


def write_tri(path, ngals, nb, ranges, values):
    with open(path, "w") as f:
        for n in ngals:
            f.write(f"{n}\n")
        f.write("%d %d %d\n" % tuple(nb))
        for r in ranges:
            f.write("%-6.3f %-6.3f %-6.3f\n" % tuple(r))
        v = np.asarray(values).reshape(nb[0] * nb[1], nb[2])
        for row in v:
            f.write("".join(f"{int(x)} " for x in row) + "\n")


def write_two(path, ngals, lo, hi, values):
    n = len(values)
    with open(path, "w") as f:
        f.write(f"{ngals[0]}\n{ngals[1]}\n{n}\n")
        f.write("%-6.3f %-6.3f %-6.3f\n" % (lo, hi, (hi - lo) / n))
        for v in values:
            f.write(f"{int(v)}\n")


RANGES = [(0.0, 4.0, 2.0), (1.0, 2.0, 1.0), (0.0, 1.0, 0.5)]


def four_files(tmp_path, cells):
    """DDD..RRR files of shape (2, 1, 2) with N_D = 10, N_R = 20: norms 120,
    900, 1900, 1140; cells[t] the four counts of term t."""
    sizes = [(10, 10, 10), (10, 10, 20), (10, 20, 20), (20, 20, 20)]
    for t, name in enumerate(("DDD", "DDR", "DRR", "RRR")):
        write_tri(tmp_path / f"{name}_x.dat", sizes[t], (2, 1, 2), RANGES, cells[t])
    return str(tmp_path / "DDD_x.dat")


def test_oracle_hand_worked_cells(tmp_path):
    # cell 0: 240/120 - 3*900/900 + 3*3800/1900 - 1140/1140 = 2 - 3 + 6 - 1 = 4
    # cell 1: RRR = 0, numerator > 0: +inf -> 0
    # cell 2: every term 0: NaN -> 0
    # cell 3: only DDR > 0: -3/0 = -inf, kept (the scripts zero +inf only)
    ddd = [240, 120, 0, 0]
    ddr = [900, 0, 0, 900]
    drr = [3800, 0, 0, 0]
    rrr = [1140, 0, 0, 0]
    z, ngals, nbins, _ = est.zeta(four_files(tmp_path, [ddd, ddr, drr, rrr]))
    assert z.reshape(-1).tolist() == [4.0, 0.0, 0.0, -math.inf]
    assert [est.triplet_norm(s, t) for t, s in enumerate([(10, 10, 10), (10, 10, 20), (10, 20, 20),
                                                          (20, 20, 20)])] == [120.0, 900.0, 1900.0, 1140.0]


def test_oracle_norms_are_unique_counts():
    for nd, nr in [(3, 4), (1000, 2500), (123457, 98765)]:
        assert est.triplet_norm((nd, nd, nd), 0) == float(math.comb(nd, 3))
        assert est.triplet_norm((nd, nd, nr), 1) == float(math.comb(nd, 2) * nr)
        assert est.triplet_norm((nd, nr, nr), 2) == float(nd * math.comb(nr, 2))
        assert est.pair_norm((nd, nd), 0) == float(math.comb(nd, 2))
        assert est.pair_norm((nd, nr), 1) == float(nd * nr)


def test_oracle_proportional_terms_give_zero(tmp_path):
    shape = np.array([3, 7, 11, 5])
    norms = [120, 900, 1900, 1140]
    z, *_ = est.zeta(four_files(tmp_path, [shape * n for n in norms]))
    assert np.all(z == 0.0)
    write_two(tmp_path / "DD_p.dat", (10, 10), 0.0, 4.0, [2 * 45, 3 * 45])
    write_two(tmp_path / "DR_p.dat", (10, 20), 0.0, 4.0, [2 * 200, 3 * 200])
    write_two(tmp_path / "RR_p.dat", (20, 20), 0.0, 4.0, [2 * 190, 3 * 190])
    lines, x = est.xi_table(str(tmp_path / "DD_p.dat"))
    assert np.all(x == 0.0) and lines == ["2.000000 0.000000\n", "4.000000 0.000000\n"]


def test_oracle_table_abscissae_match_reference_table(tmp_path):
    """plot_MICE_2pcf.py wrote examples/master_2pcf_reference.dat from 700
    bins on [0, 70): its first column is the upper bin edges as "%f"."""
    rng = np.random.default_rng(0)
    for name in ("DD", "DR", "RR"):
        write_two(tmp_path / f"{name}_m.dat", (1000, 1000), 0.0, 70.0, rng.integers(1, 100, 700))
    lines, _ = est.xi_table(str(tmp_path / "DD_m.dat"))
    ref = open(TABLE).read().splitlines()
    assert len(lines) == len(ref) == 700
    assert [ln.split()[0] for ln in lines] == [r.split()[0] for r in ref]


def test_oracle_nearest_neighbour_lookup():
    xs = np.array([0.1, 0.2, 0.3, 0.4])
    ys = np.array([10.0, 20.0, 30.0, 40.0])
    assert est.sxx2cf(np.array([0.0, 0.149, 0.26, 5.0]), xs, ys).tolist() == [10.0, 10.0, 30.0, 40.0]


# ---- the product against the oracle ----------------------------------------


def test_library_norms_and_zeta_match_oracle_on_reference_files():
    want, ngals, nbins, _ = est.zeta(DEBBIE)
    h, norms = [], []
    for t, name in enumerate(("DDD", "DDR", "DRR", "RRR")):
        g, nb, _, pts = est.read_tri(DEBBIE.replace("DDD", name))
        h.append(pts)
        norms.append(tp.triplet_norm([int(x) for x in g], t))
        assert norms[-1] == est.triplet_norm(g, t)
    got = tp.zeta(*h, norms)
    assert np.array_equal(got, want)
    assert np.count_nonzero(got) > 100


def test_library_xi_and_reduced_match_oracle(tmp_path):
    rng = np.random.default_rng(1)
    dd = rng.integers(0, 5000, 700)
    dd[:3] = 0
    write_two(tmp_path / "DD_r.dat", (1000, 1000), 0.0, 70.0, dd)
    write_two(tmp_path / "DR_r.dat", (1000, 1200), 0.0, 70.0, rng.integers(0, 6000, 700))
    rr = rng.integers(0, 5000, 700)
    rr[1] = 0
    write_two(tmp_path / "RR_r.dat", (1200, 1200), 0.0, 70.0, rr)
    lines, want = est.xi_table(str(tmp_path / "DD_r.dat"))
    norms = [est.pair_norm(est.read_two(str(tmp_path / f"{n}_r.dat"))[0], t) for t, n in enumerate(("DD", "DR", "RR"))]
    got = tp.xi(*[est.read_two(str(tmp_path / f"{n}_r.dat"))[3] for n in ("DD", "DR", "RR")], norms)
    assert np.array_equal(got, want)
    out = tmp_path / "xi.dat"
    subprocess.run([PROG, "xi", str(tmp_path / "DD_r.dat"), "-o", str(out)], check=True)
    assert out.read_text() == "".join(lines)
    # reduced amplitude over the reference's table
    z, _, nbins, hv = est.zeta(DEBBIE)
    rows = est.reduced(DEBBIE, 5, [0, 3, 7], TABLE)
    tab = np.loadtxt(TABLE)
    for q, x, zrow, den, qred in rows:
        th, d2, q2 = tp.reduced_q(zrow, est.bin2val(5, hv[0][0], hv[0][1], nbins[0]),
                                  est.bin2val(q, hv[1][0], hv[1][1], nbins[1]), hv[2][0], hv[2][2],
                                  tab[:, 0], tab[:, 1])
        assert np.array_equal(th, x)
        # numpy's cos may differ from libm's by an ulp; the sides then differ
        # by about 1e-16 relative and the nearest table entry is the same
        np.testing.assert_allclose(d2, den, rtol=1e-12)
        np.testing.assert_allclose(q2, qred, rtol=1e-12)


def test_program_zeta_and_reduced(tmp_path):
    out = tmp_path / "zeta.dat"
    subprocess.run([PROG, "zeta", DEBBIE, "-o", str(out)], check=True)
    body = out.read_text().split("\n")
    want, ngals, nbins, _ = est.zeta(DEBBIE)
    header = open(DEBBIE).read().split("\n")[:7]
    assert [ln.split() for ln in body[:7]] == [ln.split() for ln in header]
    got = np.array([float(v) for ln in body[7:] for v in ln.split()]).reshape(want.shape)
    assert np.array_equal(got, want)
    red = tmp_path / "red.dat"
    subprocess.run([PROG, "reduced", DEBBIE, "5", "0", "3", "--xi", TABLE, "-o", str(red)], check=True)
    rows = [ln.split() for ln in red.read_text().splitlines() if not ln.startswith("#")]
    assert len(rows) == 2 * nbins[2]
    want_rows = est.reduced(DEBBIE, 5, [0, 3], TABLE)
    got_q = np.array([float(r[5]) for r in rows])
    np.testing.assert_allclose(got_q, np.concatenate([r[4] for r in want_rows]), rtol=1e-12)


def test_program_rejects_bad_input(tmp_path):
    r = subprocess.run([PROG, "zeta", str(tmp_path / "DDD_missing.dat")], capture_output=True, text=True)
    assert r.returncode != 0 and "cannot read" in r.stderr
    r = subprocess.run([PROG, "reduced", DEBBIE, "999", "0", "--xi", TABLE], capture_output=True, text=True)
    assert r.returncode != 0 and "out of range" in r.stderr
    with pytest.raises(tp.TpcfError):
        tp.triplet_norm([10, 10, 10], 7)


# ---- reduced amplitude: hand-worked cases -------------------------------------
# Q = zeta / (xi12 xi23 + xi23 xi31 + xi31 xi12) with the sides of the triangle
# (s, q, theta) -> (s, s q, s sqrt(1 + q^2 - 2 q cos theta)) and xi looked up at
# the nearest table abscissa (examples/plot_MICE_3pcf.py:43-51, :9-30, :196-221).
# Each case below has sides and a denominator known exactly.

def _table(xs, ys):
    return np.asarray(xs, dtype=float), np.asarray(ys, dtype=float)


@pytest.mark.parametrize("case", [
    # (s, q, theta_lo, theta_width, table xs, table ys, want denominator)
    # 3-4-5 right triangle: theta = pi (0 + 1/2) = pi/2, sides 3, 4, 5
    (3.0, 4.0 / 3.0, 0.0, 1.0, [1, 3, 4, 5, 9], [7, 2, 3, 5, 11], 2 * 3 + 3 * 5 + 5 * 2),
    # equilateral: theta = pi (0 + 1/3) = pi/3, q = 1, every side 2
    (2.0, 1.0, 0.0, 2.0 / 3.0, [1, 2, 3], [5, 4, 9], 3 * 4 * 4),
    # collinear: theta = pi (0 + 1) = pi, sides 1, 2, 3
    (1.0, 2.0, 0.0, 2.0, [1, 2, 3, 4], [1, 10, 100, 1000], 1 * 10 + 10 * 100 + 100 * 1),
    # a side halfway between two abscissae takes the first (sides 1, 2.5, ~1.5)
    (1.0, 2.5, 0.0, 2.0 / 3.0, [1, 2, 3], [2, 3, 7], None),
])
def test_reduced_q_hand_worked_triangles(case):
    s, q, lo, width, xs, ys, want = case
    xs, ys = _table(xs, ys)
    zeta_row = np.array([6.2])
    th, den, qred = tp.reduced_q(zeta_row, s, q, lo, width, xs, ys)
    if want is None:
        # sides 1, 2.5, sqrt(1 + 6.25 - 2.5) = 2.179...: xi 2, 3 (2.5 is as far
        # from 2 as from 3: the first, 2), 3 -> 2*3 + 3*3 + 3*2 = 21
        want = 21.0
    assert th[0] == lo
    assert den[0] == want
    assert qred[0] == 6.2 / want
    # the oracle's restatement of the script's steps gives the same
    s12, s23, s31 = est.sqtheta2s1s2s3(s, q, np.pi * (np.array([lo]) + width / 2.0))
    c12, c23, c31 = (est.sxx2cf(v, xs, ys) for v in (np.atleast_1d(s12), s23, s31))
    assert (c12 * c23 + c23 * c31 + c31 * c12)[0] == want


def test_reduced_q_constant_xi(tmp_path):
    """With xi constant the denominator is 3 xi^2 for every triangle, so Q is
    zeta / (3 xi^2) along the whole theta row: checked through the oracle's
    file-level restatement (est.reduced) and the library on the same row."""
    cells = [[12, 0, 5, 8], [30, 6, 20, 0], [95, 30, 10, 40], [114, 57, 20, 80]]
    ddd = four_files(tmp_path, cells)
    table = tmp_path / "xi_const.dat"
    table.write_text("".join(f"{x:f} {0.5:f}\n" for x in np.arange(0.1, 70.0, 0.1)))
    z, _, nbins, hv = est.zeta(ddd)
    for sbin in range(nbins[0]):
        for q, x, zrow, den, qred in est.reduced(ddd, sbin, [0], str(table)):
            assert np.all(den == 0.75)
            np.testing.assert_array_equal(qred, zrow / 0.75)
            tab = np.loadtxt(table)
            th, d2, q2 = tp.reduced_q(zrow, est.bin2val(sbin, hv[0][0], hv[0][1], nbins[0]),
                                      est.bin2val(q, hv[1][0], hv[1][1], nbins[1]), hv[2][0], hv[2][2],
                                      tab[:, 0], tab[:, 1])
            assert np.all(d2 == 0.75)
            np.testing.assert_array_equal(q2, zrow / 0.75)
