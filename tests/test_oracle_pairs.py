"""The oracle's pair counts (src/2pcf_C_version.cc): the FP32 restatement
pinned byte for byte to the reference program's outputs (compiled from
source, committed under tests/golden/ref_outputs/), and the clean64 pair
counts pinned by closed forms, the union identity, a numpy brute force and
exact lattice distances."""
import itertools
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
INP = os.path.join(GOLD, "inputs")
D, R = "d200.dat", "r180.dat"
D1K, R1K = "MICE_5degx5deg_1k_randomized.dat", "flat_MICE_5degx5deg_1k.dat"

CASES = {
    "two_DD.dat": [D, D], "two_DR.dat": [D, R], "two_RR.dat": [R, R],
    "two_DD1k.dat": [D1K, D1K], "two_DR1k.dat": [D1K, R1K], "two_RR1k.dat": [R1K, R1K],
    "two_DD_vox2_001000.dat": [D, D, "-X", "2", "-x", "001000000"],
    "two_DD_vox2_000001.dat": [D, D, "-X", "2", "-x", "000001000"],
}


@pytest.fixture(scope="module", autouse=True)
def _built():
    orc.build()


@pytest.mark.parametrize("name", sorted(CASES))
def test_ref_two_matches_reference_outputs(tmp_path, name):
    out = tmp_path / "o.dat"
    subprocess.run([orc.CLI, "ref-two", *CASES[name], "-o", str(out), "--threads", "4"], cwd=INP, check=True,
                   stdout=subprocess.DEVNULL)
    assert out.read_bytes() == open(os.path.join(GOLD, "ref_outputs", name), "rb").read()


def _read7(path):
    a = np.loadtxt(path)
    return a[:, 4:7]


def _brute(segs, terms, ax):
    e = orc.edges(ax)
    T = orc.sq_thresholds(e)
    out = []
    for a, b in terms:
        h = np.zeros(ax.n, np.uint64)
        A, B = segs[a], segs[b]
        pairs = itertools.combinations(range(len(A)), 2) if a == b else itertools.product(range(len(A)), range(len(B)))
        for i, j in pairs:
            dx, dy, dz = A[i] - B[j]
            s = (dx * dx + dy * dy) + dz * dz
            m = int(np.searchsorted(T, s, side="right")) - 1
            if 0 <= m < ax.n:
                h[m] += 1
        out.append(h)
    return np.stack(out)


def test_clean_pairs_brute_force_and_lattice():
    rng = np.random.default_rng(1)
    a = rng.random((70, 3)) * 20.0
    b = rng.random((55, 3)) * 20.0 + 3.0
    for ax in (orc.axis(37, 0.0, 30.0), orc.axis(23, 0.5, 25.0, "log"), orc.axis(700, 0.0, 70.0)):
        got = orc.clean_pairs([a, b], [(0, 0), (0, 1), (1, 1)], ax)
        assert np.array_equal(got, _brute([a, b], [(0, 0), (0, 1), (1, 1)], ax))
    # integer lattice: distances sqrt(k) land exactly on the edges 1, 2, 3
    g = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"), -1).reshape(-1, 3) * 1.0
    ax = orc.axis(6, 0.0, 3.0)
    assert np.array_equal(orc.clean_pairs([g], [(0, 0)], ax), _brute([g], [(0, 0)], ax))
    # bin [1.0, 1.5) holds the unit neighbours (distance exactly 1: 3 axes x
    # 3 x 16) and the face diagonals (sqrt 2: 3 planes x 3 x 3 x 2 x 4)
    h = orc.clean_pairs([g], [(0, 0)], ax)[0]
    assert h[2] == 3 * 3 * 16 + 3 * 3 * 3 * 2 * 4
    # bin [0.5, 1.0) is empty: no distance below 1 on the lattice
    assert h[1] == 0


def test_clean_pairs_closed_forms_and_union():
    rng = np.random.default_rng(2)
    a = rng.random((300, 3)) * 10.0
    b = rng.random((260, 3)) * 10.0
    ax = orc.axis(50, 0.0, 20.0)  # every pair is in range (diagonal 17.3)
    h = orc.clean_pairs([a, b], [(0, 0), (0, 1), (1, 1)], ax)
    assert int(h[0].sum()) == math.comb(300, 2)
    assert int(h[1].sum()) == 300 * 260
    assert int(h[2].sum()) == math.comb(260, 2)
    u = orc.clean_pairs([np.vstack([a, b])], [(0, 0)], ax)[0]
    assert np.array_equal(h.sum(0), u)
    # permutation invariance
    p = rng.permutation(300)
    assert np.array_equal(orc.clean_pairs([a[p], b], [(0, 1)], ax)[0], h[1])


def test_ref_two_self_pairs_and_fp32_binning():
    """The restatement counts every point's self pair in bin 0 for the same
    file; otherwise its pairs are the clean unique pairs up to FP32 binning."""
    d = _read7(os.path.join(INP, D))
    h = orc.ref_two([d, d], 0).astype(np.int64)
    c = orc.clean_pairs([d], [(0, 0)], orc.axis(700, 0.0, 70.0))[0].astype(np.int64)
    assert int(h.sum()) - len(d) == int(c.sum()) or abs(int(h.sum()) - len(d) - int(c.sum())) <= 2
    assert h[0] >= len(d)
