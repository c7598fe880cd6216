"""The FP32 pair stage's error bound (the precision form), checked on the CPU.

pair_tiles_kernel (3pcf_amd/csrc/kernels_r123.hip) bins a pair in FP32 from
coordinates shifted to the job's bounding-box centre and rounded to FP32, and
recomputes it in FP64 when the FP32 squared distance lies within its error
bound E (or the relative recheck window) of a squared threshold.  This test
restates that arithmetic in numpy float32 operation by operation (no fused
multiply-add: the kernel is built with -ffp-contract=off) and checks, with the
relative window switched off so that E alone must hold, that every pair the
FP32 path keeps gets the FP64 definition's bin (the clean64 oracle's squared
distance ((dx*dx + dy*dy) + dz*dz) against the squared thresholds).
"""
import numpy as np

from oracle import oracle as orc

F = np.float32
U = F(2.0 ** -24)


def _fp32_path(pa, pb, T64, rel=0.0, base_scale=1.0):
    """(bin, kept) of the FP32 pair stage for point pairs pa[i], pb[i]."""
    n = len(T64) - 1
    allp = np.concatenate([pa, pb])
    lo, hi = allp.min(0), allp.max(0)
    o = 0.5 * (lo + hi)
    r = np.max(np.maximum(hi - o, o - lo))
    rmax = F(r * (1.0 + 1e-6)) + F(1e-30)
    a32 = (pa - o).astype(np.float32)
    b32 = (pb - o).astype(np.float32)
    d = a32 - b32
    dx, dy, dz = d[:, 0], d[:, 1], d[:, 2]
    s = (dx * dx + dy * dy) + dz * dz
    T32 = T64.astype(np.float32)
    m = np.searchsorted(T32, s, side="right") - 1  # largest m with T32[m] <= s
    base = F(2.002) * U * rmax * F(base_scale)
    ax, ay, az = np.abs(dx), np.abs(dy), np.abs(dz)
    ex = base + F(1.002) * U * ax
    ey = base + F(1.002) * U * ay
    ez = base + F(1.002) * U * az
    E = F(1.01) * (ex * (F(2) * ax + ex) + ey * (F(2) * ay + ey) + ez * (F(2) * az + ez)) + F(3.01) * U * s + F(1e-30)
    sd = s.astype(np.float64)
    Ed = E.astype(np.float64) + 1e-15 * sd
    near = np.zeros(len(s), bool)
    mlo = np.clip(m, 0, n)
    Tm = T64[mlo]
    near |= (m >= 0) & (np.abs(sd - Tm) <= np.maximum(Ed, 2.0 * rel * Tm))
    mhi = np.clip(m + 1, 0, n)
    Tn = T64[mhi]
    near |= (m < n) & (np.abs(sd - Tn) <= np.maximum(Ed, 2.0 * rel * Tn))
    b = np.where((m >= 0) & (m < n), m, -1)
    return b, ~near


def _fp64_bin(pa, pb, T64):
    d = pa - pb
    s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    n = len(T64) - 1
    m = np.searchsorted(T64, s, side="right") - 1
    return np.where((m >= 0) & (m < n), m, -1)


def _pairs_on_edges(e, npairs, rng, offset):
    """Pairs whose distance is an edge of the axis (up to FP64 rounding of the
    coordinates), at random directions, around `offset`."""
    pa = offset + rng.random((npairs, 3)) * 50.0
    v = rng.normal(size=(npairs, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    dist = rng.choice(e, npairs) * (1.0 + rng.uniform(-2e-7, 2e-7, npairs))
    return pa, pa + v * dist[:, None]


SPECS = [(20, 1.0, 50.0, "log"), (16, 0.0, 40.0, "lin"), (32, 2.0, 150.0, "log")]


def test_kept_pairs_match_fp64_random_and_on_edges():
    rng = np.random.default_rng(3)
    for spec in SPECS:
        e = orc.edges(orc.axis(*spec))
        T64 = orc.sq_thresholds(e)
        for offset in (0.0, 650.0):
            pa = offset + rng.random((300000, 3)) * 60.0
            pb = offset + rng.random((300000, 3)) * 60.0
            for kind, (a, b) in (("random", (pa, pb)), ("on edges", _pairs_on_edges(e, 300000, rng, offset))):
                got, kept = _fp32_path(a, b, T64)
                want = _fp64_bin(a, b, T64)
                bad = np.nonzero(kept & (got != want))[0]
                assert bad.size == 0, (spec, offset, kind, bad.size)
                if kind == "random":
                    assert kept.mean() > 0.99  # few random pairs fall within E of an edge
                else:
                    assert kept.mean() < 0.5  # pairs at an edge must go to FP64


def test_check_detects_a_bound_without_coordinate_rounding():
    """Planted mistake (in this CPU restatement only): without the
    coordinate-rounding term of E, pairs 650 Mpc/h from the origin at an edge
    are kept with a wrong bin, and the comparison catches it."""
    rng = np.random.default_rng(5)
    e = orc.edges(orc.axis(20, 1.0, 50.0, "log"))
    T64 = orc.sq_thresholds(e)
    a, b = _pairs_on_edges(e, 300000, rng, 650.0)
    want = _fp64_bin(a, b, T64)
    got, kept = _fp32_path(a, b, T64)
    assert not np.any(kept & (got != want))
    got, kept = _fp32_path(a, b, T64, base_scale=0.0)
    assert np.any(kept & (got != want))
