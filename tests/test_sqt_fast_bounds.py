"""The FP32 error bounds of the FAST (s,q,theta) and JBJ entries, checked on
the CPU.

The device kernel (3pcf_amd/csrc/kernels_sqt_fast.hip, entry32 / bracket_asc /
bracket_desc / sqt_entries) decides an entry in FP32 only when every value is
farther from every deciding edge than a stated error bound; otherwise it
recomputes the triangle in FP64.  This test restates that FP32 arithmetic
operation by operation in numpy float32 (no fused multiply-add, as the kernel
is built with -ffp-contract=off; each reciprocal is either FP32 neighbour of
the exact value, the 1-ulp accuracy of v_rcp_f32, chosen at random) and
checks, over random and deliberately near-edge triangles, that every entry
the FP32 path claims to be certain of equals the FP64 definition (the clean64
oracle's arithmetic: src/3pcf_C_version_triangle_parametrization.cc:188-236
restated with sorted squared sides, q = l2/l1 and the cosine rule; for JBJ
q = l1/l2 and (l3 - l2)/l1).  A bound derived too tight fails here.
"""
import math

import numpy as np

from oracle import oracle as orc

U = np.float32(2.0 ** -24)
F = np.float32


def _faithful_rcp(x, rng):
    """Either FP32 neighbour of the exact reciprocal of float32 x."""
    exact = 1.0 / x.astype(np.float64)
    lo = exact.astype(np.float32)
    hi = np.where(lo.astype(np.float64) < exact, np.nextafter(lo, np.float32(np.inf)),
                  np.where(lo.astype(np.float64) > exact, np.nextafter(lo, np.float32(-np.inf)), lo))
    pick = rng.random(x.shape) < 0.5
    return np.where(pick, lo, hi).astype(np.float32)


def _bin_asc(v, e):
    """Bin m with e[m] <= v < e[m+1], -1 outside (FP64 definition)."""
    m = np.searchsorted(e, v, side="right") - 1
    return np.where((v >= e[0]) & (v < e[-1]), m, -1)


def _bin_cos(c, C):
    """Bin m with C[m+1] < c <= C[m] for the descending table, -1 outside."""
    n = len(C) - 1
    # count of table entries >= c, minus one
    m = np.sum(C[None, :] >= c[:, None], axis=1) - 1
    return np.where((c <= C[0]) & (c > C[n]), m, -1)


def _bracket_asc32(v, d, e32, m, scale=1):
    """(bin, certain) of the kernel's bracket test with the exact FP32 bin as
    the hint; scale multiplies the whole margin (1 in the kernel)."""
    n = len(e32) - 1
    certain = np.ones(v.shape, bool)
    below = m < 0
    above = m >= n
    mid = ~below & ~above
    e0, en = e32[0], e32[n]
    g = F(scale)
    certain &= ~below | ((e0 - v) > g * (d + U * np.abs(e0)))
    certain &= ~above | ((v - en) > g * (d + U * np.abs(en)))
    mm = np.clip(m, 0, n - 1)
    a, b = e32[mm], e32[mm + 1]
    certain &= ~mid | (((v - a) > g * (d + U * np.abs(a))) & ((b - v) > g * (d + U * np.abs(b))))
    return np.where(mid, m, -1), certain


def _bracket_desc32(c, d, C32, m):
    n = len(C32) - 1
    certain = np.ones(c.shape, bool)
    below = m < 0
    above = m >= n
    mid = ~below & ~above
    certain &= ~below | ((c - C32[0]) > d)
    certain &= ~above | ((C32[n] - c) > d)
    mm = np.clip(m, 0, n - 1)
    certain &= ~mid | (((C32[mm] - c) > d) & ((c - C32[mm + 1]) > d))
    return np.where(mid, m, -1), certain


def _fp32_hint_asc(v, e32):
    """Largest m in [-1, n] with e32[m] <= v: the bin a perfect hint gives."""
    return np.searchsorted(e32, v, side="right") - 1


def _fp32_hint_desc(c, C32):
    return np.sum(C32[None, :] >= c[:, None], axis=1) - 1


def _entries(points, axes, rng, qscale=1, cscale=1):
    """FP64 cells and FP32-path (cells, certain) of every triangle (rows of
    points: (T, 3, 3)).  qscale and cscale multiply the kernel's q and cosine
    margins (1: the kernel's arithmetic)."""
    ea = [orc.edges(orc.axis(*a)) for a in axes]
    T = orc.sq_thresholds(ea[0])
    Eq = ea[1]
    Et = np.array([math.cos(math.pi * t) for t in ea[2]])
    nq, nt = axes[1][0], axes[2][0]
    P = points
    def sq(a, b):
        d = P[:, a] - P[:, b]
        return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    s = np.stack([sq(0, 1), sq(0, 2), sq(1, 2)], axis=1)
    sbin = _bin_asc(s, T)  # exact s bins (squared thresholds)
    # FP64 definition: sorted squared sides.
    order = np.argsort(s, axis=1, kind="stable")
    s64 = np.take_along_axis(s, order, 1)
    b64 = np.take_along_axis(sbin, order, 1)
    L = np.sqrt(s64)
    def cell(sb, qb, tb):
        ok = (sb >= 0) & (qb >= 0) & (tb >= 0)
        return np.where(ok, (sb * nq + qb) * nt + tb, -1)
    c1, c2, c3 = s64[:, 0], s64[:, 1], s64[:, 2]
    ref = np.stack([
        cell(b64[:, 0], _bin_asc(L[:, 1] / L[:, 0], Eq), _bin_cos(((c1 + c2) - c3) / ((2.0 * L[:, 0]) * L[:, 1]), Et)),
        cell(b64[:, 0], _bin_asc(L[:, 2] / L[:, 0], Eq), _bin_cos(((c1 + c3) - c2) / ((2.0 * L[:, 0]) * L[:, 2]), Et)),
        cell(b64[:, 1], _bin_asc(L[:, 2] / L[:, 1], Eq), _bin_cos(((c2 + c3) - c1) / ((2.0 * L[:, 1]) * L[:, 2]), Et)),
    ], axis=1)
    # FP32 path: lengths rounded from FP64, sorted in FP32 carrying the codes.
    l32 = np.sqrt(s).astype(np.float32)
    o32 = np.argsort(l32, axis=1, kind="stable")
    l = np.take_along_axis(l32, o32, 1)
    k = np.take_along_axis(sbin, o32, 1)
    l1, l2, l3 = l[:, 0], l[:, 1], l[:, 2]
    certain = ((l2 - l1) > F(4) * U * l2) & ((l3 - l2) > F(4) * U * l3) & (l1 > F(1e-18)) & (l3 < F(1e18))
    fQ = Eq.astype(np.float32)
    fA = Et.astype(np.float32)
    s1, s2, s3 = l1 * l1, l2 * l2, l3 * l3
    ssum = (s1 + s2) + s3
    def entry32(sb, q, num, rden):
        mq, okq = _bracket_asc32(q, F(6) * U * q, fQ, _fp32_hint_asc(q, fQ), qscale)
        c = num * rden
        dc = (F(5) * U * ssum * rden + F(10) * U * np.abs(c) + F(2) * U) * F(cscale)
        mt, okt = _bracket_desc32(c, dc, fA, _fp32_hint_desc(c, fA))
        return cell(sb, mq, mt), okq & okt
    # three reciprocals; 1/(2ab) = (0.5 * (1/a)) * (1/b)
    r1, r2, r3 = _faithful_rcp(l1, rng), _faithful_rcp(l2, rng), _faithful_rcp(l3, rng)
    h1, h2 = F(0.5) * r1, F(0.5) * r2
    e0, ok0 = entry32(k[:, 0], l2 * r1, (s1 + s2) - s3, h1 * r2)
    e1, ok1 = entry32(k[:, 0], l3 * r1, (s1 + s3) - s2, h1 * r3)
    e2, ok2 = entry32(k[:, 1], l3 * r2, (s2 + s3) - s1, h2 * r3)
    # Entries without an s bin are not computed (cell -1), their values do
    # not make the triangle unsure.
    has0 = k[:, 0] >= 0
    has1 = k[:, 1] >= 0
    fast = np.stack([np.where(has0, e0, -1), np.where(has0, e1, -1), np.where(has1, e2, -1)], axis=1)
    certain &= (~has0 | (ok0 & ok1)) & (~has1 | ok2)
    ref_no_s = np.where(b64[:, 0:1] >= 0, ref[:, 0:2], -1)
    ref = np.concatenate([ref_no_s, np.where(b64[:, 1] >= 0, ref[:, 2], -1)[:, None]], axis=1)
    return ref, fast, certain


def _mismatches(points, axes, rng, **margins):
    ref, fast, certain = _entries(points, axes, rng, **margins)
    a = np.sort(ref[certain], axis=1)
    b = np.sort(fast[certain], axis=1)
    return np.nonzero(np.any(a != b, axis=1))[0], a, b, certain


def _check(points, axes, rng):
    bad, a, b, certain = _mismatches(points, axes, rng)
    assert bad.size == 0, f"{bad.size} certain triangles disagree, first: ref {a[bad[0]]} fast {b[bad[0]]}"
    return certain.mean()


AXES = [
    [(10, 2.0, 20.0, "log"), (8, 1.0, 3.0, "lin"), (16, 0.0, 1.0, "lin")],
    [(50, 2.0, 12.0, "lin"), (16, 0.9, 4.1, "lin"), (32, 0.0, 1.0, "lin")],
    [(12, 1.0, 15.0, "log"), (6, 1.0, 4.0, "log"), (9, 0.05, 0.95, "log")],
]


def test_random_triangles_certain_entries_match_fp64():
    rng = np.random.default_rng(7)
    for axes in AXES:
        for offset in (0.0, 600.0):
            P = rng.random((200000, 3, 3)) * 25.0 + offset
            frac = _check(P, axes, rng)
            assert frac > 0.97  # near-ties are rare among random triangles


def _near_edge_triangles(axes, n, rng, offset):
    """Triangles whose q (or theta) lies within about 1e-7 (relative) of an
    edge of the axis: shortest side l1 at A-B, second side from A at angle
    theta."""
    ea = [orc.edges(orc.axis(*a)) for a in axes]
    l1 = rng.uniform(2.5, 8.0, n)
    q = rng.choice(ea[1][1:-1], n) * (1.0 + rng.uniform(-3e-7, 3e-7, n))
    t = rng.uniform(0.2, 0.8, n)
    pick = rng.random(n) < 0.5
    t = np.where(pick, rng.choice(ea[2][1:-1], n) * (1.0 + rng.uniform(-3e-7, 3e-7, n)), t)
    q = np.where(pick, rng.uniform(1.05, 2.9, n), q)
    l2 = q * l1
    th = np.pi * t
    A = np.zeros((n, 3))
    B = np.stack([l1, np.zeros(n), np.zeros(n)], axis=1)
    C = np.stack([l2 * np.cos(th), l2 * np.sin(th), np.zeros(n)], axis=1)
    # random rotation about z and a translation so coordinates round in FP64
    phi = rng.uniform(0, 2 * np.pi, n)
    R = np.stack([np.stack([np.cos(phi), -np.sin(phi), np.zeros(n)], 1),
                  np.stack([np.sin(phi), np.cos(phi), np.zeros(n)], 1),
                  np.stack([np.zeros(n), np.zeros(n), np.ones(n)], 1)], 1)
    pts = np.stack([A, B, C], axis=1) @ np.transpose(R, (0, 2, 1))
    return pts + offset + rng.random((n, 1, 3))


def test_near_edge_triangles_certain_entries_match_fp64():
    rng = np.random.default_rng(11)
    for axes in AXES[:2]:
        for offset in (0.0, 600.0):
            P = _near_edge_triangles(axes, 200000, rng, offset)
            frac = _check(P, axes, rng)
            # many of these are within the bounds of an edge: the FP32 path
            # must give them up rather than guess
            assert frac < 0.995


def test_check_detects_a_bound_too_tight():
    """Planted mistakes (in this CPU restatement only): without the q margin
    or without the cosine margin, near-edge triangles are decided wrongly and
    the comparison catches it."""
    rng = np.random.default_rng(13)
    axes = AXES[0]
    P = _near_edge_triangles(axes, 100000, rng, 600.0)
    assert _mismatches(P, axes, rng)[0].size == 0
    assert _mismatches(P, axes, rng, qscale=0)[0].size > 0
    assert _mismatches(P, axes, rng, cscale=0)[0].size > 0


def _jbj_entries(points, axes, rng, qscale=1, vscale=1):
    """FP64 cell and FP32-path (cell, certain) of the JBJ entry of every
    triangle: the shortest side's s bin, q = l1/l2 and v = (l3 - l2)/l1."""
    ea = [orc.edges(orc.axis(*a)) for a in axes]
    T = orc.sq_thresholds(ea[0])
    Eq, Ev = ea[1], ea[2]
    nq, nt = axes[1][0], axes[2][0]
    P = points
    def sq(a, b):
        d = P[:, a] - P[:, b]
        return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    s = np.stack([sq(0, 1), sq(0, 2), sq(1, 2)], axis=1)
    sbin = _bin_asc(s, T)
    s64 = np.sort(s, axis=1)
    L = np.sqrt(s64)
    b1 = _bin_asc(s64[:, 0], T)
    iq = _bin_asc(L[:, 0] / L[:, 1], Eq)
    iv = _bin_asc((L[:, 2] - L[:, 1]) / L[:, 0], Ev)
    ok = (b1 >= 0) & (iq >= 0) & (iv >= 0)
    ref = np.where(ok, (b1 * nq + iq) * nt + iv, -1)
    l32 = np.sqrt(s).astype(np.float32)
    o32 = np.argsort(l32, axis=1, kind="stable")
    l = np.take_along_axis(l32, o32, 1)
    k1 = np.take_along_axis(sbin, o32, 1)[:, 0]
    l1, l2, l3 = l[:, 0], l[:, 1], l[:, 2]
    certain = ((l2 - l1) > F(4) * U * l2) & ((l3 - l2) > F(4) * U * l3) & (l1 > F(1e-18)) & (l3 < F(1e18))
    fQ, fV = Eq.astype(np.float32), Ev.astype(np.float32)
    r1 = _faithful_rcp(l1, rng)
    q = l1 * _faithful_rcp(l2, rng)
    mq, okq = _bracket_asc32(q, F(6) * U * q, fQ, _fp32_hint_asc(q, fQ), qscale)
    v = (l3 - l2) * r1
    dv = F(2) * U * (l2 + l3) * r1 + F(6) * U * np.abs(v)
    mv, okv = _bracket_asc32(v, dv, fV, _fp32_hint_asc(v, fV), vscale)
    has = k1 >= 0
    fast = np.where(has & (mq >= 0) & (mv >= 0), (k1 * nq + mq) * nt + mv, -1)
    certain &= ~has | (okq & okv)
    return ref, fast, certain


JBJ_AXES = [
    [(20, 1.0, 30.0, "log"), (10, 0.0, 1.0, "lin"), (12, 0.0, 1.2, "lin")],
    [(12, 1.0, 15.0, "log"), (8, 0.3, 1.0, "lin"), (10, -0.2, 1.5, "lin")],
    [(10, 2.0, 12.0, "lin"), (6, 0.2, 1.0, "log"), (7, 0.01, 1.0, "log")],
]


def _jbj_near_edge(axes, n, rng, offset):
    """Triangles with q = l1/l2 or v = (l3 - l2)/l1 within about 1e-7
    (relative) of an edge, built from their three side lengths."""
    ea = [orc.edges(orc.axis(*a)) for a in axes]
    l1 = rng.uniform(2.5, 9.0, n)
    qe = ea[1][(ea[1] > 0.05) & (ea[1] < 0.999)]
    ve = ea[2][(ea[2] > 0.01) & (ea[2] < 0.95)]
    pick = rng.random(n) < 0.5
    q = np.where(pick, rng.choice(qe, n) * (1.0 + rng.uniform(-3e-7, 3e-7, n)), rng.uniform(0.35, 0.95, n))
    v = np.where(pick, rng.uniform(0.05, 0.9, n), rng.choice(ve, n) * (1.0 + rng.uniform(-3e-7, 3e-7, n)))
    l2 = l1 / q
    l3 = l2 + v * l1
    x = (l1 * l1 + l2 * l2 - l3 * l3) / (2.0 * l1)
    y = np.sqrt(np.maximum(l2 * l2 - x * x, 0.0))
    pts = np.stack([np.zeros((n, 3)), np.stack([l1, np.zeros(n), np.zeros(n)], 1),
                    np.stack([x, y, np.zeros(n)], 1)], axis=1)
    phi = rng.uniform(0, 2 * np.pi, n)
    R = np.stack([np.stack([np.cos(phi), -np.sin(phi), np.zeros(n)], 1),
                  np.stack([np.sin(phi), np.cos(phi), np.zeros(n)], 1),
                  np.stack([np.zeros(n), np.zeros(n), np.ones(n)], 1)], 1)
    return pts @ np.transpose(R, (0, 2, 1)) + offset + rng.random((n, 1, 3))


def test_jbj_certain_entries_match_fp64():
    rng = np.random.default_rng(17)
    for axes in JBJ_AXES:
        for offset in (0.0, 600.0):
            for P in (rng.random((200000, 3, 3)) * 25.0 + offset, _jbj_near_edge(axes, 200000, rng, offset)):
                ref, fast, certain = _jbj_entries(P, axes, rng)
                bad = np.nonzero(certain & (ref != fast))[0]
                assert bad.size == 0, (axes, offset, bad.size)


def test_jbj_check_detects_a_bound_too_tight():
    rng = np.random.default_rng(19)
    axes = JBJ_AXES[1]
    P = _jbj_near_edge(axes, 200000, rng, 600.0)
    for margins in ({"qscale": 0}, {"vscale": 0}):
        ref, fast, certain = _jbj_entries(P, axes, rng, **margins)
        assert np.any(certain & (ref != fast)), margins
