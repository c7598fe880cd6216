"""numpy restatement of the reference's estimator scripts.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
product's estimators (3pcf_amd/csrc/estimators.cpp, the tpcf_estimate
program); never part of the product.

* examples/plot_MICE_3pcf.py: file parsing (:76-89), per-file triplet
  normalizations (:91-107), zeta (:143-156), bin centres (bin2val, :33-40),
  triangle sides (sqtheta2s1s2s3, :43-51), nearest-neighbour 2PCF lookup
  (sxx2cf, :9-30), reduced amplitude (:176-221);
* examples/plot_MICE_2pcf.py: parsing (:36-43), normalizations (:43-50), xi
  (:63-69), the table it writes (:127-131).

The plotting is left out; every number the scripts plot or write is
returned.  Python 2 integer products are exact, then converted to float,
as the scripts' float(...) calls do.
"""
from __future__ import annotations

import numpy as np


def read_tri(path):
    """(ngals[3], nbins[3], histvals[3][3], values reshaped (ns, nq, nt))."""
    vals = np.array(open(path).read().split()).astype("str")
    ngals = vals[0:3].astype("int64")
    nbins = vals[3:6].astype("int64")
    histvals = [vals[6 + i * 3:9 + i * 3].astype("float") for i in range(3)]
    pts = vals[15:].astype("float").reshape((nbins[0], nbins[1], nbins[2]), order="C")
    return ngals, nbins, histvals, pts


def triplet_norm(ngals, term):
    a, b, c = (int(x) for x in ngals)
    if term in (0, 3):
        return float(a * (b - 1) * (c - 2)) / 6.0
    if term == 1:
        return float(a * (b - 1) * c) / 2.0
    return float(a * b * (c - 1)) / 2.0


def zeta(ddd_path):
    """(zeta (ns, nq, nt), ngals, nbins, histvals of the last file read)."""
    names = [ddd_path, ddd_path.replace("DDD", "DDR"), ddd_path.replace("DDD", "DRR"),
             ddd_path.replace("DDD", "RRR")]
    terms = []
    for t, f in enumerate(names):
        ngals, nbins, histvals, pts = read_tri(f)
        terms.append(pts / triplet_norm(ngals, t))
    ddd, ddr, drr, rrr = terms
    with np.errstate(divide="ignore", invalid="ignore"):
        tpcf = (ddd - (3 * ddr) + (3 * drr) - rrr) / rrr
    tpcf[tpcf != tpcf] = 0
    tpcf[tpcf == np.inf] = 0
    return tpcf, ngals, nbins, histvals


def bin2val(ibin, lo, hi, nbins):
    width = (hi - lo) / float(nbins)
    return lo + (width * ibin) + (width / 2.0)


def sqtheta2s1s2s3(s, q, theta):
    return s, s * q, s * np.sqrt(1 + q * q - 2 * q * np.cos(theta))


def sxx2cf(s, xpts, ypts):
    """Array in, array out; scalar in, scalar out (both branches of :9-30)."""
    if isinstance(s, np.ndarray):
        cf = np.zeros(len(s))
        for i, spt in enumerate(s):
            diff = np.abs(spt - xpts)
            m = min(diff)
            cf[i] = ypts[diff.tolist().index(m)]
        return cf
    diff = np.abs(s - xpts)
    m = min(diff)
    return ypts[diff.tolist().index(m)]


def reduced(ddd_path, sbin, qbins, table_path):
    """Rows (q bin, theta/pi values, zeta row, denominator, Q) for s bin sbin."""
    tpcf, _, nbins, histvals = zeta(ddd_path)
    slo, shi = histvals[0][0], histvals[0][1]
    qlo, qhi = histvals[1][0], histvals[1][1]
    thetalo, thetahi, thetawidth = histvals[2]
    vals = np.array(open(table_path).read().split()).astype("float")
    idx = np.arange(0, len(vals), 2)
    x2pcf, y2pcf = vals[idx], vals[idx + 1]
    x = np.arange(thetalo, thetahi, thetawidth)
    sval = bin2val(sbin, slo, shi, nbins[0])
    rows = []
    for q in qbins:
        qval = bin2val(q, qlo, qhi, nbins[1])
        thetaval = np.pi * (x + thetawidth / 2.)
        s12, s23, s31 = sqtheta2s1s2s3(sval, qval, thetaval)
        c12, c23, c31 = sxx2cf(s12, x2pcf, y2pcf), sxx2cf(s23, x2pcf, y2pcf), sxx2cf(s31, x2pcf, y2pcf)
        denominator = c12 * c23 + c23 * c31 + c31 * c12
        with np.errstate(divide="ignore", invalid="ignore"):
            rows.append((q, x, tpcf[sbin][q], denominator, tpcf[sbin][q] / denominator))
    return rows


def read_two(path):
    vals = np.array(open(path).read().split()).astype("str")
    ngals = vals[0:2].astype("int64")
    nbins = int(vals[2])
    histvals = vals[3:6].astype("float")
    return ngals, nbins, histvals, vals[6:].astype("float")


def pair_norm(ngals, term):
    a, b = int(ngals[0]), int(ngals[1])
    return float((a * a) - a) / 2. if term in (0, 2) else float(a * b)


def xi_table(dd_path):
    """Lines the 2PCF script writes: [(x + width, xi)] and the xi array."""
    names = [dd_path, dd_path.replace("DD", "DR"), dd_path.replace("DD", "RR")]
    terms = []
    for t, f in enumerate(names):
        ngals, nbins, histvals, pts = read_two(f)
        terms.append(pts / pair_norm(ngals, t))
    dd, dr, rr = terms
    with np.errstate(divide="ignore", invalid="ignore"):
        tpcf = (dd - (2 * dr) + rr) / rr
    tpcf[tpcf != tpcf] = 0
    tpcf[tpcf == np.inf] = 0
    lo, hi, width = histvals
    x = np.arange(lo, hi, width)
    return ["%f %f\n" % (xpt + width, ypt) for xpt, ypt in zip(x, tpcf)], tpcf
