"""Seeded synthetic catalogs in the shapes of the benchmark workloads.

There is no network and the reference's large catalogs are not shipped
(.MISSING_LARGE_BLOBS), so every workload is generated here from a seed:

* ``uniform_cube(n, size, seed)``            uniform points in [0, size)^3;
* ``neyman_scott(n, nclust, sigma, size, seed)``  clustered data: centres
  uniform in the cube, each point a centre plus N(0, sigma^2) per axis,
  wrapped into the cube;
* ``wedge(n, ra_deg, dec_deg, r_lo, r_hi, seed, clustered=...)``  an RA/Dec
  wedge with comoving distance uniform in [r_lo, r_hi], converted to
  cartesian (x = r cos(dec) cos(ra), y = r cos(dec) sin(ra), z = r sin(dec)),
  the same conversion examples' catalog generator applies
  (sample_data/gen_flat_MICE_data.py:53-99).

All generators use numpy's PCG64 with the given seed, so a seed names a
catalog on any machine.
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def uniform_cube(n: int, size: float = 100.0, seed: int = 0) -> np.ndarray:
    return _rng(seed).random((n, 3)) * size


def neyman_scott(n: int, nclust: int = 200, sigma: float = 2.0, size: float = 200.0, seed: int = 0) -> np.ndarray:
    g = _rng(seed)
    centres = g.random((nclust, 3)) * size
    pick = g.integers(0, nclust, n)
    pts = centres[pick] + g.normal(0.0, sigma, (n, 3))
    return np.mod(pts, size)


def wedge(n: int, ra_deg: float = 10.0, dec_deg: float = 10.0, r_lo: float = 500.0, r_hi: float = 800.0,
          seed: int = 0, clustered: bool = False, nclust: int = 200, sigma: float = 2.0) -> np.ndarray:
    """Points in an RA/Dec wedge; with clustered=True, Neyman-Scott clusters
    whose centres are drawn uniformly in the wedge."""
    g = _rng(seed)

    def draw(m):
        ra = np.deg2rad(g.random(m) * ra_deg)
        sdec = g.random(m) * np.sin(np.deg2rad(dec_deg))
        dec = np.arcsin(sdec)
        r = r_lo + (r_hi - r_lo) * g.random(m)
        return np.stack([r * np.cos(dec) * np.cos(ra), r * np.cos(dec) * np.sin(ra), r * np.sin(dec)], axis=1)

    if not clustered:
        return draw(n)
    centres = draw(nclust)
    pick = g.integers(0, nclust, n)
    return centres[pick] + g.normal(0.0, sigma, (n, 3))


def angular_voxels(nside: int = 8, total: int = 150000, ra_deg: float = 20.0, dec_deg: float = 20.0,
                   r_lo: float = 500.0, r_hi: float = 800.0, lo: int = 500, hi: int = 8000, seed: int = 0,
                   nclust_per_voxel: int = 8, sigma: float = 2.0):
    """Survey split into nside x nside RA/Dec cells of a wedge: per cell a
    clustered data catalog and a uniform random catalog of the same uneven
    (lognormal, clipped to [lo, hi], summing to total) size.  Returns a list
    of (data, randoms) (n, 3) arrays."""
    sizes = lognormal_voxel_sizes(nside * nside, total, lo, hi, seed=seed)
    out = []
    for v, n in enumerate(sizes):
        a, b = divmod(v, nside)
        g = _rng(seed * 1000 + v + 1)
        ra0, dec0 = ra_deg * a / nside, dec_deg * b / nside
        dra, ddec = ra_deg / nside, dec_deg / nside

        def draw(m):
            ra = np.deg2rad(ra0 + g.random(m) * dra)
            s0, s1 = np.sin(np.deg2rad(dec0)), np.sin(np.deg2rad(dec0 + ddec))
            dec = np.arcsin(s0 + (s1 - s0) * g.random(m))
            r = r_lo + (r_hi - r_lo) * g.random(m)
            return np.stack([r * np.cos(dec) * np.cos(ra), r * np.cos(dec) * np.sin(ra), r * np.sin(dec)], axis=1)

        centres = draw(nclust_per_voxel)
        data = centres[g.integers(0, nclust_per_voxel, int(n))] + g.normal(0.0, sigma, (int(n), 3))
        out.append((data, draw(int(n))))
    return out


def lognormal_voxel_sizes(nvox: int, total: int, lo: int = 500, hi: int = 8000, seed: int = 0) -> np.ndarray:
    """Uneven voxel occupancies: lognormal, clipped to [lo, hi], rescaled to
    sum to total (clipping again after rescaling, remainder spread)."""
    g = _rng(seed)
    raw = g.lognormal(mean=0.0, sigma=0.8, size=nvox)
    sizes = np.clip(raw / raw.sum() * total, lo, hi)
    for _ in range(50):
        sizes = np.clip(sizes * (total / sizes.sum()), lo, hi)
    sizes = np.floor(sizes).astype(np.int64)
    i = 0
    while sizes.sum() < total:
        if sizes[i % nvox] < hi:
            sizes[i % nvox] += 1
        i += 1
    while sizes.sum() > total:
        if sizes[i % nvox] > lo:
            sizes[i % nvox] -= 1
        i += 1
    return sizes
