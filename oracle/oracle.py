"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline,
never as a path that produces the product's results.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
CLI = os.path.join(HERE, "orc_cli")
REF_DIR = os.path.join(HERE, "_ref")

R123, SQT, JBJ = 0, 1, 2
LIN, LOG = 0, 1


class OrcAxis(C.Structure):
    _fields_ = [("n", C.c_int), ("lo", C.c_double), ("hi", C.c_double), ("scale", C.c_int)]


class OrcRefAxis(C.Structure):
    _fields_ = [("n", C.c_int), ("lo", C.c_double), ("hi", C.c_double)]


class OrcJob(C.Structure):
    _fields_ = [
        ("nseg", C.c_int),
        ("seg", C.c_void_p * 3),
        ("nseg_pts", C.c_int64 * 3),
        ("nterms", C.c_int),
        ("term", (C.c_int * 3) * 4),
        ("param", C.c_int),
        ("ax", OrcAxis * 3),
    ]


LIB_O1 = os.path.join(HERE, "liboracle_o1.so")

_lib = None
_lib_o1 = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(o1: bool = False) -> C.CDLL:
    """The oracle library (-O2), or with o1=True the same source built at the
    reference's -O1 (for timing the reference C path's loop)."""
    global _lib, _lib_o1
    if o1:
        if _lib_o1 is None:
            if not os.path.exists(LIB_O1):
                build()
            _lib_o1 = _declare(C.CDLL(LIB_O1))
        return _lib_o1
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _declare(C.CDLL(LIB))
    return _lib


def _declare(lib: C.CDLL) -> C.CDLL:
    """Declare the argument and result types of the oracle functions."""
    lib.orc_clean_count.argtypes = [C.POINTER(OrcJob), C.c_void_p, C.c_int]
    lib.orc_clean_count.restype = C.c_int
    lib.orc_hist_cells.argtypes = [C.POINTER(OrcJob)]
    lib.orc_hist_cells.restype = C.c_int64
    lib.orc_edges.argtypes = [C.POINTER(OrcAxis), C.c_void_p]
    lib.orc_edges.restype = C.c_int
    lib.orc_sq_thresholds.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    lib.orc_sq_thresholds.restype = None
    lib.orc_bin_sq.argtypes = [C.c_double, C.c_void_p, C.c_int]
    lib.orc_bin_sq.restype = C.c_int
    lib.orc_ref_tri.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(OrcRefAxis),
                                C.c_void_p, C.POINTER(C.c_uint64), C.c_int]
    lib.orc_ref_tri.restype = C.c_int
    lib.orc_ref_plain.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_void_p,
                                  C.POINTER(C.c_uint64), C.c_int]
    lib.orc_ref_plain.restype = C.c_int
    lib.orc_ref_which.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
    lib.orc_ref_which.restype = C.c_int
    lib.orc_ref_vox2gal.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.orc_ref_vox2gal.restype = C.c_int
    lib.orc_clean_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(OrcAxis),
                                    C.c_void_p, C.c_int]
    lib.orc_clean_pairs.restype = C.c_int
    lib.orc_ref_two.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    lib.orc_ref_two.restype = C.c_int
    return lib


def axis(n, lo, hi, scale="lin") -> OrcAxis:
    return OrcAxis(int(n), float(lo), float(hi), LOG if scale in ("log", LOG) else LIN)


def edges(ax: OrcAxis) -> np.ndarray:
    e = np.zeros(ax.n + 1)
    if load().orc_edges(C.byref(ax), e.ctypes.data):
        raise ValueError("bad axis")
    return e


def sq_thresholds(e: np.ndarray) -> np.ndarray:
    e = np.ascontiguousarray(e, dtype=np.float64)
    T = np.zeros_like(e)
    load().orc_sq_thresholds(e.ctypes.data, len(e) - 1, T.ctypes.data)
    return T


def clean_count(segments: Sequence[np.ndarray], terms, param: int, axes: Sequence[OrcAxis],
                nthreads: int = 8) -> np.ndarray:
    """clean64 histograms, shape (nterms, cells), uint64."""
    lib = load()
    job = OrcJob()
    job.nseg = len(segments)
    keep = []
    for s, pts in enumerate(segments):
        soa = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).T)
        keep.append(soa)
        job.seg[s] = soa.ctypes.data
        job.nseg_pts[s] = soa.shape[1]
    terms = list(terms)
    job.nterms = len(terms)
    for t, tr in enumerate(terms):
        for k in range(3):
            job.term[t][k] = tr[k]
    job.param = param
    for a, ax in enumerate(axes):
        job.ax[a] = ax
    cells = lib.orc_hist_cells(C.byref(job))
    out = np.zeros((len(terms), cells), dtype=np.uint64)
    if lib.orc_clean_count(C.byref(job), out.ctypes.data, nthreads):
        raise ValueError("bad oracle job")
    return out


def ref_tri(cats: Sequence[np.ndarray], which: int, variant: int, axes, kmin_legacy: int = 0,
            nthreads: int = 8, o1: bool = False):
    """refcompat32 (s,q,theta) histogram (uint32, wrapping) and numcalc."""
    lib = load(o1)
    soas = [np.ascontiguousarray(np.asarray(c, dtype=np.float32).T) for c in cats]
    ptrs = (C.c_void_p * 3)(*[s.ctypes.data for s in soas])
    ns = (C.c_int * 3)(*[s.shape[1] for s in soas])
    ax = (OrcRefAxis * 3)(*[OrcRefAxis(int(n), float(lo), float(hi)) for n, lo, hi in axes])
    cells = axes[0][0] * axes[1][0] * axes[2][0]
    h = np.zeros(cells, dtype=np.uint32)
    numcalc = C.c_uint64()
    lib.orc_ref_tri(ptrs, ns, which, variant, kmin_legacy, ax, h.ctypes.data, C.byref(numcalc), nthreads)
    return h, numcalc.value


def clean_pairs(segments: Sequence[np.ndarray], terms, ax: OrcAxis, nthreads: int = 8) -> np.ndarray:
    """clean64 pair-count histograms, shape (nterms, ax.n), uint64; a term
    (a, b) counts the unique pairs of segment a (a == b) or every pair of
    segments a and b."""
    lib = load()
    soas = [np.ascontiguousarray(np.asarray(p, dtype=np.float64).T) for p in segments]
    ptrs = (C.c_void_p * 2)(*([s.ctypes.data for s in soas] + [None] * (2 - len(soas))))
    ns = (C.c_int64 * 2)(*([s.shape[1] for s in soas] + [0] * (2 - len(soas))))
    terms = list(terms)
    tarr = (C.c_int * (2 * len(terms)))(*[v for t in terms for v in t[:2]])
    out = np.zeros((len(terms), ax.n), dtype=np.uint64)
    if lib.orc_clean_pairs(ptrs, ns, len(segments), tarr, len(terms), C.byref(ax), out.ctypes.data, nthreads):
        raise ValueError("bad oracle pair job")
    return out


def ref_two(cats: Sequence[np.ndarray], which: int, nthreads: int = 8) -> np.ndarray:
    """refcompat32 pair counts of the reference's 2pcf program (700 uint32)."""
    lib = load()
    soas = [np.ascontiguousarray(np.asarray(c, dtype=np.float32).T) for c in cats]
    ptrs = (C.c_void_p * 2)(*[s.ctypes.data for s in soas])
    ns = (C.c_int * 2)(*[s.shape[1] for s in soas])
    h = np.zeros(700, dtype=np.uint32)
    lib.orc_ref_two(ptrs, ns, which, h.ctypes.data, nthreads)
    return h
