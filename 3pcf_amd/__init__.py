"""Host-side Python binding of libtpcf, the MI355X triplet counter.

The package directory is ``3pcf_amd`` (not a Python identifier); import it
with ``importlib.import_module("3pcf_amd")``.

The library is a C ABI (``include/tpcf.h``); this module wraps it with
ctypes.  Nothing here computes histograms: every count comes from the HIP
kernels in ``libtpcf.so``.  If the library is missing or no gfx950 device is
present, the calls raise instead of falling back to anything else.

Reference interface being replaced: the command-line programs
``src/3pcf_C_version_triangle_parametrization.cc`` (main at :245-571) and
``src/3pcf_C_version.cc`` (main at :147-466): three catalogs in, one histogram
out; and ``src/2pcf_C_version.cc`` (main at :124-411) for pair counts
(:meth:`Context.count_pairs`).  :func:`count_files` mirrors that call; :class:`Context` exposes the
library's catalogs, jobs, batches and the serving queue.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TPCF_LIBRARY names another build of the library (the host-sanitized one of
# the Makefile's "san" target, for the CPU tests).
LIB_PATH = os.environ.get("TPCF_LIBRARY") or os.path.join(HERE, "libtpcf.so")
BIN_DIR = os.path.join(HERE, "bin")

R123, SQT, JBJ = 0, 1, 2
LIN, LOG = 0, 1
MODE_FAST, MODE_PLAIN, MODE_FAST64, MODE_CELLS = 0, 1, 2, 3
PARAMS = {"r123": R123, "sqt": SQT, "jbj": JBJ}
MODES = {"fast": MODE_FAST, "plain": MODE_PLAIN, "fast64": MODE_FAST64, "cells": MODE_CELLS}

# DDD, DDR, DRR, RRR over segment 0 = data, segment 1 = randoms.
FOUR_TERMS = ((0, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1))
TERM_NAMES = ("DDD", "DDR", "DRR", "RRR")
# DD, DR, RR pair terms (the third label of a pair term is ignored).
PAIR_TERMS = ((0, 0), (0, 1), (1, 1))
PAIR_TERM_NAMES = ("DD", "DR", "RR")
# Bins of src/2pcf_C_version.cc:14-16: 700 linear bins on [0, 70).
TWO_PCF_AXIS = (700, 0.0, 70.0, "lin")


class TpcfError(RuntimeError):
    """A libtpcf call returned an error status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libtpcf error {code}: {msg}")
        self.code = code


class Axis(C.Structure):
    _fields_ = [("n", C.c_int), ("lo", C.c_double), ("hi", C.c_double), ("scale", C.c_int)]


class BinSpec(C.Structure):
    _fields_ = [("param", C.c_int), ("ax", Axis * 3), ("recheck_rel", C.c_double)]


class Job(C.Structure):
    _fields_ = [
        ("nseg", C.c_int),
        ("seg", C.c_void_p * 3),
        ("seg_lo", C.c_int64 * 3),
        ("seg_hi", C.c_int64 * 3),
        ("nterms", C.c_int),
        ("term", (C.c_int * 3) * 4),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("triplets", C.c_uint64),
        ("pairs", C.c_uint64),
        ("pairs_rechecked", C.c_uint64),
        ("device_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("pair_ms", C.c_double),
        ("triplet_ms", C.c_double),
        ("queue_ms", C.c_double),
        ("mask_ops", C.c_uint64),
        ("tile_bins_mean", C.c_double),
        ("tile_fill", C.c_double),
        ("triplet_form", C.c_int32),
        ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


# Every symbol include/tpcf.h declares, with its ctypes signature.
_SIGNATURES = {
    "tpcf_version": (C.c_char_p, []),
    "tpcf_last_error": (C.c_char_p, []),
    "tpcf_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "tpcf_ctx_destroy": (C.c_int, [C.c_void_p]),
    "tpcf_ctx_set_memory_budget": (C.c_int, [C.c_void_p, C.c_uint64]),
    "tpcf_catalog_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]),
    "tpcf_catalog_destroy": (C.c_int, [C.c_void_p]),
    "tpcf_catalog_size": (C.c_int64, [C.c_void_p]),
    "tpcf_read_catalog": (C.c_int64, [C.c_char_p, C.c_int, C.c_double, C.POINTER(C.POINTER(C.c_double))]),
    "tpcf_free": (None, [C.c_void_p]),
    "tpcf_hist_cells": (C.c_int64, [C.POINTER(BinSpec)]),
    "tpcf_axis_edges": (C.c_int, [C.POINTER(Axis), C.POINTER(C.c_double)]),
    "tpcf_binspec_check": (C.c_int, [C.POINTER(BinSpec)]),
    "tpcf_count": (C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(BinSpec), C.c_int, C.c_void_p, C.POINTER(Stats)]),
    "tpcf_count_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(BinSpec), C.c_int, C.c_void_p,
                                     C.POINTER(Stats)]),
    "tpcf_count_batch": (
        C.c_int,
        [C.c_void_p, C.POINTER(Job), C.POINTER(BinSpec), C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(Stats)],
    ),
    "tpcf_submit": (
        C.c_int,
        [C.c_void_p, C.POINTER(Job), C.POINTER(BinSpec), C.c_int, C.c_void_p, C.POINTER(C.c_int64)],
    ),
    "tpcf_poll": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Stats)]),
    "tpcf_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Stats)]),
    "tpcf_job_footprint": (C.c_int64, [C.POINTER(Job), C.POINTER(BinSpec), C.c_int]),
    "tpcf_ctx_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "tpcf_batch_create": (
        C.c_int,
        [C.c_void_p, C.POINTER(Job), C.POINTER(BinSpec), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)],
    ),
    "tpcf_batch_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(Stats)]),
    "tpcf_batch_info": (C.c_char_p, [C.c_void_p]),
    "tpcf_batch_destroy": (C.c_int, [C.c_void_p]),
    "tpcf_spec_reach": (C.c_int, [C.POINTER(BinSpec), C.POINTER(C.c_double)]),
    "tpcf_survey_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_double,
                                     C.POINTER(C.c_void_p)]),
    "tpcf_survey_cells": (C.c_int64, [C.c_void_p, C.POINTER(C.c_int)]),
    "tpcf_survey_jobs": (C.c_int64, [C.c_void_p]),
    "tpcf_survey_count": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BinSpec), C.c_int, C.c_void_p, C.c_void_p,
                                    C.POINTER(Stats)]),
    "tpcf_survey_destroy": (C.c_int, [C.c_void_p]),
    "tpcf_count_pairs": (C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(Axis), C.c_int, C.c_void_p, C.POINTER(Stats)]),
    "tpcf_count_pairs_batch": (
        C.c_int,
        [C.c_void_p, C.POINTER(Job), C.POINTER(Axis), C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(Stats)],
    ),
    "tpcf_radecz_to_xyz": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "tpcf_read_radecz": (C.c_int64, [C.c_char_p, C.c_void_p, C.POINTER(C.POINTER(C.c_double))]),
    "tpcf_triplet_norm": (C.c_int, [C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_double)]),
    "tpcf_pair_norm": (C.c_int, [C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_double)]),
    "tpcf_zeta": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_double),
                            C.c_void_p]),
    "tpcf_xi": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.c_void_p]),
    "tpcf_arange": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_int64]),
    "tpcf_reduced_q": (
        C.c_int,
        [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int,
         C.c_void_p, C.c_void_p, C.c_void_p],
    ),
}

BATCH_AUTOTUNE, BATCH_GRAPH = 1, 2

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libtpcf.so (built in-tree by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> int:
    if rc < 0:
        raise TpcfError(rc, load_library().tpcf_last_error().decode())
    return rc


def axis(n: int, lo: float, hi: float, scale: str | int = "lin") -> Axis:
    s = scale if isinstance(scale, int) else (LOG if scale == "log" else LIN)
    return Axis(int(n), float(lo), float(hi), s)


def binspec(param: str | int, axes: Sequence[Axis], recheck_rel: float = 1e-5) -> BinSpec:
    p = param if isinstance(param, int) else PARAMS[param]
    spec = BinSpec()
    spec.param = p
    for i, a in enumerate(axes):
        spec.ax[i] = a
    spec.recheck_rel = recheck_rel
    return spec


def r123_spec(n: int, lo: float, hi: float, scale: str = "lin", recheck_rel: float = 1e-5) -> BinSpec:
    return binspec(R123, [axis(n, lo, hi, scale)], recheck_rel)


def hist_cells(spec: BinSpec) -> int:
    return _check(load_library().tpcf_hist_cells(C.byref(spec)))


def axis_edges(ax: Axis) -> np.ndarray:
    e = np.zeros(ax.n + 1, dtype=np.float64)
    _check(load_library().tpcf_axis_edges(C.byref(ax), e.ctypes.data_as(C.POINTER(C.c_double))))
    return e


def read_catalog(path: str, fmt: int = 7, scale: float = 1.0) -> np.ndarray:
    """Read a catalog file into an (n, 3) float64 array (library reader)."""
    lib = load_library()
    ptr = C.POINTER(C.c_double)()
    n = _check(lib.tpcf_read_catalog(path.encode(), fmt, scale, C.byref(ptr)))
    try:
        soa = np.ctypeslib.as_array(ptr, shape=(3 * n,)).copy() if n else np.zeros(0)
    finally:
        lib.tpcf_free(ptr)
    return soa.reshape(3, n).T.copy() if n else np.zeros((0, 3))


def radecz_to_xyz(ra, dec, zred, poly: Sequence[float] | None = None) -> np.ndarray:
    """(n, 3) Cartesian Mpc/h from RA, Dec (degrees) and redshift (library
    conversion; poly None = the MICE distance fit)."""
    ra, dec, zred = (np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in (ra, dec, zred))
    n = ra.size
    soa = np.zeros(3 * n)
    pp = (C.c_double * 3)(*poly) if poly is not None else None
    _check(load_library().tpcf_radecz_to_xyz(ra.ctypes.data, dec.ctypes.data, zred.ctypes.data, n,
                                             C.cast(pp, C.c_void_p) if pp is not None else None, soa.ctypes.data))
    return soa.reshape(3, n).T.copy()


def read_radecz(path: str, poly: Sequence[float] | None = None) -> np.ndarray:
    """Read an "id ra dec z ..." catalog and convert it to (n, 3) Cartesian."""
    lib = load_library()
    ptr = C.POINTER(C.c_double)()
    pp = (C.c_double * 3)(*poly) if poly is not None else None
    n = _check(lib.tpcf_read_radecz(path.encode(), C.cast(pp, C.c_void_p) if pp is not None else None,
                                    C.byref(ptr)))
    try:
        soa = np.ctypeslib.as_array(ptr, shape=(3 * n,)).copy() if n else np.zeros(0)
    finally:
        lib.tpcf_free(ptr)
    return soa.reshape(3, n).T.copy() if n else np.zeros((0, 3))


class Catalog:
    """A catalog resident in device memory."""

    def __init__(self, ctx: "Context", xyz: np.ndarray):
        xyz = np.asarray(xyz, dtype=np.float64)
        if xyz.ndim != 2 or xyz.shape[1] != 3:
            raise ValueError("catalog must be an (n, 3) array")
        soa = np.ascontiguousarray(xyz.T)
        self.ctx = ctx
        self.n = xyz.shape[0]
        self.handle = C.c_void_p()
        _check(ctx.lib.tpcf_catalog_create(ctx.handle, soa.ctypes.data, self.n, C.byref(self.handle)))

    def close(self):
        if self.handle:
            self.ctx.lib.tpcf_catalog_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_job(segments: Sequence[tuple], terms: Iterable[tuple]) -> Job:
    """segments: (Catalog, lo, hi) triples; terms: sorted label triples (pair
    terms: label pairs).  The job keeps its catalogs alive."""
    job = Job()
    job._catalogs = [seg[0] for seg in segments]
    job.nseg = len(segments)
    for s, (cat, lo, hi) in enumerate(segments):
        job.seg[s] = cat.handle.value
        job.seg_lo[s] = lo
        job.seg_hi[s] = hi
    terms = list(terms)
    job.nterms = len(terms)
    for t, tr in enumerate(terms):
        for k in range(3):
            job.term[t][k] = tr[k] if k < len(tr) else 0
    return job


class Context:
    """A libtpcf context bound to one device."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.handle = C.c_void_p()
        _check(self.lib.tpcf_ctx_create(device, C.byref(self.handle)))

    def close(self):
        if self.handle:
            self.lib.tpcf_ctx_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def catalog(self, xyz: np.ndarray) -> Catalog:
        return Catalog(self, xyz)

    def set_memory_budget(self, nbytes: int):
        _check(self.lib.tpcf_ctx_set_memory_budget(self.handle, int(nbytes)))

    def count(self, job: Job, spec: BinSpec, mode: int = MODE_FAST):
        """Histograms of every term of job: array (nterms, cells) uint64, Stats."""
        cells = hist_cells(spec)
        out = np.zeros((job.nterms, cells), dtype=np.uint64)
        st = Stats()
        _check(self.lib.tpcf_count(self.handle, C.byref(job), C.byref(spec), mode, out.ctypes.data, C.byref(st)))
        return out, st

    def count_noalloc(self, job: Job, spec: BinSpec, mode: int = MODE_FAST):
        """count() through tpcf_count_noalloc: no device allocation in the
        call (the workspace is the library's fixed 64 MiB scratch)."""
        cells = hist_cells(spec)
        out = np.zeros((job.nterms, cells), dtype=np.uint64)
        st = Stats()
        _check(self.lib.tpcf_count_noalloc(self.handle, C.byref(job), C.byref(spec), mode, out.ctypes.data,
                                           C.byref(st)))
        return out, st

    def count_batch(self, jobs: Sequence[Job], specs: Sequence[BinSpec], mode: int = MODE_FAST):
        n = len(jobs)
        JA = (Job * n)(*jobs)
        SA = (BinSpec * n)(*specs)
        outs = [np.zeros((j.nterms, hist_cells(s)), dtype=np.uint64) for j, s in zip(jobs, specs)]
        ptrs = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
        st = Stats()
        _check(self.lib.tpcf_count_batch(self.handle, JA, SA, n, mode, ptrs, C.byref(st)))
        return outs, st

    def count_pairs(self, job: Job, ax: Axis, mode: int = MODE_FAST):
        """Pair-count histograms of every pair term of job: (nterms, ax.n) uint64, Stats."""
        out = np.zeros((job.nterms, ax.n), dtype=np.uint64)
        st = Stats()
        _check(self.lib.tpcf_count_pairs(self.handle, C.byref(job), C.byref(ax), mode, out.ctypes.data, C.byref(st)))
        return out, st

    def count_pairs_batch(self, jobs: Sequence[Job], axes: Sequence[Axis], mode: int = MODE_FAST):
        n = len(jobs)
        JA = (Job * n)(*jobs)
        AA = (Axis * n)(*axes)
        outs = [np.zeros((j.nterms, a.n), dtype=np.uint64) for j, a in zip(jobs, axes)]
        ptrs = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
        st = Stats()
        _check(self.lib.tpcf_count_pairs_batch(self.handle, JA, AA, n, mode, ptrs, C.byref(st)))
        return outs, st

    def survey(self, data: np.ndarray, randoms: np.ndarray, maxsep: float) -> "Survey":
        return Survey(self, data, randoms, maxsep)

    def batch(self, jobs: Sequence[Job], specs: Sequence[BinSpec], mode: int = MODE_FAST,
              autotune: bool = True, graph: bool = True) -> "Batch":
        return Batch(self, jobs, specs, mode, (BATCH_AUTOTUNE if autotune else 0) | (BATCH_GRAPH if graph else 0))

    def memory(self) -> dict:
        """Serving workers' device memory: held now, peak, budget (bytes)."""
        out = (C.c_uint64 * 3)()
        _check(self.lib.tpcf_ctx_memory(self.handle, out))
        return {"in_use": out[0], "peak": out[1], "budget": out[2]}

    def footprint(self, job: Job, spec: BinSpec, mode: int = MODE_FAST) -> int:
        return _check(self.lib.tpcf_job_footprint(C.byref(job), C.byref(spec), mode))

    def footprint_triplets(self, job: Job, spec: BinSpec) -> int:
        """Unique triplets of the job's terms (host arithmetic on the segment sizes)."""
        c2 = lambda n: n * (n - 1) // 2
        c3 = lambda n: n * (n - 1) * (n - 2) // 6
        n = [job.seg_hi[s] - job.seg_lo[s] for s in range(job.nseg)]
        total = 0
        for t in range(job.nterms):
            a, b, c = (job.term[t][k] for k in range(3))
            if a == b == c:
                total += c3(n[a])
            elif a == b:
                total += c2(n[a]) * n[c]
            elif b == c:
                total += n[a] * c2(n[b])
            else:
                total += n[a] * n[b] * n[c]
        return total

    def submit(self, job: Job, spec: BinSpec, mode: int = MODE_FAST):
        """Enqueue a job; returns (ticket, output array).  Keep the array alive."""
        out = np.zeros((job.nterms, hist_cells(spec)), dtype=np.uint64)
        t = C.c_int64()
        _check(self.lib.tpcf_submit(self.handle, C.byref(job), C.byref(spec), mode, out.ctypes.data, C.byref(t)))
        return t.value, out

    def poll(self, ticket: int):
        st = Stats()
        r = _check(self.lib.tpcf_poll(self.handle, ticket, C.byref(st)))
        return (r == 1), st

    def wait(self, ticket: int) -> Stats:
        st = Stats()
        _check(self.lib.tpcf_wait(self.handle, ticket, C.byref(st)))
        return st


class Survey:
    """A survey (data + randoms) split into spatial cells of side >= maxsep,
    counted cell by cell (tpcf_survey_*); see include/tpcf.h."""

    def __init__(self, ctx: "Context", data: np.ndarray, randoms: np.ndarray, maxsep: float):
        self.ctx = ctx
        d = np.ascontiguousarray(np.asarray(data, dtype=np.float64).T)
        r = np.ascontiguousarray(np.asarray(randoms, dtype=np.float64).T)
        self.handle = C.c_void_p()
        _check(ctx.lib.tpcf_survey_create(ctx.handle, d.ctypes.data, d.shape[1], r.ctypes.data, r.shape[1],
                                          float(maxsep), C.byref(self.handle)))
        dims = (C.c_int * 3)()
        self.ncells = _check(ctx.lib.tpcf_survey_cells(self.handle, dims))
        self.dims = tuple(dims)
        self.njobs = _check(ctx.lib.tpcf_survey_jobs(self.handle))

    def count(self, spec: BinSpec, mode: int = MODE_FAST, per_cell: bool = False):
        """(hist (4, cells): DDD, DDR, DRR, RRR of the survey, per-cell
        histograms (ncells, 4, cells) or None, Stats)."""
        cells = hist_cells(spec)
        hist = np.zeros((4, cells), dtype=np.uint64)
        pc = np.zeros((self.ncells, 4, cells), dtype=np.uint64) if per_cell else None
        st = Stats()
        _check(self.ctx.lib.tpcf_survey_count(self.ctx.handle, self.handle, C.byref(spec), mode, hist.ctypes.data,
                                              pc.ctypes.data if pc is not None else None, C.byref(st)))
        return hist, pc, st

    def close(self):
        if self.handle:
            self.ctx.lib.tpcf_survey_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spec_reach(spec: BinSpec) -> float:
    """Longest side a counted triangle of spec can have (tpcf_spec_reach)."""
    out = C.c_double()
    _check(load_library().tpcf_spec_reach(C.byref(spec), C.byref(out)))
    return out.value


class Batch:
    """A prepared batch (tpcf_batch): staged once, replayed by run()."""

    def __init__(self, ctx: Context, jobs: Sequence[Job], specs: Sequence[BinSpec], mode: int, flags: int):
        self.ctx = ctx
        n = len(jobs)
        self._keep = list(jobs)  # the jobs (and their catalogs) live as long as the batch
        self._jobs = (Job * n)(*jobs)
        self._specs = (BinSpec * n)(*specs)
        self.outs = [np.zeros((j.nterms, hist_cells(s)), dtype=np.uint64) for j, s in zip(jobs, specs)]
        self._ptrs = (C.c_void_p * n)(*[o.ctypes.data for o in self.outs])
        self.handle = C.c_void_p()
        _check(ctx.lib.tpcf_batch_create(ctx.handle, self._jobs, self._specs, n, mode, flags, C.byref(self.handle)))

    def run(self):
        """Recompute every histogram; returns (list of arrays, Stats)."""
        st = Stats()
        _check(self.ctx.lib.tpcf_batch_run(self.handle, self._ptrs, C.byref(st)))
        return self.outs, st

    @property
    def info(self) -> str:
        return self.ctx.lib.tpcf_batch_info(self.handle).decode()

    def close(self):
        if self.handle:
            self.ctx.lib.tpcf_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- estimators (host arithmetic in libtpcf; no device needed) -------------


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def triplet_norm(ngal: Sequence[int], term: int) -> float:
    """Triplets term 0..3 (DDD, DDR, DRR, RRR) is normalized by, from the
    three catalog sizes of its histogram file."""
    out = C.c_double()
    _check(load_library().tpcf_triplet_norm((C.c_int64 * 3)(*ngal), term, C.byref(out)))
    return out.value


def pair_norm(ngal: Sequence[int], term: int) -> float:
    """Pairs term 0..2 (DD, DR, RR) is normalized by, from the two catalog
    sizes of its pair-count file."""
    out = C.c_double()
    _check(load_library().tpcf_pair_norm((C.c_int64 * 2)(*ngal), term, C.byref(out)))
    return out.value


def zeta(ddd, ddr, drr, rrr, norms: Sequence[float]) -> np.ndarray:
    """(DDD - 3 DDR + 3 DRR - RRR) / RRR of normalized histograms, any shape."""
    arrs = [_f64(a) for a in (ddd, ddr, drr, rrr)]
    out = np.zeros(arrs[0].shape, dtype=np.float64)
    _check(load_library().tpcf_zeta(*[a.ctypes.data for a in arrs], arrs[0].size, (C.c_double * 4)(*norms),
                                    out.ctypes.data))
    return out


def xi(dd, dr, rr, norms: Sequence[float]) -> np.ndarray:
    """(DD - 2 DR + RR) / RR of normalized pair counts."""
    arrs = [_f64(a) for a in (dd, dr, rr)]
    out = np.zeros(arrs[0].shape, dtype=np.float64)
    _check(load_library().tpcf_xi(*[a.ctypes.data for a in arrs], arrs[0].size, (C.c_double * 3)(*norms),
                                  out.ctypes.data))
    return out


def reduced_q(zeta_theta, s: float, q: float, theta_lo: float, theta_width: float, xs, ys):
    """(theta/pi, denominator, Q) along theta for one (s, q) from a 2PCF table."""
    z, x, y = _f64(zeta_theta), _f64(xs), _f64(ys)
    n = z.size
    th, den, qr = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(load_library().tpcf_reduced_q(z.ctypes.data, n, s, q, theta_lo, theta_width, x.ctypes.data,
                                         y.ctypes.data, x.size, th.ctypes.data, den.ctypes.data, qr.ctypes.data))
    return th, den, qr


def four_term_job(data: Catalog, randoms: Catalog) -> Job:
    """DDD, DDR, DRR, RRR of one data and one random catalog."""
    return make_job([(data, 0, data.n), (randoms, 0, randoms.n)], FOUR_TERMS)


def pair_job(data: Catalog, randoms: Catalog) -> Job:
    """DD, DR, RR pair terms of one data and one random catalog."""
    return make_job([(data, 0, data.n), (randoms, 0, randoms.n)], PAIR_TERMS)
