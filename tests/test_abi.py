"""The C-ABI library: it loads, exports every entry point include/tpcf.h
declares, and its host-side parts (bin tables, catalog reader, argument
checks, command-line handling) behave; no GPU compute is called here."""
import ctypes as C
import importlib
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
tp = importlib.import_module("3pcf_amd")
INP = os.path.join(ROOT, "tests", "golden", "inputs")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tpcf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tpcf_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(tp.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 19
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(tp._SIGNATURES)


def test_version_and_error_channel():
    lib = tp.load_library()
    assert b"gfx950" in lib.tpcf_version()
    assert lib.tpcf_binspec_check(None) == -1
    assert b"NULL" in lib.tpcf_last_error()


@pytest.mark.parametrize("spec", [(16, 0.0, 40.0, "lin"), (20, 1.0, 50.0, "log"), (32, 0.0, 1.0, "lin"),
                                  (10, 2.0, 20.0, "log"), (255, 0.01, 500.0, "log")])
def test_edge_tables_equal_oracle(spec):
    got = tp.axis_edges(tp.axis(*spec))
    want = orc.edges(orc.axis(*spec))
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("bad", [
    (tp.R123, [(0, 0.0, 1.0, "lin")]),
    (tp.R123, [(8, 2.0, 1.0, "lin")]),
    (tp.R123, [(8, 0.0, 10.0, "log")]),
    (tp.R123, [(256, 0.0, 10.0, "lin")]),
    (tp.SQT, [(10, 1.0, 2.0, "lin"), (8, 1.0, 3.0, "lin"), (8, 0.0, 1.5, "lin")]),
    (tp.SQT, [(100, 1.0, 2.0, "lin"), (40, 1.0, 3.0, "lin"), (40, 0.0, 1.0, "lin")]),
    (7, [(8, 0.0, 1.0, "lin")] * 3),
])
def test_binspec_rejections(bad):
    param, axes = bad
    spec = tp.binspec(param, [tp.axis(*a) for a in axes])
    assert tp.load_library().tpcf_binspec_check(C.byref(spec)) == -1
    assert tp.load_library().tpcf_last_error()


def test_hist_cells():
    assert tp.hist_cells(tp.r123_spec(20, 1.0, 50.0, "log")) == 8000
    s = tp.binspec(tp.SQT, [tp.axis(50, 2, 12), tp.axis(16, 0.9, 4.1), tp.axis(32, 0, 1)])
    assert tp.hist_cells(s) == 50 * 16 * 32


@pytest.mark.parametrize("name,fmt", [("MICE_5degx5deg_1k_randomized.dat", 7), ("flat_MICE_5degx5deg_1k.dat", 7),
                                      ("MICE_LRGs_10degx10deg.dat", 7), ("input_1k_0.cat", 3)])
def test_reader_matches_oracle_reader(name, fmt):
    path = os.path.join(INP, name)
    got = tp.read_catalog(path, fmt)
    lib = orc.load()
    lib.orc_read_7col.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
    lib.orc_read_7col.restype = C.c_int64
    lib.orc_read_count.argtypes = [C.c_char_p, C.c_int, C.c_double, C.POINTER(C.c_void_p)]
    lib.orc_read_count.restype = C.c_int64
    ptr = C.c_void_p()
    n = lib.orc_read_7col(path.encode(), 1, C.byref(ptr)) if fmt == 7 else \
        lib.orc_read_count(path.encode(), 1, 1.0, C.byref(ptr))
    want = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(3 * n,)).reshape(3, n).T.copy()
    assert got.shape == (n, 3) and n > 900
    assert got.tobytes() == want.tobytes()


def test_reader_scale_and_errors(tmp_path):
    f = tmp_path / "c.cat"
    f.write_text("2\n206264 0 412528\n1 2 3\n")
    xyz = tp.read_catalog(str(f), 3, 206264.0)
    assert xyz[0, 0] == 1.0 and xyz[0, 2] == 2.0
    with pytest.raises(tp.TpcfError):
        tp.read_catalog(str(tmp_path / "missing.dat"), 7)
    bad = tmp_path / "bad.dat"
    bad.write_text("1 0 0 0 1 2 3\n2 0 0 0 1 2\n")
    with pytest.raises(tp.TpcfError):
        tp.read_catalog(str(bad), 7)
    short = tmp_path / "short.cat"
    short.write_text("3\n1 2 3\n")
    with pytest.raises(tp.TpcfError):
        tp.read_catalog(str(short), 3)


def test_no_gpu_fails_loudly():
    """Without a usable gfx950 device the library refuses to run (there is no
    CPU fallback)."""
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(tp.TpcfError):
        tp.Context(0)


@pytest.mark.parametrize("prog", ["3pcf_C_version_triangle_parametrization", "3pcf_C_version",
                                  "3pcf_C_version_triangle_parametrization_JBJ"])
def test_cli_argument_errors(prog, tmp_path):
    exe = os.path.join(tp.BIN_DIR, prog)
    r = subprocess.run([exe, "only_one.dat"], capture_output=True, text=True)
    assert r.returncode == 1 and "three input files" in r.stdout
    f = os.path.join(INP, "d200.dat")
    r = subprocess.run([exe, f, f, f, "--param", "bogus"], capture_output=True, text=True)
    assert r.returncode == 2
    r = subprocess.run([exe, f, f, str(tmp_path / "missing.dat")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
    # --all-voxels needs -X and a file-name template with {index}; checked
    # before any device is touched.
    for extra in (["-o", str(tmp_path / "v.dat"), "-X", "2"], ["-o", str(tmp_path / "v_{index}.dat")]):
        r = subprocess.run([exe, f, f, f, "--all-voxels", *extra], capture_output=True, text=True)
        assert r.returncode == 2 and "--all-voxels" in r.stderr


def test_radecz_conversion_reproduces_reference_catalogs():
    """sample_data/gen_flat_MICE_data.py wrote the x y z columns of the flat
    MICE catalogs from their ra dec z columns: the library conversion gives
    the same doubles (5x5 deg file: every value; 20x20 deg file, written
    with an older libm: all but 3 of 3000 values, those within two ulps)."""
    for name, ulps in (("flat_MICE_5degx5deg_1k.dat", 0), ("flat_MICE_20degx20deg_1k.dat", 2)):
        path = os.path.join(INP, name)
        a = np.loadtxt(path)
        got = tp.radecz_to_xyz(a[:, 1], a[:, 2], a[:, 3])
        want = a[:, 4:7]
        if ulps == 0:
            assert np.array_equal(got, want)
        else:
            assert np.all(np.abs(got - want) <= ulps * np.spacing(np.abs(want)))
            assert np.count_nonzero(got != want) <= 3
        assert np.array_equal(tp.read_radecz(path), got)


def test_radecz_closed_forms_and_poly():
    # pole: dec = 90 -> (0, 0, D); ra = 90, dec = 0 -> (D, ~0, ~0)
    d = tp.radecz_to_xyz([0.0, 90.0], [90.0, 0.0], [0.0, 0.0], poly=[100.0, 0.0, 0.0])
    assert d[0].tolist() == [0.0, 0.0, 100.0]
    np.testing.assert_allclose(d[1], [100.0, 0.0, 0.0], atol=1e-12)
    # D = p0 + p1 z + p2 z^2
    e = tp.radecz_to_xyz([0.0], [90.0], [2.0], poly=[1.0, 3.0, 5.0])
    assert e[0, 2] == 1.0 + 3.0 * 2.0 + 5.0 * 4.0
