"""Pin the oracle's refcompat32 restatement to the reference's own outputs.

* examples/sandbox histograms (copied to tests/golden/sandbox): byte equality
  of the oracle command line with the reference's output files;
* outputs of the reference programs compiled from source (tests/golden/
  ref_outputs, made by tests/golden/make_fixtures.py) on small catalog slices,
  covering every file-order case (DDD, DDR, DRR, RRR, DRD), the JBJ variant and
  the (r1,r2,r3) cube program.
"""
import json
import os
import shutil
import subprocess

import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
INP = os.path.join(GOLD, "inputs")
MANIFEST = json.load(open(os.path.join(GOLD, "MANIFEST.json")))


@pytest.fixture(scope="module")
def cli():
    orc.build()
    return orc.CLI


def run_cli(cli, tmp_path, mode, files, extra=(), cwd=INP):
    out = tmp_path / "out.dat"
    cmd = [cli, mode, "--threads", "8", *extra, *files, "-o", str(out)]
    subprocess.run(cmd, cwd=cwd, check=True, stdout=subprocess.DEVNULL)
    return out.read_bytes()


@pytest.mark.parametrize("name", sorted(f for f in MANIFEST["sandbox"] if "voxel2" in f))
def test_sandbox_voxel2(cli, tmp_path, name):
    vox = name.split("_")[2]
    f = "MICE_5degx5deg_1k_randomized.dat"
    got = run_cli(cli, tmp_path, "ref-tri", [f, f, f], ["--axis", "2,25,0,1", "-X", "2", "-x", vox])
    assert got == open(os.path.join(GOLD, "sandbox", name), "rb").read()


@pytest.mark.parametrize("name", sorted(f for f in MANIFEST["sandbox"] if "voxel4" in f))
def test_sandbox_voxel4_legacy_bound(cli, tmp_path, name):
    """The voxel4 files predate the current DDD bound: they need kmin=j+1
    (src/3pcf_C_version_triangle_parametrization.cc:499, commented out now)."""
    vox = name.split("_")[2]
    f = "MICE_5degx5deg_1k_randomized.dat"
    got = run_cli(cli, tmp_path, "ref-tri", [f, f, f],
                  ["--axis", "2,25,0,1", "--kmin-legacy", "-X", "4", "-x", vox])
    assert got == open(os.path.join(GOLD, "sandbox", name), "rb").read()


def test_sandbox_debbie_full_run_plus_sum_newline(cli, tmp_path):
    """The Debbie DDD file is a full run reduced by sum_voxels.py, whose
    print adds one trailing newline (examples/sum_voxels.py:56)."""
    f = "MICE_LRGs_10degx10deg.dat"
    got = run_cli(cli, tmp_path, "ref-tri", [f, f, f], ["--axis", "2,25,0,1"])
    want = open(os.path.join(GOLD, "sandbox", "DDD_Debbie_LRG_CPU_10degx10deg_LRG_binning_SLAC_1k.dat"), "rb").read()
    assert got + b"\n" == want


@pytest.mark.parametrize("name", sorted(f for f in MANIFEST["ref_outputs"] if not f.startswith("two_")))
def test_compiled_reference_outputs(cli, tmp_path, name):
    info = MANIFEST["ref_outputs"][name]
    prog, argv = info["program"], info["argv"]
    extra = []
    if prog == "tri25":
        extra = ["--axis", "2,25,0,1"]
    if prog == "plain16":
        extra = ["--axis", "0,16,0,41.569219381653056"]
    mode = {"tri32": "ref-tri", "tri25": "ref-tri", "trij32": "ref-jbj", "plain": "ref-plain",
            "plain16": "ref-plain"}[prog]
    files = [a for a in argv if not a.startswith("-") and "." in a]
    flags = [a for a in argv if a not in files]
    got = run_cli(cli, tmp_path, mode, files, extra + flags)
    assert got == open(os.path.join(GOLD, "ref_outputs", name), "rb").read()


def test_plain_oob_counter(cli, tmp_path):
    """src/3pcf_C_version.cc:29-30 sends d >= hist_max to bin nbins+1, whose
    index overflows the cube; the restatement counts those writes instead."""
    d = tmp_path / "far.cat"
    d.write_text("3\n0 0 0\n100 0 0\n0 100 0\n")
    out = subprocess.run([cli, "ref-plain", str(d), str(d), str(d), "-o", str(tmp_path / "o.dat")],
                         check=True, capture_output=True, text=True).stdout
    assert "oob: " in out and int(out.split("oob: ")[1].split()[0]) > 0


@pytest.mark.skipif(not os.path.isdir(os.path.join(os.path.dirname(HERE), "oracle", "_ref"))
                    or not os.path.isdir("/root/reference/examples/sandbox"),
                    reason="reference sources and their build are only present in the build container")
def test_fixtures_match_reference_tree():
    import hashlib
    for name, digest in MANIFEST["sandbox"].items():
        src = os.path.join("/root/reference/examples/sandbox", name)
        if not os.path.exists(src):  # the 2PCF table sits one level up
            src = os.path.join("/root/reference/examples", name)
        assert hashlib.sha256(open(src, "rb").read()).hexdigest() == digest
