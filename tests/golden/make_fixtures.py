#!/usr/bin/env python3
"""Regenerate the golden fixtures under tests/golden/ (run in a container
where the reference is mounted at /root/reference and oracle/_ref is built by
oracle/build_ref.sh).

Fixtures are data only:
  inputs/        catalog files copied from the reference's sample_data/, plus
                 small slices of them (first lines) for the compiled-reference
                 runs below;
  sandbox/       histogram files copied from the reference's examples/sandbox/
                 (outputs of the reference's own CPU program, verified
                 byte-reproducible): all 8 DDD_voxel2 files, 8 of the 64
                 DDD_voxel4 files, the Debbie LRG DDD, DDR, DRR and RRR
                 files, and examples/master_2pcf_reference.dat (the 2PCF
                 table of the reduced-amplitude plot);
  ref_outputs/   outputs of the reference programs compiled from source
                 (oracle/_ref) on the small slices: (s,q,theta) and JBJ for
                 every file-order case, the (r1,r2,r3) cube program, and the
                 pair-count program (2pcf_C_version) on the slices and on the
                 full 1k catalogs;
  MANIFEST.json  sha256 of every fixture and the command that produced each
                 ref_outputs file.
"""
import hashlib
import json
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
REF_BIN = os.path.join(REPO, "oracle", "_ref")

SANDBOX = [f"DDD_voxel2_{a}{b}{c}_evenbinning_CPU_5degx5deg_LRG_binning_1k.dat"
           for a in ("000", "001") for b in ("000", "001") for c in ("000", "001")]
SANDBOX += [f"DDD_voxel4_{v}_evenbinning_CPU_5degx5deg_LRG_binning_1k.dat"
            for v in ("000000000", "000001002", "001001001", "001002003", "002001000", "002002003",
                      "003000002", "003003003")]
SANDBOX += [f"{t}_Debbie_LRG_CPU_10degx10deg_LRG_binning_SLAC_1k.dat" for t in ("DDD", "DDR", "DRR", "RRR")]
# The 2PCF table the reduced-amplitude plot reads (examples/plot_MICE_3pcf.py:176).
EXAMPLES = ["master_2pcf_reference.dat"]
INPUTS = ["MICE_5degx5deg_1k_randomized.dat", "MICE_LRGs_10degx10deg.dat", "flat_MICE_5degx5deg_1k.dat",
          "flat_MICE_20degx20deg_1k.dat", "input_1k_0.cat", "input_1k_1.cat"]


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def head(src, dst, n):
    with open(src) as f, open(dst, "w") as g:
        for _ in range(n):
            g.write(f.readline())


def count_slice(src, dst, n):
    """First n points of a count-header catalog, with the header rewritten."""
    with open(src) as f:
        f.readline()
        lines = [f.readline() for _ in range(n)]
    with open(dst, "w") as g:
        g.write(f"{n}\n")
        g.writelines(lines)


def main():
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "sandbox"), exist_ok=True)
    out_dir = os.path.join(HERE, "ref_outputs")
    os.makedirs(out_dir, exist_ok=True)
    manifest = {"inputs": {}, "sandbox": {}, "ref_outputs": {}}
    for f in INPUTS:
        dst = os.path.join(HERE, "inputs", f)
        shutil.copyfile(os.path.join(REF, "sample_data", f), dst)
    for f in SANDBOX:
        dst = os.path.join(HERE, "sandbox", f)
        shutil.copyfile(os.path.join(REF, "examples", "sandbox", f), dst)
        manifest["sandbox"][f] = sha(dst)
    for f in EXAMPLES:
        dst = os.path.join(HERE, "sandbox", f)
        shutil.copyfile(os.path.join(REF, "examples", f), dst)
        manifest["sandbox"][f] = sha(dst)
    inp = os.path.join(HERE, "inputs")
    head(os.path.join(inp, "MICE_5degx5deg_1k_randomized.dat"), os.path.join(inp, "d200.dat"), 200)
    head(os.path.join(inp, "flat_MICE_5degx5deg_1k.dat"), os.path.join(inp, "r180.dat"), 180)
    count_slice(os.path.join(inp, "input_1k_0.cat"), os.path.join(inp, "a60.cat"), 60)
    count_slice(os.path.join(inp, "input_1k_1.cat"), os.path.join(inp, "b50.cat"), 50)
    for f in sorted(os.listdir(inp)):
        manifest["inputs"][f] = sha(os.path.join(inp, f))

    orders = {"DDD": ("d200.dat",) * 3, "DDR": ("d200.dat", "d200.dat", "r180.dat"),
              "DRR": ("d200.dat", "r180.dat", "r180.dat"), "RRR": ("r180.dat",) * 3,
              "DRD": ("d200.dat", "r180.dat", "d200.dat")}
    runs = []
    for prog in ("tri32", "trij32"):
        for name, files in orders.items():
            runs.append((f"{prog}_{name}.dat", prog, list(files), []))
    runs.append(("tri25_DDR_vox2_001000001.dat", "tri25", list(orders["DDR"]), ["-X", "2", "-x", "001000001"]))
    for name, files in {"AAA": ("a60.cat",) * 3, "AAB": ("a60.cat", "a60.cat", "b50.cat"),
                        "ABB": ("a60.cat", "b50.cat", "b50.cat")}.items():
        runs.append((f"plain_{name}.dat", "plain", list(files), []))
        runs.append((f"plain16_{name}.dat", "plain16", list(files), []))
    two = {"DD": ("d200.dat", "d200.dat"), "DR": ("d200.dat", "r180.dat"), "RR": ("r180.dat", "r180.dat"),
           "DD1k": ("MICE_5degx5deg_1k_randomized.dat",) * 2,
           "DR1k": ("MICE_5degx5deg_1k_randomized.dat", "flat_MICE_5degx5deg_1k.dat"),
           "RR1k": ("flat_MICE_5degx5deg_1k.dat",) * 2}
    for name, files in two.items():
        runs.append((f"two_{name}.dat", "two", list(files), []))
    runs.append(("two_DD_vox2_001000.dat", "two", list(two["DD"]), ["-X", "2", "-x", "001000000"]))
    runs.append(("two_DD_vox2_000001.dat", "two", list(two["DD"]), ["-X", "2", "-x", "000001000"]))
    for out, prog, files, extra in runs:
        cmd = [os.path.join(REF_BIN, prog)] + files + extra + ["-o", os.path.join(out_dir, out)]
        subprocess.run(cmd, cwd=inp, check=True, stdout=subprocess.DEVNULL)
        manifest["ref_outputs"][out] = {"sha256": sha(os.path.join(out_dir, out)),
                                        "program": prog, "argv": files + extra}
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
