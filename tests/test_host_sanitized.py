"""Host code of libtpcf under AddressSanitizer + UndefinedBehaviorSanitizer
(CPU only; the Makefile's "san" target, built by __graft_entry__.build()).

* tests/sanitize/plan_check: the job planner (layouts, work items, reach
  bitsets, culled item merging, region plans) against brute force;
* the CPU tests that drive libtpcf's host code (readers, bin specs, edge
  tables, RA/Dec conversion, estimators, error paths) run in a child process
  against the sanitized library, with the sanitizer runtime preloaded into
  the Python interpreter.  Any sanitizer report fails the run (UBSan halts on
  the first error, ASan on any).
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "3pcf_amd", "_build_san")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(SAN, "plan_check")),
                                reason="sanitized build missing (make -C 3pcf_amd/csrc san)")


def test_planner_under_sanitizers():
    r = subprocess.run([os.path.join(SAN, "plan_check")], env=ENV, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all checks passed" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_cpu_suite_against_sanitized_library():
    rt = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rt:
        pytest.skip("clang ASan runtime not found")
    env = dict(ENV, LD_PRELOAD=rt[-1], TPCF_LIBRARY=os.path.join(SAN, "libtpcf_san.so"))
    code = ("import importlib, sys; tp = importlib.import_module('3pcf_amd'); "
            "assert tp.LIB_PATH.endswith('libtpcf_san.so'), tp.LIB_PATH; "
            "import pytest; sys.exit(pytest.main(['-q', '-p', 'no:cacheprovider', '-m', 'not gpu', "
            "'tests/test_abi.py', 'tests/test_estimators.py']))")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "passed" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
