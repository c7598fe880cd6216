"""CPU: the library's host code issues no synchronous null-stream operation.

The serving workers run on CU-masked HIP streams, which synchronize with the
legacy null stream: a synchronous hipMemcpy or hipMemset (both run on the
null stream) or a hipDeviceSynchronize in any path a worker takes makes that
worker wait for every kernel of the other workers, and later launches on
their streams wait for the copy.  Every copy and clear therefore names a
stream, and this test keeps it that way."""
import os
import re

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "3pcf_amd", "csrc")
SYNC_CALLS = re.compile(r"\b(hipMemcpy|hipMemset|hipMemcpyToSymbol|hipMemcpyFromSymbol|hipMemcpyHtoD|"
                        r"hipMemcpyDtoH|hipMemcpyDtoD|hipDeviceSynchronize)\s*\(")


def strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def test_no_synchronous_null_stream_calls():
    found = []
    for name in sorted(os.listdir(CSRC)):
        if not name.endswith((".cpp", ".hip", ".h")):
            continue
        with open(os.path.join(CSRC, name)) as f:
            code = strip_comments(f.read())
        for m in SYNC_CALLS.finditer(code):
            line = code.count("\n", 0, m.start()) + 1
            found.append(f"{name}:{line}: {m.group(1)}")
    assert not found, found


def test_pattern_catches_a_synchronous_copy():
    # the check itself: a planted synchronous copy is found, an async one is not
    assert SYNC_CALLS.search("HIP_TRY(hipMemcpy(w, ctl, sizeof w, hipMemcpyDeviceToHost));")
    assert not SYNC_CALLS.search("HIP_TRY(hipMemcpyAsync(w, ctl, sizeof w, hipMemcpyDeviceToHost, st));")
    assert not SYNC_CALLS.search("HIP_TRY(hipMemsetAsync(p, 0, n, st));")
