"""The host range planner (flowvpm.jl_amd/csrc/range_plan.cpp) under the
host sanitizers, on the CPU.  The planner splits a list over up to eight
host threads that write one shared staging buffer and issue the device
copies of their pieces; tests/sanitize/plan_driver.cpp runs it with those
copies replaced by host memcpy, on a 2.6M-pair leaf list (eight threads),
lists with bad pairs in two chunks (lowest index reported), a failing
copy, shuffled leaves with empty target ranges and a general overlapping
list, checking every plan against a brute-force reading of the list.
Built once with AddressSanitizer + UndefinedBehaviorSanitizer and once
with ThreadSanitizer; both runs must be clean."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "flowvpm.jl_amd", "csrc")
DRIVER = os.path.join(ROOT, "tests", "sanitize", "plan_driver.cpp")

SANITIZERS = {
    "asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"],
    "tsan": ["-fsanitize=thread"],
}


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("which", sorted(SANITIZERS))
def test_range_planner_sanitized(tmp_path, which):
    exe = str(tmp_path / ("plan_driver_" + which))
    cmd = (["g++", "-std=c++17", "-O1", "-g", "-pthread"] + SANITIZERS[which] +
           ["-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
            os.path.join(CSRC, "range_plan.cpp"), DRIVER, "-o", exe])
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "plan_driver OK" in r.stdout
    assert "Sanitizer" not in out and "runtime error" not in out, out[-4000:]
