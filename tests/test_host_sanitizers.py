"""The library's plain C++ host parts under the host sanitizers (CPU only):
the threaded duplicate-pruning pass of the root split (`spz_dedup.h`) and the
CLI's problem-file reader (`spz_problem.h`), built with g++ once under
AddressSanitizer + UndefinedBehaviorSanitizer and once under
ThreadSanitizer, and run on random levels and malformed files
(`tests/host/host_checks.cpp`).  The CPU oracle, the checker of every parity
test, runs on small inputs under AddressSanitizer + UndefinedBehaviorSanitizer
too (`tests/host/oracle_checks.c`; its OpenMP builder rules out
ThreadSanitizer, whose runtime does not see the OpenMP library's
synchronisation)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host", "host_checks.cpp")

SANITIZERS = {
    "address+undefined": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
    "thread": ["-fsanitize=thread"],
}


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("name", sorted(SANITIZERS))
def test_host_code_under_sanitizer(name, tmp_path):
    exe = tmp_path / "host_checks"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", *SANITIZERS[name], SRC, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and r.stdout.strip() == "OK", (r.stdout[-2000:], r.stderr[-4000:])


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    exe = tmp_path / "oracle_checks"
    src = [os.path.join(HERE, "..", "oracle", "spz_oracle.c"), os.path.join(HERE, "host", "oracle_checks.c")]
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fopenmp", *SANITIZERS["address+undefined"], *src, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and r.stdout.strip() == "OK", (r.stdout[-2000:], r.stderr[-4000:])
