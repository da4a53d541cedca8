"""bench.py's command line without a GPU: more than one GPU, a step count
below 1, a negative warmup or an unknown configuration is refused with a
message and exit code 2, before anything touches a device."""
import json
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=300, cwd=REPO)
    assert r.returncode == 2, (r.stdout, r.stderr)
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_bench_refuses_bad_step_counts_and_configs():
    for args in (["--steps", "0"], ["--warmup", "-1"], ["--config", "c9"]):
        r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], capture_output=True,
                           text=True, timeout=300, cwd=REPO)
        assert r.returncode == 2, (args, r.stdout, r.stderr)
        assert "error" in json.loads(r.stdout.strip().splitlines()[-1]), args
