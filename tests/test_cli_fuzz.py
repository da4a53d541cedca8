"""spz_solve's argument parsing and problem-file reading under random input
(CPU only): whatever the options, values and file contents, the CLI ends
with a usage error (2) or a reported failure (1) -- never a signal, and
never exit 0 here, where it cannot create a device context."""
import os
import subprocess

import pytest

hypothesis = pytest.importorskip("hypothesis")
from hypothesis import HealthCheck, given, settings, strategies as st  # noqa: E402

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "sliding-puzzle_amd", "bin", "spz_solve")

FLAGS = ["-H", "-m", "--batch", "--iters", "--path", "--csv", "--pdb-cache", "--pdb-tab", "--export-tab",
         "--lb", "--mlp-params", "--resplit", "--frontier", "--wg", "--lds-depth", "--budget", "--round-time",
         "--max-iters", "--dedup", "--no-lookahead", "--autotune", "--bogus", "-x"]
VALUES = st.one_of(st.sampled_from(["pdb", "manhattan", "first", "all", "learned", "uniform", "prevload", "0",
                                    "-1", "64", "2147483648", "1e3", "", "abc"]),
                   st.integers(-(1 << 40), 1 << 40).map(str))
CELLS = st.one_of(st.permutations(list(range(16))), st.permutations(list(range(25))),
                  st.lists(st.integers(-5, 40), max_size=30))


def torch_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(torch_gpu(), reason="with a GPU a valid command line solves and exits 0")
@settings(max_examples=150, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(args=st.lists(st.one_of(st.sampled_from(FLAGS), VALUES), max_size=8), cells=CELLS,
       text=st.sampled_from(["ints", "junk", "empty"]))
def test_cli_never_crashes(tmp_path, args, cells, text):
    p = tmp_path / "prob"
    if text == "ints":
        p.write_text(" ".join(map(str, cells)) + "\n")
    elif text == "junk":
        p.write_text("1 2 x 4\n\x00\xff")
    else:
        p.write_text("")
    r = subprocess.run([EXE, *args, str(p)], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode in (1, 2), (args, r.returncode, r.stderr[-500:])
