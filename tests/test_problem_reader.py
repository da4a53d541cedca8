"""The Python mirror's read_problem and the CLI's read_problem_file
(sliding-puzzle_amd/host/spz_problem.h) accept the same problem files and
read the same boards from them, on the reference's file layouts and on
random contents: numbers in and out of range, signs, 64-bit overflow,
tokens with trailing junk, and too few integers.  CPU only."""
import os
import shutil
import subprocess

import pytest

import spz

hypothesis = pytest.importorskip("hypothesis")
from hypothesis import HealthCheck, given, settings, strategies as st  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def c_reader(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = tmp_path_factory.mktemp("reader") / "read_problem"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", os.path.join(HERE, "host", "read_problem_main.cpp"),
                    "-o", str(exe)], check=True)
    return str(exe)


def python_view(path):
    try:
        b = spz.read_problem(path)
    except spz.SpzError as e:
        assert e.code == -1
        return "REJECT"
    w = 5 if len(b) == 25 else 4
    return " ".join([str(w)] + [str(int(c)) for c in b])


def compare(c_reader, paths):
    out = subprocess.run([c_reader, *paths], capture_output=True, text=True, check=True).stdout.splitlines()
    assert len(out) == len(paths)
    for p, line in zip(paths, out):
        assert python_view(p) == line, (open(p, "rb").read(), line)


def test_reference_layouts(c_reader, tmp_path, ref_sets):
    files = []
    for i, b in enumerate(ref_sets["korf100"][:3] + ref_sets["korf50_24puzzle"][:3]):
        p = tmp_path / f"p{i}"
        w = int(len(b) ** 0.5)
        # one row per line with a trailing space (Korf100), or one line
        rows = [" ".join(map(str, b[r * w:(r + 1) * w])) for r in range(w)]
        p.write_text(" \n".join(rows) + " \n" if i % 2 else " ".join(map(str, b)))
        files.append(str(p))
    compare(c_reader, files)
    assert all(python_view(f) != "REJECT" for f in files)


TOKENS = st.one_of(st.integers(-3, 30).map(str), st.sampled_from(["+5", "-0", "007", "12abc", "12-3", "1+2", "x",
                                                                  "--3", "9223372036854775807",
                                                                  "9223372036854775808", "-9223372036854775808",
                                                                  "1_0", "0x1", "1e3", "\x00"]))
SEPS = st.sampled_from([" ", "\n", "\t", "\r\n", "  ", "\x0b", "\x0c"])


@settings(max_examples=200, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(files=st.lists(st.tuples(st.lists(TOKENS, max_size=30), SEPS, st.booleans()), min_size=1, max_size=8))
def test_random_contents(c_reader, tmp_path, files):
    paths = []
    for i, (toks, sep, perm) in enumerate(files):
        if perm:  # mostly a valid board, sometimes with junk after it
            n2 = 25 if len(toks) % 2 else 16
            toks = [str((j * 7) % n2) for j in range(n2)] + toks[:3]
        p = tmp_path / f"f{i}"
        p.write_bytes(sep.join(toks).encode("latin-1"))
        paths.append(str(p))
    compare(c_reader, paths)
