"""PDB table I/O on the GPU: the reference's .tab files out and back in
(src/24puzzle_with_pdb.cc:77-112), and the library's own cache file.  Every
path must give the tables whose checksums tests/golden/pdb_tables.json
records (the tables the reference's PDB solver reproduces
benchmarks/pdb_expand_nodes.csv 50/50 with), and corrupt or foreign files
must be rejected, never used."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import spz

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "sliding-puzzle_amd", "bin", "spz_solve")
NAMES = ("pat24.1256712.tab", "pat24.34891314.tab")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def file_sha(path) -> str:
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


@pytest.fixture(scope="module")
def built(tmp_path_factory, golden):
    d = tmp_path_factory.mktemp("pdb")
    gold = [t["sha256"] for t in golden("pdb_tables.json")["tables"]]
    s = spz.Solver(width=5, heuristic=spz.H_PDB666)
    cache = str(d / "tables.spzpdb")
    say("build with cache", cache)
    s.build_pdb(cache)
    assert os.path.exists(cache) and not os.path.exists(cache + ".tmp")
    assert [sha(s.get_pdb(0)), sha(s.get_pdb(1))] == gold
    yield s, d, cache, gold
    s.close()


def boards24():
    return np.stack([O.walk(5, 60 + 2 * i, 5_500_000 + i) for i in range(40)])


def test_export_and_load_tab(built):
    s, d, _, gold = built
    paths = [str(d / n) for n in NAMES]
    s.export_pdb_tab(*paths)
    assert [file_sha(p) for p in paths] == gold
    b = boards24()
    with spz.Solver(width=5, heuristic=spz.H_PDB666) as t:
        t.load_pdb_tab(*paths)
        assert [sha(t.get_pdb(0)), sha(t.get_pdb(1))] == gold
        assert list(t.eval_h(b)) == list(s.eval_h(b))


def test_cache_round_trip(built):
    s, _, cache, gold = built
    b = boards24()
    with spz.Solver(width=5, heuristic=spz.H_PDB666) as t:
        t.build_pdb(cache)  # read, not built
        assert t.stats()["pdb_build_ms"] == 0
        assert [sha(t.get_pdb(0)), sha(t.get_pdb(1))] == gold
        assert list(t.eval_h(b)) == list(s.eval_h(b))


def test_corrupt_files_rejected(built, tmp_path):
    _, d, cache, _ = built
    data = open(cache, "rb").read()
    bad = tmp_path / "flipped.spzpdb"
    flipped = bytearray(data)
    flipped[64 + 12345678] ^= 1
    bad.write_bytes(bytes(flipped))
    short = tmp_path / "short.spzpdb"
    short.write_bytes(data[:-1])
    foreign = tmp_path / "foreign.spzpdb"
    foreign.write_bytes(b"\0" * len(data))
    tab = [str(d / n) for n in NAMES]
    short_tab = tmp_path / "short.tab"
    short_tab.write_bytes(open(tab[0], "rb").read()[:-7])
    with spz.Solver(width=5, heuristic=spz.H_PDB666) as t:
        for path, what in ((bad, "checksum"), (short, "size"), (foreign, "not a PDB cache")):
            with pytest.raises(spz.SpzError) as e:
                t.build_pdb(str(path))
            assert what in str(e.value), e.value
        with pytest.raises(spz.SpzError):
            t.load_pdb_tab(str(short_tab), tab[1])
        with pytest.raises(spz.SpzError):
            t.load_pdb_tab(tab[0], str(tmp_path / "missing.tab"))
        with pytest.raises(spz.SpzError):  # no tables: the search refuses to run
            t.solve(boards24()[:2])


def test_cli_export_and_load(built, tmp_path, ref_sets, ref_results):
    _, _, _, gold = built
    out = tmp_path / "tabs"
    out.mkdir()
    r = subprocess.run([CLI, "--export-tab", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert [file_sha(str(out / n)) for n in NAMES] == gold
    rows = ref_results["pdb_expand_nodes"]["rows"]
    idx = [i for i, x in enumerate(rows) if x["nodes"] < 20_000_000][:6]
    files = []
    for i in idx:
        p = tmp_path / f"prob{i:03d}"
        p.write_text(" ".join(map(str, ref_sets["yama24_50_hard_new"][i])) + " \n")
        files.append(str(p))
    r = subprocess.run([CLI, "-H", "pdb", "--pdb-tab", str(out), "--batch", "--frontier", "256", *files],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == len(idx)
    for line, i in zip(lines, idx):
        path, ln, nodes = line.split()
        assert (int(ln), int(nodes)) == (rows[i]["len"], rows[i]["nodes"]), line
