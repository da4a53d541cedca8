"""Writes the golden fixtures under tests/golden/ (run in a container that
has the reference checkout at /root/reference; the fixtures are committed,
the GPU box never reads the reference).

Fixtures (all data: boards and expected outputs):
  reference_sets.json     problem boards of the reference's benchmark sets
                          (benchmarks/korf100, yama24_50*, korf50_24puzzle)
  reference_results.json  expected outputs the reference itself holds:
                          benchmarks/manhattan_expand_nodes.csv,
                          benchmarks/pdb_expand_nodes.csv,
                          test/result_15puzzle.txt, test/result_24puzzle.txt
  iterations_*.json       per-iteration all-solutions counts printed by the
                          reference's src_horie/idas_cpu_expand.c and
                          idas_cpu_25_expand.c (built by oracle/ref.mk) on
                          seeded instances of the benchmark configurations
  pdb_tables.json         checksums of the two 6-tile tables, recorded only
                          after the reference's own PDB solver
                          (src/24puzzle_expand_with_pdb.cpp) run on them
                          reproduces benchmarks/pdb_expand_nodes.csv

Usage: python tests/golden/make_golden.py [sets|iters15|iters24|c2|pdb|all]
"""
from __future__ import annotations

import hashlib
import json
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle as O  # noqa: E402
import run_ref as R  # noqa: E402


def rd(path):
    with open(path) as f:
        return [int(x) for x in f.read().split()]


def dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, separators=(",", ":"))
        f.write("\n")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------- sets ---
def make_sets():
    sets = {}
    for name, n in [("korf100", 100), ("yama24_50", 51), ("yama24_50_easy", 51),
                    ("yama24_50_med", 51), ("yama24_50_hard", 51),
                    ("yama24_50_hard_new", 50)]:
        sets[name] = [rd(f"{REF}/benchmarks/{name}/prob{i:03d}") for i in range(n)]
    sets["korf50_24puzzle"] = [rd(f"{REF}/benchmarks/korf50_24puzzle/{i:03d}")
                               for i in range(1, 51)]
    dump("reference_sets.json", {"source": "benchmarks/<set>/probNNN", "sets": sets})

    def rows(path):
        out = []
        with open(path) as f:
            for line in f.read().splitlines():
                toks = line.split()
                if len(toks) < 2:
                    continue
                rec = {"path": toks[0], "len": int(toks[1])}
                if len(toks) > 2:
                    rec["nodes"] = int(toks[2])
                out.append(rec)
        return out

    r24 = rows(f"{REF}/test/result_24puzzle.txt")
    dump("reference_results.json", {
        "manhattan_expand_nodes": {
            "source": "benchmarks/manhattan_expand_nodes.csv",
            "rows": rows(f"{REF}/benchmarks/manhattan_expand_nodes.csv")},
        "pdb_expand_nodes": {
            "source": "benchmarks/pdb_expand_nodes.csv",
            "rows": rows(f"{REF}/benchmarks/pdb_expand_nodes.csv")},
        "result_15puzzle": {"source": "test/result_15puzzle.txt",
                            "rows": rows(f"{REF}/test/result_15puzzle.txt")},
        "result_24puzzle_len": {"source": "test/result_24puzzle.txt:1-51",
                                "rows": r24[0:51]},
        "result_24puzzle_pdb_first": {
            "source": "test/result_24puzzle.txt:52-102",
            "note": "line 52 repeats the Manhattan value of line 103",
            "rows": r24[51:102]},
        "result_24puzzle_manhattan_first": {
            "source": "test/result_24puzzle.txt:103-153", "rows": r24[102:153]},
    })


# --------------------------------------------------- per-iteration counts ---
def _horie_job(args):
    tag, idx, board, timeout = args
    try:
        by_thr = R.horie_counts_by_threshold(board, timeout=timeout)
    except Exception as e:  # timeout: recorded as missing
        return tag, idx, board, None, repr(e)
    return tag, idx, board, by_thr, None


def horie_fixture(name, config, boards, timeout, procs=7):
    jobs = [(config, i, [int(x) for x in b], timeout) for i, b in enumerate(boards)]
    recs = []
    with mp.Pool(procs) as pool:
        for tag, i, board, by_thr, err in pool.imap_unordered(_horie_job, jobs):
            if by_thr is None:
                recs.append({"idx": i, "board": board, "error": err})
                continue
            h0 = min(by_thr)
            thr = [t for t in sorted(by_thr) if (t - h0) % 2 == 0]
            counts = [by_thr[t] for t in thr]
            recs.append({"idx": i, "board": board, "h0": h0,
                         "opt": thr[-1], "thresholds": thr, "counts": counts})
            print(name, i, "opt", thr[-1], "total",
                  sum(c for c in counts if c is not None), flush=True)
    recs.sort(key=lambda r: r["idx"])
    dump(name, {"config": config,
                "source": "oracle/_ref/" + ("idas_cpu_expand" if len(boards[0]) == 16
                                            else "idas_cpu_25_expand"),
                "note": "per-iteration all-solutions counts at thresholds h0, h0+2, ...; "
                        "count(h0) read from the h0+1 line (Manhattan parity); null when "
                        "the reference solver prints no count for that iteration "
                        "(optimum found at h0: src_horie/idas_cpu_expand.c:285-290)",
                "instances": recs})


def make_iters15():
    # C1: 100 non-backtracking 40-move walks, seed = 1e6 + i
    boards = [O.walk(4, 40, 1_000_000 + i) for i in range(100)]
    horie_fixture("iterations_c1_walk40.json", "c1", boards, timeout=600)
    sets = load("reference_sets.json")["sets"]
    horie_fixture("iterations_korf100.json", "korf100", sets["korf100"][:40], timeout=300)


def make_c2():
    # C2: 100 random solvable permutations, seed = 2e6 + i
    boards = [O.random_solvable(4, 2_000_000 + i) for i in range(100)]
    horie_fixture("iterations_c2_random.json", "c2", boards, timeout=900)


def make_iters24():
    sets = load("reference_sets.json")["sets"]
    res = load("reference_results.json")["manhattan_expand_nodes"]["rows"]
    pick = [i for i, r in enumerate(res) if r["nodes"] < 300_000_000]
    boards = [sets["yama24_50_hard_new"][i] for i in pick]
    horie_fixture("iterations_hard_new_manhattan.json", "hard_new", boards, timeout=600)
    # keep the original set indices
    fx = load("iterations_hard_new_manhattan.json")
    for rec in fx["instances"]:
        rec["set_idx"] = pick[rec["idx"]]
    dump("iterations_hard_new_manhattan.json", fx)


# ------------------------------------------------------------------ pdb ---
def make_pdb():
    import numpy as np

    outdir = os.path.join(REPO, "oracle", "_ref", "pdb", "blank0")
    os.makedirs(outdir, exist_ok=True)
    paths = [os.path.join(outdir, "pat24.1256712.tab"),
             os.path.join(outdir, "pat24.34891314.tab")]
    tables = []
    for p, path in enumerate(paths):
        t = O.pdb_build(p, "blank0", 8)
        t.tofile(path)
        tables.append(t)
    sets = load("reference_sets.json")["sets"]
    gold = load("reference_results.json")["pdb_expand_nodes"]["rows"]
    got = R.yamamura("pdb_all", sets["yama24_50_hard_new"], tuple(paths))
    match = sum(1 for g, r in zip(gold, got) if (g["len"], g["nodes"]) == r)
    if match != len(gold):
        raise SystemExit(f"reference PDB solver on these tables: {match}/{len(gold)}")
    info = []
    for p, t in enumerate(tables):
        hist = np.bincount(t, minlength=256)
        info.append({"pattern": p, "tiles": [int(x) for x in
                                             ([1, 2, 5, 6, 7, 12] if p == 0
                                              else [3, 4, 8, 9, 13, 14])],
                     "entries": int(t.size),
                     "sha256": hashlib.sha256(t.tobytes()).hexdigest(),
                     "histogram": {str(v): int(c) for v, c in enumerate(hist) if c},
                     "max": int(t.max())})
    dump("pdb_tables.json", {
        "goal": "pattern tiles home, blank at cell 0; value = min over blank cells "
                "of pattern-tile moves (0-1 BFS)",
        "file_layout": "src/24puzzle_with_pdb.cc:87-112 nested-loop order",
        "pinned_by": f"oracle/_ref/24puzzle_expand_with_pdb on these tables reproduces "
                     f"benchmarks/pdb_expand_nodes.csv {match}/{len(gold)}",
        "tables": info})


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("sets", "all"):
        make_sets()
    if what in ("pdb", "all"):
        make_pdb()
    if what in ("iters15", "all"):
        make_iters15()
    if what in ("iters24", "all"):
        make_iters24()
    if what in ("c2", "all"):
        make_c2()
