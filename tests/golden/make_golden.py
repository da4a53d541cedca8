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
  pdb_c4_walk160.json,    24-puzzle PDB configurations: lengths and totals of the
  pdb_c5_walk240.json     reference's src/24puzzle_expand_with_pdb.cpp, per-iteration
                          counts of the oracle (sums checked against them)
  iterations_c3_mixed.json  the 15-puzzle batch configuration (walks and the first
                          permutations), from src_horie/idas_cpu_expand.c
  pdb_tables.json         checksums of the two 6-tile tables, recorded only
                          after the reference's own PDB solver
                          (src/24puzzle_expand_with_pdb.cpp) run on them
                          reproduces benchmarks/pdb_expand_nodes.csv

Usage: python tests/golden/make_golden.py [sets|iters15|iters24|c2|pdb|all]
       python tests/golden/make_golden.py c3 | c4 | c5 I,J,...
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


# ------------------------------------------- 24-puzzle PDB configurations ---
def _tab_paths():
    """The two tables in the reference's .tab layout, as make_pdb() writes
    them (checked there against benchmarks/pdb_expand_nodes.csv 50/50)."""
    d = os.path.join(REPO, "oracle", "_ref", "pdb", "blank0")
    paths = (os.path.join(d, "pat24.1256712.tab"), os.path.join(d, "pat24.34891314.tab"))
    if not all(os.path.exists(p) for p in paths):
        raise SystemExit("PDB tables missing: run `make_golden.py pdb` first")
    return paths


def _ref_pdb_job(args):
    """One board through the reference's src/24puzzle_expand_with_pdb.cpp:
    (optimal length, all-solutions cumulative node count, its own seconds)."""
    idx, board, paths = args
    times = []
    (ln, nodes), = R.yamamura("pdb_all", [board], paths, timeout=4 * 3600, times=times)
    return idx, ln, nodes, times[0] if times else None


def _orc_init(paths):
    import numpy as np

    O.pdb_set_tables(np.fromfile(paths[0], dtype=np.uint8), np.fromfile(paths[1], dtype=np.uint8))


def _orc_pdb_job(args):
    idx, board = args
    opt, thr, cnt = O.ida(board, "pdb", "all")
    return idx, opt, thr, cnt


def _cache_path(name, what):
    d = os.path.join(REPO, "oracle", "_ref", "cache")
    os.makedirs(d, exist_ok=True)
    return os.path.join(d, f"{name}.{what}.json")


def _load_cache(name, what):
    try:
        with open(_cache_path(name, what)) as f:
            return {int(k): v for k, v in json.load(f).items()}
    except FileNotFoundError:
        return {}


def _save_cache(name, what, d):
    with open(_cache_path(name, what), "w") as f:
        json.dump(d, f)


def pdb_fixture(name, config, boards, seeds, procs=8):
    """Lengths and totals from the reference's PDB solver, per-iteration
    counts from the oracle's restatement, for seeded 24-puzzle PDB boards.
    The reference counts in a C int (src/24puzzle_expand_with_pdb.cpp), so
    its printed total is the true total modulo 2^32, as a signed value: the
    oracle's u64 sum must equal it modulo 2^32, and the fixture's "total" is
    the oracle's sum.  Each solver's results are kept in oracle/_ref/cache
    as they come in, so an interrupted run resumes."""
    paths = _tab_paths()
    boards = [[int(x) for x in b] for b in boards]
    ref = _load_cache(name, "reference")
    todo = [(i, b, paths) for i, b in enumerate(boards) if i not in ref]
    with mp.Pool(procs) as pool:
        for idx, ln, nodes, secs in pool.imap_unordered(_ref_pdb_job, todo):
            ref[idx] = (ln, nodes, secs)
            _save_cache(name, "reference", ref)
            print(name, "reference", idx, ln, nodes, secs, flush=True)
    orc = _load_cache(name, "oracle")
    todo = [(i, b) for i, b in enumerate(boards) if i not in orc]
    with mp.Pool(procs, initializer=_orc_init, initargs=(paths,)) as pool:
        for idx, opt, thr, cnt in pool.imap_unordered(_orc_pdb_job, todo):
            orc[idx] = (opt, thr, cnt)
            _save_cache(name, "oracle", orc)
            print(name, "oracle", idx, opt, sum(cnt), flush=True)
    recs = []
    for i, b in enumerate(boards):
        ln, nodes, secs = ref[i]
        opt, thr, cnt = orc[i]
        if opt != ln or (sum(cnt) - nodes) % (1 << 32) != 0:
            raise SystemExit(f"{name} board {i}: oracle {opt}/{sum(cnt)} vs reference {ln}/{nodes}")
        recs.append({"idx": i, "seed": seeds[i], "board": b, "h0": thr[0], "opt": opt,
                     "total": sum(cnt), "reference_printed_total": nodes, "thresholds": thr,
                     "counts": cnt, "ref_seconds": secs})
    dump(name, {"config": config,
                "source": "lengths and totals: oracle/_ref/24puzzle_expand_with_pdb "
                          "(src/24puzzle_expand_with_pdb.cpp) on the tables of make_pdb(); "
                          "per-iteration counts: oracle/spz_oracle.c orc_ida (PDB, all "
                          "solutions), whose sums equal the reference's totals modulo 2^32 "
                          "(the reference counts in a C int: reference_printed_total)",
                "note": "ref_seconds: the reference solver's own per-instance time, one core "
                        "of an Intel Xeon @ 2.70 GHz, g++ -O3, 8 processes at once",
                "instances": recs})


def make_c4():
    # C4: 64 non-backtracking 160-move walks on the 5x5 board, seed = 4e6 + i
    seeds = [4_000_000 + i for i in range(64)]
    pdb_fixture("pdb_c4_walk160.json", "c4", [O.walk(5, 160, s) for s in seeds], seeds)


def make_c5(indices):
    # C5: 64 non-backtracking 240-move walks, seed = 5e6 + i; the fixture
    # holds the given instances (the smallest by node count: the whole set is
    # about 3e12 nodes, days of CPU time)
    seeds = [5_000_000 + i for i in indices]
    pdb_fixture("pdb_c5_walk240.json", "c5", [O.walk(5, 240, s) for s in seeds], seeds)
    fx = load("pdb_c5_walk240.json")
    for rec, i in zip(fx["instances"], indices):
        rec["idx"] = i
    dump("pdb_c5_walk240.json", fx)


def c3_walk_length(seed):
    """Walk length of a C3 walk instance: 20 + (first splitmix64 output of
    the seed, high 32 bits) * 61 >> 32, i.e. uniform in 20..80."""
    m = 0xFFFFFFFFFFFFFFFF
    z = (seed + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return 20 + ((z >> 32) * 61 >> 32)


def make_c3(n_random=64):
    # C3: 4096 instances; 2048 walks (seed 3e6 + i, 20..80 moves) and 2048
    # random solvable permutations (seed 3e6 + 2048 + i).  The fixture holds
    # every walk and the first n_random permutations (each about 10 s of CPU).
    boards = [O.walk(4, c3_walk_length(3_000_000 + i), 3_000_000 + i) for i in range(2048)]
    boards += [O.random_solvable(4, 3_000_000 + 2048 + i) for i in range(n_random)]
    horie_fixture("iterations_c3_mixed.json", "c3", boards, timeout=1800)
    fx = load("iterations_c3_mixed.json")
    for rec in fx["instances"]:
        rec["seed"] = 3_000_000 + rec["idx"]
    dump("iterations_c3_mixed.json", fx)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("c4",):
        make_c4()
    if what in ("c3",):
        make_c3()
    if what in ("c5",):
        make_c5([int(x) for x in sys.argv[2].split(",")])
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
