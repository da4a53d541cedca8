"""GPU pattern databases and the PDB search path.

The two canonical tables are built on the GPU by 0-1 BFS and must be
byte-identical to the tables whose checksums tests/golden/pdb_tables.json
records: those are the tables on which the reference's own PDB solver
(src/24puzzle_expand_with_pdb.cpp) reproduced benchmarks/pdb_expand_nodes.csv
50/50.  The search is then checked against that CSV and the oracle."""
import hashlib
import sys

import numpy as np
import pytest

import oracle as O
import spz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pdb_solver(golden):
    say("create 24-puzzle PDB context and build tables")
    s = spz.Solver(width=5, heuristic=spz.H_PDB666, frontier_target=256)
    s.build_pdb()
    st = s.stats()
    say("pdb build ms", st["pdb_build_ms"], "levels", st["pdb_levels0"], st["pdb_levels1"])
    info = golden("pdb_tables.json")["tables"]
    tabs = []
    for p in (0, 1):
        t = s.get_pdb(p)
        assert hashlib.sha256(t.tobytes()).hexdigest() == info[p]["sha256"], p
        tabs.append(t)
    O.pdb_set_tables(*tabs)
    yield s
    s.close()


def test_gpu_tables_match_pinned_checksums(pdb_solver):
    # the fixture already compared both tables; the oracle now uses them
    assert O.pdb_h(list(range(25))) == 0


def test_pdb_h_matches_oracle(pdb_solver):
    boards = [O.random_solvable(5, 9_000 + i) for i in range(200)]
    boards += [O.walk(5, 160, 4_000_000 + i) for i in range(100)]
    h = pdb_solver.eval_h(boards)
    for b, x in zip(boards, h):
        assert int(x) == O.pdb_h(b)


def test_pdb_expand_nodes_csv(pdb_solver, ref_sets, ref_results):
    """All 50 rows of benchmarks/pdb_expand_nodes.csv in one batch."""
    rows = ref_results["pdb_expand_nodes"]["rows"]
    say("hard_new PDB batch")
    res = pdb_solver.solve(ref_sets["yama24_50_hard_new"])
    for i, (r, row) in enumerate(zip(res, rows)):
        assert (r.opt_len, r.total_nodes) == (row["len"], row["nodes"]), i


def test_pdb_per_iteration_vs_oracle(pdb_solver, ref_sets):
    boards = [ref_sets["yama24_50_hard_new"][i] for i in (0, 1, 2, 3, 4, 6, 8)]
    boards += [O.walk(5, 100, 4_500_000 + i) for i in range(6)]
    say("pdb per-iteration")
    res = pdb_solver.solve(boards)
    for b, r in zip(boards, res):
        opt, thr, cnt = O.ida(b, heuristic="pdb")
        assert (r.opt_len, r.thresholds, r.counts) == (opt, thr, cnt)


def test_pdb_first_solution_lengths(pdb_solver, ref_sets, ref_results):
    rows = ref_results["result_24puzzle_pdb_first"]["rows"]
    with spz.Solver(width=5, heuristic=spz.H_PDB666, mode=spz.MODE_FIRST,
                    frontier_target=64) as s:
        s.build_pdb()
        res = s.solve(ref_sets["yama24_50_med"])
    for i, (r, row) in enumerate(zip(res, rows)):
        assert r.opt_len == row["len"], i


def test_pdb_root_item_with_four_children(pdb_solver):
    """The root as the only frontier item: it can have four passing
    children, one more than a lane's sibling-value cache holds.  Short walks
    with the blank away from the edges, long rounds, few lanes (one lane
    then visits all four children in one round): every per-iteration count
    equals the oracle's."""
    boards = []
    for seed in range(6_100_000, 6_100_400):
        b = O.walk(5, 8 + seed % 17, seed)
        r, c = divmod(int(np.argmax(np.asarray(b) == 0)), 5)
        if 0 < r < 4 and 0 < c < 4:
            boards.append(b)
        if len(boards) == 48:
            break
    expect = [O.ida(b, heuristic="pdb") for b in boards]
    for kw in (dict(lanes=256, workgroup_size=64), dict()):
        say("pdb root items", kw)
        with spz.Solver(width=5, heuristic=spz.H_PDB666, frontier_target=1, round_budget=4096,
                        round_time_us=-1, **kw) as s:
            s.build_pdb()
            res = s.solve(boards)
        for k, (r, (opt, thr, cnt)) in enumerate(zip(res, expect)):
            assert (r.opt_len, r.thresholds, r.counts) == (opt, thr, cnt), k


def test_pdb_schedule_invariance(pdb_solver, ref_sets, ref_results):
    rows = ref_results["pdb_expand_nodes"]["rows"]
    idx = [i for i, r in enumerate(rows) if r["nodes"] < 5_000_000][:12]
    boards = [ref_sets["yama24_50_hard_new"][i] for i in idx]
    for kw in (dict(frontier_target=1), dict(frontier_target=4096, round_budget=16),
               dict(workgroup_size=64, lanes=4096)):
        say("pdb schedule", kw)
        with spz.Solver(width=5, heuristic=spz.H_PDB666, **kw) as s:
            s.build_pdb()
            res = s.solve(boards)
        for i, r in zip(idx, res):
            assert (r.opt_len, r.total_nodes) == (rows[i]["len"], rows[i]["nodes"]), i
