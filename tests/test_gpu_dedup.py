"""Duplicate pruning in the root split (config dedup): a top-tree node whose
board a kept node of its instance holds with a path-max f no larger is
dropped.  Node counts then differ from the reference's by design: the pruned
counts are checked bit-exact against the oracle's restatement of the pruned
split (orc_ida_split with dedup, whose unpruned form equals the reference-
pinned orc_ida), and through properties: optimal lengths equal the golden
ones, every iteration runs with the same thresholds, every per-iteration
count is at most the exact one, and recorded solution paths stay optimal."""
import sys

import pytest

import oracle as O
import spz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve(boards, **kw):
    with spz.Solver(**kw) as s:
        if kw.get("heuristic") == spz.H_PDB666:
            s.build_pdb()
        res = s.solve(boards)
        paths = s.get_paths() if kw.get("record_paths") else None
    return res, paths


def check_dedup(exact, dedup):
    pruned = 0
    for a, b in zip(exact, dedup):
        assert (b.opt_len, b.n_iters, b.thresholds) == (a.opt_len, a.n_iters, a.thresholds)
        assert all(y <= x for x, y in zip(a.counts, b.counts)), (a.counts, b.counts)
        pruned += a.total_nodes - b.total_nodes
    return pruned


def test_dedup_15_puzzle(golden):
    fx = golden("iterations_korf100.json")["instances"][:10]
    boards = [r["board"] for r in fx] + [O.walk(4, 60, 5_700_000 + i) for i in range(10)]
    exact, _ = solve(boards, width=4, frontier_target=4096)
    dd, _ = solve(boards, width=4, frontier_target=4096, dedup=True)
    pruned = check_dedup(exact, dd)
    say("nodes pruned", pruned, "of", sum(r.total_nodes for r in exact))
    assert pruned > 0
    for rec, r in zip(fx, dd):
        assert r.opt_len == rec["opt"]


def test_dedup_24_puzzle_pdb(ref_sets, ref_results):
    rows = ref_results["pdb_expand_nodes"]["rows"]
    idx = [i for i, r in enumerate(rows) if r["nodes"] < 30_000_000][:8]
    boards = [ref_sets["yama24_50_hard_new"][i] for i in idx]
    kw = dict(width=5, heuristic=spz.H_PDB666, frontier_target=2048)
    exact, _ = solve(boards, **kw)
    dd, _ = solve(boards, dedup=True, **kw)
    say("nodes pruned", check_dedup(exact, dd))
    for k, i in enumerate(idx):
        assert exact[k].total_nodes == rows[i]["nodes"]
        assert dd[k].opt_len == rows[i]["len"]


def test_dedup_paths_stay_optimal(golden):
    fx = golden("iterations_c1_walk40.json")["instances"][:30]
    boards = [r["board"] for r in fx]
    res, paths = solve(boards, width=4, frontier_target=4096, dedup=True, record_paths=True)
    step = {0: (-1, 0), 1: (0, 1), 2: (0, -1), 3: (1, 0)}
    for rec, r, pth in zip(fx, res, paths):
        assert r.opt_len == rec["opt"] == len(pth)
        b = list(rec["board"])
        z = b.index(0)
        for d in pth:
            rr, cc = divmod(z, 4)
            r2, c2 = rr + step[d][0], cc + step[d][1]
            assert 0 <= r2 < 4 and 0 <= c2 < 4
            t = r2 * 4 + c2
            b[z], b[t] = b[t], 0
            z = t
        assert b == list(range(16))


def test_dedup_is_deterministic():
    """Which duplicate a level keeps does not depend on the split kernel's
    output order: repeated solves give the same counts."""
    boards = [O.walk(4, 70, 5_800_000 + i) for i in range(12)]
    a, _ = solve(boards, width=4, frontier_target=4096, dedup=True)
    b, _ = solve(boards, width=4, frontier_target=4096, dedup=True)
    assert [(r.opt_len, r.counts) for r in a] == [(r.opt_len, r.counts) for r in b]


@pytest.mark.parametrize("target", [64, 4096])
def test_dedup_counts_match_oracle(golden, target):
    """Every per-iteration count of the pruned split against the oracle."""
    fx = golden("iterations_korf100.json")["instances"][:6]
    boards = [r["board"] for r in fx] + [O.walk(4, 50, 5_800_000 + i) for i in range(14)]
    dd, _ = solve(boards, width=4, frontier_target=target, dedup=True)
    for b, r in zip(boards, dd):
        opt, thr, cnt = O.ida_split(b, target=target, dedup=True)
        assert (r.opt_len, r.thresholds, r.counts) == (opt, thr, cnt)


def test_dedup_24_puzzle_manhattan_matches_oracle():
    boards = [O.walk(5, 60, 5_900_000 + i) for i in range(8)]
    dd, _ = solve(boards, width=5, frontier_target=1024, dedup=True)
    for b, r in zip(boards, dd):
        opt, thr, cnt = O.ida_split(b, target=1024, dedup=True)
        assert (r.opt_len, r.thresholds, r.counts) == (opt, thr, cnt)
