"""Heavy-node re-split (config resplit, the reference's static load balancing
src/15puzzle_block_parallel.cu:490-563) changes only which items an
iteration queues: a frontier node expected to be heavy is expanded when the
iteration starts and its children (or grandchildren) that fit are queued in
its place.  Lengths and every per-iteration count stay those of the
reference's solver; recorded paths stay optimal."""
import sys

import numpy as np
import pytest

import oracle as O
import spz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_counts(rec, r):
    assert r.opt_len == rec["opt"], rec["idx"]
    assert r.thresholds == rec["thresholds"], rec["idx"]
    for a, b in zip(r.counts, rec["counts"]):
        if b is not None:
            assert a == b, (rec["idx"], r.counts, rec["counts"])


def replay(board, moves, width):
    b = list(board)
    z = b.index(0)
    for d in moves:
        r, c = divmod(z, width)
        nr, nc = r + (-1, 0, 0, 1)[d], c + (0, 1, -1, 0)[d]
        assert 0 <= nr < width and 0 <= nc < width
        nz = nr * width + nc
        b[z], b[nz] = b[nz], 0
        z = nz
    return b == list(range(width * width))


@pytest.mark.parametrize("policy,factor", [("prevload", 1), ("learned", 1), ("prevload", 3)])
def test_resplit_15_puzzle(golden, policy, factor):
    fx = golden("iterations_korf100.json")["instances"][:10]
    recs = [r for r in fx if "error" not in r]
    boards = [r["board"] for r in recs]
    lb = spz.LB_PREVLOAD if policy == "prevload" else spz.LB_LEARNED
    with spz.Solver(width=4, frontier_target=256, resplit=factor, record_paths=True) as s:
        if lb == spz.LB_LEARNED:
            s.mlp_init(11)  # any parameters: the policy changes the schedule only
        s.set_lb(lb)
        res = s.solve(boards)
        st = s.stats()
        paths = s.get_paths()
    say(policy, factor, "re-split nodes", st["resplit_nodes"], "items", st["resplit_items"])
    assert st["resplit_nodes"] > 0 and st["resplit_items"] > st["resplit_nodes"]
    for rec, r, pth in zip(recs, res, paths):
        check_counts(rec, r)
        assert pth is not None and len(pth) == r.opt_len and replay(rec["board"], pth, 4)


def test_resplit_all_frontier_nodes(golden):
    """resplit = 1 under the learned policy with every prediction equal
    (zero parameters): no node exceeds the mean, nothing is split; the
    result is the same either way."""
    fx = golden("iterations_c1_walk40.json")["instances"][:30]
    recs = [r for r in fx if "error" not in r]
    with spz.Solver(width=4, frontier_target=64, resplit=1) as s:
        s.mlp_set_params(np.zeros(s.mlp_param_count(), dtype=np.float32))
        s.set_lb(spz.LB_LEARNED)
        res = s.solve([r["board"] for r in recs])
        assert s.stats()["resplit_nodes"] == 0
    for rec, r in zip(recs, res):
        check_counts(rec, r)


def test_resplit_24_puzzle_pdb(ref_sets, ref_results):
    rows = ref_results["pdb_expand_nodes"]["rows"]
    idx = [i for i, r in enumerate(rows) if r["nodes"] < 30_000_000][:10]
    boards = [ref_sets["yama24_50_hard_new"][i] for i in idx]
    with spz.Solver(width=5, heuristic=spz.H_PDB666, frontier_target=512, resplit=1,
                    record_paths=True) as s:
        s.build_pdb()
        for lb in (spz.LB_PREVLOAD, spz.LB_LEARNED):
            if lb == spz.LB_LEARNED:
                s.mlp_init(5)
            s.set_lb(lb)
            res = s.solve(boards)
            st = s.stats()
            paths = s.get_paths()
            say("pdb lb", lb, "re-split nodes", st["resplit_nodes"])
            if lb == spz.LB_PREVLOAD:  # random parameters predict nearly uniform sizes
                assert st["resplit_nodes"] > 0
            for i, r, pth in zip(idx, res, paths):
                assert (r.opt_len, r.total_nodes) == (rows[i]["len"], rows[i]["nodes"]), i
                assert replay(ref_sets["yama24_50_hard_new"][i], pth, 5)


def test_resplit_24_puzzle_manhattan():
    boards = [O.walk(5, 70, 5_400_000 + i) for i in range(6)]
    with spz.Solver(width=5, frontier_target=256, resplit=1, lookahead=False) as s:
        s.set_lb(spz.LB_PREVLOAD)
        res = s.solve(boards)
        assert s.stats()["resplit_nodes"] > 0
    for b, r in zip(boards, res):
        opt, thr, cnt = O.ida(b)
        assert (r.opt_len, r.thresholds, r.counts) == (opt, thr, cnt)
