"""Solution paths (config record_paths) on the GPU.  A recorded path is
checked by replaying it on the CPU: every move legal, the last board the goal,
the length the instance's optimum (from the oracle fixtures).  Any optimal
path is a correct answer, so paths are not compared with the reference's
(whose DFS order picks one particular path)."""
import sys

import numpy as np
import pytest

import oracle as O
import spz

pytestmark = pytest.mark.gpu

STEP = {0: (-1, 0), 1: (0, 1), 2: (0, -1), 3: (1, 0)}  # up, right, left, down (blank moves)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def replay(board, moves, w):
    """Apply blank moves to a board; None if a move leaves the board."""
    b = list(board)
    z = b.index(0)
    for d in moves:
        dr, dc = STEP[d]
        r, c = divmod(z, w)
        r2, c2 = r + dr, c + dc
        if not (0 <= r2 < w and 0 <= c2 < w):
            return None
        t = r2 * w + c2
        b[z], b[t] = b[t], 0
        z = t
    return b


def check_paths(boards, res, paths, w, opt=None):
    goal = list(range(w * w))
    for i, (b, r, pth) in enumerate(zip(boards, res, paths)):
        assert pth is not None, i
        assert len(pth) == r.opt_len, (i, len(pth), r.opt_len)
        if opt is not None:
            assert r.opt_len == opt[i], i
        assert replay(b, pth, w) == goal, i


@pytest.mark.parametrize("mode", [spz.MODE_ALL, spz.MODE_FIRST])
def test_paths_15_puzzle_walks(golden, mode):
    fx = golden("iterations_c1_walk40.json")["instances"][:40]
    boards = [r["board"] for r in fx]
    with spz.Solver(width=4, mode=mode, record_paths=True) as s:
        res = s.solve(boards)
        paths = s.get_paths()
    check_paths(boards, res, paths, 4, [r["opt"] for r in fx])


def test_paths_korf100_and_results_unchanged(golden):
    fx = golden("iterations_korf100.json")["instances"][:12]
    boards = [r["board"] for r in fx]
    with spz.Solver(width=4) as s:
        plain = [(r.opt_len, r.counts) for r in s.solve(boards)]
    with spz.Solver(width=4, record_paths=True) as s:
        res = s.solve(boards)
        paths = s.get_paths()
    assert [(r.opt_len, r.counts) for r in res] == plain
    check_paths(boards, res, paths, 4, [r["opt"] for r in fx])


def test_paths_shallow_and_solved():
    """A solved board (empty path) and walks short enough that the goal lies
    in the root split's top tree, recorded by the split itself."""
    boards = [list(range(16))] + [O.walk(4, L, 4_200_000 + L) for L in (1, 2, 3, 5, 8, 12)]
    with spz.Solver(width=4, record_paths=True, frontier_target=2048) as s:
        res = s.solve(boards)
        paths = s.get_paths()
    assert res[0].opt_len == 0 and paths[0] == []
    check_paths(boards, res, paths, 4)


def test_paths_through_overflow_and_donation(golden):
    """An 8-level LDS stack forces children into queue items (overflow) and a
    wide machine forces donations: the items' paths stay exact."""
    fx = golden("iterations_c1_walk40.json")["instances"][40:60]
    boards = [r["board"] for r in fx]
    with spz.Solver(width=4, record_paths=True, lds_stack_depth=8, frontier_target=16) as s:
        res = s.solve(boards)
        st = s.stats()
        paths = s.get_paths()
    say("overflow items", st["overflow_items"], "donated", st["donated"])
    assert st["overflow_items"] > 0
    check_paths(boards, res, paths, 4, [r["opt"] for r in fx])


def test_paths_24_puzzle_pdb(ref_sets, ref_results):
    rows = ref_results["pdb_expand_nodes"]["rows"]
    idx = [i for i, r in enumerate(rows) if r["nodes"] < 20_000_000][:10]
    boards = [ref_sets["yama24_50_hard_new"][i] for i in idx]
    boards += [O.walk(5, 100, 7_300_000 + i) for i in range(6)]
    with spz.Solver(width=5, heuristic=spz.H_PDB666, record_paths=True, frontier_target=512) as s:
        s.build_pdb()
        res = s.solve(boards)
        paths = s.get_paths()
    for k, i in enumerate(idx):
        assert res[k].opt_len == rows[i]["len"]
    check_paths(boards, res, paths, 5)


def test_paths_need_recording():
    with spz.Solver(width=4) as s:
        s.solve([O.walk(4, 20, 4_300_000)])
        with pytest.raises(spz.SpzError):
            s.get_paths()
