"""Lookahead (an instance's next iteration started while its current one
drains) changes only the schedule: on the same boards, lengths and every
per-iteration count match the solver that runs one iteration at a time, and
the counts still match the reference's golden values."""
import sys

import numpy as np
import pytest

import oracle as O
import spz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def results(boards, max_iters=64, **kw):
    with spz.Solver(**kw) as s:
        if kw.get("heuristic") == spz.H_PDB666:
            s.build_pdb()
        res = s.solve(boards, max_iters=max_iters)
        return [(r.opt_len, r.n_iters, r.counts) for r in res], s.stats()


def test_lookahead_15_puzzle_batch(golden):
    """Walks and reference instances of mixed depth in one batch."""
    fx = golden("iterations_korf100.json")["instances"][:16]
    boards = [r["board"] for r in fx] + [O.walk(4, L, 5_100_000 + L) for L in range(10, 90, 2)]
    say("15-puzzle batch of", len(boards))
    on, st_on = results(boards, width=4)
    off, st_off = results(boards, width=4, lookahead=False)
    say("lookahead starts", st_on["lookahead_starts"], "off", st_off["lookahead_starts"])
    assert on == off
    assert st_on["lookahead_starts"] > 0 and st_off["lookahead_starts"] == 0
    for rec, (opt, _, counts) in zip(fx, on):
        assert opt == rec["opt"]
        for a, b in zip(counts, rec["counts"]):
            if b is not None:
                assert a == b, rec["idx"]


def test_lookahead_single_instance_per_launch(golden):
    """One instance at a time (its iterations are what overlaps)."""
    fx = golden("iterations_korf100.json")["instances"][16:22]
    for rec in fx:
        on, st = results([rec["board"]], width=4)
        off, _ = results([rec["board"]], width=4, lookahead=False)
        say("instance", rec["idx"], "lookahead starts", st["lookahead_starts"])
        assert on == off, rec["idx"]
        assert on[0][0] == rec["opt"]


def test_lookahead_24_puzzle_pdb(ref_sets, ref_results):
    rows = ref_results["pdb_expand_nodes"]["rows"]
    idx = [i for i, r in enumerate(rows) if r["nodes"] < 30_000_000][:8]
    boards = [ref_sets["yama24_50_hard_new"][i] for i in idx]
    boards += [O.walk(5, 120, 5_200_000 + i) for i in range(8)]
    say("24-puzzle PDB batch of", len(boards))
    kw = dict(width=5, heuristic=spz.H_PDB666, frontier_target=1024)
    on, st = results(boards, **kw)
    off, _ = results(boards, lookahead=False, **kw)
    say("lookahead starts", st["lookahead_starts"])
    assert on == off
    for k, i in enumerate(idx):
        assert on[k][0] == rows[i]["len"]


def test_lookahead_first_solution_mode(golden):
    """First-solution mode: the last iteration stops at a goal, so its count
    depends on the schedule; lengths and every earlier count do not.  Few
    lanes: lookahead starts only after an iteration of one node per lane."""
    fx = golden("iterations_c1_walk40.json")["instances"][:40]
    boards = [r["board"] for r in fx] + [O.walk(4, 70, 5_300_000 + i) for i in range(20)]
    on, st = results(boards, width=4, mode=spz.MODE_FIRST, lanes=8192)
    off, _ = results(boards, width=4, mode=spz.MODE_FIRST, lanes=8192, lookahead=False)
    say("lookahead starts", st["lookahead_starts"])
    assert st["lookahead_starts"] > 0
    for (o1, n1, c1), (o2, n2, c2) in zip(on, off):
        assert (o1, n1) == (o2, n2)
        assert c1[:n1 - 1] == c2[:n2 - 1]


def test_lookahead_iteration_limit():
    """An iteration cap ends instances unsolved the same way with and
    without lookahead (no count of an iteration past the cap).  Random
    boards (iterations of 10^3..10^7 nodes) on few lanes, so that lookahead
    starts up to the cap."""
    boards = [O.random_solvable(4, 5_400_000 + i) for i in range(24)]
    on, st = results(boards, max_iters=6, width=4, lanes=4096)
    off, _ = results(boards, max_iters=6, width=4, lanes=4096, lookahead=False)
    say("lookahead starts", st["lookahead_starts"])
    assert on == off
    assert st["lookahead_starts"] > 0
    assert any(o == -1 for o, _, _ in on)
    assert all(n <= 6 for _, n, _ in on)


def test_lookahead_skips_small_iterations(golden):
    """Instances whose last iteration counted fewer nodes than there are
    lanes start no lookahead past their first iteration: 40-move walks
    (a few hundred thousand nodes each) on the default lane count get at
    most one lookahead start each, and their counts stay the reference's."""
    fx = golden("iterations_c1_walk40.json")["instances"]
    on, st = results([r["board"] for r in fx], width=4)
    say("lookahead starts", st["lookahead_starts"], "instances", len(fx))
    assert st["lookahead_starts"] <= len(fx)
    for rec, (opt, _, counts) in zip(fx, on):
        assert opt == rec["opt"]
        for a, b in zip(counts, rec["counts"]):
            if b is not None:
                assert a == b, rec["idx"]


def test_lookahead_off_under_load_policy():
    """The previous-load dispatch order needs one iteration at a time."""
    boards = [O.walk(4, 50, 5_500_000 + i) for i in range(16)]
    with spz.Solver(width=4, lb_policy=spz.LB_PREVLOAD) as s:
        s.solve(boards)
        assert s.stats()["lookahead_starts"] == 0


def test_round_time_limit_keeps_results(ref_sets, ref_results):
    """PDB rounds that end by time (default), by step budget (off) and by a
    short time limit give the same lengths and per-iteration counts."""
    rows = ref_results["pdb_expand_nodes"]["rows"]
    idx = [i for i, r in enumerate(rows) if r["nodes"] < 30_000_000][:6]
    boards = [ref_sets["yama24_50_hard_new"][i] for i in idx]
    boards += [O.walk(5, 120, 5_600_000 + i) for i in range(6)]
    kw = dict(width=5, heuristic=spz.H_PDB666, frontier_target=1024)
    base, _ = results(boards, **kw)
    off, _ = results(boards, round_time_us=-1, **kw)
    short, st = results(boards, round_time_us=300, **kw)
    say("rounds with a 300 us limit", st["rounds"])
    assert base == off == short
    for k, i in enumerate(idx):
        assert base[k][0] == rows[i]["len"]
