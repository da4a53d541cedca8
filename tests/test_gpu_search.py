"""GPU parity of the HIP IDA* path (libspz through its C ABI) against the
reference's golden counts and the CPU oracle.  Bit-exact: optimal lengths
and every per-iteration node count (all-solutions mode)."""
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


@pytest.fixture(scope="module")
def s15():
    say("create 15-puzzle context")
    with spz.Solver(width=4) as s:
        yield s


@pytest.fixture(scope="module")
def s24():
    say("create 24-puzzle Manhattan context")
    with spz.Solver(width=5, frontier_target=256) as s:
        yield s


def test_goal_and_trivial_boards(s15):
    say("goal / one-move boards")
    goal = list(range(16))
    r = s15.solve([goal])[0]
    oo = O.ida(goal)
    assert (r.opt_len, r.counts) == (oo[0], oo[2]) == (0, [2])
    one = O.walk(4, 1, 5)
    r = s15.solve([one])[0]
    oo = O.ida(one)
    assert (r.opt_len, r.counts) == (oo[0], oo[2])


def test_c1_walks_batched_vs_reference_solver(s15, golden):
    """Config 1 instances (100 seeded 40-move walks), one batch, against the
    per-iteration counts printed by the reference solver."""
    fx = golden("iterations_c1_walk40.json")
    recs = [r for r in fx["instances"] if "error" not in r]
    say("c1 batch", len(recs))
    res = s15.solve([r["board"] for r in recs])
    for rec, r in zip(recs, res):
        check_counts(rec, r)


def test_c1_walks_vs_oracle_fresh_seeds(s15):
    boards = [O.walk(4, 40, 1_000_000 + 500 + i) for i in range(64)]
    say("c1 fresh seeds")
    res = s15.solve(boards)
    for b, r in zip(boards, res):
        opt, thr, cnt = O.ida(b)
        assert (r.opt_len, r.thresholds, r.counts) == (opt, thr, cnt)


def test_korf100_prefix_one_per_launch(s15, golden, ref_results):
    """Reference Korf100 boards one instance per launch (root split into 2048
    frontier nodes), against reference per-iteration counts and lengths."""
    fx = golden("iterations_korf100.json")
    lens = ref_results["result_15puzzle"]["rows"]
    for rec in fx["instances"][:12]:
        if "error" in rec:
            continue
        say("korf100", rec["idx"])
        r = s15.solve([rec["board"]])[0]
        check_counts(rec, r)
        assert r.opt_len == lens[rec["idx"]]["len"]


def test_schedule_invariance(golden):
    """Counts do not depend on the split, lanes, round budget or workgroup."""
    fx = golden("iterations_korf100.json")
    recs = [r for r in fx["instances"][:6] if "error" not in r]
    boards = [r["board"] for r in recs]
    for kw in (dict(frontier_target=1), dict(frontier_target=64, lanes=2048, round_budget=64),
               dict(frontier_target=8192, workgroup_size=64), dict(workgroup_size=128, round_budget=7)):
        say("schedule", kw)
        with spz.Solver(width=4, **kw) as s:
            res = s.solve(boards)
        for rec, r in zip(recs, res):
            check_counts(rec, r)


def test_first_solution_mode_lengths(golden):
    fx = golden("iterations_c1_walk40.json")
    recs = [r for r in fx["instances"][:40] if "error" not in r]
    say("first-solution mode")
    with spz.Solver(width=4, mode=spz.MODE_FIRST) as s:
        res = s.solve([r["board"] for r in recs])
    for rec, r in zip(recs, res):
        assert r.opt_len == rec["opt"]
        # non-final iterations are identical to the all-solutions counts
        assert r.counts[:-1] == rec["counts"][:-1] or None in rec["counts"]


def test_24puzzle_manhattan_vs_csv(s24, ref_sets, ref_results):
    """benchmarks/manhattan_expand_nodes.csv rows (cumulative all-solutions)."""
    rows = ref_results["manhattan_expand_nodes"]["rows"]
    idx = [i for i, r in enumerate(rows) if r["nodes"] < 400_000_000]
    say("24-puzzle manhattan batch", len(idx))
    res = s24.solve([ref_sets["yama24_50_hard_new"][i] for i in idx])
    for i, r in zip(idx, res):
        assert (r.opt_len, r.total_nodes) == (rows[i]["len"], rows[i]["nodes"]), i


def test_24puzzle_manhattan_per_iteration(s24, golden):
    fx = golden("iterations_hard_new_manhattan.json")
    recs = [r for r in fx["instances"] if "error" not in r][:20]
    say("24-puzzle per-iteration", len(recs))
    res = s24.solve([r["board"] for r in recs])
    for rec, r in zip(recs, res):
        check_counts(rec, r)


def test_bad_input_rejected(s15):
    with pytest.raises(spz.SpzError):
        s15.solve([[0, 2, 1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15]])  # tiles 1,2 swapped
    with pytest.raises(spz.SpzError):
        s15.solve([[0] * 16])


def test_repeat_solves_same_context(golden):
    """Back-to-back solves on one context with instances that freeze their
    root split at different depths: buffers from the previous solve must not
    leak into the next (regression test for the split-freeze ordering)."""
    fx = golden("iterations_c1_walk40.json")
    recs = [r for r in fx["instances"] if "error" not in r]
    kx = golden("iterations_korf100.json")["instances"][:8]
    with spz.Solver(width=4, frontier_target=4096) as s:
        for rep in range(3):
            say("repeat", rep)
            res = s.solve([r["board"] for r in kx] + [r["board"] for r in recs])
            for rec, r in zip(kx + recs, res):
                check_counts(rec, r)


def test_queue_overflow_reported(golden):
    """A work queue too small for the batch (1024 items against 20 x 2048
    frontier nodes) fails the solve with a capacity error instead of
    returning counts from overwritten items; the same context then solves
    correctly with the automatic size."""
    fx = golden("iterations_c1_walk40.json")["instances"][:20]
    boards = [r["board"] for r in fx]
    with spz.Solver(width=4, queue_items=1024) as s:
        with pytest.raises(spz.SpzError) as e:
            s.solve(boards)
        assert "queue" in str(e.value)
        s.reconfigure(queue_items=0)
        for rec, r in zip(fx, s.solve(boards)):
            check_counts(rec, r)
