"""GPU parity on the benchmark configurations themselves, at their full
sizes: every instance of C2, C3, C4 and C5 is solved by libspz (through its
C ABI) the way bench.py solves it and compared with the reference's own CPU
solvers, bit-exact.

  C2  15-puzzle, 100 random solvable permutations, one instance per launch,
      2048 frontier nodes: every per-iteration count against
      src_horie/idas_cpu_expand.c (iterations_c2_random.json, 77 G nodes).
  C3  15-puzzle, 4096 mixed instances in one batch: the 2048 walks and 64 of
      the permutations against idas_cpu_expand; the other permutations, too
      long for the CPU, through the properties every result must have.
  C4  24-puzzle PDB, 64 walks of 160 moves, 4096 frontier nodes, batched:
      lengths and totals of all 64 against src/24puzzle_expand_with_pdb.cpp,
      per-iteration counts against the oracle (pdb_c4_walk160.json).
  C5  24-puzzle PDB, 64 walks of 240 moves, 8192 frontier nodes: the 36
      smallest (37 G of 2.9 T nodes) against the reference PDB solver, all
      64 through properties.
Boards are regenerated from their seeds by the product's own generators and
checked equal to the fixture's boards first.
"""
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


def check_reference_total(rec, r):
    """Length and total against the reference's PDB solver, whose C int
    counter prints the total modulo 2^32 as a signed value, and the exact
    u64 total (the oracle's, equal to it modulo 2^32)."""
    assert r.opt_len == rec["opt"], rec["idx"]
    assert (r.total_nodes - rec["reference_printed_total"]) % (1 << 32) == 0, rec["idx"]
    assert r.total_nodes == rec["total"], rec["idx"]


def check_properties(board, r, manhattan):
    """What every exact result must satisfy, whatever its size: solved,
    optimal length >= h0 and of h0's parity (every move changes the blank
    cell's row + column parity), thresholds h0, h0 + 2, ..., counts positive
    and non-decreasing (a threshold's tree contains the previous one's)."""
    assert r.status == spz.ST_SOLVED and r.opt_len >= 0
    assert r.opt_len >= r.h0 and (r.opt_len - r.h0) % 2 == 0
    assert r.thresholds == [r.h0 + 2 * k for k in range(r.n_iters)]
    assert r.thresholds[-1] == r.opt_len
    assert all(c > 0 for c in r.counts)
    assert all(a <= b for a, b in zip(r.counts, r.counts[1:]))
    assert sum(r.counts) == r.total_nodes
    if manhattan:
        assert r.h0 == O.manhattan(board)


def test_c2_random_one_per_launch(golden):
    fx = golden("iterations_c2_random.json")
    recs = fx["instances"]
    assert len(recs) == 100 and all("error" not in r for r in recs)
    with spz.Solver(width=4, frontier_target=2048) as s:
        for rec in recs:
            b = spz.gen_random(4, 2_000_000 + rec["idx"])
            assert list(b) == rec["board"]
            r = s.solve([b])[0]
            check_counts(rec, r)
        say("c2: 100 of 100 one per launch")


def c3_boards():
    import bench

    _, _, boards = bench.instances("c3")
    return np.stack(boards)


def test_c3_full_batch(golden):
    fx = golden("iterations_c3_mixed.json")
    boards = c3_boards()
    assert len(boards) == 4096
    say("c3: 4096 instances, one batch")
    with spz.Solver(width=4, frontier_target=16) as s:
        res = s.solve(boards)
    checked = 0
    for rec in fx["instances"]:
        assert list(boards[rec["idx"]]) == rec["board"]
        if "error" in rec:
            continue
        check_counts(rec, res[rec["idx"]])
        checked += 1
    assert checked >= 2048 + 60
    for b, r in zip(boards, res):
        check_properties(b, r, True)
    say("c3:", checked, "against the reference solver, 4096 by properties")


@pytest.fixture(scope="module")
def pdb24():
    say("create 24-puzzle PDB context")
    with spz.Solver(width=5, heuristic=spz.H_PDB666, frontier_target=4096) as s:
        s.build_pdb()
        yield s


@pytest.mark.parametrize("lb,resplit", [("uniform", 0), ("learned", 0), ("prevload", 0),
                                        ("learned", 2), ("prevload", 2)])
def test_c4_full_batch(pdb24, golden, lb, resplit):
    fx = golden("pdb_c4_walk160.json")
    recs = fx["instances"]
    assert len(recs) == 64
    boards = np.stack([spz.gen_walk(5, 160, rec["seed"]) for rec in recs])
    for b, rec in zip(boards, recs):
        assert list(b) == rec["board"]
    pdb24.reconfigure(frontier_target=4096, resplit=resplit)
    if lb == "learned":
        # any parameters: the policy changes the schedule, never the result
        pdb24.mlp_init(7)
        pdb24.set_lb(spz.LB_LEARNED)
    else:
        pdb24.set_lb(spz.LB_PREVLOAD if lb == "prevload" else spz.LB_UNIFORM)
    say("c4: 64 instances,", lb, "resplit", resplit)
    res = pdb24.solve(boards)
    if resplit and lb == "prevload":  # random parameters predict nearly equal sizes
        assert pdb24.stats()["resplit_nodes"] > 0
    pdb24.set_lb(spz.LB_UNIFORM)
    pdb24.reconfigure(resplit=0)
    for rec, r in zip(recs, res):
        check_reference_total(rec, r)
        check_counts(rec, r)
    say("c4:", sum(r.total_nodes for r in res), "nodes, all 64 bit-exact")


def test_c5_full_batch(pdb24, golden):
    fx = golden("pdb_c5_walk240.json")
    boards = np.stack([spz.gen_walk(5, 240, 5_000_000 + i) for i in range(64)])
    pdb24.reconfigure(frontier_target=8192)
    say("c5: 64 instances")
    res = pdb24.solve(boards)
    pdb24.reconfigure(frontier_target=4096)
    for rec in fx["instances"]:
        b = boards[rec["idx"]]
        assert list(b) == rec["board"]
        r = res[rec["idx"]]
        check_reference_total(rec, r)
        check_counts(rec, r)
    h = pdb24.eval_h(boards)
    for b, r, hb in zip(boards, res, h):
        check_properties(b, r, False)
        assert r.h0 == hb
    say("c5:", len(fx["instances"]), "against the reference solver, 64 by properties")
