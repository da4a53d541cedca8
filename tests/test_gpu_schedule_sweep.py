"""Schedules drawn at random (seeded) against the oracle: frontier size,
lane count, workgroup size, round step budget, round time limit and
lookahead vary together, on 15-puzzle and 24-puzzle Manhattan walks and
24-puzzle PDB walks.  Every schedule must give the oracle's lengths,
thresholds and per-iteration counts: the schedule decides only where and
when nodes are searched, never which."""
import json
import random
import sys

import pytest

import oracle as O
import spz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sets():
    s = {
        "m4": [O.walk(4, 20 + (i * 7) % 50, 7_000_000 + i) for i in range(32)],
        "m5": [O.walk(5, 30 + (i * 5) % 50, 7_200_000 + i) for i in range(24)],
        "p5": [O.walk(5, 24 + (i * 9) % 76, 7_300_000 + i) for i in range(32)],
    }
    with spz.Solver(width=5, heuristic=spz.H_PDB666) as sv:
        sv.build_pdb()
        O.pdb_set_tables(sv.get_pdb(0), sv.get_pdb(1))
    expect = {k: [O.ida(b, heuristic="pdb" if k == "p5" else "manhattan") for b in v] for k, v in s.items()}
    return s, expect


def draw(rng):
    kw = dict(frontier_target=rng.choice([1, 2, 4, 16, 64, 512, 2048]),
              lanes=rng.choice([0, 256, 1024, 4096, 65536]),
              round_budget=rng.choice([0, 16, 64, 128, 1024, 4096]),
              round_time_us=rng.choice([0, -1, 300]),
              lookahead=rng.choice([True, False]),
              workgroup_size=rng.choice([64, 128, 256]))
    if kw["lanes"] % kw["workgroup_size"]:
        kw["lanes"] = 0
    return kw


@pytest.mark.parametrize("case", range(24))
def test_random_schedule_matches_oracle(sets, case):
    boards, expect = sets
    rng = random.Random(1000 + case)
    k = ("m4", "m5", "p5", "p5")[case % 4]
    kw = draw(rng)
    say("schedule", case, k, json.dumps(kw))
    heur = spz.H_PDB666 if k == "p5" else spz.H_MANHATTAN
    with spz.Solver(width=4 if k == "m4" else 5, heuristic=heur, **kw) as s:
        if heur == spz.H_PDB666:
            s.build_pdb()
        res = s.solve(boards[k])
    for i, (r, (opt, thr, cnt)) in enumerate(zip(res, expect[k])):
        assert (r.opt_len, r.thresholds, r.counts) == (opt, thr, cnt), (k, i, kw)
