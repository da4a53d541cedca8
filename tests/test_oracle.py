"""Pins the CPU oracle (oracle/spz_oracle.c) before it is trusted as the
checker of the HIP solver: against the reference's golden files, against
per-iteration counts printed by the reference's own solvers (fixtures made
by tests/golden/make_golden.py), and against independent brute force."""
import hashlib
import heapq
import itertools
from collections import defaultdict

import numpy as np
import pytest

import oracle as O

DIRS = [(-1, 0), (0, 1), (0, -1), (1, 0)]  # up, right, left, down (blank moves)


def nbr(w, cell, d):
    r, c = divmod(cell, w)
    dr, dc = DIRS[d]
    r2, c2 = r + dr, c + dc
    if 0 <= r2 < w and 0 <= c2 < w:
        return r2 * w + c2
    return -1


def md(w, board):
    board = [int(x) for x in board]
    return sum(abs(i // w - t // w) + abs(i % w - t % w) for i, t in enumerate(board) if t)


# ------------------------------------------------------------------ basics
def test_manhattan_closed_form():
    rng = np.random.default_rng(1)
    for w in (3, 4, 5):
        for _ in range(50):
            b = rng.permutation(w * w)
            assert O.manhattan(b) == md(w, list(b))
    assert O.manhattan(list(range(16))) == 0
    assert O.manhattan(list(range(25))) == 0


def test_solvability_matches_reachability():
    """The parity rule equals the set reached by BFS from the goal (2x3 is
    not square; use 3x3 full space: 9!/2 reachable states)."""
    w = 3
    goal = tuple(range(9))
    seen = {goal}
    frontier = [goal]
    while frontier:
        nxt = []
        for s in frontier:
            b = s.index(0)
            for d in range(4):
                q = nbr(w, b, d)
                if q < 0:
                    continue
                t = list(s)
                t[b], t[q] = t[q], t[b]
                t = tuple(t)
                if t not in seen:
                    seen.add(t)
                    nxt.append(t)
        frontier = nxt
    assert len(seen) == 181440
    rng = np.random.default_rng(2)
    for _ in range(400):
        p = tuple(int(x) for x in rng.permutation(9))
        assert O.is_solvable(p) == (p in seen)


def test_generators_deterministic_and_valid():
    for w in (4, 5):
        a = O.walk(w, 40, 1_000_007)
        b = O.walk(w, 40, 1_000_007)
        assert (a == b).all() and O.is_solvable(a)
        r = O.random_solvable(w, 2_000_003)
        assert sorted(r.tolist()) == list(range(w * w)) and O.is_solvable(r)
    # a walk of L moves has Manhattan distance <= L and parity of L
    for i in range(50):
        b = O.walk(4, 40, 1_000_000 + i)
        assert O.manhattan(b) <= 40 and (O.manhattan(b) - 40) % 2 == 0


# ----------------------------------------------------- brute-force counts
def dp_counts(w, board, limit):
    """Independent restatement of one IDA* iteration's count: propagate the
    number of distinct move sequences (no immediate reversal) reaching each
    (state, last move) layer by layer; a node is expanded when its f fits the
    limit (f is monotone, so ancestors fit too) and then contributes one
    count per generated child, times its multiplicity."""
    layer = {(tuple(int(x) for x in board), -1): 1}
    total, g, found = 0, 0, False
    while layer:
        nxt = defaultdict(int)
        for (s, prev), mult in layer.items():
            h = md(w, s)
            if h == 0:
                found = True
            if g + h > limit:
                continue
            b = s.index(0)
            for d in range(4):
                if prev >= 0 and d == 3 - prev:
                    continue
                q = nbr(w, b, d)
                if q < 0:
                    continue
                total += mult
                t = list(s)
                t[b], t[q] = t[q], t[b]
                nxt[(tuple(t), d)] += mult
        layer = nxt
        g += 1
    return total, found


@pytest.mark.parametrize("seed", range(8))
def test_counts_match_layered_dp_8puzzle(seed):
    b = O.random_solvable(3, 77_000 + seed)
    r, thr, cnt = O.ida(b)
    for L, c in zip(thr, cnt):
        tot, found = dp_counts(3, list(b), L)
        assert tot == c
    assert r == thr[-1]


@pytest.mark.parametrize("seed", range(6))
def test_counts_match_layered_dp_15puzzle_walks(seed):
    b = O.walk(4, 24, 31_000 + seed)
    r, thr, cnt = O.ida(b)
    for L, c in zip(thr, cnt):
        assert dp_counts(4, list(b), L)[0] == c


def test_optimal_length_vs_bfs_8puzzle():
    # BFS distances from the goal over the full 3x3 space
    goal = tuple(range(9))
    dist = {goal: 0}
    frontier = [goal]
    while frontier:
        nxt = []
        for s in frontier:
            b = s.index(0)
            for d in range(4):
                q = nbr(3, b, d)
                if q < 0:
                    continue
                t = list(s)
                t[b], t[q] = t[q], t[b]
                t = tuple(t)
                if t not in dist:
                    dist[t] = dist[s] + 1
                    nxt.append(t)
        frontier = nxt
    for seed in range(40):
        b = O.random_solvable(3, 5_000 + seed)
        assert O.ida(b)[0] == dist[tuple(int(x) for x in b)]


def test_parity_invariant_and_first_mode():
    b = O.walk(4, 40, 1_000_003)
    r_all, thr, c_all = O.ida(b, mode="all")
    r_first, thr_f, c_first = O.ida(b, mode="first")
    assert r_all == r_first
    assert c_all[:-1] == c_first[:-1]  # non-final iterations agree
    assert c_first[-1] <= c_all[-1]


# ------------------------------------------------- reference golden files
def test_manhattan_expand_nodes_csv(ref_sets, ref_results):
    """benchmarks/manhattan_expand_nodes.csv: all-solutions cumulative totals
    of yama24_50_hard_new (rows small enough for a unit test)."""
    rows = ref_results["manhattan_expand_nodes"]["rows"]
    done = 0
    for i, row in enumerate(rows):
        if row["nodes"] > 20_000_000:
            continue
        r, thr, c = O.ida(ref_sets["yama24_50_hard_new"][i])
        assert (r, sum(c)) == (row["len"], row["nodes"]), i
        done += 1
    assert done >= 15


def test_result_24puzzle_manhattan_first(ref_sets, ref_results):
    """test/result_24puzzle.txt:103-153 -- first-solution cumulative counts
    (operator order up, right, left, down) of yama24_50_med."""
    rows = ref_results["result_24puzzle_manhattan_first"]["rows"]
    done = 0
    for i, row in enumerate(rows):
        if row["nodes"] > 20_000_000:
            continue
        r, thr, c = O.ida(ref_sets["yama24_50_med"][i], mode="first")
        assert (r, sum(c)) == (row["len"], row["nodes"]), i
        done += 1
    assert done >= 30


def test_result_24puzzle_lengths(ref_sets, ref_results):
    rows = ref_results["result_24puzzle_len"]["rows"]
    for i, row in enumerate(rows[:30]):
        r, _, _ = O.ida(ref_sets["yama24_50"][i], mode="first")
        assert r == row["len"], i


def test_result_15puzzle_lengths(ref_sets, ref_results, golden):
    """test/result_15puzzle.txt lengths for the Korf100 prefix the per-iteration
    fixture covers."""
    rows = ref_results["result_15puzzle"]["rows"]
    fx = golden("iterations_korf100.json")["instances"]
    for rec in fx:
        if "error" in rec:
            continue
        assert rec["opt"] == rows[rec["idx"]]["len"]


@pytest.mark.parametrize("name", ["iterations_c1_walk40.json", "iterations_korf100.json",
                                  "iterations_hard_new_manhattan.json"])
def test_per_iteration_counts_vs_reference_solver(golden, name):
    """Per-iteration counts printed by the reference's src_horie/idas_cpu*_expand.c."""
    fx = golden(name)
    n = 0
    for rec in fx["instances"]:
        if "error" in rec or sum(x or 0 for x in rec["counts"]) > 60_000_000:
            continue
        r, thr, c = O.ida(rec["board"])
        assert r == rec["opt"]
        assert thr == rec["thresholds"]
        for a, b in zip(c, rec["counts"]):
            assert b is None or a == b, rec["idx"]
        n += 1
    assert n >= 10


def test_manhattan_parity_iteration(golden):
    """count(h0+1) == count(h0): the reference solver's h0+1 line repeats h0."""
    fx = golden("iterations_c1_walk40.json")
    for rec in fx["instances"][:20]:
        r, thr, c = O.ida(rec["board"], max_iters=1)
        assert rec["counts"][0] is None or c[0] == rec["counts"][0]


# -------------------------------------------------------------------- PDB
def bf_pdb(w, tiles, blank_goal0):
    """Brute-force Dijkstra over (pattern cells, blank) with unit cost for a
    pattern-tile move and zero for other blank moves; min over blank."""
    n2 = w * w
    home = tuple(tiles)
    dist = {}
    pq = []
    for b in range(n2):
        if b in home:
            continue
        if not blank_goal0 or b == 0:
            dist[(home, b)] = 0
            heapq.heappush(pq, (0, home, b))
    while pq:
        d, cells, b = heapq.heappop(pq)
        if dist.get((cells, b), 1 << 30) < d:
            continue
        for dd in range(4):
            q = nbr(w, b, dd)
            if q < 0:
                continue
            if q in cells:
                j = cells.index(q)
                c2 = list(cells)
                c2[j] = b
                c2, cost = tuple(c2), 1
            else:
                c2, cost = cells, 0
            if d + cost < dist.get((c2, q), 1 << 30):
                dist[(c2, q)] = d + cost
                heapq.heappush(pq, (d + cost, c2, q))
    best = {}
    for (cells, b), d in dist.items():
        best[cells] = min(best.get(cells, 1 << 30), d)
    return best


@pytest.mark.parametrize("w,tiles", [(3, [1, 2]), (3, [1, 3, 5]), (3, [4, 8, 2]), (4, [1, 2, 5])])
@pytest.mark.parametrize("goal", ["blank0", "any_blank"])
def test_pdb_bfs_small_vs_dijkstra(w, tiles, goal):
    t = O.pdb_build_small(w, tiles, goal)
    best = bf_pdb(w, tiles, goal == "blank0")
    assert len(best) == len(t)
    for cells, d in best.items():
        assert t[O.rank_small(w * w, list(cells))] == d


def test_pdb_rank_is_file_order():
    """rank == position in the nested-loop order of src/24puzzle_with_pdb.cc:92-104
    (checked on a prefix and by bijection on random tuples)."""
    k = 0
    for s in itertools.permutations(range(25), 6):
        assert O.pdb_rank(s) == k
        k += 1
        if k > 3000:
            break
    rng = np.random.default_rng(3)
    for _ in range(200):
        s = [int(x) for x in rng.choice(25, 6, replace=False)]
        assert O.pdb_unrank(O.pdb_rank(s)) == s
    assert O.pdb_rank([24, 23, 22, 21, 20, 19]) == O.PDB_ENTRIES - 1


def test_pdb_lookup_spec_matches_reference_tables():
    """The eight lookups derived from symmetries equal the tile lists and
    maps the reference writes out in hash0..hashref3
    (src/24puzzle_with_pdb.cc:114-189)."""
    rf = [0, 5, 10, 15, 20, 1, 6, 11, 16, 21, 2, 7, 12, 17, 22, 3, 8, 13, 18, 23, 4, 9, 14, 19, 24]
    rot90 = [20, 15, 10, 5, 0, 21, 16, 11, 6, 1, 22, 17, 12, 7, 2, 23, 18, 13, 8, 3, 24, 19, 14, 9, 4]
    rot90rf = [20, 21, 22, 23, 24, 15, 16, 17, 18, 19, 10, 11, 12, 13, 14, 5, 6, 7, 8, 9, 0, 1, 2, 3, 4]
    rot180 = list(range(24, -1, -1))
    rot180rf = [24, 19, 14, 9, 4, 23, 18, 13, 8, 3, 22, 17, 12, 7, 2, 21, 16, 11, 6, 1, 20, 15, 10, 5, 0]
    ident = list(range(25))
    expect = [
        (0, ident, [1, 2, 5, 6, 7, 12]),
        (1, ident, [3, 4, 8, 9, 13, 14]),
        (1, rot180, [21, 20, 16, 15, 11, 10]),
        (1, rot90, [19, 24, 18, 23, 17, 22]),
        (0, rf, [5, 10, 1, 6, 11, 12]),
        (1, rf, [15, 20, 16, 21, 17, 22]),
        (1, rot180rf, [9, 4, 8, 3, 7, 2]),
        (1, rot90rf, [23, 24, 18, 19, 13, 14]),
    ]
    for k, (tab, m, tiles) in enumerate(expect):
        assert O.pdb_lookup_spec(k) == (tab, m, tiles), k


@pytest.fixture(scope="module")
def pdb_tables(golden):
    info = golden("pdb_tables.json")["tables"]
    tabs = []
    for p in (0, 1):
        t = O.pdb_build(p, "blank0", 0)
        assert hashlib.sha256(t.tobytes()).hexdigest() == info[p]["sha256"]
        tabs.append(t)
    O.pdb_set_tables(*tabs)
    return tabs


def test_pdb_tables_checksums_and_invariants(pdb_tables, golden):
    """Table bytes equal the ones the reference's own PDB solver reproduced
    benchmarks/pdb_expand_nodes.csv with (pdb_tables.json); plus invariants:
    value 0 only at the goal placement, parity of the pattern tiles'
    Manhattan distance, value >= that distance."""
    info = golden("pdb_tables.json")["tables"]
    rng = np.random.default_rng(5)
    for p, t in enumerate(pdb_tables):
        hist = np.bincount(t, minlength=256)
        assert {str(v): int(c) for v, c in enumerate(hist) if c} == info[p]["histogram"]
        assert int((t == 0).sum()) == 1
        tiles = info[p]["tiles"]
        for _ in range(300):
            r = int(rng.integers(0, O.PDB_ENTRIES))
            cells = O.pdb_unrank(r)
            m = sum(abs(c // 5 - tl // 5) + abs(c % 5 - tl % 5) for c, tl in zip(cells, tiles))
            assert t[r] >= m and (t[r] - m) % 2 == 0


def test_pdb_heuristic_is_not_consistent(pdb_tables):
    """The 6-6-6-6 PDB heuristic (the larger of two additive partitions,
    each entry the fewest pattern-tile moves over every blank position) can
    change by more than 1 in one move, so a node's path-max f is not its
    g + h and the BFS-layer check of duplicate pruning covers Manhattan
    only.  A board where sliding one tile raises h by 3:"""
    b = np.array([5, 1, 2, 3, 4, 10, 6, 7, 12, 9, 15, 11, 13, 8, 14, 16, 21, 17, 0, 19, 20, 22, 23, 18, 24],
                 dtype=np.uint8)
    c = b.copy()
    c[18], c[13] = c[13], 0  # the blank (cell 18) moves up to cell 13
    assert O.pdb_h(c) - O.pdb_h(b) == 3
    assert abs(O.manhattan(c) - O.manhattan(b)) == 1


def test_pdb_expand_nodes_csv(pdb_tables, ref_sets, ref_results):
    rows = ref_results["pdb_expand_nodes"]["rows"]
    for i, row in enumerate(rows):
        if row["nodes"] > 2_000_000:
            continue
        r, thr, c = O.ida(ref_sets["yama24_50_hard_new"][i], heuristic="pdb")
        assert (r, sum(c)) == (row["len"], row["nodes"]), i


def test_pdb_first_solution_counts(pdb_tables, ref_sets, ref_results):
    """test/result_24puzzle.txt:53-102 (line 52 repeats the Manhattan count of
    line 103 and is not a PDB result)."""
    rows = ref_results["result_24puzzle_pdb_first"]["rows"]
    for i, row in enumerate(rows):
        if i == 0 or row["nodes"] > 2_000_000:
            continue
        r, thr, c = O.ida(ref_sets["yama24_50_med"][i], heuristic="pdb", mode="first")
        assert (r, sum(c)) == (row["len"], row["nodes"]), i


def test_split_restatement_equals_plain_ida():
    """orc_ida_split without pruning is orc_ida's count regrouped (top-tree
    children with path-max f <= L plus the frontier nodes' subtrees): equal
    for every frontier size, 15- and 24-puzzle; with pruning the lengths and
    thresholds are kept and no count rises."""
    boards = [O.walk(4, 40, 1_000_000 + i) for i in range(12)]
    boards += [O.walk(4, 60, 2_100_000 + i) for i in range(3)]
    boards += [O.walk(5, 40, 4_100_000 + i) for i in range(4)]
    for b in boards:
        exact = O.ida(b, max_iters=20)
        if exact[0] < 0:
            continue
        for target in (1, 7, 64, 777, 5000):
            assert O.ida_split(b, target=target, max_iters=20) == exact
            opt, thr, cnt = O.ida_split(b, target=target, dedup=True, max_iters=20)
            assert (opt, thr) == exact[:2]
            assert all(x <= y for x, y in zip(cnt, exact[2]))
        # a shallow depth cap freezes the split early: same counts
        assert O.ida_split(b, target=5000, max_depth=3, max_iters=20) == exact


def pruned_split_brute_force(board, target, max_depth=24):
    """Duplicate-pruned split counts of an 8-puzzle from first principles.
    Manhattan is consistent, so a node's path-max f is g + h and a board
    kept at depth d covers every deeper copy: the kept level d is exactly
    BFS layer d of the puzzle graph, each board once, with previous move the
    smallest direction from a parent in layer d-1.  The frontier is the first
    layer of at least `target` boards (or layer max_depth); iteration L counts
    the children of every top-layer board with d + h <= L plus the plain
    IDA* tree below every frontier board with D + h <= L.  Returns (optimal
    length, thresholds, counts) for thresholds h0, h0+2, ... up to the
    optimum (the BFS distance of the goal)."""
    w = 3
    root = tuple(int(x) for x in board)
    goal = tuple(range(9))

    def moved(b, d):
        z = b.index(0)
        nb = nbr(w, z, d)
        if nb < 0:
            return None
        c = list(b)
        c[z], c[nb] = c[nb], 0
        return tuple(c)

    layers, dist = [[root]], {root: 0}
    D = opt = None
    while D is None or opt is None:
        d = len(layers) - 1
        if D is None and (len(layers[d]) >= target or d >= max_depth):
            D = d
        if opt is None and goal in dist:
            opt = dist[goal]
        if D is not None and opt is not None:
            break
        nxt = []
        for b in layers[d]:
            for m in range(4):
                c = moved(b, m)
                if c is not None and c not in dist:
                    dist[c] = d + 1
                    nxt.append(c)
        layers.append(nxt)

    def prev_of(b, d):
        # the kept copy's move: smallest direction m with a parent p in
        # layer d-1 such that moving p's blank in direction m gives b
        return min(m for m in range(4) if (p := moved(b, 3 - m)) is not None and dist.get(p) == d - 1)

    def tree(b, g, prev, L):
        h = md(w, b)
        if g + h > L:
            return 0
        n = 0
        for m in range(4):
            if prev is not None and m == 3 - prev:
                continue
            c = moved(b, m)
            if c is not None:
                n += 1 + tree(c, g + 1, m, L)
        return n

    h0 = md(w, root)
    thr, cnt = [], []
    for L in range(h0, opt + 1, 2):
        n = 0
        for d in range(D):
            for b in layers[d]:
                if d + md(w, b) <= L:
                    n += sum(moved(b, m) is not None for m in range(4)) - (1 if d else 0)
        for b in layers[D]:
            if D + md(w, b) <= L:
                n += tree(b, D, prev_of(b, D) if D else None, L)
        thr.append(L)
        cnt.append(n)
    return opt, thr, cnt


def test_pruned_split_matches_bfs_brute_force():
    """orc_ida_split with duplicate pruning against the brute force above on
    8-puzzle walks, over frontier sizes that stop the split at depths 0 to
    about 11 and a depth cap: a missed duplicate, a copy kept by the wrong
    previous move or a count taken before pruning each changes a count."""
    checked = 0
    for i in range(8):
        b = O.walk(3, 12 + 2 * i, 8_300_000 + i)
        for target in (1, 3, 10, 40, 150, 600, 2000):
            want = pruned_split_brute_force(b, target)
            assert O.ida_split(b, target=target, dedup=True, max_iters=30) == want, (i, target)
            checked += 1
        want = pruned_split_brute_force(b, 10_000, max_depth=5)
        assert O.ida_split(b, target=10_000, max_depth=5, dedup=True, max_iters=30) == want, i
    assert checked == 56
