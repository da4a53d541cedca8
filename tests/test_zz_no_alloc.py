"""The allocation-free entry point spz_eval_h_noalloc (static device scratch,
no hipMalloc) against spz_eval_h, the entry point it is a variant of, and the
oracle: Manhattan on both widths and the PDB heuristic, two calls in a row
with the second smaller (the scratch keeps the first call's contents), and
a batch above the scratch's capacity rejected with an error."""
import sys

import numpy as np
import pytest

import oracle as O
import spz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def boards_of(width, n, seed):
    return np.stack([O.walk(width, 10 + (i * 7) % 90, seed + i) if i % 3 else O.random_solvable(width, seed + i)
                     for i in range(n)])


@pytest.mark.parametrize("width,heur", [(4, spz.H_MANHATTAN), (5, spz.H_MANHATTAN), (5, spz.H_PDB666)])
def test_noalloc_matches_eval_h(width, heur):
    with spz.Solver(width=width, heuristic=heur, frontier_target=16) as s:
        if heur == spz.H_PDB666:
            s.build_pdb()
        for n, seed in ((3000, 7_100_000), (700, 7_200_000)):
            say("noalloc", width, heur, n)
            b = boards_of(width, n, seed)
            got = s.eval_h_noalloc(b)
            assert list(got) == list(s.eval_h(b))
            if heur == spz.H_MANHATTAN:
                assert list(got) == [O.manhattan(x) for x in b]
        # the largest batch the scratch holds, then one board
        b = np.repeat(boards_of(width, 16, 7_300_000), 4096, axis=0)
        assert len(b) == spz.NOALLOC_MAX_BOARDS
        assert list(s.eval_h_noalloc(b)) == list(s.eval_h(b))
        assert list(s.eval_h_noalloc(b[5:6])) == list(s.eval_h(b[5:6]))


def test_noalloc_rejects_oversize_and_bad_boards():
    with spz.Solver(width=4) as s:
        b = np.repeat(boards_of(4, 1, 7_400_000), spz.NOALLOC_MAX_BOARDS + 1, axis=0)
        with pytest.raises(spz.SpzError) as e:
            s.eval_h_noalloc(b)
        assert "at most" in str(e.value)
        with pytest.raises(spz.SpzError):
            s.eval_h_noalloc([[0, 2, 1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15]])
        assert s.eval_h_noalloc(np.zeros((0, 16), dtype=np.uint8)).size == 0
    with spz.Solver(width=5, heuristic=spz.H_PDB666) as s:
        with pytest.raises(spz.SpzError):  # no tables yet
            s.eval_h_noalloc(boards_of(5, 2, 7_500_000))
