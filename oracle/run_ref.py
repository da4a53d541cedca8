"""Runs the reference's own CPU solvers (built by oracle/ref.mk into
oracle/_ref/) on chosen boards.  TEST INFRASTRUCTURE ONLY: used in this
container to write golden fixtures and to cross-check the oracle.

The reference solvers read hard-coded relative paths
(`../benchmarks/<set>/probNNN`, `../result/*.csv`, `../pdb/*.tab`;
src/24puzzle.cc:170-190, src/24puzzle_with_pdb.cc:78-79).  Their sources are
not edited: each run gets a scratch directory tree in which those relative
paths resolve to the boards we want solved.
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile
from typing import Dict, List, Optional, Sequence, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "_ref")


def have(binary: str) -> bool:
    return os.path.exists(os.path.join(REF_BIN, binary))


def _board_text(board: Sequence[int]) -> str:
    return " ".join(str(int(x)) for x in board) + " \n"


_HORIE_LINE = re.compile(r"f_limit=\s*(\d+), expanded nodes = (-?\d+)")


def horie_iterations(board: Sequence[int], timeout: float = 600) -> List[Tuple[int, int]]:
    """Per-iteration (f_limit, count) lines of src_horie/idas_cpu*_expand.c.

    The first printed line reports 0 instead of the first iteration's count
    (src_horie/idas_cpu_expand.c:285-288 prints `nodes_expanded`, not
    `nodes_expanded_first`), so it is returned as-is and callers drop it.
    """
    binary = "idas_cpu_expand" if len(board) == 16 else "idas_cpu_25_expand"
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "prob")
        with open(path, "w") as f:
            f.write(_board_text(board))
        out = subprocess.run([os.path.join(REF_BIN, binary), path],
                             capture_output=True, text=True, timeout=timeout,
                             check=True).stdout
    return [(int(a), int(b)) for a, b in _HORIE_LINE.findall(out)]


def horie_counts_by_threshold(board: Sequence[int], timeout: float = 600) -> Dict[int, int]:
    """Map threshold -> all-solutions count from Horie's solver.

    Its schedule is h0, h0+1, then +1/+2 (src_horie/idas_cpu_expand.c:284-309);
    with Manhattan the h0+1 iteration expands exactly the h0 tree (every f
    has the parity of h0), so count(h0) is read from the h0+1 line.
    """
    lines = horie_iterations(board, timeout)
    res: Dict[int, int] = {}
    if not lines:
        return res
    h0 = lines[0][0]
    if len(lines) == 1:
        # solved in the first iteration: its count is the buggy 0 line
        return {h0: None}
    for f, c in lines[1:]:
        res[f] = c
    if h0 + 1 in res:
        res[h0] = res[h0 + 1]
    return res


def yamamura_batch(binary: str, set_name: str, n_slots: int,
                   boards: Sequence[Sequence[int]], pad_board: Sequence[int],
                   pdb_files: Optional[Tuple[str, str]] = None,
                   timeout: float = 3600, times: Optional[list] = None) -> List[Tuple[int, Optional[int]]]:
    """Run a Yamamura solver whose main loops over
    ../benchmarks/<set_name>/prob000..prob{n_slots-1}.  `boards` fill the
    first slots, `pad_board` the rest.  Returns [(len, nodes|None)] for the
    given boards, parsed from its `<path> <len> [nodes]` lines."""
    if len(boards) > n_slots:
        raise ValueError("too many boards for this solver's loop")
    with tempfile.TemporaryDirectory() as root:
        bdir = os.path.join(root, "benchmarks", set_name)
        os.makedirs(bdir)
        os.makedirs(os.path.join(root, "result"))
        os.makedirs(os.path.join(root, "run"))
        for i in range(n_slots):
            b = boards[i] if i < len(boards) else pad_board
            with open(os.path.join(bdir, "prob%03d" % i), "w") as f:
                f.write(_board_text(b))
        if pdb_files:
            os.makedirs(os.path.join(root, "pdb"))
            os.symlink(os.path.abspath(pdb_files[0]),
                       os.path.join(root, "pdb", "pat24.1256712.tab"))
            os.symlink(os.path.abspath(pdb_files[1]),
                       os.path.join(root, "pdb", "pat24.34891314.tab"))
        out = subprocess.run([os.path.join(REF_BIN, binary)], cwd=os.path.join(root, "run"),
                             capture_output=True, text=True, timeout=timeout,
                             check=True).stdout
        if times is not None:
            # per-instance seconds the solver writes to ../result/*.csv
            for name in sorted(os.listdir(os.path.join(root, "result"))):
                with open(os.path.join(root, "result", name)) as f:
                    vals = [float(x) for x in f.read().split()]
                times.extend(vals[:len(boards)])
    res = []
    for line in out.splitlines():
        toks = line.split()
        if not toks or "/prob" not in toks[0]:
            continue
        idx = int(toks[0].rsplit("prob", 1)[1])
        if idx >= len(boards):
            continue
        ln = int(toks[1])
        nodes = int(toks[2]) if len(toks) > 2 else None
        res.append((ln, nodes))
    return res


# (binary, set directory it loops over, number of slots)
YAMAMURA = {
    "manhattan_all": ("24puzzle_expand", "yama24_50_med", 51),
    "manhattan_first": ("24puzzle", "yama24_50_med", 51),
    "pdb_all": ("24puzzle_expand_with_pdb", "yama24_50_hard_new", 50),
    "pdb_first": ("24puzzle_with_pdb", "yama24_50_med", 51),
}


def yamamura(kind: str, boards, pdb_files=None, timeout: float = 3600, times=None):
    binary, set_name, slots = YAMAMURA[kind]
    goal = list(range(25))
    return yamamura_batch(binary, set_name, slots, boards, goal, pdb_files, timeout, times)


def horie_timed(board: Sequence[int], timeout: float = 600) -> Tuple[float, Optional[int], Optional[int]]:
    """Wall seconds (process start included, a few ms), the all-solutions
    total of the thresholds h0, h0+2, ... and the nodes Horie's solver
    actually generated.  Its schedule runs h0 and then h0+1, which with
    Manhattan expands the same tree (src_horie/idas_cpu_expand.c:284-309),
    and it prints 0 for the first: the h0+1 line is count(h0), so the total
    is the sum of every line after the first, and the work is that plus
    count(h0) once more.  (None, None) when it solved the board in its first
    iteration (no count printed)."""
    import time

    t = time.perf_counter()
    lines = horie_iterations(board, timeout)
    dt = time.perf_counter() - t
    if len(lines) < 2:
        return dt, None, None
    total = sum(c for _, c in lines[1:])
    return dt, total, total + lines[1][1]
