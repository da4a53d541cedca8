"""ctypes layer over liboracle.so -- the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
(libspz) never loads it.  The C source (spz_oracle.c) cites the reference
lines each function restates.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
PDB_ENTRIES = 127_512_000

MODES = {"all": 0, "first": 1}
HEURISTICS = {"manhattan": 0, "pdb": 1}
PDB_GOALS = {"any_blank": 0, "blank0": 1, "no_blank": 2}

_lib = None
_tables_keepalive = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.orc_manhattan.argtypes = [ctypes.c_int, u8p]
        L.orc_is_solvable.argtypes = [ctypes.c_int, u8p]
        L.orc_walk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, u8p]
        L.orc_random_solvable.argtypes = [ctypes.c_int, ctypes.c_uint64, u8p]
        L.orc_rng_u64.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.orc_rng_u64.restype = ctypes.c_uint64
        L.orc_ida.argtypes = [ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                              ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int)]
        L.orc_ida_split.argtypes = [ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_pdb_rank.argtypes = [u8p]
        L.orc_pdb_rank.restype = ctypes.c_uint32
        L.orc_pdb_unrank.argtypes = [ctypes.c_uint32, u8p]
        L.orc_pdb_build.argtypes = [ctypes.c_int, ctypes.c_int, u8p, ctypes.c_int]
        L.orc_pdb_build_tiles.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int]
        L.orc_pdb_set_tables.argtypes = [u8p, u8p]
        L.orc_pdb_h.argtypes = [u8p]
        L.orc_pdb_lookups.argtypes = [u8p, ctypes.POINTER(ctypes.c_int)]
        L.orc_pdb_lookup_spec.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                          u8p, u8p]
        L.orc_pdb_build_small.argtypes = [ctypes.c_int, ctypes.c_int, u8p,
                                          ctypes.c_int, u8p]
        L.orc_rank_small.argtypes = [ctypes.c_int, ctypes.c_int, u8p]
        L.orc_rank_small.restype = ctypes.c_uint32
        _lib = L
    return _lib


def _u8(a) -> Tuple[np.ndarray, ctypes.POINTER(ctypes.c_uint8)]:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.uint8))
    return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def width_of(board) -> int:
    n = len(board)
    w = int(round(n ** 0.5))
    if w * w != n:
        raise ValueError("board is not square")
    return w


def manhattan(board) -> int:
    arr, p = _u8(board)
    return lib().orc_manhattan(width_of(arr), p)


def is_solvable(board) -> bool:
    arr, p = _u8(board)
    return bool(lib().orc_is_solvable(width_of(arr), p))


def walk(width: int, steps: int, seed: int) -> np.ndarray:
    out = np.zeros(width * width, dtype=np.uint8)
    lib().orc_walk(width, steps, seed, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out


def random_solvable(width: int, seed: int) -> np.ndarray:
    out = np.zeros(width * width, dtype=np.uint8)
    lib().orc_random_solvable(width, seed,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out


def rng_u64(seed: int, nth: int) -> int:
    return int(lib().orc_rng_u64(seed, nth))


def ida(board, heuristic: str = "manhattan", mode: str = "all",
        max_iters: int = 64) -> Tuple[int, List[int], List[int]]:
    """Returns (optimal length or -1/-2, thresholds, per-iteration counts)."""
    arr, p = _u8(board)
    counts = (ctypes.c_uint64 * max_iters)()
    thr = (ctypes.c_int * max_iters)()
    nit = ctypes.c_int(0)
    r = lib().orc_ida(width_of(arr), p, HEURISTICS[heuristic], MODES[mode],
                      max_iters, counts, thr, ctypes.byref(nit))
    n = nit.value
    return r, [thr[i] for i in range(n)], [counts[i] for i in range(n)]


def ida_split(board, heuristic: str = "manhattan", target: int = 2048, max_depth: int = 24,
              dedup: bool = False, max_iters: int = 64) -> Tuple[int, List[int], List[int]]:
    """orc_ida's iterations over a root split of `target` frontier nodes, with
    or without duplicate pruning of the top tree (all-solutions mode)."""
    arr, p = _u8(board)
    counts = (ctypes.c_uint64 * max_iters)()
    thr = (ctypes.c_int * max_iters)()
    nit = ctypes.c_int(0)
    r = lib().orc_ida_split(width_of(arr), p, HEURISTICS[heuristic], target, max_depth,
                            1 if dedup else 0, max_iters, counts, thr, ctypes.byref(nit))
    n = nit.value
    return r, [thr[i] for i in range(n)], [counts[i] for i in range(n)]


def pdb_rank(cells) -> int:
    arr, p = _u8(cells)
    return int(lib().orc_pdb_rank(p))


def pdb_unrank(rank: int) -> List[int]:
    out = np.zeros(6, dtype=np.uint8)
    lib().orc_pdb_unrank(rank, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return [int(x) for x in out]


def pdb_build(pattern: int, goal: str = "any_blank", n_threads: int = 0) -> np.ndarray:
    table = np.empty(PDB_ENTRIES, dtype=np.uint8)
    r = lib().orc_pdb_build(pattern, PDB_GOALS[goal],
                            table.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                            n_threads)
    if r != 0:
        raise MemoryError("orc_pdb_build failed")
    return table


def pdb_build_tiles(tiles, goal: str = "any_blank", n_threads: int = 0) -> np.ndarray:
    arr, p = _u8(tiles)
    table = np.empty(PDB_ENTRIES, dtype=np.uint8)
    r = lib().orc_pdb_build_tiles(p, PDB_GOALS[goal],
                                  table.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                  n_threads)
    if r != 0:
        raise MemoryError("orc_pdb_build_tiles failed")
    return table


def pdb_build_small(width: int, tiles, goal: str = "any_blank") -> np.ndarray:
    arr, p = _u8(tiles)
    n2, k = width * width, len(arr)
    size = 1
    for i in range(k):
        size *= n2 - i
    table = np.empty(size, dtype=np.uint8)
    r = lib().orc_pdb_build_small(width, k, p, PDB_GOALS[goal],
                                  table.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    if r != 0:
        raise MemoryError("orc_pdb_build_small failed")
    return table


def rank_small(n2: int, cells) -> int:
    arr, p = _u8(cells)
    return int(lib().orc_rank_small(n2, len(arr), p))


def pdb_set_tables(t0: np.ndarray, t1: np.ndarray) -> None:
    global _tables_keepalive
    t0 = np.ascontiguousarray(t0, dtype=np.uint8)
    t1 = np.ascontiguousarray(t1, dtype=np.uint8)
    _tables_keepalive = (t0, t1)
    lib().orc_pdb_set_tables(t0.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                             t1.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))


def pdb_h(board) -> int:
    arr, p = _u8(board)
    return int(lib().orc_pdb_h(p))


def pdb_lookups(board) -> List[int]:
    arr, p = _u8(board)
    out = (ctypes.c_int * 8)()
    lib().orc_pdb_lookups(p, out)
    return list(out)


def pdb_lookup_spec(k: int):
    tab = ctypes.c_int(0)
    m = np.zeros(25, dtype=np.uint8)
    t = np.zeros(6, dtype=np.uint8)
    lib().orc_pdb_lookup_spec(k, ctypes.byref(tab),
                              m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                              t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return tab.value, [int(x) for x in m], [int(x) for x in t]
