"""spz -- Python binding of libspz, the MI355X IDA* solver (ctypes over the C ABI
declared in include/spz.h).

The reference is a set of solver processes: read a problem file of N*N
integers (0 = blank), run IDA* with thresholds h0, h0+2, ..., print
"<path> <len> <nodes>" (src/24puzzle.cc:72-87,146-198).  This module offers
the same surface for a whole batch on the GPU:

    with Solver(width=4) as s:
        res = s.solve(boards)            # list of Result
        res[0].opt_len, res[0].counts    # per-iteration node counts

The native library must be built (`python sliding-puzzle_amd/build.py`);
there is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes
import re
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libspz.so")

H_MANHATTAN, H_PDB666 = 0, 1
MODE_ALL, MODE_FIRST = 0, 1
LB_UNIFORM, LB_LEARNED, LB_PREVLOAD = 0, 1, 2
ST_SOLVED, ST_ITER_LIMIT = 0, 1
NOALLOC_MAX_BOARDS = 65536  # SPZ_NOALLOC_MAX_BOARDS
MAX_ITERS = 128

ERRORS = {-1: "bad argument", -2: "HIP error", -3: "no PDB", -4: "queue capacity",
          -5: "I/O error", -6: "unsolvable", -7: "stalled", -8: "timeout",
          -9: "autotune result mismatch", -10: "host failure"}


class SpzError(RuntimeError):
    """A libspz call failed; `code` is its SPZ_E_* return code."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("heuristic", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("frontier_target", ctypes.c_int32),
                ("max_split_depth", ctypes.c_int32), ("workgroup_size", ctypes.c_int32),
                ("lanes", ctypes.c_int32), ("round_budget", ctypes.c_int32),
                ("lds_stack_depth", ctypes.c_int32),
                ("lb_policy", ctypes.c_int32), ("device", ctypes.c_int32),
                ("time_kernels", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("log_samples", ctypes.c_int32),
                ("queue_items", ctypes.c_int64), ("record_paths", ctypes.c_int32),
                ("lookahead", ctypes.c_int32), ("round_time_us", ctypes.c_int32),
                ("dedup", ctypes.c_int32), ("resplit", ctypes.c_int32)]


class _Result(ctypes.Structure):
    _fields_ = [("opt_len", ctypes.c_int32), ("h0", ctypes.c_int32),
                ("n_iters", ctypes.c_int32), ("status", ctypes.c_int32),
                ("total_nodes", ctypes.c_uint64)]


@dataclass
class Result:
    opt_len: int
    h0: int
    n_iters: int
    status: int
    total_nodes: int
    thresholds: List[int] = field(default_factory=list)
    counts: List[int] = field(default_factory=list)


_lib = None

# every entry point declared in include/spz.h
EXPORTS = ["spz_default_config", "spz_create", "spz_build_pdb", "spz_load_pdb_tab", "spz_export_pdb_tab",
           "spz_get_pdb", "spz_solve_batch", "spz_eval_h", "spz_set_lb", "spz_last_stats",
           "spz_last_error", "spz_destroy", "spz_pdb_geometry", "spz_gen_walk",
           "spz_gen_random", "spz_gather_probe", "spz_mlp_param_count", "spz_mlp_init",
           "spz_mlp_set_params", "spz_mlp_get_params", "spz_mlp_forward", "spz_mlp_grad",
           "spz_mlp_train", "spz_get_samples", "spz_reconfigure", "spz_get_config",
           "spz_autotune", "spz_mlp_predict", "spz_get_paths", "spz_eval_h_noalloc"]


def lib() -> ctypes.CDLL:
    """Load libspz.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SpzError(f"{LIB_PATH} is missing: build it with sliding-puzzle_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp, u8p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8)
    L.spz_default_config.argtypes = [ctypes.POINTER(Config), ctypes.c_int32, ctypes.c_int32]
    L.spz_default_config.restype = None
    L.spz_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Config)]
    L.spz_build_pdb.argtypes = [vp, ctypes.c_char_p]
    L.spz_export_pdb_tab.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
    L.spz_load_pdb_tab.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
    L.spz_get_pdb.argtypes = [vp, ctypes.c_int32, u8p]
    L.spz_solve_batch.argtypes = [vp, u8p, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.POINTER(_Result), ctypes.POINTER(ctypes.c_uint64)]
    L.spz_eval_h.argtypes = [vp, u8p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.spz_eval_h_noalloc.argtypes = [vp, u8p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.spz_set_lb.argtypes = [vp, ctypes.c_int32]
    L.spz_last_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.spz_last_error.argtypes = [vp]
    L.spz_last_error.restype = ctypes.c_char_p
    L.spz_destroy.argtypes = [vp]
    L.spz_destroy.restype = None
    L.spz_gather_probe.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double)]
    f32p, i32p = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32)
    L.spz_mlp_param_count.argtypes = []
    L.spz_mlp_param_count.restype = ctypes.c_int32
    L.spz_mlp_init.argtypes = [vp, ctypes.c_uint64]
    L.spz_mlp_set_params.argtypes = [vp, f32p]
    L.spz_mlp_get_params.argtypes = [vp, f32p]
    L.spz_mlp_forward.argtypes = [vp, u8p, i32p, ctypes.c_int32, f32p]
    L.spz_mlp_predict.argtypes = [vp, u8p, i32p, ctypes.c_int32, f32p]
    L.spz_get_paths.argtypes = [vp, u8p, ctypes.c_int32, i32p]
    L.spz_mlp_grad.argtypes = [vp, u8p, i32p, f32p, ctypes.c_int32, f32p, f32p]
    L.spz_mlp_train.argtypes = [vp, u8p, i32p, f32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                ctypes.c_float, ctypes.c_uint64, f32p]
    L.spz_get_samples.argtypes = [vp, u8p, i32p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int32, i32p]
    L.spz_reconfigure.argtypes = [vp, ctypes.POINTER(Config)]
    L.spz_get_config.argtypes = [vp, ctypes.POINTER(Config)]
    L.spz_get_config.restype = None
    L.spz_autotune.argtypes = [vp, u8p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                               ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_int32)]
    L.spz_pdb_geometry.argtypes = [u8p, u8p]
    L.spz_pdb_geometry.restype = None
    L.spz_gen_walk.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, u8p]
    L.spz_gen_walk.restype = ctypes.c_int32
    L.spz_gen_random.argtypes = [ctypes.c_int32, ctypes.c_uint64, u8p]
    L.spz_gen_random.restype = ctypes.c_int32
    _lib = L
    return L


def _u8p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def gen_walk(width: int, steps: int, seed: int) -> np.ndarray:
    out = np.zeros(max(width, 0) ** 2, dtype=np.uint8)
    if lib().spz_gen_walk(width, steps, seed, _u8p(out)) != 0:
        raise ValueError(f"spz_gen_walk: width {width} not in 2..5 or steps {steps} < 0")
    return out


def gen_random(width: int, seed: int) -> np.ndarray:
    out = np.zeros(max(width, 0) ** 2, dtype=np.uint8)
    if lib().spz_gen_random(width, seed, _u8p(out)) != 0:
        raise ValueError(f"spz_gen_random: width {width} not in 2..5")
    return out


def pdb_geometry():
    tiles = np.zeros(48, dtype=np.uint8)
    maps = np.zeros(200, dtype=np.uint8)
    lib().spz_pdb_geometry(_u8p(tiles), _u8p(maps))
    return tiles.reshape(8, 6), maps.reshape(8, 25)


def read_problem(path: str, width: Optional[int] = None) -> np.ndarray:
    """Parse a reference problem file: N*N whitespace-separated integers,
    0 = blank (src/24puzzle.cc:72-87), by the CLI's rules: integers are read
    up to the first other token; 25 or more make a 5x5 board, 16 or more a
    4x4 one (the rest is ignored); a given width must match; every cell must
    lie in 0..N*N-1.  Any other file raises SpzError (SPZ_E_ARG)."""
    vals = []
    with open(path, "rb") as f:
        data = f.read()
    # as a C++ stream reads 64-bit integers: each read skips whitespace and
    # takes an optionally signed run of digits; reading stops at the first
    # read that finds none or overflows ("12-3" gives 12 and -3, "12abc" 12)
    for tok in data.split():
        pos, ok = 0, True
        while pos < len(tok) and len(vals) < 25:
            m = re.match(rb"[+-]?[0-9]+", tok[pos:])
            if not m or not -(1 << 63) <= int(m.group()) < (1 << 63):
                ok = False
                break
            vals.append(int(m.group()))
            pos += m.end()
        if not ok or len(vals) == 25:
            break
    w = 5 if len(vals) >= 25 else 4 if len(vals) >= 16 else 0
    if w == 0 or (width is not None and width != w):
        raise SpzError(f"{path}: not a {width or 4}x{width or 4} or 5x5 board "
                       f"({len(vals)} leading integers)", -1)
    cells = vals[:w * w]
    if any(c < 0 or c >= w * w for c in cells):
        raise SpzError(f"{path}: cells must lie in 0..{w * w - 1}", -1)
    return np.array(cells, dtype=np.uint8)


def format_line(path: str, res: Result) -> str:
    """The reference's result line: "<path> <len> <cumulative nodes>"
    (src/24puzzle_expand.cc:154, src/24puzzle_expand_with_pdb.cpp)."""
    return f"{path} {res.opt_len} {res.total_nodes}"


class Solver:
    """One device context (see spz_create).  Not thread-safe."""

    def __init__(self, width: int = 4, heuristic: int = H_MANHATTAN, mode: int = MODE_ALL,
                 frontier_target: int = 2048, max_split_depth: int = 24,
                 workgroup_size: int = 256, lanes: int = 0, round_budget: int = 0,
                 lb_policy: int = LB_UNIFORM, device: int = 0, queue_items: int = 0,
                 time_kernels: bool = False, lds_stack_depth: int = 0, log_samples: int = 0,
                 use_graph: bool = True, record_paths: bool = False, lookahead: bool = True,
                 round_time_us: int = 0, dedup: bool = False, resplit: int = 0):
        L = lib()
        cfg = Config()
        L.spz_default_config(ctypes.byref(cfg), width, heuristic)
        cfg.mode = mode
        cfg.frontier_target = frontier_target
        cfg.max_split_depth = max_split_depth
        cfg.workgroup_size = workgroup_size
        cfg.lanes = lanes
        cfg.round_budget = round_budget
        cfg.lb_policy = lb_policy
        cfg.device = device
        cfg.queue_items = queue_items
        cfg.time_kernels = 1 if time_kernels else 0
        cfg.lds_stack_depth = lds_stack_depth
        cfg.log_samples = log_samples
        cfg.use_graph = 1 if use_graph else 0
        cfg.record_paths = 1 if record_paths else 0
        cfg.lookahead = 1 if lookahead else 0
        cfg.round_time_us = round_time_us
        cfg.dedup = 1 if dedup else 0
        cfg.resplit = resplit
        self.cfg = cfg
        self.width = width
        self._ctx = ctypes.c_void_p()
        rc = L.spz_create(ctypes.byref(self._ctx), ctypes.byref(cfg))
        if rc != 0:
            msg = L.spz_last_error(self._ctx).decode() if self._ctx else ERRORS.get(rc, str(rc))
            if self._ctx:
                L.spz_destroy(self._ctx)
            self._ctx = None
            raise SpzError(f"spz_create failed ({rc}): {msg}", rc)

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = lib().spz_last_error(self._ctx).decode()
            raise SpzError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}", rc)

    def close(self):
        if self._ctx:
            lib().spz_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_pdb(self, cache_path: Optional[str] = None):
        self._check(lib().spz_build_pdb(self._ctx, cache_path.encode() if cache_path else None),
                    "spz_build_pdb")

    def load_pdb_tab(self, path0: str, path1: str):
        """Load the reference's pat24.1256712.tab and pat24.34891314.tab."""
        self._check(lib().spz_load_pdb_tab(self._ctx, path0.encode(), path1.encode()),
                    "spz_load_pdb_tab")

    def export_pdb_tab(self, path0: str, path1: str):
        self._check(lib().spz_export_pdb_tab(self._ctx, path0.encode(), path1.encode()),
                    "spz_export_pdb_tab")

    def get_pdb(self, pattern: int) -> np.ndarray:
        out = np.empty(127_512_000, dtype=np.uint8)
        self._check(lib().spz_get_pdb(self._ctx, pattern, _u8p(out)), "spz_get_pdb")
        return out

    def gather_probe(self, n_gathers: int = 1 << 30, span_bytes: int = 0):
        """Random byte gathers over the first span_bytes of the direct tables
        (0: all of them).  Returns (ms, gathers per second)."""
        ms, gps = ctypes.c_double(), ctypes.c_double()
        self._check(lib().spz_gather_probe(self._ctx, n_gathers, span_bytes, ctypes.byref(ms), ctypes.byref(gps)),
                    "spz_gather_probe")
        return ms.value, gps.value

    # ---- learned load balancer -------------------------------------------
    @staticmethod
    def mlp_param_count() -> int:
        return int(lib().spz_mlp_param_count())

    def mlp_init(self, seed: int = 1):
        self._check(lib().spz_mlp_init(self._ctx, seed), "spz_mlp_init")

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.mlp_param_count()
        self._check(lib().spz_mlp_set_params(self._ctx, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "spz_mlp_set_params")

    def mlp_get_params(self) -> np.ndarray:
        p = np.zeros(self.mlp_param_count(), dtype=np.float32)
        self._check(lib().spz_mlp_get_params(self._ctx, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "spz_mlp_get_params")
        return p

    def _rows(self, boards, thr):
        b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint8).reshape(-1, self.width ** 2))
        t = np.ascontiguousarray(np.asarray(thr, dtype=np.int32).reshape(-1))
        assert len(b) == len(t)
        return b, t, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))

    def mlp_forward(self, boards, thr) -> np.ndarray:
        b, t, tp = self._rows(boards, thr)
        out = np.zeros(len(b), dtype=np.float32)
        self._check(lib().spz_mlp_forward(self._ctx, _u8p(b), tp, len(b),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "spz_mlp_forward")
        return out

    def get_paths(self, max_len: int = 256):
        """One optimal solution per instance of the last solve (record_paths):
        a list of move lists (0 up, 1 right, 2 left, 3 down: the blank's
        moves), None where no path was recorded."""
        n = getattr(self, "_last_n", 0)
        moves = np.zeros((n, max_len), dtype=np.uint8)
        lens = np.zeros(n, dtype=np.int32)
        self._check(lib().spz_get_paths(self._ctx, _u8p(moves), max_len,
                                        lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                    "spz_get_paths")
        return [None if ln < 0 else [int(x) for x in moves[i, :ln]] for i, ln in enumerate(lens)]

    def mlp_predict(self, boards, thr0) -> np.ndarray:
        """Predictions of the learned load balancer's fused kernel: row i
        holds boards[i] at thresholds thr0[i], thr0[i] + 2, ... (16)."""
        b, t, tp = self._rows(boards, thr0)
        out = np.zeros((len(b), 16), dtype=np.float32)
        self._check(lib().spz_mlp_predict(self._ctx, _u8p(b), tp, len(b),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "spz_mlp_predict")
        return out

    def mlp_grad(self, boards, thr, target):
        b, t, tp = self._rows(boards, thr)
        tg = np.ascontiguousarray(target, dtype=np.float32)
        loss = ctypes.c_float()
        g = np.zeros(self.mlp_param_count(), dtype=np.float32)
        self._check(lib().spz_mlp_grad(self._ctx, _u8p(b), tp,
                                       tg.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(b),
                                       ctypes.byref(loss),
                                       g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "spz_mlp_grad")
        return loss.value, g

    def mlp_train(self, boards, thr, target, epochs=4, batch=1024, lr=1e-3, seed=1) -> np.ndarray:
        b, t, tp = self._rows(boards, thr)
        tg = np.ascontiguousarray(target, dtype=np.float32)
        el = np.zeros(epochs, dtype=np.float32)
        self._check(lib().spz_mlp_train(self._ctx, _u8p(b), tp,
                                        tg.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(b),
                                        epochs, batch, lr, seed,
                                        el.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "spz_mlp_train")
        return el

    def get_samples(self):
        n = ctypes.c_int32()
        self._check(lib().spz_get_samples(self._ctx, None, None, None, 0, ctypes.byref(n)),
                    "spz_get_samples")
        k = n.value
        b = np.zeros((max(k, 1), self.width ** 2), dtype=np.uint8)
        t = np.zeros(max(k, 1), dtype=np.int32)
        ld = np.zeros(max(k, 1), dtype=np.uint64)
        self._check(lib().spz_get_samples(self._ctx, _u8p(b), t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          ld.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), k,
                                          ctypes.byref(n)), "spz_get_samples")
        k = n.value
        return b[:k], t[:k], ld[:k]

    def config(self) -> dict:
        c = Config()
        lib().spz_get_config(self._ctx, ctypes.byref(c))
        return {f: getattr(c, f) for f, _ in Config._fields_}

    def reconfigure(self, **kw):
        c = Config()
        lib().spz_get_config(self._ctx, ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        self._check(lib().spz_reconfigure(self._ctx, ctypes.byref(c)), "spz_reconfigure")

    def autotune(self, boards, max_iters: int = 64, passes: int = 1):
        """Tunes this context in place on the sample boards; returns
        (best config dict, best ms, candidates tried)."""
        b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint8).reshape(-1, self.width ** 2))
        best, ms, n = Config(), ctypes.c_double(), ctypes.c_int32()
        self._check(lib().spz_autotune(self._ctx, _u8p(b), len(b), max_iters, passes, ctypes.byref(best),
                                       ctypes.byref(ms), ctypes.byref(n)), "spz_autotune")
        return {f: getattr(best, f) for f, _ in Config._fields_}, ms.value, n.value

    def set_lb(self, policy: int):
        self._check(lib().spz_set_lb(self._ctx, policy), "spz_set_lb")

    def eval_h(self, boards) -> np.ndarray:
        b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint8).reshape(-1, self.width ** 2))
        h = np.zeros(len(b), dtype=np.int32)
        self._check(lib().spz_eval_h(self._ctx, _u8p(b), len(b),
                                     h.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                    "spz_eval_h")
        return h

    def eval_h_noalloc(self, boards) -> np.ndarray:
        """eval_h through the allocation-free entry point (at most 65536
        boards; calls on one GPU share its scratch and must not overlap)."""
        b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint8).reshape(-1, self.width ** 2))
        h = np.zeros(len(b), dtype=np.int32)
        self._check(lib().spz_eval_h_noalloc(self._ctx, _u8p(b), len(b),
                                             h.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                    "spz_eval_h_noalloc")
        return h

    def solve(self, boards, max_iters: int = 64) -> List[Result]:
        b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint8).reshape(-1, self.width ** 2))
        n = len(b)
        out = (_Result * max(n, 1))()
        ic = np.zeros((max(n, 1), max_iters), dtype=np.uint64)
        self._check(lib().spz_solve_batch(self._ctx, _u8p(b), n, max_iters, out,
                                          ic.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))),
                    "spz_solve_batch")
        self._last_n = n
        res = []
        for i in range(n):
            r = out[i]
            k = r.n_iters
            res.append(Result(r.opt_len, r.h0, k, r.status, int(r.total_nodes),
                              [r.h0 + 2 * j for j in range(k)],
                              [int(x) for x in ic[i, :k]]))
        return res

    def stats(self) -> dict:
        s = (ctypes.c_double * 32)()
        self._check(lib().spz_last_stats(self._ctx, s), "spz_last_stats")
        return {"rounds": s[0], "solve_ms": s[1], "donated": s[2], "lanes": s[3],
                "pdb_levels0": s[4], "pdb_levels1": s[5], "pdb_build_ms": s[6],
                "overflow_items": s[7], "round_kernel_ms": s[8], "round_launches": s[9],
                "split_ms": s[10], "instances": s[11], "nodes": s[12],
                "setup_ms": s[13], "rounds_wall_ms": s[14], "graph": s[15],
                "prof_idle": s[16], "prof_expand": s[17], "prof_pop": s[18], "prof_wave_steps": s[19],
                "prof_sibling_moves": s[20], "prof_reads_first": s[21], "prof_reads_second": s[22],
                "lookahead_starts": s[23], "resplit_nodes": s[24], "resplit_items": s[25]}
