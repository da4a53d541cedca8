"""Benchmark of the IDA* hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4]

A step is one solve of the configuration's whole batch of synthetic
instances (seeded generators; seed = config_id * 10^6 + instance index),
inputs already on the host in packed form; the per-step result (optimal
lengths and node counts) is checked to be identical across steps.  The
timed steps replay the search rounds from the solver's captured hipGraph;
a separate step with every round launch timed by HIP events gives the
kernel's average duration (the roofline).  The default policy is the
learned predictor with re-splitting of nodes predicted above 2 x the mean.
Prints one JSON line.  Configurations:
  c1  15-puzzle Manhattan, 100 walks of 40 moves
  c2  15-puzzle Manhattan, 100 random solvable permutations, one instance
      per launch, root split into 2048 frontier nodes
  c3  15-puzzle Manhattan, 4096 mixed instances (2048 walks of 20-80 moves,
      2048 random permutations) in one batched launch sequence
  c4  24-puzzle 6-6-6-6 PDB, 64 walks of 160 moves, batched, 4096 frontier
      nodes per instance
  c5  as c4 with 240-move walks, 8192 frontier nodes
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "sliding-puzzle_amd"))

import numpy as np  # noqa: E402

import spz  # noqa: E402

HBM_PEAK_GBS = 8000.0
# VALU peak of MI355X: the published 157.3 TFLOP/s FP32 vector rate is
# 256 CUs x 2.4 GHz x 128 lanes x 2 (FMA), i.e. 128 lane-operations, two
# wave64 instructions, per CU and cycle.  (The counters agree: the search
# kernels' SQ_ACTIVE_INST_VALU quad-cycles exceed the SIMDs' elapsed
# quad-cycles, so a wave64 instruction holds a SIMD for fewer than 4 cycles.)
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12  # int32 lane-ops/s
VALU_ISSUE_PEAK = 256 * 2 * 2.4e9  # wave64 VALU instructions/s


def instances(config: str, n: int | None = None):
    """Seeded synthetic boards of a configuration."""
    if config == "c1":
        b = [spz.gen_walk(4, 40, 1_000_000 + i) for i in range(n or 100)]
        return 4, spz.H_MANHATTAN, b
    if config == "c2":
        b = [spz.gen_random(4, 2_000_000 + i) for i in range(n or 100)]
        return 4, spz.H_MANHATTAN, b
    if config == "c3":
        n = n or 4096
        half = n // 2
        b = []
        for i in range(half):
            seed = 3_000_000 + i
            L = 20 + int((int(spz_rng_u64(seed)) >> 32) * 61 >> 32)
            b.append(spz.gen_walk(4, L, seed))
        b += [spz.gen_random(4, 3_000_000 + half + i) for i in range(n - half)]
        return 4, spz.H_MANHATTAN, b
    if config == "c4":
        b = [spz.gen_walk(5, 160, 4_000_000 + i) for i in range(n or 64)]
        return 5, spz.H_PDB666, b
    if config == "c5":
        b = [spz.gen_walk(5, 240, 5_000_000 + i) for i in range(n or 64)]
        return 5, spz.H_PDB666, b
    raise SystemExit(f"unknown config {config}")


def spz_rng_u64(seed: int) -> int:
    """First splitmix64 output of `seed` (walk-length draw of config c3)."""
    z = (seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


FRONTIER = {"c1": 2048, "c2": 2048, "c3": 16, "c4": 4096, "c5": 8192}
WORKLOAD = {
    "c1": "15-puzzle Manhattan, 100 walks x 40 moves",
    "c2": "15-puzzle Manhattan, 100 random solvable permutations, one instance per launch",
    "c3": "15-puzzle Manhattan, 4096 mixed (2048 walks 20-80 + 2048 random), batched",
    "c4": "24-puzzle 6-6-6-6 PDB, 64 walks x 160 moves, batched, 4096 frontier nodes/instance",
    "c5": "24-puzzle 6-6-6-6 PDB, 64 walks x 240 moves, batched, 8192 frontier nodes/instance",
}


# a span of the direct tables that every XCD's L2 (4 MB) holds: gathers over
# it measure the L2-hit side of the random-read ceiling
L2_RESIDENT_SPAN = 2 << 20


def profile_figure(config: str, kind: str, what: str, key: str):
    """A per-node figure of k_round measured by rocprofv3 --pmc passes on this
    workload and kept in its profiles/ summary k_round_<config>_<kind>_<what>.md
    (kind "pdb" or "manhattan"; what = "traffic": calibrated FETCH_SIZE +
    WRITE_SIZE bytes and L2 misses; "instructions": VALU wave instructions,
    VALU busy, active lanes, LDS bank conflicts)."""
    path = os.path.join(REPO, "profiles", f"k_round_{config}_{kind}_{what}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def train_predictor(solver, config, width, n_train=None):
    """The learned load balancer's setup (not timed): solve a separate set of
    training instances of the same configuration on the same context in
    frontier order with the sample log on, train the MLP on the GPU on
    (frontier node, threshold, log2(1 + subtree size)), and measure its
    held-out error against the uniform predictor (one constant, the training
    mean)."""
    t0 = time.perf_counter()
    _, heur, tb = instances(config, n_train)
    # training boards: seeds disjoint from the benchmark's (offset 10^7)
    if config in ("c4", "c5"):
        L = 160 if config == "c4" else 240
        base = (4 if config == "c4" else 5) * 1_000_000 + 10_000_000
        tb = [spz.gen_walk(5, L, base + i) for i in range(n_train or 64)]
    solver.set_lb(spz.LB_UNIFORM)
    solver.reconfigure(log_samples=2_000_000)
    solver.solve(np.stack(tb))
    sb, st, sl = solver.get_samples()
    solver.reconfigure(log_samples=0)
    t_log = time.perf_counter()
    y = np.log2(sl.astype(np.float64) + 1).astype(np.float32)
    rng = np.random.default_rng(1)
    idx = rng.permutation(len(sb))
    cut = int(0.9 * len(idx))
    tr, te = idx[:cut], idx[cut:]
    solver.mlp_init(1)
    losses = solver.mlp_train(sb[tr], st[tr], y[tr], epochs=6, batch=1024, lr=1e-3, seed=3)
    t_train = time.perf_counter()
    pred = solver.mlp_forward(sb[te], st[te])
    mse = float(np.mean((pred - y[te]) ** 2))
    mse_u = float(np.mean((y[tr].mean() - y[te]) ** 2))
    return {"samples": int(len(sb)), "train_instances": len(tb), "epochs": 6,
            "epoch_loss": [round(float(v), 4) for v in losses],
            "heldout_mse_log2": mse, "uniform_predictor_mse_log2": mse_u,
            "setup_s": time.perf_counter() - t0, "log_solve_s": t_log - t0,
            "train_s": t_train - t_log}


def progress(*a):
    """Progress of the run on stderr (the JSON line alone goes to stdout)."""
    print("bench:", *a, file=sys.stderr, flush=True)


STAT_KEYS = ("round_kernel_ms", "round_launches", "solve_ms", "setup_ms", "split_ms",
             "rounds_wall_ms", "rounds", "donated")


def solve_step(solver, boards, per_instance: bool, max_iters: int):
    """One step: the whole batch in one solve, or (per_instance) one solve
    per board.  Returns the results and the solver statistics summed over
    every solve of the step."""
    acc = dict.fromkeys(STAT_KEYS, 0.0)

    def add():
        st = solver.stats()
        for k in STAT_KEYS:
            acc[k] += st[k]

    if per_instance:
        out = []
        for i, b in enumerate(boards):
            out += solver.solve([b], max_iters=max_iters)
            add()
            if (i + 1) % 512 == 0:
                progress(f"{i + 1} of {len(boards)} instances one per launch")
        return out, acc
    out = solver.solve(boards, max_iters=max_iters)
    add()
    return out, acc


def cpu_cores() -> int:
    """Host cores the CPU baseline may use: this process's CPU set, at most
    16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def lpt_groups(items, weights, k):
    """Longest-processing-time-first assignment of items to k groups."""
    groups = [[] for _ in range(k)]
    load = [0] * k
    for i in sorted(items, key=lambda i: -weights[i]):
        g = min(range(k), key=lambda j: load[j])
        groups[g].append(i)
        load[g] += weights[i]
    return [g for g in groups if g]


def cpu_baseline(width, heur, boards, res, solver, budget_s=15.0):
    """CPU baseline on the host cores, on a bounded sample of the workload:
    its instances in increasing order of their node count, until about
    budget_s seconds of work per core.  One solver process per core, the
    sample split between them longest-first.  Solver: the reference's own
    (oracle/_ref: src/24puzzle_expand_with_pdb.cpp for the PDB configs,
    src_horie/idas_cpu_expand.c for the 15-puzzle), else the oracle's C port
    on one core.  Each instance's result is checked against the GPU's.
    nodes_per_s_all_cores = sample nodes / the longest process's solve time
    (table loading excluded); nodes_per_s_one_core = sample nodes / the sum
    of the per-instance times (each core while all run); value = the
    all-core rate / the workload's mean nodes per instance."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    import run_ref as R

    n = len(boards)
    nodes_of = [r.total_nodes for r in res]
    order = sorted(range(n), key=lambda i: nodes_of[i])
    mean_nodes = sum(nodes_of) / n
    cores = cpu_cores()
    est_rate = 30e6 if heur == spz.H_PDB666 else 60e6  # nodes/s of the CPU solvers, one core
    pick, acc = [], 0
    for i in order:
        if pick and acc + nodes_of[i] > est_rate * budget_s * cores:
            break
        pick.append(i)
        acc += nodes_of[i]
    wall0 = time.perf_counter()
    if heur == spz.H_PDB666 and R.have("24puzzle_expand_with_pdb"):
        groups = lpt_groups(pick, nodes_of, cores)
        with tempfile.TemporaryDirectory() as d:
            paths = (os.path.join(d, "pat24.1256712.tab"), os.path.join(d, "pat24.34891314.tab"))
            solver.export_pdb_tab(*paths)

            def run(g):
                times = []
                out = R.yamamura("pdb_all", [list(map(int, boards[i])) for i in g], paths, times=times)
                return g, out, times

            with ThreadPoolExecutor(len(groups)) as ex:
                runs = list(ex.map(run, groups))
        secs_sum, makespan, nodes = 0.0, 0.0, 0
        for g, out, times in runs:
            for (ln, nd), i in zip(out, g):
                assert (ln, nd) == (res[i].opt_len, res[i].total_nodes), i
                nodes += nd
            secs_sum += sum(times)
            makespan = max(makespan, sum(times))
        kind, who = "reference", "src/24puzzle_expand_with_pdb.cpp (oracle/_ref, g++ -O3)"
        used = len(groups)
    elif heur == spz.H_MANHATTAN and width == 4 and R.have("idas_cpu_expand"):
        def run(i):
            return i, R.horie_timed(list(map(int, boards[i])))

        t = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            runs = list(ex.map(run, sorted(pick, key=lambda i: -nodes_of[i])))
        makespan = time.perf_counter() - t
        secs_sum, nodes = 0.0, 0
        for i, (dt, total, work) in runs:
            if total is None:  # solved in its first iteration: no count printed
                continue
            assert total == res[i].total_nodes, i
            secs_sum += dt
            nodes += work  # the nodes the process generated (its h0 tree twice)
        kind, who = "reference", "src_horie/idas_cpu_expand.c (oracle/_ref, gcc -O3)"
        used = cores
    else:
        import oracle as O

        if heur == spz.H_PDB666:
            O.pdb_set_tables(solver.get_pdb(0), solver.get_pdb(1))
        t0, nodes = time.perf_counter(), 0
        for i in pick:
            r, thr, c = O.ida(list(map(int, boards[i])),
                              "pdb" if heur == spz.H_PDB666 else "manhattan", "all")
            assert (r, sum(c)) == (res[i].opt_len, res[i].total_nodes), i
            nodes += sum(c)
        secs_sum = makespan = time.perf_counter() - t0
        kind, who, used = "port", "oracle/spz_oracle.c (C port)", 1
    wall = time.perf_counter() - wall0
    one = nodes / secs_sum if secs_sum > 0 else 0.0
    allc = nodes / makespan if makespan > 0 else 0.0
    return {"value": allc / mean_nodes if mean_nodes else None, "unit": "instances/s",
            "cores": used, "kind": kind,
            "sample": f"{who}, one process per core on {used} cores, on {len(pick)} of the "
                      f"workload's {n} instances (smallest node counts first; "
                      f"{100.0 * nodes / sum(nodes_of):.2f}% of its nodes): {nodes / 1e6:.1f} M "
                      f"nodes, longest process {makespan:.1f} s, {secs_sum:.1f} core-s; value = "
                      f"the all-core rate / the workload's mean of {mean_nodes / 1e6:.1f} M nodes "
                      f"per instance",
            "nodes_per_s_all_cores": allc, "nodes_per_s_one_core": one,
            "instances_per_s_one_core": one / mean_nodes if mean_nodes else None,
            "sample_fraction_of_nodes": nodes / sum(nodes_of), "wall_s": wall}


POLICIES = {"uniform": spz.LB_UNIFORM, "learned": spz.LB_LEARNED, "prevload": spz.LB_PREVLOAD}


def timed_pass(solver, boards, per_instance, max_iters):
    """One step with every search-round launch timed by HIP events on the
    solver's stream (the rounds then run eagerly, not from the hipGraph).
    Returns the results, the summed statistics and the step's wall time."""
    solver.reconfigure(time_kernels=1)
    t = time.perf_counter()
    res, acc = solve_step(solver, boards, per_instance, max_iters)
    dt = time.perf_counter() - t
    solver.reconfigure(time_kernels=0)
    return res, acc, dt


def same(a, b):
    return [(r.opt_len, r.counts) for r in a] == [(r.opt_len, r.counts) for r in b]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default=os.environ.get("SPZ_BENCH_CONFIG", "c4"))
    ap.add_argument("--n", type=int, default=None, help="instances (default: the config's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--max-iters", type=int, default=64)
    ap.add_argument("--round-budget", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--lds-depth", type=int, default=0)
    ap.add_argument("--wg", type=int, default=256)
    ap.add_argument("--lb", default="learned", choices=list(POLICIES),
                    help="dispatch policy of the timed solver (default: the learned predictor)")
    ap.add_argument("--resplit", type=int, default=2,
                    help="re-split frontier nodes whose estimate exceeds K x the mean (0: off)")
    ap.add_argument("--no-ablation", action="store_true",
                    help="skip the one-step comparison solves under the other policies")
    ap.add_argument("--eager", action="store_true",
                    help="time kernel launches in the timed steps too (no hipGraph replay; "
                         "for profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one plain device-to-device copy of MIB MiB (a kernel of "
                         "known traffic that calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; --gpus above 1 is not supported "
                                   "(instance batches shard across GPUs with no exchange; not built here)"}))
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0 or args.config not in FRONTIER:
        print(json.dumps({"error": f"need --steps >= 1, --warmup >= 0 and --config one of {sorted(FRONTIER)}"}))
        sys.exit(2)

    if args.copy_probe:
        import torch

        a = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    width, heur, boards = instances(args.config, args.n)
    boards = np.stack(boards)
    per_instance = args.config == "c2"
    solver = spz.Solver(width=width, heuristic=heur, frontier_target=FRONTIER[args.config],
                        round_budget=args.round_budget, lanes=args.lanes, time_kernels=args.eager,
                        lds_stack_depth=args.lds_depth, workgroup_size=args.wg, resplit=args.resplit)
    pdb_ms = None
    if heur == spz.H_PDB666:
        t = time.time()
        solver.build_pdb()
        pdb_ms = (time.time() - t) * 1e3

    learned = None
    if args.lb == "learned":
        learned = train_predictor(solver, args.config, width)
        progress("predictor trained")
    solver.set_lb(POLICIES[args.lb])

    ref = None
    for _ in range(args.warmup):
        ref, _ = solve_step(solver, boards, per_instance, args.max_iters)
        progress("warmup step done")
    nodes_total = 0
    brk = dict.fromkeys(STAT_KEYS, 0.0)
    graph_used = True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, acc = solve_step(solver, boards, per_instance, args.max_iters)
        for k in brk:
            brk[k] += acc[k] / args.steps
        graph_used = graph_used and solver.stats()["graph"] == 1.0
        nodes_total += sum(r.total_nodes for r in res)
        if ref is not None:
            assert same(res, ref)
        ref = res
        progress("timed step done")
    elapsed = time.perf_counter() - t0
    n = len(boards)
    ms_per_step = elapsed / args.steps * 1e3
    inst_per_s = n * args.steps / elapsed
    nodes_per_s = nodes_total / elapsed
    unsolved = sum(1 for r in ref if r.opt_len < 0)

    # roofline of the dominant kernel (the search round): a separate step
    # with every launch timed by HIP events (the timed steps above replay
    # the hipGraph, whose launches cannot be timed one by one)
    if args.eager:
        kres, kacc, kdt = res, acc, elapsed / args.steps
    else:
        kres, kacc, kdt = timed_pass(solver, boards, per_instance, args.max_iters)
        assert same(kres, ref)
        progress("kernel-timing step done")
    knodes = sum(r.total_nodes for r in kres)
    launches = int(kacc["round_launches"])
    nodes_per_launch = knodes / max(launches, 1)
    avg_launch_s = (kacc["round_kernel_ms"] / max(launches, 1)) / 1e3
    pcfg = args.config
    if heur == spz.H_PDB666:
        # algorithmic: 2 table gathers per generated node (one per
        # partition), each accounted as a 64 B granule
        achieved = nodes_per_launch * 128 / avg_launch_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "achieved_is": "algorithmic: 2 gathers x 64 B per generated node / launch time"}
        tb = profile_figure(pcfg, "pdb", "traffic", "traffic_bytes_per_node")
        if tb is not None:
            roof["traffic"] = tb * nodes_per_launch
            roof["traffic_source"] = f"calibrated FETCH_SIZE+WRITE_SIZE bytes per generated node from " \
                                     f"profiles/k_round_{pcfg}_pdb_traffic.md x this run's nodes per launch"
            roof["traffic_frac_of_peak"] = tb * nodes_per_launch / avg_launch_s / 1e9 / HBM_PEAK_GBS
        miss = profile_figure(pcfg, "pdb", "traffic", "l2_misses_per_node")
        if miss is not None:
            roof["l2_misses_per_node"] = miss
            roof["l2_miss_rate_per_s"] = miss * nodes_per_launch / avg_launch_s
    else:
        # 16 int32 lane-ops per generated node (integer VALU issue bound)
        achieved = nodes_per_launch * 16 / avg_launch_s / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tlane-op/s",
                "frac": achieved / VALU_PEAK_TOPS, "traffic": None,
                "achieved_is": "algorithmic: 16 int32 lane-ops per generated node / launch time"}
    kind = "pdb" if heur == spz.H_PDB666 else "manhattan"
    vi = profile_figure(pcfg, kind, "instructions", "valu_wave_instructions_per_node")
    if vi is not None:
        # VALU issue: SQ_INSTS_VALU wave instructions against two wave64
        # instructions per CU and cycle
        rate = vi * nodes_per_launch / avg_launch_s
        roof["valu_issue"] = {
            "wave_instructions_per_node": vi, "achieved": rate / 1e9, "peak": VALU_ISSUE_PEAK / 1e9,
            "unit": "G wave-instructions/s", "frac": rate / VALU_ISSUE_PEAK,
            "source": f"SQ_INSTS_VALU per generated node from profiles/k_round_{pcfg}_{kind}_instructions.md "
                      f"x this run's nodes per launch / its HIP-event launch time"}
    for key in ("valu_busy", "active_lane_fraction", "lds_bank_conflict_rate"):
        v = profile_figure(pcfg, kind, "instructions", key)
        if v is not None:
            roof[key] = v
    roof["kernel"] = "k_round"
    roof["avg_launch_ms"] = avg_launch_s * 1e3
    roof["launches"] = launches
    roof["nodes_per_launch"] = nodes_per_launch

    out = {
        "metric": "instances/sec + nodes/sec, batched+learned-LB vs plain GPU IDA*, exact optimal length",
        "value": inst_per_s, "unit": "instances/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded)",
        "config": {"workload": WORKLOAD[args.config], "config": args.config, "instances": n,
                   "frontier_target": FRONTIER[args.config], "parallelism": "single GPU"},
        "nodes_per_s": nodes_per_s, "nodes_per_step": nodes_total / args.steps,
        "unsolved_within_max_iters": unsolved,
        "timed_path": "hipGraph replay of the search rounds" if graph_used and not args.eager
        else "eager launches (kernel timing on)",
        "eager_timed_step_instances_per_s": n / kdt,
        "roofline": roof,
        "breakdown_per_step": dict(brk, lanes=solver.stats()["lanes"]),
    }
    out["lb"] = args.lb
    out["resplit"] = args.resplit
    if learned is not None:
        out["learned_lb"] = learned
    out["lookahead"] = bool(solver.config()["lookahead"]) and args.lb != "prevload"
    if not args.no_ablation and not per_instance:
        # the same batch under the other dispatch policies (one step each; the
        # previous-load policy runs one iteration at a time, without lookahead)
        for name in ("uniform", "prevload", "learned"):
            if name == args.lb or (name == "learned" and learned is None):
                continue
            solver.set_lb(POLICIES[name])
            t = time.perf_counter()
            r2 = solver.solve(boards, max_iters=args.max_iters)
            dt = time.perf_counter() - t
            assert same(r2, ref)
            out[f"{name}_lb_instances_per_s"] = len(boards) / dt
            out[f"{name}_lb_donated"] = solver.stats()["donated"]
            progress(f"{name} policy done")
        # the timed solver's policy, iterations strictly one at a time
        solver.set_lb(POLICIES[args.lb])
        out[f"{args.lb}_lb_donated"] = brk["donated"]
        solver.reconfigure(lookahead=0)
        t = time.perf_counter()
        r2 = solver.solve(boards, max_iters=args.max_iters)
        dt = time.perf_counter() - t
        assert same(r2, ref)
        out["no_lookahead_instances_per_s"] = len(boards) / dt
        progress("no-lookahead solve done")
        solver.reconfigure(lookahead=1)
    if pdb_ms is not None:
        out["pdb_build_ms"] = pdb_ms
        _, gps = solver.gather_probe(1 << 30)
        _, gps_l2 = solver.gather_probe(1 << 30, span_bytes=L2_RESIDENT_SPAN)
        roof["measured_gather_peak"] = {
            "gathers_per_s": gps, "GB_per_s_at_64B": gps * 64 / 1e9,
            "what": "uniform random u8 gathers over the two 244 MB direct tables (spz_gather_probe)",
            "l2_resident_gathers_per_s": gps_l2,
            "l2_resident_what": f"the same over the tables' first {L2_RESIDENT_SPAN >> 20} MB (L2-resident)"}
        if "l2_miss_rate_per_s" in roof:
            roof["measured_gather_peak"]["l2_miss_rate_frac_of_it"] = roof["l2_miss_rate_per_s"] / gps
        hits = profile_figure(pcfg, "pdb", "traffic", "l2_hits_per_node")
        if hits is not None and "l2_miss_rate_per_s" in roof:
            # the table reads' share of the memory system's random-read
            # capacity: each L2 hit takes 1/gps_l2 s of it, each miss 1/gps
            hit_rate = hits * nodes_per_launch / avg_launch_s
            roof["gather_capacity"] = {
                "l2_hits_per_node": hits, "l2_misses_per_node": roof["l2_misses_per_node"],
                "reads_per_s": hit_rate + roof["l2_miss_rate_per_s"],
                "frac": hit_rate / gps_l2 + roof["l2_miss_rate_per_s"] / gps,
                "what": "L2 hits / L2-resident gather rate + L2 misses / whole-table gather rate, "
                        "per second of launch time (1.0 = the measured random-read ceiling at the kernel's hit mix)"}

    if not args.no_plain and not per_instance:
        # plain GPU IDA*: one instance per launch, uniform static split, one
        # iteration at a time (the reference's plain method)
        plain = spz.Solver(width=width, heuristic=heur, frontier_target=FRONTIER[args.config],
                           lookahead=False)
        if heur == spz.H_PDB666:
            plain.build_pdb()
        t = time.perf_counter()
        pres, _ = solve_step(plain, boards, True, args.max_iters)
        pt = time.perf_counter() - t
        assert same(pres, ref)
        out["plain_instances_per_s"] = n / pt
        progress("plain pass done")
        out["speedup_vs_plain"] = inst_per_s / (n / pt)
        if not args.no_ablation:
            # the plain method with lookahead added (not the reference's method;
            # shows what the batch adds beyond lookahead)
            plain.reconfigure(lookahead=1)
            t = time.perf_counter()
            pres, _ = solve_step(plain, boards, True, args.max_iters)
            pt = time.perf_counter() - t
            assert same(pres, ref)
            out["plain_lookahead_instances_per_s"] = n / pt
            out["speedup_vs_plain_lookahead"] = inst_per_s / (n / pt)
        plain.close()

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(width, heur, boards, ref, solver)
        progress("cpu baseline done")
    print(json.dumps(out), flush=True)
    solver.close()


if __name__ == "__main__":
    main()
