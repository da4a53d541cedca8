/*
 * spz.h -- C ABI of libspz, the MI355X IDA* solver for the 15- and 24-puzzle.
 *
 * The reference has no library API: every solver is a process that reads
 * problem files and prints "<path> <len> [nodes]" (src/24puzzle.cc:167-198,
 * src_horie/idas_cpu.c:370-389).  These entry points replace that solver
 * surface; each cites the reference code whose behaviour it takes over.
 *
 * Boards: row-major uint8 arrays of width*width cells, 0 = blank, goal has
 * tile t at cell t (src/24puzzle.cc:89-108).  Node counts follow the
 * reference: every in-bounds child other than the reverse of the previous
 * move is counted when its parent is expanded (src/24puzzle.cc:119-131);
 * thresholds are h0, h0+2, ... (src/24puzzle.cc:146-152).  Counts are u64.
 *
 * Errors: functions return SPZ_OK (0) or a negative SPZ_E* code and never
 * exit or trap (the reference exits on CUDA errors,
 * src/15puzzle_block_parallel.cu:42-57).  spz_last_error() describes the
 * last failure.  A context is not thread-safe; use one per host thread.
 */
#ifndef SPZ_H
#define SPZ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPZ_OK 0
#define SPZ_E_ARG -1      /* bad argument (non-permutation board, sizes) */
#define SPZ_E_HIP -2      /* HIP runtime error */
#define SPZ_E_NOPDB -3    /* PDB heuristic requested before spz_build_pdb */
#define SPZ_E_CAPACITY -4 /* device work queue full (raise queue_items) */
#define SPZ_E_IO -5       /* file read/write failed */
#define SPZ_E_UNSOLVABLE -6
#define SPZ_E_STALL -7    /* internal: work vanished with instances active */
#define SPZ_E_TIMEOUT -8  /* wall-clock limit (env SPZ_TIMEOUT_S, default 900) */
#define SPZ_E_MISMATCH -9 /* spz_autotune: a candidate configuration changed the result */
#define SPZ_E_HOST -10    /* host-side failure (out of host memory, a thread that cannot
                             start); no C++ exception leaves the library */

/* heuristics */
#define SPZ_H_MANHATTAN 0 /* src/24puzzle.cc:89-108 */
#define SPZ_H_PDB666 1    /* src/24puzzle_with_pdb.cc:114-194 (width 5) */

/* count modes */
#define SPZ_MODE_ALL 0   /* exhaust the final iteration: src/24puzzle_expand.cc */
#define SPZ_MODE_FIRST 1 /* stop at a goal: src/24puzzle.cc (final count partial) */

/* dispatch order of frontier nodes (the reference's static LB is
 * src/15puzzle_block_parallel.cu:490-563) */
#define SPZ_LB_UNIFORM 0 /* frontier order */
#define SPZ_LB_LEARNED 1 /* longest predicted subtree first (MLP, needs parameters) */
#define SPZ_LB_PREVLOAD 2 /* longest subtree of the previous iteration first (the
                            reference's load-based assignment,
                            src_horie/idas_dynamic.cu:1527-1614) */

/* per-instance status */
#define SPZ_ST_SOLVED 0
#define SPZ_ST_ITER_LIMIT 1 /* max_iters iterations without a goal */

typedef struct spz_ctx spz_ctx;

typedef struct spz_config {
    int32_t width;           /* 4 (15-puzzle) or 5 (24-puzzle) */
    int32_t heuristic;       /* SPZ_H_* */
    int32_t mode;            /* SPZ_MODE_* */
    int32_t frontier_target; /* root split: per instance, expand the search
                                tree breadth-first until a level holds at
                                least this many nodes (reference: 2048,
                                src/15puzzle_block_parallel.cu:150-187);
                                1 = no split; at most 1048576 */
    int32_t max_split_depth; /* cap on that split depth (0 = 24) */
    int32_t workgroup_size;  /* 64, 128 or 256 lanes (0 = 256) */
    int32_t lanes;           /* DFS lanes in flight (0 = fill the device;
                                at most 2^24) */
    int32_t round_budget;    /* DFS steps per lane per round (0 = auto:
                                2048 for Manhattan when the batch has work
                                for every lane, else 1024; at most 2^20) */
    int32_t lds_stack_depth; /* DFS stack levels per lane kept in LDS (0 =
                                80 for width 4, 96 for width 5 Manhattan, 64
                                for the PDB search; at most 160);
                                deeper subtrees spill to the work queue */
    int32_t lb_policy;       /* SPZ_LB_* */
    int32_t device;          /* HIP device ordinal */
    int32_t time_kernels;    /* 1: time every search-round launch with HIP
                                events on the solver's stream (stats 8-9);
                                the rounds then run without the hipGraph */
    int32_t use_graph;       /* 1: replay the search rounds from a captured
                                hipGraph (default); 0: launch them one by one */
    int32_t log_samples;     /* >0: keep up to this many (frontier node,
                                threshold, subtree count) samples per solve
                                (spz_get_samples); training data of the MLP */
    int64_t queue_items;     /* capacity in items of each of the two work-queue
                                rings (current iterations, lookahead), rounded
                                up to a power of two (0 = auto; at most
                                2^30) */
    int32_t record_paths;    /* 1: record one optimal solution path per instance
                                (spz_get_paths); needs max_split_depth <= 32
                                and at most 2^27 queue items a ring */
    int32_t lookahead;       /* 1 (default): while lanes are idle, start an
                                instance's next iteration beside the one that
                                is draining; counts and lengths are unchanged
                                (each iteration counts apart, results resolve
                                in order), work of a next iteration that turns
                                out not to be needed is dropped.  Off under
                                SPZ_LB_PREVLOAD and sample logging.  0: one
                                iteration at a time, as in the reference
                                (src/24puzzle.cc:146-152) */
    int32_t round_time_us;   /* a search round also ends this many
                                microseconds after its start (whichever comes
                                first with round_budget steps), so that waves
                                that step faster under load do not wait for
                                the slowest; 0 = auto (4500 for the PDB search;
                                Manhattan 3500 when the batch has work for
                                every lane, else 700), < 0 = off */
    int32_t dedup;           /* 1: duplicate pruning in the root split: a
                                top-tree node whose board an earlier kept node
                                of its instance holds with a path-max f no
                                larger is dropped before it is expanded (the
                                closed-list root distribution of
                                src_horie/idas_smem.cu:1057-1149).  Optimal
                                lengths are unchanged; node counts become
                                smaller than the reference's, so parity runs
                                leave it off (default 0) */
    int32_t resplit;         /* heavy-node re-split under the SPZ_LB_PREVLOAD
                                and SPZ_LB_LEARNED policies: when an iteration
                                starts, a frontier node whose estimated
                                subtree (previous iteration's count, or the
                                prediction) exceeds resplit x the mean over
                                the nodes that fit is queued as its children
                                that fit (beyond 8 x that: grandchildren), the
                                reference's static load balancing
                                (src/15puzzle_block_parallel.cu:490-563,
                                divide_root_set :325-375).  Counts and lengths
                                are unchanged.  0 = off (default) */
} spz_config;

typedef struct spz_result {
    int32_t opt_len; /* optimal length, -1 if not found within max_iters */
    int32_t h0;      /* heuristic value of the start board */
    int32_t n_iters; /* iterations run (thresholds h0, h0+2, ...) */
    int32_t status;  /* SPZ_ST_* */
    uint64_t total_nodes; /* sum of the per-iteration counts */
} spz_result;

/* Fill *cfg with defaults for a width/heuristic (no device work). */
void spz_default_config(spz_config *cfg, int32_t width, int32_t heuristic);

/* Create a context on cfg->device.  Replaces the per-process globals and
 * cudaMalloc of src/15puzzle_block_parallel.cu:424-452. */
int spz_create(spz_ctx **ctx, const spz_config *cfg);

/* Change the launch configuration of a context (workgroup size, lanes,
 * LDS stack depth, root split, round budget, policy ...); width, heuristic
 * and device stay fixed and built PDBs are kept.  No reference counterpart
 * (the reference fixes these at compile time, src/15puzzle_block_parallel.cu:18-38). */
int spz_reconfigure(spz_ctx *ctx, const spz_config *cfg);
void spz_get_config(spz_ctx *ctx, spz_config *cfg);

/* Autotuner (no reference counterpart; the reference's shared-memory sweep
 * was done by hand, main.tex:742): coordinate descent over workgroup size
 * 64/128/256, root-split frontier target (= split depth), LDS stack depth,
 * round budget, lookahead on/off, the round time limit on/off and, under
 * SPZ_LB_PREVLOAD or SPZ_LB_LEARNED, the re-split factor 0/1/2/4, timing
 * spz_solve_batch on the sample boards.  Every candidate must reproduce the
 * first one's lengths, thresholds and per-iteration counts (in first-solution
 * mode all but the final, partial count); otherwise SPZ_E_MISMATCH.  With
 * duplicate pruning the frontier target stays fixed (pruned counts depend on
 * the split depth).  The context keeps the fastest configuration, also
 * returned in *best. */
int spz_autotune(spz_ctx *ctx, const uint8_t *boards, int32_t n, int32_t max_iters, int32_t passes,
                 spz_config *best, double *best_ms, int32_t *tried);

/* Build the 6-6-6-6 pattern databases on the GPU by backward 0-1 BFS from
 * the goal (blank at cell 0).  Replaces loading Korf's pat24.*.tab files
 * (src/24puzzle_with_pdb.cc:77-112; src/24puzzle_block_parallel_with_pdb.cu:
 * 164-230,943-944).  cache_path: if non-NULL and the file exists, the two
 * canonical tables are read from it instead (a file of this library: magic,
 * sizes and per-table checksums are checked, SPZ_E_IO if any differs); if it
 * does not exist they are written there after the build (through a
 * temporary file renamed into place). */
int spz_build_pdb(spz_ctx *ctx, const char *cache_path);

/* Load the reference's two table files instead of building them: the
 * pattern-database input of src/24puzzle_with_pdb.cc:77-112 (PatternDataBase::
 * input reading ../pdb/pat24.1256712.tab and ../pdb/pat24.34891314.tab).  Each
 * file must be exactly 127,512,000 bytes (one per placement, nested-loop
 * order) with 0 at the goal placement; SPZ_E_IO otherwise. */
int spz_load_pdb_tab(spz_ctx *ctx, const char *path_1256712, const char *path_34891314);

/* Write the two canonical tables in the reference's .tab layout (one byte
 * per placement, nested-loop order, src/24puzzle_with_pdb.cc:87-112). */
int spz_export_pdb_tab(spz_ctx *ctx, const char *path_1256712,
                       const char *path_34891314);

/* Copy a canonical table (pattern 0: tiles 1,2,5,6,7,12; pattern 1: tiles
 * 3,4,8,9,13,14) to host memory (127,512,000 bytes). */
int spz_get_pdb(spz_ctx *ctx, int32_t pattern, uint8_t *out);

/* Solve n boards.  Replaces ida_star() of src/24puzzle_expand.cc:146-164
 * (all-solutions) / src/24puzzle.cc:146-162 (first-solution) and the GPU
 * driver src/15puzzle_block_parallel.cu:429-565, for a whole batch in one
 * device-resident search.  out[i] receives instance i's result;
 * iter_counts (may be NULL) receives n*max_iters u64 per-iteration counts,
 * row i = thresholds h0_i, h0_i+2, ... (unused entries zero).  max_iters in
 * 1..SPZ_MAX_ITERS, n at most 2^24; SPZ_E_ARG otherwise. */
#define SPZ_MAX_ITERS 128
int spz_solve_batch(spz_ctx *ctx, const uint8_t *boards, int32_t n,
                    int32_t max_iters, spz_result *out, uint64_t *iter_counts);

/* Heuristic value of boards (Manhattan, or PDB after spz_build_pdb),
 * computed on the device; h must hold n int32. */
int spz_eval_h(spz_ctx *ctx, const uint8_t *boards, int32_t n, int32_t *h);

/* spz_eval_h that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, allocation-free harnesses).  The boards and
 * values pass through a static device array of the library (1.9 MB per GPU,
 * held until the process ends): at most 65536 boards per call (SPZ_E_ARG
 * beyond).  Every call on a GPU shares that array, so calls that overlap on
 * one GPU (other threads, other contexts) race; serialise them. */
#define SPZ_NOALLOC_MAX_BOARDS 65536
int spz_eval_h_noalloc(spz_ctx *ctx, const uint8_t *boards, int32_t n, int32_t *h);

/* Change the dispatch policy (SPZ_LB_*) between solves.  The policy only
 * orders the frontier nodes an iteration queues; results never depend on
 * it.  Replaces the host re-split of heavy roots between iterations
 * (src/15puzzle_block_parallel.cu:490-563). */
int spz_set_lb(spz_ctx *ctx, int32_t lb_policy);

/* Learned load balancing (no reference counterpart): an MLP
 * 640 -> 256 -> 256 -> 1 (ReLU, bf16 weights/activations, fp32 biases and
 * accumulation) predicting log2(1 + subtree size) of a frontier node at a
 * threshold.  Features: one-hot of tile t at cell c (index t*width^2 + c),
 * h/32 at 625, (threshold - h)/32 at 626.  Parameters are fp32, flat:
 * W1[640][256], b1[256], W2[256][256], b2[256], w3[256], b3. */
int32_t spz_mlp_param_count(void);
int spz_mlp_init(spz_ctx *ctx, uint64_t seed);
int spz_mlp_set_params(spz_ctx *ctx, const float *params);
int spz_mlp_get_params(spz_ctx *ctx, float *params);
/* forward pass on the GPU: out[i] for boards[i] at threshold thr[i] */
int spz_mlp_forward(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr, int32_t n, float *out);
/* the frontier predictor of SPZ_LB_LEARNED (one fused MFMA kernel): out[i*16 + m]
 * for boards[i] at threshold thr0[i] + 2m, m = 0..15 (out holds 16 n floats) */
int spz_mlp_predict(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr0, int32_t n, float *out);
/* loss = mean (y - target)^2 and its gradient w.r.t. the parameters */
int spz_mlp_grad(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr, const float *target, int32_t n,
                 float *loss, float *grad);
/* Adam training on the GPU; epoch_loss (may be NULL) gets the mean batch
 * loss of each epoch */
int spz_mlp_train(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr, const float *target, int32_t n,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss);
/* Samples logged by the last solve (config log_samples): boards (n x
 * width^2), thresholds and subtree counts.  *n_out = number copied; with
 * boards == NULL, the number available. */
int spz_get_samples(spz_ctx *ctx, uint8_t *boards, int32_t *thr, uint64_t *loads, int32_t max, int32_t *n_out);

/* Solution paths of the last solve (config record_paths): lens[i] = length of
 * the path recorded for instance i (its optimal length; -1 if none, e.g. not
 * solved within max_iters or longer than 256 moves); moves[i*max_len + j],
 * j < min(lens[i], max_len) = the j-th move of the blank: 0 up, 1 right,
 * 2 left, 3 down -- the direction codes the reference's stack_dump prints
 * (src_horie/idas_cpu.c:50-57; src_horie/idas_dynamic.cu:1494-1504 prints
 * them as U R L D).  No reference counterpart for the prefix above a GPU
 * work item (the reference leaves it out, idas_dynamic.cu:1496-1497). */
int spz_get_paths(spz_ctx *ctx, uint8_t *moves, int32_t max_len, int32_t *lens);

/* Statistics of the last solve (stats must hold SPZ_STATS_LEN doubles):
 *  0 rounds run            1 solve wall ms        2 items donated
 *  3 lanes                 4,5 PDB BFS levels     6 PDB build ms
 *  7 stack-overflow items  8 round-kernel ms (sum, when time_kernels)
 *  9 round launches timed  10 split ms            11 instances
 *  12 nodes (all instances, all iterations)
 *  13 setup ms (upload, root split, first restart)  14 rounds wall ms
 *  15 1 if the rounds ran from a hipGraph
 *  16-19 with env SPZ_PROFILE=1: lane-steps idle, expanding, only popping,
 *        and wave-steps (x64 = lane-steps) of the search rounds;
 *  20-22 with SPZ_PROFILE=1 (PDB): moves into a sibling (2 table reads each),
 *        children's first-partition reads, second-partition reads
 *  23 iterations started as lookahead
 *  24 frontier nodes re-split (config resplit)  25 the queue items they became
 *  26-31 reserved (0)
 * With env SPZ_TRACE=1 each poll of the round loop prints elapsed ms, busy
 * lanes, active instances and queued items to stderr. */
#define SPZ_STATS_LEN 32
int spz_last_stats(spz_ctx *ctx, double *stats);

const char *spz_last_error(spz_ctx *ctx);
void spz_destroy(spz_ctx *ctx);

/* Diagnostics: uniformly random byte gathers over the first span_bytes of
 * the two direct tables (span_bytes <= 0: all 2 x 25^6 bytes) -- the
 * search's access pattern without the search.  Over the whole tables nearly
 * every gather misses L2; over a span of a few MB nearly every one hits:
 * the two measured random-access ceilings of the PDB kernel.  n_gathers in
 * 1..2^36.  No reference counterpart. */
int spz_gather_probe(spz_ctx *ctx, int64_t n_gathers, int64_t span_bytes, double *ms, double *gathers_per_s);

/* Diagnostics: the eight PDB lookups as (tiles in slot order, cell map),
 * derived from the cell symmetries; the reference writes them out as
 * hash0..hashref3 (src/24puzzle_with_pdb.cc:114-189).  tiles48 = 8x6,
 * maps200 = 8x25 (either may be NULL). */
void spz_pdb_geometry(uint8_t *tiles48, uint8_t *maps200);

/* Seeded instance generators (splitmix64-seeded xoshiro256**):
 * a non-backtracking walk of `steps` blank moves from the goal, or a
 * uniformly random solvable permutation.  The reference's generator
 * (src/make_24puzzle_problems.cc:54-77) backtracks; these do not.  `out`
 * takes width*width cells; SPZ_E_ARG (nothing written) for a null `out`, a
 * width outside 2..5 or negative steps. */
int32_t spz_gen_walk(int32_t width, int32_t steps, uint64_t seed, uint8_t *out);
int32_t spz_gen_random(int32_t width, uint64_t seed, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif /* SPZ_H */
