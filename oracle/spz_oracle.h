/*
 * spz_oracle.h -- CPU oracle for the sliding-puzzle IDA* hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP solver
 * (libspz) is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * It is a plain-C restatement of the reference's CPU solvers:
 *   - Manhattan IDA*, all-solutions ("expand") and first-solution modes:
 *       src/24puzzle_expand.cc:111-164, src/24puzzle.cc:111-162,
 *       src/15puzzle_expand.cc:111-164 (same algorithm, width 4);
 *       per-iteration counts as printed by src_horie/idas_cpu_expand.c:226-310.
 *   - 6-6-6-6 additive PDB heuristic with reflection max:
 *       src/24puzzle_with_pdb.cc:114-194; table file layout :87-112.
 *   - PDB generation (absent from the reference, .gitignore:1-5): 0-1 BFS
 *       over (pattern placement, blank cell), Korf & Felner 2002 semantics,
 *       value = min over blank cells of the number of pattern-tile moves.
 *   - instance generators (seeded walks / solvable permutations).
 *
 * Conventions (src/24puzzle.cc:89-108): board[i] = tile at cell i, row-major,
 * 0 = blank, goal has tile t at cell t (blank at cell 0).  Operators are tried
 * in the order up, right, left, down; the reverse of the previous move is
 * never generated (src/24puzzle.cc:26,126).
 */
#ifndef SPZ_ORACLE_H
#define SPZ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MODE_ALL   0   /* exhaust the final iteration (reference *_expand) */
#define ORC_MODE_FIRST 1   /* stop at the first goal (reference 24puzzle.cc) */

#define ORC_PDB_ENTRIES 127512000u /* 25!/19! */

/* heuristic selectors for orc_ida */
#define ORC_H_MANHATTAN 0
#define ORC_H_PDB       1

/* PDB goal semantics for the builder:
 *  0: blank tracked, goal = pattern tiles home with the blank in ANY free cell
 *  1: blank tracked, goal = pattern tiles home with the blank at cell 0
 *  2: blank not tracked: a pattern tile may step into any cell not holding
 *     another pattern tile                                                  */
#define ORC_PDB_GOAL_ANY_BLANK  0
#define ORC_PDB_GOAL_BLANK0     1
#define ORC_PDB_GOAL_NO_BLANK   2

int orc_manhattan(int width, const uint8_t *board);
int orc_is_solvable(int width, const uint8_t *board);
int orc_is_permutation(int width, const uint8_t *board);

/* Seeded instance generators (splitmix64-seeded xoshiro256**). */
void orc_walk(int width, int steps, uint64_t seed, uint8_t *out);
void orc_random_solvable(int width, uint64_t seed, uint8_t *out);
uint64_t orc_rng_u64(uint64_t seed, int nth); /* nth output of the stream */

/*
 * IDA* with thresholds h0, h0+2, ... (src/24puzzle.cc:146-152).
 * counts[k], thresholds[k] receive iteration k (at most max_iters).
 * Returns the optimal length (or the first-found depth in FIRST mode, which is
 * optimal too), -1 if max_iters iterations found nothing, -2 on bad input.
 * *n_iters receives the number of iterations run.
 * heuristic ORC_H_PDB needs orc_pdb_set_tables() first (width 5 only).
 */
int orc_ida(int width, const uint8_t *board, int heuristic, int mode,
            int max_iters, uint64_t *counts, int *thresholds, int *n_iters);

/*
 * The same iterations computed over a root split (top tree expanded level by
 * level until a level holds >= target nodes or reaches max_depth; counts =
 * top-tree generated children with path-max f <= L + depth-first counts below
 * the frontier nodes with path-max f <= L), all-solutions mode.  dedup = 0
 * equals orc_ida; dedup = 1 prunes duplicate boards from the top tree (a node
 * whose board a kept node -- the root, a shallower node, or a same-level copy
 * with smaller (path-max f, previous move) -- holds with a path-max f no larger
 * is dropped before its level is expanded), the closed list of
 * src_horie/idas_smem.cu:1057-1149.  Pruned counts: pinned for Manhattan
 * (8-puzzle BFS brute force); parity unpinned under the PDB heuristic beyond
 * kept lengths and counts no larger than the unpruned ones.
 */
int orc_ida_split(int width, const uint8_t *board, int heuristic, int target, int max_depth,
                  int dedup, int max_iters, uint64_t *counts, int *thresholds, int *n_iters);

/* --- PDB (24-puzzle, 6-6-6-6) ------------------------------------------ */
/* rank of 6 distinct cells in the reference's nested-loop file order */
uint32_t orc_pdb_rank(const uint8_t cells[6]);
void orc_pdb_unrank(uint32_t rank, uint8_t cells[6]);
/* tiles of canonical pattern p (0: 1,2,5,6,7,12  1: 3,4,8,9,13,14) */
void orc_pdb_pattern_tiles(int pattern, uint8_t tiles[6]);
/* build canonical table (ORC_PDB_ENTRIES bytes) for pattern 0 or 1 */
int orc_pdb_build(int pattern, int goal_kind, uint8_t *table, int n_threads);
/* build a table for an arbitrary 6-tile pattern given as tile list (slot
 * order); used for the 4-table self-consistency check */
int orc_pdb_build_tiles(const uint8_t tiles[6], int goal_kind, uint8_t *table,
                        int n_threads);
/* register the two canonical tables used by ORC_H_PDB (pointers are kept) */
void orc_pdb_set_tables(const uint8_t *t0, const uint8_t *t1);
/* PDB heuristic of a 5x5 board: max(regular sum, reflected sum) */
int orc_pdb_h(const uint8_t *board);
/* the 8 individual lookups: out[0..3] regular, out[4..7] reflected */
void orc_pdb_lookups(const uint8_t *board, int out[8]);
/* The 8 lookups as (table id, cell map, tile list).  For lookup k:
 * value = table[tab[k]][rank(map[k][pos(tiles[k][0..5])])]              */
void orc_pdb_lookup_spec(int k, int *tab, uint8_t map[25], uint8_t tiles[6]);
/* generic small-board builder (width w, k pattern tiles; table of
 * (w*w)!/(w*w-k)! bytes) and its rank, for the brute-force tests */
int orc_pdb_build_small(int w, int k, const uint8_t *tiles, int goal_kind,
                        uint8_t *table);
uint32_t orc_rank_small(int n2, int k, const uint8_t *s);

#ifdef __cplusplus
}
#endif
#endif
