/*
 * spz_oracle.c -- CPU oracle (test infrastructure only; see spz_oracle.h).
 *
 * Every function states the reference lines it restates.  Nothing here is
 * used by the product library; the HIP solver is checked against it.
 */
#include "spz_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* geometry and moves                                                        */
/* ------------------------------------------------------------------------ */

/* Directions in the reference's trial order (src/24puzzle.cc:23-26:
 * order = {1,0,2,3} over dx/dy = up, right, left, down).  The reverse of d is
 * 3-d (src_horie/idas_cpu.c:7-12 uses the same numbering). */
enum { D_UP = 0, D_RIGHT = 1, D_LEFT = 2, D_DOWN = 3 };

static int neighbor(int w, int cell, int d)
{
    int r = cell / w, c = cell % w;
    switch (d) {
    case D_UP:    return r > 0 ? cell - w : -1;
    case D_RIGHT: return c < w - 1 ? cell + 1 : -1;
    case D_LEFT:  return c > 0 ? cell - 1 : -1;
    default:      return r < w - 1 ? cell + w : -1;
    }
}

static int iabs(int x) { return x < 0 ? -x : x; }

static int cell_dist(int w, int a, int b)
{
    return iabs(a / w - b / w) + iabs(a % w - b % w);
}

/* src/24puzzle.cc:89-98 (get_md_sum) with md from :100-108 (set_md) */
int orc_manhattan(int width, const uint8_t *board)
{
    int n2 = width * width, h = 0;
    for (int i = 0; i < n2; ++i)
        if (board[i] != 0)
            h += cell_dist(width, i, board[i]);
    return h;
}

int orc_is_permutation(int width, const uint8_t *board)
{
    int n2 = width * width;
    int seen[25] = {0};
    if (width < 2 || width > 5)
        return 0;
    for (int i = 0; i < n2; ++i) {
        if (board[i] >= n2 || seen[board[i]])
            return 0;
        seen[board[i]] = 1;
    }
    return 1;
}

/* Reachability of the blank-at-0 goal: every move is a transposition of two
 * cells (flips the permutation parity) and moves the blank one cell (flips
 * the parity of the blank's row+col).  The goal has both even, so a board is
 * solvable iff parity(permutation of all n2 cells) == parity(row+col of blank);
 * the invariant is also sufficient for w >= 2 (Johnson & Story 1879). */
int orc_is_solvable(int width, const uint8_t *board)
{
    int n2 = width * width, inv = 0, blank = -1;
    if (!orc_is_permutation(width, board))
        return 0;
    for (int i = 0; i < n2; ++i) {
        if (board[i] == 0)
            blank = i;
        for (int j = i + 1; j < n2; ++j)
            if (board[i] > board[j])
                ++inv;
    }
    return (inv & 1) == ((blank / width + blank % width) & 1);
}

/* ------------------------------------------------------------------------ */
/* PRNG: splitmix64 seeding xoshiro256** (Blackman & Vigna, public domain)    */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i)
        r->s[i] = splitmix64(&x);
}

static uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform integer in [0, k): top 32 bits scaled by k */
static int rng_below(rng_t *r, int k)
{
    return (int)(((rng_next(r) >> 32) * (uint64_t)k) >> 32);
}

uint64_t orc_rng_u64(uint64_t seed, int nth)
{
    rng_t r;
    uint64_t v = 0;
    rng_seed(&r, seed);
    for (int i = 0; i <= nth; ++i)
        v = rng_next(&r);
    return v;
}

/* Non-backtracking random walk of `steps` blank moves from the goal.  Each
 * step is uniform over the legal moves other than the reverse of the previous
 * one (candidates enumerated in the up, right, left, down order). */
void orc_walk(int width, int steps, uint64_t seed, uint8_t *out)
{
    int n2 = width * width, blank = 0, prev = -1;
    rng_t r;
    rng_seed(&r, seed);
    for (int i = 0; i < n2; ++i)
        out[i] = (uint8_t)i;
    for (int s = 0; s < steps; ++s) {
        int cand[4], nc = 0;
        for (int d = 0; d < 4; ++d) {
            if (prev >= 0 && d == 3 - prev)
                continue;
            if (neighbor(width, blank, d) >= 0)
                cand[nc++] = d;
        }
        int d = cand[rng_below(&r, nc)];
        int nb = neighbor(width, blank, d);
        out[blank] = out[nb];
        out[nb] = 0;
        blank = nb;
        prev = d;
    }
}

/* Fisher-Yates shuffle of 0..n2-1, redrawn until solvable. */
void orc_random_solvable(int width, uint64_t seed, uint8_t *out)
{
    int n2 = width * width;
    rng_t r;
    rng_seed(&r, seed);
    for (;;) {
        for (int i = 0; i < n2; ++i)
            out[i] = (uint8_t)i;
        for (int i = n2 - 1; i > 0; --i) {
            int j = rng_below(&r, i + 1);
            uint8_t t = out[i];
            out[i] = out[j];
            out[j] = t;
        }
        if (orc_is_solvable(width, out))
            return;
    }
}

/* ------------------------------------------------------------------------ */
/* PDB: index, lookups, tables                                               */
/* ------------------------------------------------------------------------ */

/* Rank of k distinct cells of n2 in the reference's nested-loop file order
 * (src/24puzzle_with_pdb.cc:87-112): s0 outermost, each loop over the cells
 * not yet used, in increasing order.  rank = sum r_i * (n2-1-i)!/(n2-k)!,
 * r_i = s_i - #{j<i : s_j < s_i}. */
static uint32_t rank_k(int n2, int k, const uint8_t *s)
{
    uint32_t rank = 0;
    for (int i = 0; i < k; ++i) {
        int ri = s[i];
        for (int j = 0; j < i; ++j)
            if (s[j] < s[i])
                --ri;
        rank = rank * (uint32_t)(n2 - i) + (uint32_t)ri;
    }
    return rank;
}

static void unrank_k(int n2, int k, uint32_t rank, uint8_t *s)
{
    int r[8];
    for (int i = k - 1; i >= 0; --i) {
        r[i] = (int)(rank % (uint32_t)(n2 - i));
        rank /= (uint32_t)(n2 - i);
    }
    uint32_t used = 0;
    for (int i = 0; i < k; ++i) {
        int c = -1, left = r[i];
        for (int v = 0; v < n2; ++v) {
            if (used & (1u << v))
                continue;
            if (left-- == 0) {
                c = v;
                break;
            }
        }
        s[i] = (uint8_t)c;
        used |= 1u << c;
    }
}

static uint32_t table_size_k(int n2, int k)
{
    uint32_t sz = 1;
    for (int i = 0; i < k; ++i)
        sz *= (uint32_t)(n2 - i);
    return sz;
}

uint32_t orc_pdb_rank(const uint8_t cells[6]) { return rank_k(25, 6, cells); }
void orc_pdb_unrank(uint32_t rank, uint8_t cells[6]) { unrank_k(25, 6, rank, cells); }

/* canonical patterns, tiles in slot order (src/24puzzle_with_pdb.cc:116,122) */
static const uint8_t PAT_A[6] = {1, 2, 5, 6, 7, 12};
static const uint8_t PAT_B[6] = {3, 4, 8, 9, 13, 14};

void orc_pdb_pattern_tiles(int pattern, uint8_t tiles[6])
{
    memcpy(tiles, pattern == 0 ? PAT_A : PAT_B, 6);
}

/* Cell symmetries of the 5x5 board, (r,c) -> ... .  The reference tabulates
 * them as rf, rot90, rot90rf, rot180, rot180rf (src/24puzzle_with_pdb.cc:33-46);
 * here they are derived from the geometry. */
enum { SYM_ID, SYM_ROT180, SYM_ROT90, SYM_RF, SYM_ROT180RF, SYM_ROT90RF };

static int sym_cell(int sym, int cell)
{
    int r = cell / 5, c = cell % 5;
    switch (sym) {
    case SYM_ID:       return cell;
    case SYM_ROT180:   return (4 - r) * 5 + (4 - c);
    case SYM_ROT90:    return (4 - c) * 5 + r;
    case SYM_RF:       return c * 5 + r;
    case SYM_ROT180RF: return (4 - c) * 5 + (4 - r);
    default:           return (4 - r) * 5 + c; /* SYM_ROT90RF */
    }
}

/* The eight lookups of src/24puzzle_with_pdb.cc:114-189: table, symmetry,
 * canonical pattern.  Lookup k reads canonical table tab with the positions
 * of the tiles sym^-1(canonical tile) mapped through sym. */
static const int LK_TAB[8] = {0, 1, 1, 1, 0, 1, 1, 1};
static const int LK_SYM[8] = {SYM_ID, SYM_ID, SYM_ROT180, SYM_ROT90,
                              SYM_RF, SYM_RF, SYM_ROT180RF, SYM_ROT90RF};

static int sym_inverse_cell(int sym, int cell)
{
    for (int x = 0; x < 25; ++x)
        if (sym_cell(sym, x) == cell)
            return x;
    return -1;
}

void orc_pdb_lookup_spec(int k, int *tab, uint8_t map[25], uint8_t tiles[6])
{
    const uint8_t *canon = LK_TAB[k] == 0 ? PAT_A : PAT_B;
    *tab = LK_TAB[k];
    for (int c = 0; c < 25; ++c)
        map[c] = (uint8_t)sym_cell(LK_SYM[k], c);
    for (int j = 0; j < 6; ++j)
        tiles[j] = (uint8_t)sym_inverse_cell(LK_SYM[k], canon[j]);
}

static const uint8_t *g_tab[2];

void orc_pdb_set_tables(const uint8_t *t0, const uint8_t *t1)
{
    g_tab[0] = t0;
    g_tab[1] = t1;
}

/* the eight lookup specifications, derived once */
static int g_spec_ready;
static int g_spec_tab[8];
static uint8_t g_spec_map[8][25], g_spec_tiles[8][6];

static void spec_init(void)
{
    if (g_spec_ready)
        return;
    for (int k = 0; k < 8; ++k)
        orc_pdb_lookup_spec(k, &g_spec_tab[k], g_spec_map[k], g_spec_tiles[k]);
    g_spec_ready = 1;
}

static void pdb_lookups_pos(const uint8_t *pos, int out[8])
{
    spec_init();
    for (int k = 0; k < 8; ++k) {
        uint8_t cells[6];
        for (int j = 0; j < 6; ++j)
            cells[j] = g_spec_map[k][pos[g_spec_tiles[k][j]]];
        out[k] = g_tab[g_spec_tab[k]][rank_k(25, 6, cells)];
    }
}

void orc_pdb_lookups(const uint8_t *board, int out[8])
{
    uint8_t pos[25];
    for (int i = 0; i < 25; ++i)
        pos[board[i]] = (uint8_t)i;
    pdb_lookups_pos(pos, out);
}

/* src/24puzzle_with_pdb.cc:191-194: max(regular sum, reflected sum) */
static int pdb_h_pos(const uint8_t *pos)
{
    int v[8];
    pdb_lookups_pos(pos, v);
    int a = v[0] + v[1] + v[2] + v[3];
    int b = v[4] + v[5] + v[6] + v[7];
    return a > b ? a : b;
}

int orc_pdb_h(const uint8_t *board)
{
    uint8_t pos[25];
    for (int i = 0; i < 25; ++i)
        pos[board[i]] = (uint8_t)i;
    return pdb_h_pos(pos);
}

/* ------------------------------------------------------------------------ */
/* PDB generation: level-synchronous 0-1 BFS                                 */
/* ------------------------------------------------------------------------ */

/* Abstract state = (placement of the k pattern tiles, blank cell).  Moving
 * the blank onto a pattern tile costs 1, onto a free cell costs 0.  Level d
 * holds the states at distance d; a state set is closed under 0-cost moves
 * by flooding the blank through the free cells.  The stored value of a
 * placement is the first level at which any of its blank cells is reached,
 * i.e. the minimum over blank cells.  With goal_kind NO_BLANK the blank is
 * dropped and a pattern tile may step into any free cell at cost 1. */

static uint32_t grid_dilate(int w, uint32_t m)
{
    uint32_t full = (w * w == 32) ? 0xffffffffu : ((1u << (w * w)) - 1u);
    uint32_t col0 = 0, colw = 0;
    for (int r = 0; r < w; ++r) {
        col0 |= 1u << (r * w);
        colw |= 1u << (r * w + w - 1);
    }
    uint32_t out = m | ((m << 1) & ~col0) | ((m >> 1) & ~colw) | (m << w) |
                   (m >> w);
    return out & full;
}

static uint32_t grid_flood(int w, uint32_t seed, uint32_t free_mask)
{
    uint32_t m = seed & free_mask;
    for (;;) {
        uint32_t m2 = grid_dilate(w, m) & free_mask;
        if (m2 == m)
            return m;
        m = m2;
    }
}

/* generic builder: width w, k pattern tiles (tiles[] in slot order) */
static int pdb_build_generic(int w, int k, const uint8_t *tiles, int goal_kind,
                             uint8_t *table, int n_threads)
{
    int n2 = w * w;
    uint32_t size = table_size_k(n2, k);
    uint32_t full = (n2 == 32) ? 0xffffffffu : ((1u << n2) - 1u);
    uint32_t *seen = (uint32_t *)calloc(size, sizeof(uint32_t));
    uint32_t *cur = (uint32_t *)calloc(size, sizeof(uint32_t));
    uint32_t *nxt = (uint32_t *)calloc(size, sizeof(uint32_t));
    if (!seen || !cur || !nxt) {
        free(seen);
        free(cur);
        free(nxt);
        return -1;
    }
#ifdef _OPENMP
    omp_set_num_threads(n_threads > 0 ? n_threads : omp_get_num_procs());
#else
    (void)n_threads;
#endif
    memset(table, 0xff, size);

    /* level 0 */
    uint8_t home[8];
    uint32_t occ0 = 0;
    for (int j = 0; j < k; ++j) {
        home[j] = tiles[j];
        occ0 |= 1u << tiles[j];
    }
    uint32_t p0 = rank_k(n2, k, home);
    uint32_t free0 = full & ~occ0;
    uint32_t m0;
    if (goal_kind == ORC_PDB_GOAL_ANY_BLANK)
        m0 = free0;
    else if (goal_kind == ORC_PDB_GOAL_BLANK0)
        m0 = grid_flood(w, 1u, free0);
    else
        m0 = 1u; /* NO_BLANK: bit 0 = "placement reached" */
    cur[p0] = m0;
    seen[p0] = m0;
    table[p0] = 0;

    for (int d = 0; d < 255; ++d) {
        /* expand level d */
#pragma omp parallel for schedule(dynamic, 65536)
        for (int64_t pi = 0; pi < (int64_t)size; ++pi) {
            uint32_t m = cur[pi];
            if (!m)
                continue;
            uint8_t cells[8];
            unrank_k(n2, k, (uint32_t)pi, cells);
            uint32_t occ = 0;
            for (int j = 0; j < k; ++j)
                occ |= 1u << cells[j];
            if (goal_kind == ORC_PDB_GOAL_NO_BLANK) {
                for (int j = 0; j < k; ++j)
                    for (int dd = 0; dd < 4; ++dd) {
                        int q = neighbor(w, cells[j], dd);
                        if (q < 0 || (occ & (1u << q)))
                            continue;
                        uint8_t c2[8];
                        memcpy(c2, cells, (size_t)k);
                        c2[j] = (uint8_t)q;
                        uint32_t p2 = rank_k(n2, k, c2);
                        if (!seen[p2])
                            __atomic_fetch_or(&nxt[p2], 1u, __ATOMIC_RELAXED);
                    }
                continue;
            }
            while (m) {
                int b = __builtin_ctz(m);
                m &= m - 1;
                for (int dd = 0; dd < 4; ++dd) {
                    int q = neighbor(w, b, dd);
                    if (q < 0 || !(occ & (1u << q)))
                        continue;
                    int j = 0;
                    while (cells[j] != q)
                        ++j;
                    uint8_t c2[8];
                    memcpy(c2, cells, (size_t)k);
                    c2[j] = (uint8_t)b; /* tile steps into the blank */
                    uint32_t p2 = rank_k(n2, k, c2);
                    uint32_t bit = 1u << q; /* blank is where the tile was */
                    if (!(seen[p2] & bit))
                        __atomic_fetch_or(&nxt[p2], bit, __ATOMIC_RELAXED);
                }
            }
        }
        /* close level d+1 under 0-cost moves */
        int64_t nnew = 0;
#pragma omp parallel for schedule(static) reduction(+ : nnew)
        for (int64_t pi = 0; pi < (int64_t)size; ++pi) {
            uint32_t m = nxt[pi];
            nxt[pi] = 0;
            if (!m) {
                cur[pi] = 0;
                continue;
            }
            if (goal_kind != ORC_PDB_GOAL_NO_BLANK) {
                uint8_t cells[8];
                unrank_k(n2, k, (uint32_t)pi, cells);
                uint32_t occ = 0;
                for (int j = 0; j < k; ++j)
                    occ |= 1u << cells[j];
                m = grid_flood(w, m, full & ~occ);
            }
            m &= ~seen[pi];
            cur[pi] = m;
            if (m) {
                if (table[pi] == 0xff)
                    table[pi] = (uint8_t)(d + 1);
                seen[pi] |= m;
                ++nnew;
            }
        }
        if (nnew == 0)
            break;
    }
    free(seen);
    free(cur);
    free(nxt);
    return 0;
}

int orc_pdb_build_tiles(const uint8_t tiles[6], int goal_kind, uint8_t *table,
                        int n_threads)
{
    return pdb_build_generic(5, 6, tiles, goal_kind, table, n_threads);
}

int orc_pdb_build(int pattern, int goal_kind, uint8_t *table, int n_threads)
{
    return orc_pdb_build_tiles(pattern == 0 ? PAT_A : PAT_B, goal_kind, table,
                               n_threads);
}

/* small-board entry used by the brute-force tests */
int orc_pdb_build_small(int w, int k, const uint8_t *tiles, int goal_kind,
                        uint8_t *table)
{
    return pdb_build_generic(w, k, tiles, goal_kind, table, 1);
}
uint32_t orc_rank_small(int n2, int k, const uint8_t *s) { return rank_k(n2, k, s); }

/* ------------------------------------------------------------------------ */
/* IDA*                                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int w, heur, mode, limit;
    uint8_t b[25];
    uint8_t pos[25]; /* pos[tile] = cell, kept for the PDB */
    int blank, h;
    uint64_t count;
    int found, found_depth;
} ida_t;

/* recursive depth-first iteration: src/24puzzle_expand.cc:111-144 (ALL) and
 * src/24puzzle.cc:111-144 (FIRST).  The goal test precedes the f-limit test;
 * every in-bounds, non-reverse child is counted before its own f test. */
static int ida_dfs(ida_t *s, int g, int prev)
{
    if (s->h == 0) {
        s->found = 1;
        s->found_depth = g;
        if (s->mode == ORC_MODE_FIRST)
            return 1;
    }
    if (g + s->h > s->limit)
        return 0;
    for (int d = 0; d < 4; ++d) {
        if (prev >= 0 && d == 3 - prev)
            continue;
        int nb = neighbor(s->w, s->blank, d);
        if (nb < 0)
            continue;
        ++s->count;
        int old_blank = s->blank, old_h = s->h;
        uint8_t t = s->b[nb];
        s->b[old_blank] = t;
        s->b[nb] = 0;
        s->pos[t] = (uint8_t)old_blank;
        s->pos[0] = (uint8_t)nb;
        s->blank = nb;
        if (s->heur == ORC_H_MANHATTAN)
            s->h = old_h - cell_dist(s->w, nb, t) + cell_dist(s->w, old_blank, t);
        else
            s->h = pdb_h_pos(s->pos);
        int r = ida_dfs(s, g + 1, d);
        s->b[nb] = t;
        s->b[old_blank] = 0;
        s->pos[t] = (uint8_t)nb;
        s->pos[0] = (uint8_t)old_blank;
        s->blank = old_blank;
        s->h = old_h;
        if (r)
            return 1;
    }
    return 0;
}

/* threshold loop: src/24puzzle.cc:146-152 (limit = h0, h0+2, ...) */
int orc_ida(int width, const uint8_t *board, int heuristic, int mode,
            int max_iters, uint64_t *counts, int *thresholds, int *n_iters)
{
    ida_t s;
    *n_iters = 0;
    if (width < 2 || width > 5 || !orc_is_solvable(width, board))
        return -2;
    if (heuristic == ORC_H_PDB && (width != 5 || !g_tab[0] || !g_tab[1]))
        return -2;
    memset(&s, 0, sizeof s);
    s.w = width;
    s.heur = heuristic;
    s.mode = mode;
    for (int i = 0; i < width * width; ++i) {
        s.b[i] = board[i];
        s.pos[board[i]] = (uint8_t)i;
        if (board[i] == 0)
            s.blank = i;
    }
    int h0 = heuristic == ORC_H_MANHATTAN ? orc_manhattan(width, board)
                                          : pdb_h_pos(s.pos);
    for (int it = 0; it < max_iters; ++it) {
        s.limit = h0 + 2 * it;
        s.h = h0;
        s.count = 0;
        s.found = 0;
        ida_dfs(&s, 0, -1);
        counts[it] = s.count;
        thresholds[it] = s.limit;
        *n_iters = it + 1;
        if (s.found)
            return s.found_depth;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* IDA* over a root split, optionally with duplicate pruning                 */
/* ------------------------------------------------------------------------ */
/* The counts of orc_ida computed the way a block-parallel solver splits the
 * search (src/15puzzle_block_parallel.cu:150-187): the top of the tree is
 * expanded breadth-first, level by level, until a level holds at least
 * `target` nodes (or is at depth max_depth); that level is the frontier.
 * Iteration L then counts the generated children of every top node whose
 * path-max f (the largest f on its path from the root) is <= L, plus the
 * depth-first count below every frontier node whose path-max f is <= L.
 * Without pruning this equals orc_ida's count of every iteration.
 * With dedup, before a level is expanded, a node is dropped when a kept
 * node of the top tree (the root, a shallower node, or a copy in the same
 * level with a smaller (path-max f, previous move)) holds its board with a
 * path-max f no larger: the closed list of the reference's A* root
 * distribution (src_horie/idas_smem.cu:1057-1149), whose counts no
 * reference file holds.  Pinned for Manhattan by an 8-puzzle brute force
 * (a consistent heuristic keeps exactly the BFS layers); under the PDB max,
 * which is not consistent, parity unpinned beyond the properties (lengths
 * kept, no count above the unpruned one). */
typedef struct {
    uint8_t b[25];
    uint8_t g, h, F, prev; /* prev: 0..3, 4 = none */
} tnode_t;

static int g_n2; /* board size for the comparators (the oracle is single-threaded here) */

static int cmp_board_f_prev(const void *x, const void *y)
{
    const tnode_t *a = (const tnode_t *)x, *b = (const tnode_t *)y;
    int c = memcmp(a->b, b->b, (size_t)g_n2);
    if (c)
        return c;
    if (a->F != b->F)
        return (int)a->F - (int)b->F;
    return (int)a->prev - (int)b->prev;
}

typedef struct {
    uint8_t b[25];
    uint8_t F;
} seen_t;

static int cmp_seen(const void *x, const void *y)
{
    return memcmp(((const seen_t *)x)->b, ((const seen_t *)y)->b, (size_t)g_n2);
}

static int node_h(int w, int heur, const uint8_t *b)
{
    if (heur == ORC_H_MANHATTAN)
        return orc_manhattan(w, b);
    uint8_t pos[25];
    for (int i = 0; i < 25; ++i)
        pos[b[i]] = (uint8_t)i;
    return pdb_h_pos(pos);
}

/* prune a level in place against the sorted closed list; returns the new
 * length (the closed list grows) */
static size_t prune_level(tnode_t *lv, size_t n, seen_t **seen, size_t *ns, size_t *cap)
{
    qsort(lv, n, sizeof *lv, cmp_board_f_prev);
    /* the list is sorted up to here; this level's new boards are appended
     * after it (each board occurs once in the level) and sorted in at the end */
    const size_t sorted = *ns;
    size_t out = 0;
    for (size_t i = 0; i < n; ++i) {
        if (i > 0 && memcmp(lv[i].b, lv[i - 1].b, (size_t)g_n2) == 0)
            continue; /* not the board's best copy in this level */
        seen_t key;
        memcpy(key.b, lv[i].b, 25);
        seen_t *hit = (seen_t *)bsearch(&key, *seen, sorted, sizeof key, cmp_seen);
        if (hit) {
            if (hit->F <= lv[i].F)
                continue; /* a kept copy covers it */
            hit->F = lv[i].F;
        } else {
            if (*ns == *cap) {
                *cap = *cap ? 2 * *cap : 1024;
                *seen = (seen_t *)realloc(*seen, *cap * sizeof(seen_t));
            }
            key.F = lv[i].F;
            (*seen)[(*ns)++] = key;
        }
        lv[out++] = lv[i];
    }
    qsort(*seen, *ns, sizeof(seen_t), cmp_seen);
    return out;
}

int orc_ida_split(int width, const uint8_t *board, int heuristic, int target, int max_depth,
                  int dedup, int max_iters, uint64_t *counts, int *thresholds, int *n_iters)
{
    *n_iters = 0;
    if (width < 2 || width > 5 || !orc_is_solvable(width, board) || target < 1 || max_depth < 0)
        return -2;
    if (heuristic == ORC_H_PDB && (width != 5 || !g_tab[0] || !g_tab[1]))
        return -2;
    const int n2 = width * width;
    g_n2 = n2;
    size_t cap = 16, n = 1;
    tnode_t *lv = (tnode_t *)calloc(cap, sizeof(tnode_t));
    memcpy(lv[0].b, board, (size_t)n2);
    const int h0 = node_h(width, heuristic, board);
    lv[0].g = 0;
    lv[0].h = (uint8_t)h0;
    lv[0].F = (uint8_t)h0;
    lv[0].prev = 4;
    seen_t *seen = NULL;
    size_t ns = 0, scap = 0;
    if (dedup) {
        scap = 1024;
        seen = (seen_t *)calloc(scap, sizeof(seen_t));
        memcpy(seen[0].b, board, (size_t)n2);
        seen[0].F = (uint8_t)h0;
        ns = 1;
    }
    uint64_t topgen[256];
    memset(topgen, 0, sizeof topgen);
    int top_goal = 0x7fffffff;
    for (int depth = 0; depth <= max_depth; ++depth) {
        if (dedup && depth > 0)
            n = prune_level(lv, n, &seen, &ns, &scap);
        if (n >= (size_t)target || depth >= max_depth)
            break;
        size_t ncap = 4 * n, nn = 0;
        tnode_t *nx = (tnode_t *)calloc(ncap, sizeof(tnode_t));
        for (size_t i = 0; i < n; ++i) {
            const tnode_t *u = &lv[i];
            int blank = 0;
            while (u->b[blank] != 0)
                ++blank;
            if (u->h == 0 && u->F < top_goal)
                top_goal = u->F;
            for (int d = 0; d < 4; ++d) {
                if (u->prev < 4 && d == 3 - u->prev)
                    continue;
                int nb = neighbor(width, blank, d);
                if (nb < 0)
                    continue;
                topgen[u->F] += 1;
                tnode_t *c = &nx[nn++];
                *c = *u;
                c->b[blank] = u->b[nb];
                c->b[nb] = 0;
                c->g = (uint8_t)(u->g + 1);
                c->h = (uint8_t)node_h(width, heuristic, c->b);
                int f = c->g + c->h;
                c->F = (uint8_t)(f > u->F ? f : u->F);
                c->prev = (uint8_t)d;
            }
        }
        free(lv);
        lv = nx;
        n = nn;
    }
    int result = -1;
    for (int it = 0; it < max_iters; ++it) {
        const int L = h0 + 2 * it;
        uint64_t cnt = 0;
        for (int f = 0; f < 256 && f <= L; ++f)
            cnt += topgen[f];
        int found = top_goal <= L ? top_goal : 0x7fffffff;
        for (size_t i = 0; i < n; ++i) {
            if (lv[i].F > L)
                continue;
            ida_t s;
            memset(&s, 0, sizeof s);
            s.w = width;
            s.heur = heuristic;
            s.mode = ORC_MODE_ALL;
            s.limit = L;
            for (int c = 0; c < n2; ++c) {
                s.b[c] = lv[i].b[c];
                s.pos[lv[i].b[c]] = (uint8_t)c;
                if (lv[i].b[c] == 0)
                    s.blank = c;
            }
            s.h = lv[i].h;
            ida_dfs(&s, lv[i].g, lv[i].prev < 4 ? lv[i].prev : -1);
            cnt += s.count;
            if (s.found && s.found_depth < found)
                found = s.found_depth;
        }
        counts[it] = cnt;
        thresholds[it] = L;
        *n_iters = it + 1;
        if (found != 0x7fffffff) {
            result = found;
            break;
        }
    }
    free(lv);
    free(seen);
    return result;
}
