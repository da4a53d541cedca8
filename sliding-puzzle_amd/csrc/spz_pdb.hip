// spz_pdb.hip -- 6-6-6-6 pattern databases built on the GPU.
//
// The reference ships no tables and no generator (.gitignore:1-5); its
// solvers load Korf's pat24.1256712.tab / pat24.34891314.tab
// (src/24puzzle_with_pdb.cc:77-112).  Here each canonical table is computed
// by a level-synchronous 0-1 BFS backward from the goal:
//   state = (cells of the 6 pattern tiles, blank cell); the blank stepping
//   onto a pattern tile costs 1 (a pattern-tile move), onto any other cell 0.
//   Level 0 is the goal (pattern tiles home, blank at cell 0) closed under
//   0-cost moves; a table entry is the first level that reaches any blank
//   cell of that placement, i.e. the minimum over blank cells.
// Per placement the BFS keeps three 25-bit masks of blank cells (seen, in
// the current level, reached for the next level), indexed by the placement's
// rank in the reference's file order (25!/19! = 127,512,000 placements).
// The finished compact table is then spread into direct 25^6 addressing so
// that the search updates an index with one multiply-add per move.
#include "spz_board.h"

#include <string.h>

namespace spz {

constexpr uint32_t kPlacements = 127512000u; // 25*24*23*22*21*20
constexpr uint32_t kFull25 = (1u << 25) - 1u;

__host__ __device__ inline int sym_cell(int sym, int cell) {
    int r = cell / 5, c = cell % 5;
    switch (sym) {
    case 0: return cell;                     // identity
    case 1: return (4 - r) * 5 + (4 - c);    // rot180
    case 2: return (4 - c) * 5 + r;          // rot90
    case 3: return c * 5 + r;                // reflection about the diagonal
    case 4: return (4 - c) * 5 + (4 - r);    // rot180 o reflection
    default: return (4 - r) * 5 + c;         // rot90 o reflection
    }
}

// lookup k -> (canonical table, symmetry); the reference's hash0..hashref3
// (src/24puzzle_with_pdb.cc:114-189) in order.
static const int kLkTab[8] = {0, 1, 1, 1, 0, 1, 1, 1};
static const int kLkSym[8] = {0, 0, 1, 2, 3, 3, 4, 5};
static const uint8_t kPatA[6] = {1, 2, 5, 6, 7, 12};
static const uint8_t kPatB[6] = {3, 4, 8, 9, 13, 14};

void pdb_geo_build(PdbGeo *g) {
    memset(g, 0, sizeof(*g));
    static const uint32_t pw[6] = {9765625u, 390625u, 15625u, 625u, 25u, 1u};
    for (int k = 0; k < 8; ++k) {
        const uint8_t *canon = kLkTab[k] == 0 ? kPatA : kPatB;
        for (int c = 0; c < 25; ++c) g->map[k][c] = (uint8_t)sym_cell(kLkSym[k], c);
        for (int j = 0; j < 6; ++j) {
            // tile whose home maps onto the canonical tile's home
            int tile = -1;
            for (int x = 0; x < 25; ++x)
                if (sym_cell(kLkSym[k], x) == canon[j]) tile = x;
            g->tiles[k][j] = (uint8_t)tile;
            if (k < 4) {
                g->reg_k[tile] = (uint8_t)k;
                g->reg_mul[tile] = pw[j];
            } else {
                g->ref_k[tile] = (uint8_t)k;
                g->ref_mul[tile] = pw[j];
            }
        }
    }
    for (int t = 1; t < 25; ++t)
        for (int c = 0; c < 25; ++c) {
            g->w[t][c].x = (uint32_t)g->map[g->reg_k[t]][c] * g->reg_mul[t] | ((uint32_t)g->reg_k[t] << 28);
            g->w[t][c].y = (uint32_t)g->map[g->ref_k[t]][c] * g->ref_mul[t] | ((uint32_t)(g->ref_k[t] - 4) << 28);
        }
}

// ---------------------------------------------------------------------------
// rank / unrank of 6 distinct cells in the nested-loop file order
// (src/24puzzle_with_pdb.cc:92-104): digit i is the cell's position among
// the cells not used by slots < i.
__device__ __forceinline__ uint32_t pl_rank(const int (&s)[6]) {
    uint32_t used = 0, rank = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        uint32_t below = used & ((1u << s[i]) - 1u);
        rank = rank * (uint32_t)(25 - i) + (uint32_t)(s[i] - __builtin_popcount(below));
        used |= 1u << s[i];
    }
    return rank;
}

// rank of canonical pattern t's goal placement (each tile on its own cell),
// the host-side twin of pl_rank
uint32_t pdb_goal_rank(int t) {
    const uint8_t *pat = t == 0 ? kPatA : kPatB;
    uint32_t used = 0, rank = 0;
    for (int i = 0; i < 6; ++i) {
        const uint32_t below = used & ((1u << pat[i]) - 1u);
        rank = rank * (uint32_t)(25 - i) + (uint32_t)(pat[i] - __builtin_popcount(below));
        used |= 1u << pat[i];
    }
    return rank;
}

__device__ __forceinline__ void pl_unrank(uint32_t rank, int (&s)[6]) {
    int dig[6];
#pragma unroll
    for (int i = 5; i >= 0; --i) {
        dig[i] = (int)(rank % (uint32_t)(25 - i));
        rank /= (uint32_t)(25 - i);
    }
    uint32_t used = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        uint32_t free_ = ~used & kFull25;
        for (int k = 0; k < dig[i]; ++k) free_ &= free_ - 1u;
        s[i] = __builtin_ctz(free_);
        used |= 1u << s[i];
    }
}

__device__ __forceinline__ uint32_t dilate5(uint32_t m) {
    const uint32_t col0 = 0x0108421u, col4 = 0x1084210u;
    return (m | ((m << 1) & ~col0) | ((m >> 1) & ~col4) | (m << 5) | (m >> 5)) & kFull25;
}

__device__ __forceinline__ uint32_t flood5(uint32_t seed, uint32_t free_) {
    uint32_t m = seed & free_;
    for (;;) {
        uint32_t m2 = dilate5(m) & free_;
        if (m2 == m) return m;
        m = m2;
    }
}

__device__ __forceinline__ int nb5(int cell, int d) {
    int r = cell / 5, c = cell - r * 5;
    if (d == DIR_UP) return r > 0 ? cell - 5 : -1;
    if (d == DIR_RIGHT) return c < 4 ? cell + 1 : -1;
    if (d == DIR_LEFT) return c > 0 ? cell - 1 : -1;
    return r < 4 ? cell + 5 : -1;
}

// level d -> candidates of level d+1
__global__ void __launch_bounds__(256) k_bfs_expand(const uint32_t *__restrict__ cur, const uint32_t *__restrict__ seen,
                                                    uint32_t *nxt) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kPlacements) return;
    uint32_t m = cur[p];
    if (!m) return;
    int s[6];
    pl_unrank(p, s);
    uint32_t occ = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) occ |= 1u << s[j];
    while (m) {
        int b = __builtin_ctz(m);
        m &= m - 1u;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            int q = nb5(b, d);
            if (q < 0 || !(occ & (1u << q))) continue;
            int s2[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) s2[j] = (s[j] == q) ? b : s[j]; // the tile steps into the blank
            uint32_t p2 = pl_rank(s2);
            uint32_t bit = 1u << q; // the blank is where the tile was
            if (!(seen[p2] & bit)) atomicOr(&nxt[p2], bit);
        }
    }
}

// close the candidates under 0-cost blank moves; record first levels
__global__ void __launch_bounds__(256) k_bfs_close(uint32_t *cur, uint32_t *seen, uint32_t *nxt, uint8_t *val,
                                                   int level, unsigned long long *n_new) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t hit = 0;
    if (p < kPlacements) {
        uint32_t m = nxt[p];
        if (m) {
            nxt[p] = 0;
            int s[6];
            pl_unrank(p, s);
            uint32_t occ = 0;
#pragma unroll
            for (int j = 0; j < 6; ++j) occ |= 1u << s[j];
            m = flood5(m, kFull25 & ~occ) & ~seen[p];
            if (m) {
                if (val[p] == 0xffu) val[p] = (uint8_t)level;
                seen[p] |= m;
                hit = 1;
            }
        }
        cur[p] = m;
    }
    uint64_t b = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(n_new, (unsigned long long)__popcll(b));
}

__global__ void k_bfs_seed(uint32_t *cur, uint32_t *seen, uint8_t *val, int pat) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int s[6];
    const int A[6] = {1, 2, 5, 6, 7, 12}, B[6] = {3, 4, 8, 9, 13, 14};
    uint32_t occ = 0;
    for (int j = 0; j < 6; ++j) {
        s[j] = pat == 0 ? A[j] : B[j];
        occ |= 1u << s[j];
    }
    uint32_t p = pl_rank(s);
    uint32_t m = flood5(1u, kFull25 & ~occ); // blank at cell 0, closed under 0-cost moves
    cur[p] = m;
    seen[p] = m;
    val[p] = 0;
}

// compact (file order) -> direct 25^6 addressing; invalid tuples get 0
__global__ void __launch_bounds__(256) k_pdb_spread(const uint8_t *__restrict__ compact, uint8_t *direct) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kPdbDirect) return;
    uint32_t x = (uint32_t)i;
    int s[6];
#pragma unroll
    for (int j = 5; j >= 0; --j) {
        s[j] = (int)(x % 25u);
        x /= 25u;
    }
    uint32_t used = 0;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        ok = ok && !(used & (1u << s[j]));
        used |= 1u << s[j];
    }
    direct[i] = ok ? compact[pl_rank(s)] : (uint8_t)0;
}

// Random-gather probe: every thread reads `per_thread` bytes at uniformly
// random positions of the two direct tables (independent addresses, so many
// are in flight), the access pattern of the PDB search without its search.
__global__ void __launch_bounds__(256) k_gather_probe(const uint8_t *__restrict__ t, uint64_t span, int per_thread,
                                                      uint64_t seed, uint32_t *out) {
    uint64_t x = seed ^ ((uint64_t)(blockIdx.x * blockDim.x + threadIdx.x) * 0x9e3779b97f4a7c15ull);
    uint32_t acc = 0;
    for (int k = 0; k < per_thread; k += 4) {
        uint32_t i[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            x ^= x >> 12;
            x ^= x << 25;
            x ^= x >> 27;
            uint64_t r = x * 0x2545f4914f6cdd1dull;
            i[j] = (uint32_t)(((r >> 32) * span) >> 32); // < span
        }
        acc += t[i[0]] + t[i[1]] + t[i[2]] + t[i[3]];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// span < 2^32 bytes (the two direct tables are 2 x 25^6)
hipError_t launch_gather_probe(const uint8_t *t, uint64_t span, int blocks, int per_thread, uint64_t seed, uint32_t *out,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_gather_probe, dim3(blocks), dim3(256), 0, s, t, span, per_thread, seed, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host driver: builds canonical table `pat` (0 or 1) into `compact` (device,
// kPlacements bytes).  Scratch: 3 x kPlacements u32.  Returns levels run.
hipError_t pdb_bfs_build(int pat, uint8_t *compact, uint32_t *work, unsigned long long *d_cnt, hipStream_t s,
                         int *levels_out) {
    uint32_t *seen = work, *cur = work + kPlacements, *nxt = work + 2ull * kPlacements;
    hipError_t e;
    if ((e = hipMemsetAsync(work, 0, 3ull * kPlacements * 4, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(compact, 0xff, kPlacements, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bfs_seed, dim3(1), dim3(64), 0, s, cur, seen, compact, pat);
    dim3 grid((kPlacements + 255) / 256);
    int level = 0;
    for (; level < 254; ++level) {
        hipLaunchKernelGGL(k_bfs_expand, grid, dim3(256), 0, s, cur, seen, nxt);
        if ((e = hipMemsetAsync(d_cnt, 0, 8, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_bfs_close, grid, dim3(256), 0, s, cur, seen, nxt, compact, level + 1, d_cnt);
        unsigned long long h = 0;
        if ((e = hipMemcpyAsync(&h, d_cnt, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        if (h == 0) break;
    }
    if (levels_out) *levels_out = level + 1;
    return hipGetLastError();
}

hipError_t pdb_spread(const uint8_t *compact, uint8_t *direct, hipStream_t s) {
    dim3 grid((unsigned)((kPdbDirect + 255) / 256));
    hipLaunchKernelGGL(k_pdb_spread, grid, dim3(256), 0, s, compact, direct);
    return hipGetLastError();
}

} // namespace spz
