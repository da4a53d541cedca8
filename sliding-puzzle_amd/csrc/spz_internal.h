// spz_internal.h -- device-side data layout shared by the search kernels,
// the PDB builder and the host driver of libspz.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spz {

// Blank-move directions in the reference's trial order: up, right, left,
// down (src/24puzzle.cc:23-26, order = {1,0,2,3}); reverse(d) = 3 - d
// (src_horie/idas_cpu.c:7-12).
enum : int { DIR_UP = 0, DIR_RIGHT = 1, DIR_LEFT = 2, DIR_DOWN = 3, DIR_NONE = 4 };

// Blank displacement of move d (up, right, left, down): row change, column
// change and, on a width-W board, cell change, each the signed byte d of a
// constant -- a shift and a sign extension, no branches.
__device__ __forceinline__ int dir_byte(uint32_t packed, int d) { return (int)(int8_t)(packed >> (8 * d)); }
__device__ __forceinline__ int dir_dr(int d) { return dir_byte(0x010000FFu, d); }
__device__ __forceinline__ int dir_dc(int d) { return dir_byte(0x00FF0100u, d); }
template <int W> __device__ __forceinline__ int dir_cell(int d) {
    return dir_byte(((uint32_t)W << 24) | 0x00FF0100u | (uint32_t)(uint8_t)(-W), d);
}

// Lane states of the DFS engine.
// lane states; ST_DESC (children f-tested, descent into the first passing one
// pending) never outlives a round
constexpr int kQCtl = 16; // u64 stride between the two rings' counters

enum : int { ST_IDLE = 0, ST_FRESH = 1, ST_POP = 2, ST_DRAINED = 3, ST_DESC = 4 };

constexpr int kMaxIters = 128;      // threshold iterations kept per instance
constexpr int kFBins = 256;         // f-value histogram bins of the top tree
constexpr int kNoGoal = 0x7fffffff; // found-depth sentinel

// A search node in the work queue (32 bytes).  The board is packed:
//  width 4: b[0] = row0 | row1<<16, b[1] = row2 | row3<<16 (4-bit cells)
//  width 5: b[0..1] = u64 of cells 0-11 (5 bits each, bits 0-59) with the low
//  4 bits of cell 24 at bits 60-63; b[2..3] = u64 of cells 12-23 with the high
//  bit of cell 24 at bit 60
// meta = blank | g<<8 | h<<16 | prev<<24; x0 = largest f on the root path
// (root split), x1 = index of the frontier node the item descends from.
struct alignas(16) Item {
    uint32_t b[4];
    uint32_t meta;
    uint32_t inst;
    uint32_t x0, x1;
};

// A logged training sample of the learned load balancer: a frontier node,
// the threshold of an iteration, and the nodes generated below it then.
struct Sample {
    Item item;
    int32_t thr;
    uint32_t pad;
    unsigned long long load;
};

constexpr int kPathWords = 16;            // path storage: 16 words of 16 moves
constexpr int kPathMax = 16 * kPathWords; // longest recordable path (256 moves)
constexpr int kPredK = 16; // thresholds predicted per frontier node

__host__ __device__ inline uint32_t meta_pack(int blank, int g, int h, int prev) {
    return (uint32_t)blank | ((uint32_t)g << 8) | ((uint32_t)h << 16) | ((uint32_t)prev << 24);
}
__host__ __device__ inline int meta_blank(uint32_t m) { return (int)(m & 0xff); }
__host__ __device__ inline int meta_g(uint32_t m) { return (int)((m >> 8) & 0xff); }
__host__ __device__ inline int meta_h(uint32_t m) { return (int)((m >> 16) & 0xff); }
__host__ __device__ inline int meta_prev(uint32_t m) { return (int)(m >> 24); }

// Everything the kernels touch, passed by value.
struct KParams {
    // work queue: two rings of qcap (power of two) items each, ring 0 for
    // the current iterations, ring 1 for lookahead iterations (ring r at
    // q + r * qcap); q_ctl[kQCtl*r + 0..2] = head, limit, tail of ring r
    // (each ring's counters on a 128-byte line of their own: every idle
    // wave's grab is an atomic on them)
    Item *q;
    unsigned long long qmask;
    unsigned long long *q_ctl;

    // instances
    int n_inst;
    int max_iters;
    int mode;
    const Item *roots;           // [n_inst]
    // An instance runs its current iteration and, with lookahead, the next
    // one beside it.  Iteration j lives in slot j & 1; the arrays marked
    // [2 n_inst] are indexed instance * 2 + slot ("slot index"), which is
    // also what items and lanes carry as their instance.
    int *L;                      // [2 n_inst] threshold of the slot's iteration
    int *itn;                    // [2 n_inst] iteration index held by the slot
    unsigned long long *cnt;     // [2 n_inst] count of the slot's iteration
    unsigned int *pend;          // [2 n_inst] its items not finished
    int *found;                  // [2 n_inst] min goal depth (kNoGoal if none)
    int *drained;                // [2 n_inst] 1 once pend reached zero
    int *dead;                   // [2 n_inst] 1: the slot's work is not needed
                                 // (instance done, or a goal in the older slot)
    int *it;                     // [n_inst] current (oldest unresolved) iteration
    int *nlive;                  // [n_inst] iterations started, not resolved (0-2)
    const int *Lbase;            // [n_inst] threshold of iteration 0 (h of the root)
    int *listed;                 // [n_inst] 1 while on the restart list
    int *want_next;              // [n_inst] lookahead requested for the next restart
    int lookahead;               // 1: start iteration it+1 while `it` drains
    int *done;
    int *opt;
    unsigned long long *iter_counts; // [n_inst][max_iters]
    int *restart_list;
    unsigned int *ctl;           // 0 n_restart, 1 active, 2 busy, 3 queue overflow,
                                 // 4 donate_per_lane, 5 budget, 6 rounds,
                                 // 7 donated, 8 overflow exports, 9/10 big/small
                                 // budget, 12 busy at last round end,
                                 // 13 lookahead iterations started,
                                 // 14/15 frontier nodes re-split / the items
                                 // they became,
                                 // 19 donate_per_lane of lanes on lookahead
                                 // iterations, 34 how many are busy (own
                                 // cache line: every wave adds to it)

    // top tree of the root split (per instance)
    const Item *frontier;        // [n_inst][capl]
    const unsigned int *nfront;  // [n_inst]
    const unsigned int *topgen;  // [n_inst][kFBins] generated children by f
    const int *top_goal;         // [n_inst] min depth of a goal in the top tree
    int capl;

    // load balancing: nodes generated below each frontier node in the
    // current iteration, dispatch policy, learned predictions, sample log
    unsigned long long *load;    // [n_inst][capl]
    int lb;                      // 0 frontier order, 1 learned, 2 previous load
    int resplit;                 // heavy-node re-split factor (0: off)
    const float *pred;           // [n_inst][capl][kPredK] predicted log2(1+load)
    Sample *samples;
    unsigned *n_samples;
    unsigned max_samples;

    // lane contexts, structure of arrays over n_lanes lanes
    int n_lanes;
    int maxd;          // LDS stack levels per lane in use
    unsigned long long *prof; // null, or [4]: lane-steps idle, expanding, popping only, wave-steps
    unsigned long long round_ticks; // big rounds end this many wall-clock ticks after a wave's start (0: off)
    unsigned long long *wtime; // null, or [2][waves]: start and end clock of each wave in round wtime_round
    unsigned wtime_round;
    uint32_t *c_st;    // state | sp<<8
    uint32_t *c_rows;  // [5][n_lanes]
    uint32_t *c_root;  // [4][n_lanes] packed item root
    uint32_t *c_meta;  // blank,g,h,prev of the current node
    uint32_t *c_rmeta; // meta of the item root
    uint32_t *c_inst;
    uint32_t *c_cnt;
    uint32_t *c_origin; // frontier node the lane's item descends from
    uint32_t *c_pdb;   // [10][n_lanes]: 8 lookup indices + 2 packed value words
    uint8_t *c_stack;  // [maxd][n_lanes]

    // pattern databases: the two canonical tables in direct 25^6 addressing,
    // laid end to end (table 1 at kPdbDirect)
    const uint8_t *pdb;

    // solution paths (record_paths; all null when off): moves of the blank
    // from the instance root, 2 bits each, kPathWords words per path
    uint32_t *qpath;                 // [2][qcap][kPathWords] paths of the queued items
    uint32_t *c_ipath;               // [kPathWords][n_lanes] path of each lane's item root
    const unsigned long long *fpath; // [n_inst][capl] paths of the frontier nodes
    uint32_t *sol;                   // [2 n_inst][kPathWords] one goal path per slot
    int *sol_len;                    // [2 n_inst] its length, -1 until recorded
};

} // namespace spz
