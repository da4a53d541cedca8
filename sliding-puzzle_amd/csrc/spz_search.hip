// spz_search.hip -- the IDA* hot path on CDNA4 (gfx950).
//
// One search = many instances (a batch of boards), each running its own
// threshold sequence h0, h0+2, ... (src/24puzzle.cc:146-152).  An iteration
// of an instance is:
//   * its top tree (root split, src/15puzzle_block_parallel.cu:150-187, done
//     once per solve by k_split_level): the count of every top node whose f
//     fits the threshold is added from an f-histogram, and
//   * one work item per frontier node whose f fits the threshold; items are
//     searched depth-first by single lanes (k_round), which may donate their
//     shallowest unexplored children back to the queue when lanes go idle
//     (device-wide work stealing; the reference steals only inside a block,
//     src_horie/idas_dynamic.cu:199-247).
// The count of iteration L is  sum_{top u: F(u)<=L} gen(u)
//                            + sum_{frontier v: F(v)<=L} subtree_count(v),
// where gen(u) counts every in-bounds non-reverse child of an expanded node
// before its f-test (src/24puzzle.cc:119-131) and F(u) is the largest f on
// u's path from the root.  IDA* expands a node in iteration L iff every node
// on its path has f <= L, i.e. F <= L; for Manhattan (consistent) F is the
// node's own f, the PDB max is not consistent, hence the path maximum.  So
// the counts are exact whatever the schedule.  With lookahead an instance
// runs its next iteration beside the current one; each keeps its own count
// and k_restart resolves them in order (k_spec, k_restart).
#include <hip/hip_ext.h>

#include "spz_board.h"

namespace spz {

// stack levels a lane may pop in one step: a wave runs the pop loop as often
// as its deepest popping lane needs, and each pop is a chain of dependent LDS
// reads, so more pops per step lengthen every lane's step (measured on
// MI355X: 2 beats 1, 3, 4, 6, 8 and 12 for both heuristics)
constexpr int kPops = 2;
// sibling-value cache of the PDB search: per lane, the deepest kSibLevels
// stack levels (ring, tagged with the level); see k_round
constexpr int kSibLevels = 8;


__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t lo = __builtin_amdgcn_mbcnt_lo(~0u, 0u);
    uint32_t n = __builtin_amdgcn_mbcnt_hi(~0u, lo);
    return n == 0 ? 0ull : (~0ull >> (64 - n));
}

__device__ __forceinline__ unsigned long long shfl_u64(unsigned long long v, int src) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl(lo, src);
    hi = __shfl(hi, src);
    return ((unsigned long long)hi << 32) | lo;
}

// The last item of a slot's iteration has ended.  Called by exactly one lane
// (the one whose pend decrement reached zero): it marks the slot and puts the
// instance on the restart list; k_restart resolves the instance's iterations
// in order between rounds.
__device__ void iteration_drained(const KParams &p, int isl) {
    p.drained[isl] = 1;
    const int inst = isl >> 1;
    if (atomicExch(&p.listed[inst], 1) == 0) {
        const unsigned k = atomicAdd(&p.ctl[0], 1u);
        p.restart_list[k] = inst;
    }
}

// A goal in slot isl makes a newer iteration in the other slot unnecessary
// (this one is the instance's last): mark that slot dead.  Lanes drop items
// of a dead slot unsearched.
__device__ __forceinline__ void kill_newer(const KParams &p, int isl) {
    const int o = isl ^ 1;
    if (p.itn[o] > p.itn[isl]) p.dead[o] = 1;
}

// The work queue is two rings of equal size: ring 0 holds work of the
// instances' current iterations, ring 1 that of lookahead iterations; idle
// lanes take from ring 0 first.  Slot of position pos of ring r:
__device__ __forceinline__ size_t q_slot(const KParams &p, int r, unsigned long long pos) {
    return (size_t)r * (size_t)(p.qmask + 1ull) + (size_t)(pos & p.qmask);
}

// Does the slot index isl belong to a lookahead iteration (newer than the
// instance's current one)?  Its work then goes to ring 1.
__device__ __forceinline__ int ahead_ring(const KParams &p, int isl) {
    return (p.lookahead && p.itn[isl] > p.it[isl >> 1]) ? 1 : 0;
}

// Append one item.  pos is a reserved position of ring r.
template <int W, int HEUR, class BD>
__device__ __forceinline__ void put_item(const KParams &p, int r, unsigned long long pos, const Pos<W, HEUR, BD> &s,
                                         int g, int prev, int inst, uint32_t origin) {
    Item it;
    s.bd.pack(it.b);
    it.meta = meta_pack(s.cell(), g, s.hs.h, prev);
    it.inst = (uint32_t)inst;
    it.x0 = 0;
    it.x1 = origin;
    p.q[q_slot(p, r, pos)] = it;
}

// Reserve k slots at the tail of ring r (counters q_ctl[kQCtl*r + 0..2]:
// head, limit, tail).  Items in [min(head, limit), tail) are not taken yet; a tail
// more than the ring's capacity past them would overwrite some, so that sets
// the error word (ctl[3]) and the host fails the solve with SPZ_E_CAPACITY
// rather than return counts from a corrupt queue.  head is read as it
// stands, possibly behind, which only makes the test stricter.
__device__ __forceinline__ unsigned long long q_reserve(const KParams &p, int r, unsigned long long k) {
    unsigned long long *c = p.q_ctl + kQCtl * r;
    const unsigned long long pos = atomicAdd(&c[2], k);
    const unsigned long long h = *(volatile unsigned long long *)&c[0], lim = c[1];
    const unsigned long long live_from = h < lim ? h : lim;
    if (pos + k - live_from > p.qmask + 1ull) atomicOr(&p.ctl[3], 1u);
    return pos;
}

// Solution paths: move i of a path sits in bits 2(i%16) of word i/16; a
// path's words are `stride` apart (1 in the queue and solution arrays, the
// lane count in the per-lane array).  Bits past a path's length are zero, so
// a copied prefix needs no clearing.  The loops stay rolled: this code runs
// rarely, and unrolled it would hold a whole path in registers and cost the
// search kernel occupancy.
// Paths are addressed as base[off + word * stride] with 32-bit offsets
// (every path array holds fewer than 2^32 words), so the base stays scalar.
__device__ __forceinline__ void path_put(uint32_t *base, uint32_t off, uint32_t stride, int i, int d) {
    if (i >= kPathMax) return; // too long to record; no goal through it is recorded
    uint32_t &x = base[off + (uint32_t)(i >> 4) * stride];
    const int sh = 2 * (i & 15);
    x = (x & ~(3u << sh)) | ((uint32_t)d << sh);
}
__device__ __forceinline__ void path_copy(const uint32_t *sb, uint32_t so, uint32_t ss, uint32_t *db, uint32_t dof,
                                          uint32_t dstr) {
#pragma unroll 1
    for (uint32_t k = 0; k < (uint32_t)kPathWords; ++k) db[dof + k * dstr] = sb[so + k * ss];
}
__device__ __forceinline__ void path_from_u64(unsigned long long v, uint32_t *db, uint32_t dof) {
    db[dof] = (uint32_t)v;
    db[dof + 1] = (uint32_t)(v >> 32);
#pragma unroll 1
    for (uint32_t k = 2; k < (uint32_t)kPathWords; ++k) db[dof + k] = 0u;
}

template <int W, int HEUR, class BD>
__device__ __forceinline__ void pos_from_item(Pos<W, HEUR, BD> &s, const uint32_t (&b)[4], uint32_t meta,
                                              const PdbCtx &P) {
    s.bd.unpack(b);
    s.set_cell(meta_blank(meta));
    init_heur<W, HEUR>(s, P, meta);
}

// register copy of a lane position (the search kernel's 5x5 board is in LDS)
template <int W, int HEUR, class BD>
__device__ __forceinline__ Pos<W, HEUR> reg_pos(const Pos<W, HEUR, BD> &s) {
    Pos<W, HEUR> c;
    if constexpr (W == 5)
        c.bd = s.bd.regs();
    else
        c.bd = s.bd;
    if constexpr (W == 4) {
        c.bl = s.bl;
    } else {
        c.br = s.br;
        c.bc = s.bc;
    }
    c.hs = s.hs;
    return c;
}

// ---------------------------------------------------------------------------
// k_round: every lane is an independent DFS engine.  A round runs `budget`
// wave-uniform steps; each step a lane either takes an item, pops one stack
// level, or expands the node it stands on.  Busy lanes keep their state
// across rounds in the c_* context arrays.  The stack holds one byte per
// level: the move taken from that level (2 bits) and the moves still to
// try there (4-bit mask of children that passed the f-test).
// The 15-puzzle kernel is held to 64 VGPRs (8 waves per SIMD: its LDS use is
// small, and it fits without scratch); the 24-puzzle kernels are bounded by
// their LDS (stack and boards) at 3 waves per SIMD.
template <int W, int HEUR>
__global__ void __launch_bounds__(256, (W == 4 ? 8 : 1)) k_round(KParams p, const PdbGeo *geo_g) {
    // every wave of the grid starts within microseconds of the others
    const long long t_start = (p.wtime || p.round_ticks) ? wall_clock64() : 0;
    const int MAXD = p.maxd;
    extern __shared__ uint8_t lds[];
    __shared__ PdbGeo geo_s;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int gl = blockIdx.x * blockDim.x + threadIdx.x;
    const int NL = p.n_lanes;
    using SE = StackEntry<HEUR>;
    using ST = typename SE::T;
    // level k at stk[k * sstride]; an LDS-qualified pointer keeps the
    // address arithmetic in 32 bits
    __attribute__((address_space(3))) ST *stk = (__attribute__((address_space(3))) ST *)lds + wave * 64 + lane;
    ST *cstk = (ST *)p.c_stack;
    const int sstride = nw * 64;

    PdbCtx P{&geo_s, p.pdb};
    if constexpr (HEUR == 1) {
        const uint32_t *src = (const uint32_t *)geo_g;
        uint32_t *dst = (uint32_t *)&geo_s;
        for (int i = threadIdx.x; i < (int)(sizeof(PdbGeo) / 4); i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }

    const unsigned long long limit = p.q_ctl[1], limit1 = p.q_ctl[kQCtl + 1];
    // a ring seen empty stays empty for the round (heads only grow): the
    // wave then skips its atomic
    bool q0_empty = p.q_ctl[0] >= limit, q1_empty = !p.lookahead || p.q_ctl[kQCtl] >= limit1;
    const int budget = (int)p.ctl[5];
    const int mode = p.mode;

    uint32_t stw = p.c_st[gl];
    int st = (int)(stw & 0xff);
    int sp = (int)(stw >> 8);
    // stack levels, then (width 5) the lanes' boards
    Pos<W, HEUR, typename HotBoard<W>::T> s;
    // Sibling-value cache (PDB): when a lane enters the first passing child
    // of a node with at most two more, the values its f-test read for them
    // go into slot level % kSibLevels (tag = level in bits 24-31, 6+6 bits
    // per sibling); moving into such a sibling later then reads no table.
    // An entry overwritten by a deeper level (or from an earlier round: LDS
    // does not persist) fails its tag and the move reads the tables as
    // before.
    uint32_t *sibc = nullptr;
    if constexpr (HEUR == 1) {
        sibc = (uint32_t *)(lds + (size_t)MAXD * sstride * sizeof(ST) + (size_t)blockDim.x * kLdsBoardBytes) +
               threadIdx.x;
#pragma unroll
        for (int k = 0; k < kSibLevels; ++k) sibc[k * blockDim.x] = 0xffffffffu;
    }
    if constexpr (W == 5) {
        s.bd.p = lds + (size_t)MAXD * sstride * sizeof(ST) + threadIdx.x * kLdsBoardBytes;
    } else {
        s.bd.unpack({0u, 0u, 0u, 0u});
    }
    s.set_cell(0);
    s.hs.h = 0;
    uint32_t root[4] = {0, 0, 0, 0};
    uint32_t rmeta = 0;
    // inst is the item's slot index (instance * 2 + iteration slot)
    int g = 0, prev = DIR_NONE, inst = 0, L = 0;
    uint32_t cnt = 0, origin = 0;

    if (st == ST_FRESH || st == ST_POP) {
        uint32_t b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) b[k] = p.c_rows[k * NL + gl];
        s.bd.unpack(b);
#pragma unroll
        for (int k = 0; k < 4; ++k) root[k] = p.c_root[k * NL + gl];
        uint32_t m = p.c_meta[gl];
        rmeta = p.c_rmeta[gl];
        inst = (int)p.c_inst[gl];
        cnt = p.c_cnt[gl];
        origin = p.c_origin[gl];
        s.set_cell(meta_blank(m));
        g = meta_g(m);
        prev = meta_prev(m);
        s.hs.h = meta_h(m);
        if constexpr (HEUR == 1) {
            s.hs.i0 = p.c_pdb[0 * NL + gl];
            s.hs.i1 = p.c_pdb[1 * NL + gl];
            s.hs.i2 = p.c_pdb[2 * NL + gl];
            s.hs.i3 = p.c_pdb[3 * NL + gl];
            s.hs.j0 = p.c_pdb[4 * NL + gl];
            s.hs.j1 = p.c_pdb[5 * NL + gl];
            s.hs.j2 = p.c_pdb[6 * NL + gl];
            s.hs.j3 = p.c_pdb[7 * NL + gl];
            s.hs.v0 = p.c_pdb[8 * NL + gl];
            s.hs.v1 = p.c_pdb[9 * NL + gl];
            uint32_t v0 = s.hs.v0, v1 = s.hs.v1;
            s.hs.A = (int)((v0 & 0xff) + ((v0 >> 8) & 0xff) + ((v0 >> 16) & 0xff) + (v0 >> 24));
            s.hs.B = (int)((v1 & 0xff) + ((v1 >> 8) & 0xff) + ((v1 >> 16) & 0xff) + (v1 >> 24));
        }
        L = p.L[inst];
        if (p.lookahead && p.dead[inst]) sp = 0, st = ST_POP; // unwinds and finishes next step
        for (int k = 0; k < sp; ++k) stk[k * sstride] = cstk[(size_t)k * NL + gl];
    } else {
        st = ST_IDLE;
    }

    auto start_item = [&](const Item &it) {
        pos_from_item<W, HEUR>(s, it.b, it.meta, P);
#pragma unroll
        for (int k = 0; k < 4; ++k) root[k] = it.b[k];
        rmeta = it.meta;
        inst = (int)it.inst;
        origin = it.x1;
        g = meta_g(it.meta);
        prev = meta_prev(it.meta);
        L = p.L[inst];
        sp = 0;
        cnt = 0;
        st = ST_FRESH;
        if (mode == 1 && p.found[inst] != kNoGoal) st = ST_POP; // first-solution: abandon
        if (p.lookahead && p.dead[inst]) st = ST_POP;
    };

    // move down into child d (stack level sl, R = passing children still to
    // visit there); for the PDB the child's values are known or read
    uint32_t dpass = 0, dval = 0, dval2 = 0; // pending descent (ST_DESC)
    auto enter_child = [&](int d, uint32_t Rr, int sl, bool kn, uint32_t kvv) {
        uint32_t ov = 0;
        if constexpr (HEUR == 1)
            ov = move_pdb<W>(s, d, P, kn, kvv);
        else
            ov = move_md(s, d);
        stk[sl * sstride] = SE::make(d, Rr, ov);
        sp = sl + 1;
        ++g;
        prev = d;
        st = ST_FRESH;
    };
    // path of a new queue item: the item root's path, the moves of stack
    // levels 0..nst-1, then move last_d
    auto item_path = [&](int r, unsigned long long qpos, int nst, int last_d) {
        const uint32_t qo = (uint32_t)q_slot(p, r, qpos) * kPathWords;
        path_copy(p.c_ipath, (uint32_t)gl, (uint32_t)NL, p.qpath, qo, 1);
        const int gr = meta_g(rmeta);
        for (int k = 0; k < nst; ++k) path_put(p.qpath, qo, 1, gr + k, SE::dir(stk[k * sstride]));
        path_put(p.qpath, qo, 1, gr + nst, last_d);
    };

    uint32_t pf_idle = 0, pf_exp = 0, pf_pop = 0, pf_steps = 0;
    uint32_t pf_sib = 0, pf_rd1 = 0, pf_rd2 = 0; // table reads: sibling moves, first, second partition
    // the restored context is complete before the loop, so the compiler's
    // wait tracking does not carry its loads into every iteration (where it
    // would wait on a move's table reads before issuing the f-test's)
    __builtin_amdgcn_s_waitcnt(0);
    for (int step = 0; step < budget; ++step) {
        // a round with a time limit ends between two steps once it is up
        // (the clock is a scalar read, wave-uniform)
        if (p.round_ticks && (step & 3) == 3 && (unsigned long long)(wall_clock64() - t_start) >= p.round_ticks)
            break;
        // ---- idle lanes take queue items: one atomic per wave and ring,
        // ring 0 first, none on a ring the wave has seen empty ----
        {
            uint64_t need = __ballot(st == ST_IDLE);
            if (need) {
                int n = __popcll(need);
                int leader = __ffsll((long long)need) - 1;
                unsigned long long base = 0;
                if (!q0_empty) {
                    if (lane == leader) base = atomicAdd(&p.q_ctl[0], (unsigned long long)n);
                    base = shfl_u64(base, leader);
                    const unsigned long long my = base + __popcll(need & lanemask_lt());
                    if (st == ST_IDLE && my < limit) {
                        start_item(p.q[q_slot(p, 0, my)]);
                        if (p.qpath)
                            path_copy(p.qpath, (uint32_t)q_slot(p, 0, my) * kPathWords, 1, p.c_ipath, (uint32_t)gl,
                                      (uint32_t)NL);
                    }
                    q0_empty = base + (unsigned long long)n >= limit;
                }
                // lanes ring 0 could not serve try the lookahead ring
                const uint64_t need1 = q1_empty ? 0ull : __ballot(st == ST_IDLE);
                if (need1) {
                    const int n1 = __popcll(need1);
                    const int leader1 = __ffsll((long long)need1) - 1;
                    unsigned long long base1 = 0;
                    if (lane == leader1) base1 = atomicAdd(&p.q_ctl[kQCtl], (unsigned long long)n1);
                    base1 = shfl_u64(base1, leader1);
                    const unsigned long long my = base1 + __popcll(need1 & lanemask_lt());
                    if (st == ST_IDLE && my < limit1) {
                        start_item(p.q[q_slot(p, 1, my)]);
                        if (p.qpath)
                            path_copy(p.qpath, (uint32_t)q_slot(p, 1, my) * kPathWords, 1, p.c_ipath, (uint32_t)gl,
                                      (uint32_t)NL);
                    }
                    q1_empty = base1 + (unsigned long long)n1 >= limit1;
                }
                if (st == ST_IDLE) st = ST_DRAINED;
            }
        }
        if (!__ballot(st == ST_FRESH || st == ST_POP || st == ST_DESC)) break;

        // ---- pop: undo up to kPops moves, until a level with a passing
        // child not visited yet ----
        bool finish = false;
        int down = -1;      // child entered this step (-1: none)
        int slot = 0;       // its stack level
        uint32_t R = 0;     // passing children still to visit at that level
        bool known = false; // its lookup values are known (PDB)
        uint32_t kv = 0;
#pragma unroll 1
        for (int pk = 0; pk < kPops && st == ST_POP; ++pk) {
            if (sp == 0) {
                finish = true;
                break;
            }
            uint32_t e = stk[(sp - 1) * sstride];
            int d = SE::dir(e);
            uint32_t Re = SE::rest(e);
            if constexpr (HEUR == 1)
                do_move_known<W>(s, 3 - d, P, SE::vals(e));
            else
                unmove_md(s, d, SE::vals(e));
            --g;
            if (Re) {
                down = __builtin_ctz(Re);
                R = Re & (Re - 1);
                slot = sp - 1;
                st = ST_FRESH;
                if constexpr (HEUR == 1) {
                    uint32_t &c = sibc[(slot % kSibLevels) * blockDim.x];
                    const uint32_t cv = c;
                    if ((cv >> 24) == (uint32_t)slot) {
                        known = true;
                        kv = (cv & 63u) | (((cv >> 6) & 63u) << 8);
                        c = (cv & 0xff000000u) | ((cv >> 12) & 0xfffu); // the sibling after it moves up
                    }
                }
            } else {
                --sp;
            }
        }
        // ---- or enter the first passing child of last step's expansion ----
        if (st == ST_DESC) {
            down = __builtin_ctz(dpass);
            R = dpass & (dpass - 1);
            slot = sp;
            known = true;
            kv = (((down >= 2) ? dval2 : dval) >> (16 * (down & 1))) & 0xffffu;
            if constexpr (HEUR == 1) {
                if (R) {
                    // values of the next two passing children, 6 + 6 bits
                    // each.  An entry holds two siblings: a node with three
                    // more (an item rooted at an instance's root, all four
                    // moves passing) voids its level's entry instead, so
                    // that none of them takes values left by an older node
                    // of this depth, or the zeros shifted in after two
                    uint32_t ent = 0xffffffffu;
                    if (__builtin_popcount(R) <= 2) {
                        const int s1 = __builtin_ctz(R);
                        const uint32_t R2 = R & (R - 1);
                        const int s2 = R2 ? __builtin_ctz(R2) : s1;
                        const uint32_t v1 = (((s1 >= 2) ? dval2 : dval) >> (16 * (s1 & 1))) & 0xffffu;
                        const uint32_t v2 = (((s2 >= 2) ? dval2 : dval) >> (16 * (s2 & 1))) & 0xffffu;
                        const uint32_t p1 = (v1 & 63u) | (((v1 >> 8) & 63u) << 6);
                        const uint32_t p2 = (v2 & 63u) | (((v2 >> 8) & 63u) << 6);
                        ent = ((uint32_t)slot << 24) | p1 | (p2 << 12);
                    }
                    sibc[(slot % kSibLevels) * blockDim.x] = ent;
                }
            }
        }
        // ---- one move down per lane and step, both cases on one path.  For
        // the PDB the move's values are applied after the f-test's table
        // reads are issued, so both sets of reads are in flight together ----
        PendMove pm;
        bool pmv = false;
        if (down >= 0) {
            if constexpr (HEUR == 1) {
                uint32_t ov;
                if (p.prof) pf_sib += known ? 0u : 1u;
                pm = move_pdb_begin<W>(s, down, P, known, kv, ov);
                pmv = true;
                stk[slot * sstride] = SE::make(down, R, ov);
                sp = slot + 1;
                ++g;
                prev = down;
                st = ST_FRESH;
            } else {
                enter_child(down, R, slot, known, kv);
            }
        }

        // ---- expand, first half (PDB): the children's lookup indices and
        // table reads; they need the node's indices, not its values ----
        uint32_t gen = 0;
        // per child slot: lookups kr/kf, the index of the second lookup, the
        // first and second values read, the direction
        int ckr[4], ckf[4], cv1[4], cv2[4], cdir[4];
        uint32_t cidx2[4];
        int ng = 0;
        bool regfirst = true; // read the regular lookup first (else the reflected)
        if (st == ST_FRESH) {
            gen = legal_moves(s) & ~(prev < 4 ? (1u << (3 - prev)) : 0u);
            if constexpr (HEUR == 1) {
                // lane-slot j reads the lane's j-th generated child, so the
                // wave runs as many slots as its largest family (3 below the
                // root), not one per direction.  Only one lookup per child
                // is read here: that of the partition whose sum is larger at
                // the node (before a pending move's values: a heuristic
                // choice); the other is read only for children the first
                // does not already cut off (the reference's early cut,
                // src_horie/idas_pdb.cu:228-231, regular sum first).
                ng = __builtin_popcount(gen);
                regfirst = s.hs.A >= s.hs.B;
                const int bl = s.cell();
                const PdbGeo &G = *P.geo;
                uint32_t rem = gen;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    ckr[j] = ckf[j] = cv1[j] = cv2[j] = cdir[j] = 0;
                    cidx2[j] = 0;
                    if (!__ballot(j < ng)) continue;
                    if (j < ng) {
                        const int d = __builtin_ctz(rem);
                        rem &= rem - 1;
                        const int from = bl + dir_cell<W>(d);
                        const int t = s.bd.tile(from);
                        const uint2 wa = G.w[t][from], wb = G.w[t][bl];
                        const int kr = (int)(wa.x >> 28), kf = (int)(wa.y >> 28);
                        const uint32_t ir = sel4(s.hs.i0, s.hs.i1, s.hs.i2, s.hs.i3, kr) + (wb.x - wa.x);
                        const uint32_t iff = sel4(s.hs.j0, s.hs.j1, s.hs.j2, s.hs.j3, kf) + (wb.y - wa.y);
                        cdir[j] = d;
                        ckr[j] = kr;
                        ckf[j] = kf;
                        cv1[j] = pdb_read(P, regfirst ? ir : iff);
                        cidx2[j] = regfirst ? iff : ir;
                    }
                }
            }
        }
        if constexpr (HEUR == 1) {
            if (pmv) move_pdb_finish(s.hs, pm);
        }

        // ---- expand, second half: goal test, count, f-tests ----
        bool expanded = false;
        if (st == ST_FRESH) {
            const int h = s.hs.h;
            bool xp = true;
            if (h == 0) {
                atomicMin(&p.found[inst], g);
                if (p.lookahead) kill_newer(p, inst);
                // the first goal of the slot's iteration records its path
                // (the first iteration with a goal has threshold = optimum,
                // and every goal it finds has that depth)
                if (p.sol && g <= kPathMax && atomicCAS(&p.sol_len[inst], -1, g) == -1) {
                    const uint32_t so = (uint32_t)inst * kPathWords;
                    path_copy(p.c_ipath, (uint32_t)gl, (uint32_t)NL, p.sol, so, 1);
                    const int gr = meta_g(rmeta);
                    for (int k = 0; k < sp; ++k) path_put(p.sol, so, 1, gr + k, SE::dir(stk[k * sstride]));
                }
                if (mode == 1) {
                    st = ST_POP;
                    xp = false;
                }
            }
            if (xp) {
                expanded = true;
                cnt += (uint32_t)__builtin_popcount(gen);
                uint32_t pass = gen;
                uint32_t cvals = 0, cvals2 = 0; // PDB values of children 0,1 / 2,3 (16 bits each)
                if constexpr (HEUR == 0) {
                    // Manhattan is consistent: h changes by exactly +-1 per
                    // move and f keeps the threshold's parity, so with slack
                    // >= 2 every child fits and with slack 0 only children
                    // with h-1 do.
                    const int slack = L - g - h;
                    if (slack <= 2) {
                        const uint32_t near = nearer_moves<W>(s.bd, s.cell());
                        if (slack < 2) pass = gen & near;
                        // A child at slack 0 (slack 0 and nearer, or slack 2
                        // and farther) whose own children all fail the
                        // f-test is a leaf: its children are counted here
                        // and it is not entered.  Its children's tiles are
                        // all on the current board (entering it would only
                        // move the blank there).  With h >= 2 no such child
                        // is a goal, and a leaf has no goal child (one would
                        // be nearer, and pass).  One child per expansion is
                        // tested, the first such: a wave would run the test
                        // as often as its lane with the most.
                        const uint32_t cand = h < 2 ? 0u : (slack < 2 ? pass : gen & ~near);
                        if (cand) {
                            const int d = __builtin_ctz(cand);
                            const int cc = s.cell() + dir_cell<W>(d);
                            const uint32_t cgen = legal_cell<W>(cc) & ~(1u << (3 - d));
                            if (!(cgen & nearer_moves<W>(s.bd, cc))) {
                                cnt += (uint32_t)__builtin_popcount(cgen);
                                pass &= ~(1u << d);
                            }
                        }
                    }
                } else {
                    // the PDB max is not consistent (lookups take the minimum
                    // over blank cells): h may jump by 3, so every generated
                    // child gets its own f-test
                    pass = 0;
                    const HState<1> &hs = s.hs;
                    // the first partition's sum of each child; the second
                    // lookup is read only where that sum fits
                    uint32_t need = 0;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (j < ng) {
                            const int s1 = regfirst ? hs.A - byte_of(hs.v0, ckr[j]) + cv1[j]
                                                    : hs.B - byte_of(hs.v1, ckf[j]) + cv1[j];
                            if (g + 1 + s1 <= L) {
                                need |= 1u << j;
                                cv2[j] = pdb_read(P, cidx2[j]);
                            }
                            if (p.prof) {
                                pf_rd1 += 1u;
                                pf_rd2 += (g + 1 + s1 <= L) ? 1u : 0u;
                            }
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if ((need >> j) & 1u) {
                            const int d = cdir[j];
                            const int vr = regfirst ? cv1[j] : cv2[j];
                            const int vf = regfirst ? cv2[j] : cv1[j];
                            const int A = hs.A - byte_of(hs.v0, ckr[j]) + vr;
                            const int B = hs.B - byte_of(hs.v1, ckf[j]) + vf;
                            const int hc = A > B ? A : B;
                            pass |= (g + 1 + hc <= L) ? (1u << d) : 0u;
                            const uint32_t v = ((uint32_t)vr | ((uint32_t)vf << 8)) << (16u * (uint32_t)(d & 1));
                            if (d < 2)
                                cvals |= v;
                            else
                                cvals2 |= v;
                        }
                    }
                }
                if (!pass) {
                    st = ST_POP;
                } else if (sp < MAXD) {
                    // enter the first passing child next step
                    st = ST_DESC;
                    dpass = pass;
                    dval = cvals;
                    dval2 = cvals2;
                } else {
                    // stack full: the passing children become queue items
                    int k = __builtin_popcount(pass);
                    atomicAdd(&p.pend[inst], (unsigned)k);
                    const int r = ahead_ring(p, inst);
                    unsigned long long pos = q_reserve(p, r, (unsigned long long)k);
                    atomicAdd(&p.ctl[8], (unsigned)k);
                    while (pass) {
                        int d = __builtin_ctz(pass);
                        pass &= pass - 1;
                        if (p.qpath) item_path(r, pos, sp, d);
                        Pos<W, HEUR> c = reg_pos(s);
                        if constexpr (HEUR == 1)
                            move_board_only(c, d); // the item's lane recomputes its lookups
                        else
                            do_move<W, HEUR>(c, d, P);
                        put_item<W, HEUR>(p, r, pos++, c, g + 1, d, inst, origin);
                    }
                    st = ST_POP;
                }
            }
        }

        // ---- items whose subtree is done (after the expansion, so that its
        // fence does not wait on the step's table reads) ----
        if (finish) {
            if (cnt) {
                atomicAdd(&p.cnt[inst], (unsigned long long)cnt);
                atomicAdd(&p.load[(size_t)(inst >> 1) * p.capl + origin], (unsigned long long)cnt);
            }
            __threadfence();
            unsigned old = atomicSub(&p.pend[inst], 1u);
            if (old == 1u) iteration_drained(p, inst);
            cnt = 0;
            st = ST_IDLE;
        }
        if (p.prof) {
            pf_steps += 1;
            pf_idle += (st == ST_IDLE || st == ST_DRAINED) ? 1u : 0u;
            pf_exp += expanded ? 1u : 0u;
            pf_pop += (!expanded && (st == ST_POP || st == ST_DESC)) ? 1u : 0u;
        }
    }

    // a descent still pending is made now, so the saved context is a node
    if (st == ST_DESC) {
        const int dn = __builtin_ctz(dpass);
        enter_child(dn, dpass & (dpass - 1), sp, true, (((dn >= 2) ? dval2 : dval) >> (16 * (dn & 1))) & 0xffffu);
    }

    // ---- round end: donate shallow open nodes, then save the context ----
    const bool busy = (st == ST_FRESH || st == ST_POP);
    // lanes on lookahead iterations donate into ring 1, under their own quota
    const int ring = busy ? ahead_ring(p, inst) : 0;
    const int per_lane = (int)(ring ? p.ctl[19] : p.ctl[4]);
    int want = 0;
    if (busy && per_lane > 0) {
        int avail = 0;
        for (int k = 0; k < sp && avail < per_lane; ++k) avail += __builtin_popcount(SE::rest(stk[k * sstride]));
        want = avail < per_lane ? avail : per_lane;
    }
    if (__ballot(want > 0)) {
        // per ring: wave prefix sum of `want`, one queue reservation per wave
        unsigned long long pos = 0;
        for (int r = 0; r < 2; ++r) {
            const int wr = ring == r ? want : 0;
            if (!__ballot(wr > 0)) continue;
            int incl = wr;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                int v = __shfl_up(incl, off);
                if (lane >= off) incl += v;
            }
            int total = __shfl(incl, 63);
            unsigned long long base = 0;
            if (lane == 0) {
                base = q_reserve(p, r, (unsigned long long)total);
                atomicAdd(&p.ctl[7], (unsigned)total);
            }
            base = shfl_u64(base, 0);
            if (wr > 0) pos = base + (unsigned long long)(incl - wr);
        }
        if (want > 0) {
            atomicAdd(&p.pend[inst], (unsigned)want);
            if (p.qpath) {
                // the donated items' paths first, in the order of the pass
                // below (a pass of its own keeps the replayed positions out
                // of its registers)
                unsigned long long pp = pos;
                int emitted = 0;
                for (int k = 0; k < sp && emitted < want; ++k) {
                    uint32_t R = SE::rest(stk[k * sstride]);
                    while (R && emitted < want) {
                        const int d2 = __builtin_ctz(R);
                        R &= R - 1;
                        item_path(ring, pp++, k, d2);
                        ++emitted;
                    }
                }
            }
            // replay from the item root along the taken moves (for the PDB the
            // board alone: the lane that takes an item recomputes its lookups)
            Pos<W, HEUR> c;
            if constexpr (HEUR == 1) {
                c.bd.unpack(root);
                c.set_cell(meta_blank(rmeta));
                c.hs.h = 0;
            } else {
                pos_from_item<W, HEUR>(c, root, rmeta, P);
            }
            int cg = meta_g(rmeta);
            int emitted = 0;
            for (int k = 0; k < sp && emitted < want; ++k) {
                uint32_t e = stk[k * sstride];
                int d = SE::dir(e);
                uint32_t R = SE::rest(e);
                while (R && emitted < want) {
                    int d2 = __builtin_ctz(R);
                    R &= R - 1;
                    Pos<W, HEUR> x = c;
                    if constexpr (HEUR == 1)
                        move_board_only(x, d2);
                    else
                        do_move<W, HEUR>(x, d2, P);
                    put_item<W, HEUR>(p, ring, pos++, x, cg + 1, d2, inst, origin);
                    ++emitted;
                }
                stk[k * sstride] = SE::make(d, R, SE::vals(e));
                if (emitted < want) {
                    if constexpr (HEUR == 1)
                        move_board_only(c, d);
                    else
                        do_move<W, HEUR>(c, d, P);
                    ++cg;
                }
            }
        }
    }

    // keep the per-lane u32 count far from wrapping on long-lived items
    if (busy && cnt >= (1u << 30)) {
        atomicAdd(&p.cnt[inst], (unsigned long long)cnt);
        atomicAdd(&p.load[(size_t)(inst >> 1) * p.capl + origin], (unsigned long long)cnt);
        cnt = 0;
    }
    if (p.prof) {
        // wave sums of this round's lane-step states
        for (int off = 32; off > 0; off >>= 1) {
            pf_idle += __shfl_xor(pf_idle, off);
            pf_exp += __shfl_xor(pf_exp, off);
            pf_pop += __shfl_xor(pf_pop, off);
            pf_sib += __shfl_xor(pf_sib, off);
            pf_rd1 += __shfl_xor(pf_rd1, off);
            pf_rd2 += __shfl_xor(pf_rd2, off);
        }
        if (lane == 0) {
            atomicAdd(&p.prof[0], (unsigned long long)pf_idle);
            atomicAdd(&p.prof[1], (unsigned long long)pf_exp);
            atomicAdd(&p.prof[2], (unsigned long long)pf_pop);
            atomicAdd(&p.prof[3], (unsigned long long)pf_steps);
            atomicAdd(&p.prof[4], (unsigned long long)pf_sib);
            atomicAdd(&p.prof[5], (unsigned long long)pf_rd1);
            atomicAdd(&p.prof[6], (unsigned long long)pf_rd2);
        }
    }
    if (p.wtime && lane == 0 && p.ctl[6] == p.wtime_round) {
        const int wg = gl >> 6; // global wave index
        p.wtime[2 * wg] = (unsigned long long)t_start;
        p.wtime[2 * wg + 1] = (unsigned long long)wall_clock64();
    }
    uint64_t busy_mask = __ballot(busy);
    if (lane == 0 && busy_mask) atomicAdd(&p.ctl[2], (unsigned)__popcll(busy_mask));
    const uint64_t ahead_mask = p.lookahead ? __ballot(busy && ring == 1) : 0ull;
    if (lane == 0 && ahead_mask) atomicAdd(&p.ctl[34], (unsigned)__popcll(ahead_mask));
    if (busy) {
        p.c_st[gl] = (uint32_t)st | ((uint32_t)sp << 8);
        uint32_t b[4];
        s.bd.pack(b);
#pragma unroll
        for (int k = 0; k < 4; ++k) p.c_rows[k * NL + gl] = b[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) p.c_root[k * NL + gl] = root[k];
        p.c_meta[gl] = meta_pack(s.cell(), g, s.hs.h, prev);
        p.c_rmeta[gl] = rmeta;
        p.c_inst[gl] = (uint32_t)inst;
        p.c_cnt[gl] = cnt;
        p.c_origin[gl] = origin;
        if constexpr (HEUR == 1) {
            p.c_pdb[0 * NL + gl] = s.hs.i0;
            p.c_pdb[1 * NL + gl] = s.hs.i1;
            p.c_pdb[2 * NL + gl] = s.hs.i2;
            p.c_pdb[3 * NL + gl] = s.hs.i3;
            p.c_pdb[4 * NL + gl] = s.hs.j0;
            p.c_pdb[5 * NL + gl] = s.hs.j1;
            p.c_pdb[6 * NL + gl] = s.hs.j2;
            p.c_pdb[7 * NL + gl] = s.hs.j3;
            p.c_pdb[8 * NL + gl] = s.hs.v0;
            p.c_pdb[9 * NL + gl] = s.hs.v1;
        }
        for (int k = 0; k < sp; ++k) cstk[(size_t)k * NL + gl] = stk[k * sstride];
    } else {
        p.c_st[gl] = ST_IDLE;
    }
}

// largest f on a frontier node's root path (its own f included)
__device__ __forceinline__ int path_f(const Item &it) {
    int f = meta_g(it.meta) + meta_h(it.meta);
    return (int)it.x0 > f ? (int)it.x0 : f;
}

// ---------------------------------------------------------------------------
// k_restart: between rounds, one block per instance on the restart list (a
// slot of it drained, it is new, or lookahead was asked for).  Thread 0
// resolves the instance's iterations in order: while the current iteration
// has drained, its count is recorded and it ends the instance (goal found,
// or iteration limit) or the next one becomes current -- already running if
// it was started as lookahead.  An iteration is started when the instance has
// none running, or as lookahead beside the current one when k_spec asked for
// it.  Starting iteration j (threshold Lbase + 2j) in slot j & 1:
//   1. logs (frontier node, previous threshold, nodes generated below it)
//      samples of the iteration that ended (learned-LB training data; only
//      without lookahead, where one iteration runs at a time),
//   2. sets the slot's count to the top-tree count of the threshold from the
//      f-histogram and notes a goal in the top tree,
//   3. queues the frontier nodes whose path-f fits, ordered by a key in
//      decreasing order (counting sort over 64 half-log2 buckets): frontier
//      order (uniform), last iteration's load (the reference's load-based
//      static LB, src/15puzzle_block_parallel.cu:490-563), or the learned
//      prediction -- longest expected subtree first.  With re-splitting
//      (resplit > 0, previous-load or learned policy), a node whose
//      estimate exceeds resplit x the mean of the nodes that fit is queued
//      as its children (two levels: grandchildren) instead, as the
//      reference re-splits heavy roots into ceil(load / mean) pieces,
//   4. clears the per-frontier-node loads (a split node's starts at the
//      children the split counted).
// An iteration with no frontier work is drained at once.  Lanes run no
// kernel meanwhile, so the resolution needs no atomics on instance state.

// Estimated subtree size of frontier node i in iteration it: the previous
// iteration's count below it (previous-load policy) or the learned
// prediction of log2(1 + size) (learned policy); 0 for frontier order.
__device__ __forceinline__ float lb_log2_est(const KParams &p, int inst, unsigned i, int it) {
    if (p.lb == 2) return it > 0 ? log2f((float)p.load[(size_t)inst * p.capl + i] + 1.f) : 0.f;
    if (p.lb == 1) return p.pred[((size_t)inst * p.capl + i) * kPredK + (it < kPredK ? it : kPredK - 1)];
    return 0.f;
}
// dispatch bucket of a key in half-log2 steps (64 buckets, largest first)
__device__ __forceinline__ int key_bucket(float key) {
    int b = (int)(key * 2.f);
    return b < 0 ? 0 : (b > 63 ? 63 : b);
}
// bucket of an item `level` split levels below a node of bucket b: each level
// divides the estimate by about 3 (the children of a node), 3 half-log2 steps
__device__ __forceinline__ int split_bucket(int b, int level) {
    int c = b - 3 * level;
    return c < 0 ? 0 : c;
}

// Re-split of a heavy frontier node (the reference's static load balancing,
// src/15puzzle_block_parallel.cu:490-563 with divide_root_set :325-375): the
// node is expanded here, its count of generated children added to the
// iteration (the lane that would take it counts the same children,
// src/24puzzle.cc:119-131), and its children whose f fits the threshold are
// queued in its place -- or, at two levels, their children.  A goal node is
// queued itself (its lane records the goal).  emit(pos, g, prev, level, d1,
// d2) receives each item; returns the generated-children count.
template <int W, int HEUR, class Emit>
__device__ unsigned split_node(const PdbCtx &P, const Item &x, int L, int levels, Emit &&emit) {
    Pos<W, HEUR> s;
    pos_from_item<W, HEUR>(s, x.b, x.meta, P);
    const int g = meta_g(x.meta), prev = meta_prev(x.meta);
    unsigned gen_total = 0;
    uint32_t gen = legal_moves(s);
    gen &= ~(prev < 4 ? (1u << (3 - prev)) : 0u);
    gen_total += (unsigned)__builtin_popcount(gen);
#pragma unroll 1
    while (gen) {
        const int d = __builtin_ctz(gen);
        gen &= gen - 1;
        Pos<W, HEUR> c = s;
        do_move<W, HEUR>(c, d, P);
        if (g + 1 + c.hs.h > L) continue;
        if (levels < 2 || c.hs.h == 0) {
            emit(c, g + 1, d, 1, d, 0);
            continue;
        }
        uint32_t gen2 = legal_moves(c);
        gen2 &= ~(1u << (3 - d));
        gen_total += (unsigned)__builtin_popcount(gen2);
#pragma unroll 1
        while (gen2) {
            const int d2 = __builtin_ctz(gen2);
            gen2 &= gen2 - 1;
            Pos<W, HEUR> c2 = c;
            do_move<W, HEUR>(c2, d2, P);
            if (g + 2 + c2.hs.h > L) continue;
            emit(c2, g + 2, d2, 2, d, d2);
        }
    }
    return gen_total;
}

// thread 0 of k_restart: resolve drained iterations; returns the index of an
// iteration to start, or -1
__device__ int resolve_instance(const KParams &p, int inst) {
    for (;;) {
        if (p.done[inst]) return -1;
        const int cur = p.it[inst];
        const int nl = p.nlive[inst];
        if (nl == 0) {
            p.nlive[inst] = 1;
            return cur;
        }
        const int isc = inst * 2 + (cur & 1);
        if (!p.drained[isc]) {
            if (p.want_next[inst] && nl == 1 && cur + 1 < p.max_iters && p.found[isc] == kNoGoal) {
                p.want_next[inst] = 0;
                p.nlive[inst] = 2;
                atomicAdd(&p.ctl[13], 1u);
                return cur + 1;
            }
            return -1;
        }
        p.iter_counts[(size_t)inst * p.max_iters + cur] = p.cnt[isc];
        if (p.found[isc] != kNoGoal || cur + 1 >= p.max_iters) {
            p.opt[inst] = p.found[isc] != kNoGoal ? p.found[isc] : -1;
            p.done[inst] = 1;
            p.dead[inst * 2] = 1;
            p.dead[inst * 2 + 1] = 1;
            atomicSub(&p.ctl[1], 1u);
            return -1;
        }
        p.it[inst] = cur + 1;
        p.nlive[inst] = nl - 1;
    }
}

template <int W, int HEUR>
__global__ void __launch_bounds__(256) k_restart(KParams p, const PdbGeo *geo_g) {
    __shared__ unsigned long long s_cnt;
    __shared__ unsigned s_bucket[64];
    __shared__ unsigned long long s_base;
    __shared__ int s_j, s_ring;
    __shared__ float s_est;
    __shared__ unsigned s_nfit;
    __shared__ PdbGeo geo_s;
    const unsigned n_list = p.ctl[0];
    if (blockIdx.x >= n_list) return;
    const int inst = p.restart_list[blockIdx.x];
    const unsigned nf = min(p.nfront[inst], (unsigned)p.capl);
    const Item *F = p.frontier + (size_t)inst * p.capl;
    unsigned long long *load = p.load + (size_t)inst * p.capl;
    // heavy-node re-split: only under a policy with estimates
    const bool resplit = p.resplit > 0 && p.lb != 0;
    PdbCtx P{&geo_s, p.pdb};
    if constexpr (HEUR == 1) {
        if (resplit) {
            const uint32_t *src = (const uint32_t *)geo_g;
            uint32_t *dst = (uint32_t *)&geo_s;
            for (int i = threadIdx.x; i < (int)(sizeof(PdbGeo) / 4); i += blockDim.x) dst[i] = src[i];
        }
    }

    for (;;) {
        if (threadIdx.x == 0) {
            s_j = resolve_instance(p, inst);
            s_ring = s_j > p.it[inst] ? 1 : 0; // a lookahead start goes to ring 1
        }
        __syncthreads();
        const int it = s_j;
        if (it < 0) break;
        const int isl = inst * 2 + (it & 1);
        const int L = p.Lbase[inst] + 2 * it;
        // 1. samples of the iteration that ended
        if (it > 0 && p.max_samples > 0 && !p.lookahead) {
            for (unsigned i = threadIdx.x; i < nf; i += blockDim.x) {
                if (path_f(F[i]) <= L - 2) {
                    unsigned k = atomicAdd(p.n_samples, 1u);
                    if (k < p.max_samples) {
                        Sample sm;
                        sm.item = F[i];
                        sm.thr = L - 2;
                        sm.pad = 0;
                        sm.load = load[i];
                        p.samples[k] = sm;
                    }
                }
            }
        }
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_est = 0.f;
            s_nfit = 0;
        }
        for (int b = threadIdx.x; b < 64; b += blockDim.x) s_bucket[b] = 0;
        __syncthreads();
        // 2. top tree: sum of gen over path-f <= L; with re-splitting, the
        // mean estimated subtree size of the frontier nodes that fit
        unsigned long long c = 0;
        for (int f = threadIdx.x; f < kFBins; f += blockDim.x)
            if (f <= L) c += p.topgen[(size_t)inst * kFBins + f];
        atomicAdd(&s_cnt, c);
        if (resplit) {
            float e = 0.f;
            unsigned k = 0;
            for (unsigned i = threadIdx.x; i < nf; i += blockDim.x)
                if (path_f(F[i]) <= L) {
                    e += exp2f(lb_log2_est(p, inst, i, it)) - 1.f;
                    ++k;
                }
            atomicAdd(&s_est, e);
            atomicAdd(&s_nfit, k);
        }
        __syncthreads();
        // a node is split one level when its estimate exceeds resplit x the
        // mean, two levels beyond 8 x that
        const float cut1 = resplit && s_nfit ? (float)p.resplit * s_est / (float)s_nfit : 3.4e38f;
        const float cut2 = 8.f * cut1;
        auto levels_of = [&](unsigned i, float key) -> int {
            if (!resplit) return 0;
            const float e = exp2f(key) - 1.f;
            return e > cut2 ? 2 : (e > cut1 ? 1 : 0);
        };
        // 3. bucket sizes of the items (frontier nodes that fit, or the items
        // a split node becomes); split nodes' generated children are counted
        for (unsigned i = threadIdx.x; i < nf; i += blockDim.x) {
            const Item x = F[i];
            if (path_f(x) > L) continue;
            const float key = lb_log2_est(p, inst, i, it);
            const int b = key_bucket(key);
            const int lv = levels_of(i, key);
            if (lv == 0 || meta_h(x.meta) == 0) {
                atomicAdd(&s_bucket[b], 1u);
                continue;
            }
            const unsigned gn = split_node<W, HEUR>(P, x, L, lv, [&](const Pos<W, HEUR> &, int, int, int level, int, int) {
                atomicAdd(&s_bucket[split_bucket(b, level)], 1u);
            });
            atomicAdd(&s_cnt, (unsigned long long)gn);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned off = 0;
            for (int b = 63; b >= 0; --b) { // largest keys first
                unsigned n = s_bucket[b];
                s_bucket[b] = off;
                off += n;
            }
            s_base = off ? q_reserve(p, s_ring, (unsigned long long)off) : 0ull;
            p.L[isl] = L;
            p.itn[isl] = it;
            p.found[isl] = p.top_goal[inst] <= L ? p.top_goal[inst] : kNoGoal;
            p.cnt[isl] = s_cnt;
            p.pend[isl] = off;
            p.drained[isl] = off == 0 ? 1 : 0;
            p.dead[isl] = 0;
        }
        __syncthreads();
        // 4. the items to the queue; the per-node loads of the new iteration
        // start from zero, or from the children a split counted
        for (unsigned i = threadIdx.x; i < nf; i += blockDim.x) {
            Item x = F[i];
            unsigned long long ld = 0;
            if (path_f(x) <= L) {
                const float key = lb_log2_est(p, inst, i, it);
                const int b = key_bucket(key);
                const int lv = levels_of(i, key);
                if (lv == 0 || meta_h(x.meta) == 0) {
                    unsigned k = atomicAdd(&s_bucket[b], 1u);
                    x.inst = (uint32_t)isl;
                    x.x1 = i;
                    const size_t slot = q_slot(p, s_ring, s_base + k);
                    p.q[slot] = x;
                    if (p.qpath) path_from_u64(p.fpath[(size_t)inst * p.capl + i], p.qpath, (uint32_t)slot * kPathWords);
                } else {
                    const int g0 = meta_g(x.meta);
                    unsigned made = 0;
                    ld = split_node<W, HEUR>(P, x, L, lv, [&](const Pos<W, HEUR> &cp, int g, int prev, int level, int d1, int d2) {
                        ++made;
                        unsigned k = atomicAdd(&s_bucket[split_bucket(b, level)], 1u);
                        Item ci;
                        cp.bd.pack(ci.b);
                        ci.meta = meta_pack(cp.cell(), g, cp.hs.h, prev);
                        ci.inst = (uint32_t)isl;
                        ci.x0 = 0;
                        ci.x1 = i;
                        const size_t slot = q_slot(p, s_ring, s_base + k);
                        p.q[slot] = ci;
                        if (p.qpath) {
                            const uint32_t qo = (uint32_t)slot * kPathWords;
                            path_from_u64(p.fpath[(size_t)inst * p.capl + i], p.qpath, qo);
                            path_put(p.qpath, qo, 1, g0, d1);
                            if (level == 2) path_put(p.qpath, qo, 1, g0 + 1, d2);
                        }
                    });
                    atomicAdd(&p.ctl[14], 1u);
                    atomicAdd(&p.ctl[15], made);
                }
            }
            load[i] = ld;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        p.want_next[inst] = 0;
        p.listed[inst] = 0;
    }
}

// k_spec (lookahead): after a round, while idle lanes outnumber the queued
// items by more than 1/32 of the lanes, put instances that run a single
// iteration (no goal yet, and a previous iteration, if any, of at least one
// node per lane) on the restart list with lookahead asked for, until their
// frontiers would fill the idle lanes.  Instances are taken from a
// start that rotates with the round number.  Only the schedule changes: each
// iteration's count is its own, and the result is resolved in order.
__global__ void __launch_bounds__(256) k_spec(KParams p) {
    __shared__ unsigned long long s_room, s_used;
    if (threadIdx.x == 0) {
        unsigned long long qlen = 0;
        for (int r = 0; r < 2; ++r) {
            const unsigned long long *c = p.q_ctl + kQCtl * r;
            const unsigned long long head = c[0], lim = c[1], tail = c[2];
            qlen += tail - (head < lim ? head : lim);
        }
        const unsigned long long lanes = (unsigned long long)p.n_lanes, busy = p.ctl[2];
        const unsigned long long taken = busy + qlen;
        s_room = (taken < lanes && lanes - taken >= lanes / 32) ? lanes - taken : 0ull;
        s_used = 0;
    }
    __syncthreads();
    if (s_room == 0) return;
    const int n = p.n_inst;
    const int start = (int)(p.ctl[6] % (unsigned)n);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        int inst = start + k;
        if (inst >= n) inst -= n;
        if (p.done[inst] || p.nlive[inst] != 1 || p.listed[inst]) continue;
        const int cur = p.it[inst];
        const int isc = inst * 2 + (cur & 1);
        if (cur + 1 >= p.max_iters || p.found[isc] != kNoGoal || p.drained[isc]) continue;
        // an instance whose last finished iteration counted fewer nodes than
        // there are lanes has iterations that end within a round or two:
        // lookahead would mostly run work that its goal then drops (100
        // 40-move 15-puzzles: 3200 instances/s with it, 5300 without)
        if (cur > 0 && p.iter_counts[(size_t)inst * p.max_iters + cur - 1] < (unsigned long long)p.n_lanes) continue;
        const unsigned long long est = min(p.nfront[inst], (unsigned)p.capl);
        if (atomicAdd(&s_used, est) >= s_room) continue;
        if (atomicExch(&p.listed[inst], 1) == 0) {
            p.want_next[inst] = 1;
            const unsigned k2 = atomicAdd(&p.ctl[0], 1u);
            p.restart_list[k2] = inst;
        }
    }
}

// k_control: between rounds.  Publishes the items appended to both rings
// during the last round, sizes the donation quotas from the number of idle
// lanes, and picks the next round's step budget.  ctl[9] = budget used when
// the queue can keep every lane busy, ctl[10] = budget while it cannot.
// Lanes on current iterations (ctl[4]) donate to fill the idle lanes not
// covered by ring 0; lanes on lookahead iterations (ctl[19]) only where both
// rings together fall short, at most kAheadDonate nodes per lane and round,
// and their items wait behind ring 0's.  A lookahead iteration's work is
// dropped when the current one ends the search; spread wide it also holds
// lanes, round after round, that the current iteration's tail needs (one
// random 15-puzzle per launch: caps of 1 / 2 / 4 / 8 / 32 gave 43.0 / 44.7 /
// 44.1 / 42.7 / 42.3 instances/s, and no donation at all 38.7).
constexpr unsigned kAheadDonate = 2;
__device__ __forceinline__ unsigned donate_quota(const KParams &p, unsigned long long idle,
                                                 unsigned long long queued, unsigned long long donors) {
    if (queued >= idle || donors == 0) return 0;
    unsigned long long per = (idle - queued + donors - 1) / donors;
    // donations of one round stay below a quarter of a ring
    const unsigned long long cap = ((p.qmask + 1ull) / 4ull) / donors;
    if (per > cap) per = cap;
    return (unsigned)(per > 32ull ? 32ull : per);
}

__global__ void k_control(KParams p) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long qlen[2];
    for (int r = 0; r < 2; ++r) {
        unsigned long long *c = p.q_ctl + kQCtl * r;
        unsigned long long head = c[0], limit = c[1], tail = c[2];
        if (head > limit) head = limit;
        c[0] = head;
        c[1] = tail;
        qlen[r] = tail - head;
    }
    const unsigned busy = p.ctl[2], busy1 = p.ctl[34];
    p.ctl[2] = 0;
    p.ctl[34] = 0;
    p.ctl[12] = busy; // lanes busy at the end of the last round, for the host
    p.ctl[0] = 0;     // restart list consumed
    const unsigned long long lanes = (unsigned long long)p.n_lanes;
    const unsigned long long idle = lanes > busy ? lanes - busy : 0ull;
    p.ctl[4] = donate_quota(p, idle, qlen[0], busy - busy1);
    p.ctl[19] = min(donate_quota(p, idle, qlen[0] + qlen[1], busy1), kAheadDonate);
    p.ctl[5] = (qlen[0] + qlen[1] + busy >= lanes / 2) ? p.ctl[9] : p.ctl[10];
    p.ctl[6] += 1;
}

// ---------------------------------------------------------------------------
// k_split_level: one level of the root split, all instances at once.  A level
// that already holds >= target nodes (or is at max depth) becomes the
// instance's frontier; otherwise every node is expanded without pruning:
// its generated children are counted into the top-tree histogram under the
// node's f, and the children form the next level.
template <int W, int HEUR>
__global__ void __launch_bounds__(256) k_split_level(KParams p, const PdbGeo *geo_g, const Item *lvl_in,
                                                     const unsigned *n_in, Item *lvl_out, unsigned *n_out,
                                                     unsigned *nfront, unsigned *topgen, int *top_goal,
                                                     int *frozen, int depth, int max_depth, unsigned target,
                                                     const unsigned long long *lp_in, unsigned long long *lp_out,
                                                     unsigned long long *fp_out) {
    __shared__ PdbGeo geo_s;
    PdbCtx P{&geo_s, p.pdb};
    if constexpr (HEUR == 1) {
        const uint32_t *src = (const uint32_t *)geo_g;
        uint32_t *dst = (uint32_t *)&geo_s;
        for (int i = threadIdx.x; i < (int)(sizeof(PdbGeo) / 4); i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int inst = (int)(gid / (size_t)p.capl);
    const unsigned j = (unsigned)(gid % (size_t)p.capl);
    if (inst >= p.n_inst) return;
    // frozen[inst] = 1 + the depth at which the instance froze; it is written
    // during that launch, so this launch only honours earlier freezes
    const int fz = frozen[inst];
    if (fz != 0 && fz - 1 < depth) return;
    const unsigned n = min(n_in[inst], (unsigned)p.capl);
    const bool freeze = (n >= target) || (depth >= max_depth);
    if (freeze) {
        if (j < n) {
            ((Item *)p.frontier)[(size_t)inst * p.capl + j] = lvl_in[(size_t)inst * p.capl + j];
            if (fp_out) fp_out[(size_t)inst * p.capl + j] = lp_in[(size_t)inst * p.capl + j];
        }
        if (j == 0) {
            nfront[inst] = n;
            frozen[inst] = depth + 1;
        }
        return;
    }
    if (j >= n) return;
    Item it = lvl_in[(size_t)inst * p.capl + j];
    Pos<W, HEUR> s;
    pos_from_item<W, HEUR>(s, it.b, it.meta, P);
    int g = meta_g(it.meta), prev = meta_prev(it.meta);
    // F = largest f on the path from the root (x0 carries the parent's F):
    // the node is expanded in iteration L iff F <= L, heuristic consistent
    // or not
    int f = g + s.hs.h;
    if ((int)it.x0 > f) f = (int)it.x0;
    if (s.hs.h == 0) {
        atomicMin(&top_goal[inst], f);
        // levels run in depth order, so the first goal recorded is a
        // shallowest one; it goes to both iteration slots (the iteration that
        // ends the instance, in either slot, finds no shallower goal)
        if (p.sol)
            for (int sl = inst * 2; sl < inst * 2 + 2; ++sl)
                if (atomicCAS(&p.sol_len[sl], -1, g) == -1)
                    path_from_u64(lp_in[(size_t)inst * p.capl + j], p.sol, (uint32_t)sl * kPathWords);
    }
    uint32_t gen = legal_moves(s) & ~(prev < 4 ? (1u << (3 - prev)) : 0u);
    atomicAdd(&topgen[(size_t)inst * kFBins + (f < kFBins ? f : kFBins - 1)], (unsigned)__builtin_popcount(gen));
    unsigned k = atomicAdd(&n_out[inst], (unsigned)__builtin_popcount(gen));
    while (gen) {
        int d = __builtin_ctz(gen);
        gen &= gen - 1;
        Pos<W, HEUR> c = s;
        do_move<W, HEUR>(c, d, P);
        Item ci;
        c.bd.pack(ci.b);
        ci.meta = meta_pack(c.cell(), g + 1, c.hs.h, d);
        ci.inst = (uint32_t)inst;
        ci.x0 = (uint32_t)f;
        ci.x1 = 0;
        if (k < (unsigned)p.capl) {
            lvl_out[(size_t)inst * p.capl + k] = ci;
            // split depth <= 32 when paths are recorded: move g fits the u64
            if (lp_out) lp_out[(size_t)inst * p.capl + k] = lp_in[(size_t)inst * p.capl + j] | ((unsigned long long)d << (2 * g));
        }
        ++k;
    }
}

// heuristic of packed items (spz_eval_h and the root h0)
template <int W, int HEUR>
__global__ void k_eval_h(KParams p, const PdbGeo *geo_g, Item *items, int n) {
    __shared__ PdbGeo geo_s;
    PdbCtx P{&geo_s, p.pdb};
    if constexpr (HEUR == 1) {
        const uint32_t *src = (const uint32_t *)geo_g;
        uint32_t *dst = (uint32_t *)&geo_s;
        for (int i = threadIdx.x; i < (int)(sizeof(PdbGeo) / 4); i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Item it = items[i];
    Board<W> bd;
    bd.unpack(it.b);
    int h = 0;
    if constexpr (HEUR == 0) {
        for (int c = 0; c < W * W; ++c) {
            int t = bd.tile(c);
            if (t) h += abs(c / W - trow<W>(t)) + abs(c % W - tcol<W>(t));
        }
    } else {
        HState<1> hs;
        pdb_init(P, hs, bd);
        h = hs.h;
    }
    it.meta = (it.meta & 0xff00ffffu) | ((uint32_t)h << 16);
    items[i] = it;
}

// heuristic of n boards given as bytes (row-major, W*W per board) into
// h[n]; the allocation-free evaluation entry point's kernel
template <int W, int HEUR>
__global__ void k_eval_h_bytes(const uint8_t *bytes, int n, const PdbGeo *geo_g, const uint8_t *pdb, int *h) {
    __shared__ PdbGeo geo_s;
    PdbCtx P{&geo_s, pdb};
    if constexpr (HEUR == 1) {
        const uint32_t *src = (const uint32_t *)geo_g;
        uint32_t *dst = (uint32_t *)&geo_s;
        for (int i = threadIdx.x; i < (int)(sizeof(PdbGeo) / 4); i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t *b = bytes + (size_t)i * (W * W);
    Board<W> bd;
    if constexpr (W == 4) {
        bd.b = 0;
#pragma unroll
        for (int c = 0; c < 16; ++c) bd.b |= (uint64_t)b[c] << (4 * c);
    } else {
        bd.lo = bd.hi = 0;
        bd.top = b[24];
#pragma unroll
        for (int c = 0; c < 12; ++c) bd.lo |= (uint64_t)b[c] << (5 * c);
#pragma unroll
        for (int c = 12; c < 24; ++c) bd.hi |= (uint64_t)b[c] << (5 * (c - 12));
    }
    int v = 0;
    if constexpr (HEUR == 0) {
        for (int c = 0; c < W * W; ++c) {
            const int t = bd.tile(c);
            if (t) v += abs(c / W - trow<W>(t)) + abs(c % W - tcol<W>(t));
        }
    } else {
        HState<1> hs;
        pdb_init(P, hs, bd);
        v = hs.h;
    }
    h[i] = v;
}

hipError_t launch_eval_h_bytes(int width, int heur, const uint8_t *bytes, int n, const PdbGeo *geo,
                               const uint8_t *pdb, int *h, hipStream_t s) {
    dim3 grid((unsigned)((n + 255) / 256));
    if (width == 4)
        hipLaunchKernelGGL((k_eval_h_bytes<4, 0>), grid, dim3(256), 0, s, bytes, n, geo, pdb, h);
    else if (heur == 0)
        hipLaunchKernelGGL((k_eval_h_bytes<5, 0>), grid, dim3(256), 0, s, bytes, n, geo, pdb, h);
    else
        hipLaunchKernelGGL((k_eval_h_bytes<5, 1>), grid, dim3(256), 0, s, bytes, n, geo, pdb, h);
    return hipGetLastError();
}

// roots -> slot 0 of each instance's level-0 region
__global__ void k_scatter_roots(const Item *roots, Item *lvl, int capl, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lvl[(size_t)i * capl] = roots[i];
}

// ---------------------------------------------------------------------------
// host-side launchers
size_t round_lds_bytes(int width, int maxd, int wg, int heur);

template <int W, int HEUR>
// With t0/t1 the runtime records the two events with the dispatch itself, so
// their interval is the kernel's own duration (no launch gap before it).
static hipError_t launch_round_t(const KParams &p, const PdbGeo *geo, int wg, hipStream_t s, hipEvent_t t0,
                                 hipEvent_t t1) {
    size_t lds = round_lds_bytes(W, p.maxd, wg, HEUR);
    if (t0 && t1)
        hipExtLaunchKernelGGL((k_round<W, HEUR>), dim3(p.n_lanes / wg), dim3(wg), (uint32_t)lds, s, t0, t1, 0u, p,
                              geo);
    else
        hipLaunchKernelGGL((k_round<W, HEUR>), dim3(p.n_lanes / wg), dim3(wg), lds, s, p, geo);
    return hipGetLastError();
}

hipError_t launch_round(int width, int heur, const KParams &p, const PdbGeo *geo, int wg, hipStream_t s,
                        hipEvent_t t0, hipEvent_t t1) {
    if (width == 4) return launch_round_t<4, 0>(p, geo, wg, s, t0, t1);
    if (heur == 0) return launch_round_t<5, 0>(p, geo, wg, s, t0, t1);
    return launch_round_t<5, 1>(p, geo, wg, s, t0, t1);
}

// dynamic LDS of k_round: the stack levels, then (width 5) the boards, then
// (PDB) the sibling-value cache
size_t round_lds_bytes(int width, int maxd, int wg, int heur) {
    return (size_t)maxd * wg * (heur == 1 ? 2 : 1) + (width == 5 ? (size_t)kLdsBoardBytes * wg : 0) +
           (heur == 1 ? (size_t)kSibLevels * 4 * wg : 0);
}

const void *round_kernel_ptr(int width, int heur) {
    if (width == 4) return (const void *)&k_round<4, 0>;
    if (heur == 0) return (const void *)&k_round<5, 0>;
    return (const void *)&k_round<5, 1>;
}

hipError_t launch_restart(int width, int heur, const KParams &p, const PdbGeo *geo, hipStream_t s) {
    if (width == 4)
        hipLaunchKernelGGL((k_restart<4, 0>), dim3(p.n_inst), dim3(256), 0, s, p, geo);
    else if (heur == 0)
        hipLaunchKernelGGL((k_restart<5, 0>), dim3(p.n_inst), dim3(256), 0, s, p, geo);
    else
        hipLaunchKernelGGL((k_restart<5, 1>), dim3(p.n_inst), dim3(256), 0, s, p, geo);
    return hipGetLastError();
}

hipError_t launch_spec(const KParams &p, hipStream_t s) {
    hipLaunchKernelGGL(k_spec, dim3(1), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_control(const KParams &p, hipStream_t s) {
    hipLaunchKernelGGL(k_control, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_split_level(int width, int heur, const KParams &p, const PdbGeo *geo, const Item *lvl_in,
                              const unsigned *n_in, Item *lvl_out, unsigned *n_out, unsigned *nfront,
                              unsigned *topgen, int *top_goal, int *frozen, int depth, int max_depth,
                              unsigned target, const unsigned long long *lp_in, unsigned long long *lp_out,
                              unsigned long long *fp_out, hipStream_t s) {
    size_t total = (size_t)p.n_inst * p.capl;
    dim3 grid((unsigned)((total + 255) / 256));
    if (width == 4)
        hipLaunchKernelGGL((k_split_level<4, 0>), grid, dim3(256), 0, s, p, geo, lvl_in, n_in, lvl_out, n_out, nfront,
                           topgen, top_goal, frozen, depth, max_depth, target, lp_in, lp_out, fp_out);
    else if (heur == 0)
        hipLaunchKernelGGL((k_split_level<5, 0>), grid, dim3(256), 0, s, p, geo, lvl_in, n_in, lvl_out, n_out, nfront,
                           topgen, top_goal, frozen, depth, max_depth, target, lp_in, lp_out, fp_out);
    else
        hipLaunchKernelGGL((k_split_level<5, 1>), grid, dim3(256), 0, s, p, geo, lvl_in, n_in, lvl_out, n_out, nfront,
                           topgen, top_goal, frozen, depth, max_depth, target, lp_in, lp_out, fp_out);
    return hipGetLastError();
}

hipError_t launch_scatter_roots(const Item *roots, Item *lvl, int capl, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_roots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, roots, lvl, capl, n);
    return hipGetLastError();
}

hipError_t launch_eval_h(int width, int heur, const KParams &p, const PdbGeo *geo, Item *items, int n,
                         hipStream_t s) {
    dim3 grid((unsigned)((n + 255) / 256));
    if (width == 4)
        hipLaunchKernelGGL((k_eval_h<4, 0>), grid, dim3(256), 0, s, p, geo, items, n);
    else if (heur == 0)
        hipLaunchKernelGGL((k_eval_h<5, 0>), grid, dim3(256), 0, s, p, geo, items, n);
    else
        hipLaunchKernelGGL((k_eval_h<5, 1>), grid, dim3(256), 0, s, p, geo, items, n);
    return hipGetLastError();
}

} // namespace spz
