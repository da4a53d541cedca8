// spz_board.h -- per-lane puzzle state for the DFS engine (device code).
//
// A board lives in scalar registers only (no arrays, so nothing is placed in
// scratch): the 15-puzzle as one u64 of 4-bit cells, the 24-puzzle as two
// u64 of twelve 5-bit cells (cells 0-11, 12-23) plus one word for cell 24.
// The heuristic state rides along: Manhattan h updated by +-1 per move
// (src/24puzzle.cc:129-131), or the eight PDB lookup indices/values updated
// incrementally, re-reading only the two lookups that contain the moved tile
// (src_horie/idas_pdb.cu:220-255).
#pragma once
#include "spz_internal.h"

namespace spz {

template <int W> struct Geo;
// MAXD: upper bound of the per-lane LDS stack depth (the depth in use is a
// launch parameter); DEFD: default depth.  A lane whose stack is full turns
// the children it would push into queue items, so depth never limits
// correctness.
template <> struct Geo<4> {
    static constexpr int N2 = 16;
    static constexpr int MAXD = 160;
    static constexpr int DEFD = 80;
};
template <> struct Geo<5> {
    static constexpr int N2 = 25;
    static constexpr int MAXD = 160;
    static constexpr int DEFD = 96;
};
// default LDS stack depth of the PDB search (two-byte entries)
constexpr int kDefDepthPdb = 64;

// goal row / column of tile t (tile t belongs at cell t)
template <int W> __host__ __device__ __forceinline__ int trow(int t) {
    return W == 4 ? (t >> 2) : ((t * 13) >> 6); // exact for t < 25
}
template <int W> __host__ __device__ __forceinline__ int tcol(int t) { return t - W * trow<W>(t); }

template <int W> struct Board;

template <> struct Board<4> {
    uint64_t b; // cell c at bits 4c
    __host__ __device__ __forceinline__ int tile(int c) const { return (int)((b >> (4 * c)) & 15u); }
    // tile t slides from cell `from` into the blank at `to`
    __host__ __device__ __forceinline__ void slide(int t, int from, int to) {
        b += ((uint64_t)t << (4 * to)) - ((uint64_t)t << (4 * from));
    }
    __host__ __device__ __forceinline__ void pack(uint32_t (&o)[4]) const {
        o[0] = (uint32_t)b;
        o[1] = (uint32_t)(b >> 32);
        o[2] = 0;
        o[3] = 0;
    }
    __host__ __device__ __forceinline__ void unpack(const uint32_t (&o)[4]) {
        b = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
    }
};

template <> struct Board<5> {
    uint64_t lo, hi; // cells 0-11 at bits 5c, cells 12-23 at bits 5(c-12)
    uint32_t top;    // cell 24
    __host__ __device__ __forceinline__ int tile(int c) const {
        uint32_t a = (uint32_t)((lo >> (5 * (c < 12 ? c : 0))) & 31u);
        uint32_t h = (uint32_t)((hi >> (5 * (c >= 12 && c < 24 ? c - 12 : 0))) & 31u);
        return (int)(c < 12 ? a : (c < 24 ? h : top));
    }
    __host__ __device__ __forceinline__ void add(int c, int v) {
        lo += (c < 12) ? ((uint64_t)v << (5 * (c < 12 ? c : 0))) : 0ull;
        hi += (c >= 12 && c < 24) ? ((uint64_t)v << (5 * (c >= 12 && c < 24 ? c - 12 : 0))) : 0ull;
        top += (c == 24) ? (uint32_t)v : 0u;
    }
    __host__ __device__ __forceinline__ void slide(int t, int from, int to) {
        add(to, t);
        add(from, -t);
    }
    // queue format: lo/hi low 60 bits, cell 24 split over their top 4 bits
    __host__ __device__ __forceinline__ void pack(uint32_t (&o)[4]) const {
        uint64_t l = lo | ((uint64_t)(top & 15u) << 60);
        uint64_t h = hi | ((uint64_t)(top >> 4) << 60);
        o[0] = (uint32_t)l;
        o[1] = (uint32_t)(l >> 32);
        o[2] = (uint32_t)h;
        o[3] = (uint32_t)(h >> 32);
    }
    __host__ __device__ __forceinline__ void unpack(const uint32_t (&o)[4]) {
        uint64_t l = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
        uint64_t h = (uint64_t)o[2] | ((uint64_t)o[3] << 32);
        lo = l & ((1ull << 60) - 1);
        hi = h & ((1ull << 60) - 1);
        top = (uint32_t)(l >> 60) | ((uint32_t)(h >> 60) << 4);
    }
};

// host helper: board from a row-major byte array
// 5x5 board of the search kernel, held in LDS: one byte per cell, each lane
// its own kLdsBoardBytes (7-dword) region, so reading the tile next to the
// blank is one LDS load and a slide two byte stores.  With an odd dword
// stride the lanes of a wave start in different banks (bytes of several
// lanes interleaved in one dword would make lanes reading different cells
// collide in one bank).  Copying it copies the reference, not the cells:
// regs() gives a register copy.
constexpr int kLdsBoardBytes = 28;
struct LdsBoard5 {
    uint8_t *p;  // this lane's cell 0
    __device__ __forceinline__ int tile(int c) const { return (int)p[c]; }
    __device__ __forceinline__ void slide(int t, int from, int to) {
        p[to] = (uint8_t)t;
        p[from] = 0;
    }
    __device__ __forceinline__ void unpack(const uint32_t (&o)[4]) {
        Board<5> b;
        b.unpack(o);
#pragma unroll
        for (int c = 0; c < 25; ++c) p[c] = (uint8_t)b.tile(c);
    }
    __device__ __forceinline__ Board<5> regs() const {
        Board<5> b;
        b.lo = b.hi = 0ull;
        b.top = 0u;
#pragma unroll
        for (int c = 0; c < 25; ++c) b.add(c, tile(c));
        return b;
    }
    __device__ __forceinline__ void pack(uint32_t (&o)[4]) const { regs().pack(o); }
};

// board type of the search kernel's lanes
template <int W> struct HotBoard;
template <> struct HotBoard<4> { using T = Board<4>; };
template <> struct HotBoard<5> { using T = LdsBoard5; };

template <int W> inline void board_from_bytes(const uint8_t *b, Board<W> &o);
template <> inline void board_from_bytes<4>(const uint8_t *b, Board<4> &o) {
    o.b = 0;
    for (int c = 0; c < 16; ++c) o.b |= (uint64_t)b[c] << (4 * c);
}
template <> inline void board_from_bytes<5>(const uint8_t *b, Board<5> &o) {
    o.lo = o.hi = 0;
    o.top = b[24];
    for (int c = 0; c < 12; ++c) o.lo |= (uint64_t)b[c] << (5 * c);
    for (int c = 12; c < 24; ++c) o.hi |= (uint64_t)b[c] << (5 * (c - 12));
}

// ---------------------------------------------------------------------------
// PDB lookup geometry.  Lookup k (0..3 regular partition, 4..7 reflected)
// reads canonical table tab(k) (0 for k = 0,4; 1 otherwise) at the direct
// 25^6 index sum_j map_k(pos(tile_kj)) * 25^(5-j) (src/24puzzle_with_pdb.cc:
// 114-189 writes the same eight hash functions out by hand).
struct PdbGeo {
    uint8_t map[8][25];   // cell map of lookup k
    uint8_t tiles[8][6];  // tiles of lookup k in slot order
    uint8_t reg_k[25];    // regular lookup containing tile t (0..3)
    uint8_t ref_k[25];    // reflected lookup containing tile t (4..7)
    uint8_t pad[2];
    uint32_t reg_mul[25]; // 25^(5-slot) of tile t in reg_k
    uint32_t ref_mul[25];
    // index weights of tile t on cell c in its regular (.x) and reflected (.y)
    // lookup, map_k(c) * 25^(5-slot) (< 2^28), with k (0..3) in bits 28-29: a
    // move's index change is w[t][to] - w[t][from] (the k bits cancel).  The
    // pair is one 8-byte LDS read.
    uint2 w[25][25];
};
constexpr uint32_t kWeightMask = 0x0fffffffu;

// Build the geometry on the host from the cell symmetries (the reference's
// rf / rot90 / rot180 tables, src/24puzzle_with_pdb.cc:33-46).
void pdb_geo_build(PdbGeo *g);

__host__ __device__ inline uint32_t pdb_tab_of(int k) { return (k == 0 || k == 4) ? 0u : 1u; }

constexpr uint64_t kPdbDirect = 244140625ull; // 25^6

// Per-lane heuristic state.
template <int HEUR> struct HState;

template <> struct HState<0> {
    int h;
};

template <> struct HState<1> {
    int h;
    int A, B;                   // regular / reflected sums, h = max(A, B)
    // direct indices into the two tables laid end to end (table 1 starts at
    // kPdbDirect; lookups 0 and 4 read table 0, the others table 1)
    uint32_t i0, i1, i2, i3;    // regular lookups 0..3
    uint32_t j0, j1, j2, j3;    // reflected lookups 4..7
    uint32_t v0, v1;            // lookup values, bytes 0..3 = lookups 0..3 / 4..7
};

struct PdbCtx {
    const PdbGeo *geo;  // LDS copy
    const uint8_t *tab; // table 0 at 0, table 1 at kPdbDirect
};

__device__ __forceinline__ uint32_t sel4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, int k) {
    uint32_t lo = (k & 1) ? b : a;
    uint32_t hi = (k & 1) ? d : c;
    return (k & 2) ? hi : lo;
}
__device__ __forceinline__ void add4(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, int k, uint32_t v) {
    a += (k == 0) ? v : 0u;
    b += (k == 1) ? v : 0u;
    c += (k == 2) ? v : 0u;
    d += (k == 3) ? v : 0u;
}
__device__ __forceinline__ int byte_of(uint32_t w, int k) { return (int)((w >> (8 * k)) & 0xffu); }
__device__ __forceinline__ uint32_t set_byte(uint32_t w, int k, int v) {
    return (w & ~(0xffu << (8 * k))) | ((uint32_t)v << (8 * k));
}

// idx carries its table's offset, so every lookup is one load from one base
__device__ __forceinline__ int pdb_read(const PdbCtx &P, uint32_t idx) { return (int)P.tab[idx]; }

// Heuristic of the child reached by moving tile t from cell a to cell b,
// without changing the state (f-tests at expansion); the two values read
// are returned packed (regular | reflected << 8) for reuse by pdb_apply_known.
__device__ __forceinline__ int pdb_child_h(const PdbCtx &P, const HState<1> &s, int t, int a, int b,
                                           uint32_t &vals) {
    const PdbGeo &G = *P.geo;
    const uint2 wa = G.w[t][a], wb = G.w[t][b];
    const int kr = (int)(wa.x >> 28), kf = (int)(wa.y >> 28);
    uint32_t ir = sel4(s.i0, s.i1, s.i2, s.i3, kr) + (wb.x - wa.x);
    uint32_t iff = sel4(s.j0, s.j1, s.j2, s.j3, kf) + (wb.y - wa.y);
    int vr = pdb_read(P, ir);
    int vf = pdb_read(P, iff);
    vals = (uint32_t)vr | ((uint32_t)vf << 8);
    int A = s.A - byte_of(s.v0, kr) + vr;
    int B = s.B - byte_of(s.v1, kf) + vf;
    return A > B ? A : B;
}

// pdb_apply with the two new values already known (no gathers)
__device__ __forceinline__ void pdb_apply_known(const PdbCtx &P, HState<1> &s, int t, int a, int b,
                                                uint32_t vals) {
    const PdbGeo &G = *P.geo;
    const uint2 wa = G.w[t][a], wb = G.w[t][b];
    const int kr = (int)(wa.x >> 28), kf = (int)(wa.y >> 28);
    const uint32_t dr = wb.x - wa.x, df = wb.y - wa.y;
    add4(s.i0, s.i1, s.i2, s.i3, kr, dr);
    add4(s.j0, s.j1, s.j2, s.j3, kf, df);
    int vr = (int)(vals & 0xffu), vf = (int)((vals >> 8) & 0xffu);
    s.A += vr - byte_of(s.v0, kr);
    s.B += vf - byte_of(s.v1, kf);
    s.v0 = set_byte(s.v0, kr, vr);
    s.v1 = set_byte(s.v1, kf, vf);
    s.h = s.A > s.B ? s.A : s.B;
}

__device__ __forceinline__ void pdb_apply(const PdbCtx &P, HState<1> &s, int t, int a, int b) {
    const PdbGeo &G = *P.geo;
    const uint2 wa = G.w[t][a], wb = G.w[t][b];
    const int kr = (int)(wa.x >> 28), kf = (int)(wa.y >> 28);
    const uint32_t dr = wb.x - wa.x, df = wb.y - wa.y;
    add4(s.i0, s.i1, s.i2, s.i3, kr, dr);
    add4(s.j0, s.j1, s.j2, s.j3, kf, df);
    int vr = pdb_read(P, sel4(s.i0, s.i1, s.i2, s.i3, kr));
    int vf = pdb_read(P, sel4(s.j0, s.j1, s.j2, s.j3, kf));
    s.A += vr - byte_of(s.v0, kr);
    s.B += vf - byte_of(s.v1, kf);
    s.v0 = set_byte(s.v0, kr, vr);
    s.v1 = set_byte(s.v1, kf, vf);
    s.h = s.A > s.B ? s.A : s.B;
}

// Full (non-incremental) initialisation of the PDB state from a board:
// every non-blank cell adds map_k(cell) * 25^(5-slot) to the index of the
// regular and the reflected lookup containing its tile.
template <class BD> __device__ __forceinline__ void pdb_init(const PdbCtx &P, HState<1> &s, const BD &bd) {
    const PdbGeo &G = *P.geo;
    s.i0 = s.j0 = 0u; // lookups 0 and 4: table 0
    s.i1 = s.i2 = s.i3 = (uint32_t)kPdbDirect;
    s.j1 = s.j2 = s.j3 = (uint32_t)kPdbDirect;
#pragma unroll
    for (int c = 0; c < 25; ++c) {
        int t = bd.tile(c);
        if (t != 0) {
            const uint2 wc = G.w[t][c];
            add4(s.i0, s.i1, s.i2, s.i3, (int)(wc.x >> 28), wc.x & kWeightMask);
            add4(s.j0, s.j1, s.j2, s.j3, (int)(wc.y >> 28), wc.y & kWeightMask);
        }
    }
    int a0 = pdb_read(P, s.i0), a1 = pdb_read(P, s.i1), a2 = pdb_read(P, s.i2), a3 = pdb_read(P, s.i3);
    int b0 = pdb_read(P, s.j0), b1 = pdb_read(P, s.j1), b2 = pdb_read(P, s.j2), b3 = pdb_read(P, s.j3);
    s.v0 = (uint32_t)a0 | ((uint32_t)a1 << 8) | ((uint32_t)a2 << 16) | ((uint32_t)a3 << 24);
    s.v1 = (uint32_t)b0 | ((uint32_t)b1 << 8) | ((uint32_t)b2 << 16) | ((uint32_t)b3 << 24);
    s.A = a0 + a1 + a2 + a3;
    s.B = b0 + b1 + b2 + b3;
    s.h = s.A > s.B ? s.A : s.B;
}

// Lane position: board + blank + heuristic state.  The blank is kept as its
// row and column, or (15-puzzle) as its cell, from which row and column are
// a shift and a mask.
template <int W, int HEUR, class BD = Board<W>> struct Pos {
    BD bd;
    int br, bc;
    HState<HEUR> hs;
    __device__ __forceinline__ int cell() const { return br * W + bc; }
    __device__ __forceinline__ int row() const { return br; }
    __device__ __forceinline__ int col() const { return bc; }
    __device__ __forceinline__ void set_cell(int c) {
        br = c / W;
        bc = c - br * W;
    }
    __device__ __forceinline__ void step(int d) {
        br += dir_dr(d);
        bc += dir_dc(d);
    }
};
template <int HEUR, class BD> struct Pos<4, HEUR, BD> {
    BD bd;
    int bl;
    HState<HEUR> hs;
    __device__ __forceinline__ int cell() const { return bl; }
    __device__ __forceinline__ int row() const { return bl >> 2; }
    __device__ __forceinline__ int col() const { return bl & 3; }
    __device__ __forceinline__ void set_cell(int c) { bl = c; }
    __device__ __forceinline__ void step(int d) { bl += dir_cell<4>(d); }
};

// moves the blank can make from cell c (bit d: direction d stays on the board)
template <int W> __device__ __forceinline__ uint32_t legal_moves(int r, int c) {
    return (r > 0 ? 1u : 0u) | (c < W - 1 ? 2u : 0u) | (c > 0 ? 4u : 0u) | (r < W - 1 ? 8u : 0u);
}
// 15-puzzle: the same masks, 4 bits per cell, from one 64-bit constant
constexpr uint64_t legal4_table() {
    uint64_t t = 0;
    for (int c = 0; c < 16; ++c) {
        const int r = c >> 2, k = c & 3;
        const uint64_t m = (r > 0 ? 1u : 0u) | (k < 3 ? 2u : 0u) | (k > 0 ? 4u : 0u) | (r < 3 ? 8u : 0u);
        t |= m << (4 * c);
    }
    return t;
}
template <int W> __device__ __forceinline__ uint32_t legal_cell(int c) {
    if constexpr (W == 4)
        return (uint32_t)(legal4_table() >> (4 * c)) & 15u;
    else
        return legal_moves<W>(trow<W>(c), tcol<W>(c));
}
template <int W, int HEUR, class BD> __device__ __forceinline__ uint32_t legal_moves(const Pos<W, HEUR, BD> &s) {
    if constexpr (W == 4)
        return legal_cell<4>(s.bl);
    else
        return legal_moves<W>(s.br, s.bc);
}

// Manhattan: the moves of a blank at cell c that slide a tile nearer its
// goal (h - 1), bit d for direction d.  Bits of moves off the board are
// meaningless: callers mask them with the legal moves.
template <int W, class BD> __device__ __forceinline__ uint32_t nearer_moves(const BD &bd, int c) {
    const int r = trow<W>(c), k = tcol<W>(c);
    int tU, tD, tL, tR;
    if constexpr (W == 4) {
        // the board rotated so that cell c - 4 is at bits 0-3: the four
        // neighbours sit at bits 0, 12, 20 and 32 (off-board ones wrap to
        // some cell)
        const uint32_t sh = (uint32_t)(4 * c - 16) & 63u;
        const uint64_t w = (bd.b >> sh) | (bd.b << ((64u - sh) & 63u));
        const uint32_t lo = (uint32_t)w;
        tU = (int)(lo & 15u);
        tL = (int)((lo >> 12) & 15u);
        tR = (int)((lo >> 20) & 15u);
        tD = (int)((uint32_t)(w >> 32) & 15u);
    } else {
        tU = bd.tile(r > 0 ? c - W : c);
        tD = bd.tile(r < W - 1 ? c + W : c);
        tL = bd.tile(k > 0 ? c - 1 : c);
        tR = bd.tile(k < W - 1 ? c + 1 : c);
    }
    return (trow<W>(tU) >= r ? 1u : 0u) | (tcol<W>(tR) <= k ? 2u : 0u) | (tcol<W>(tL) >= k ? 4u : 0u) |
           (trow<W>(tD) <= r ? 8u : 0u);
}

// Manhattan move of the blank in direction d.  The tile slides one cell
// along one axis, so h changes by exactly 1: down by 1 when the tile's goal
// row (vertical move) or column (horizontal move) lies at or beyond the cell
// it slides into.  Returns 1 when h went down, which is all that undoing the
// move needs (unmove_md).
template <int W, class BD>
__device__ __forceinline__ uint32_t move_md(Pos<W, 0, BD> &s, int d) {
    const int to = s.cell(), from = to + dir_cell<W>(d);
    const int t = s.bd.tile(from);
    const bool horiz = ((d ^ (d >> 1)) & 1) != 0; // right or left
    int key, at;
    if constexpr (W == 4) {
        const int sh = horiz ? 0 : 2;
        key = (t >> sh) & 3;
        at = (to >> sh) & 3;
    } else {
        key = horiz ? tcol<W>(t) : trow<W>(t);
        at = horiz ? s.col() : s.row();
    }
    // up / left: the tile moves to a larger row / column; right / down: smaller
    const int ahead = (d & 1) ? at - key : key - at;
    s.hs.h += ahead >= 0 ? -1 : 1;
    s.bd.slide(t, from, to);
    s.step(d);
    return ahead >= 0 ? 1u : 0u;
}

// Undo of move_md(s, d) that returned `nearer`: the blank moves back.
template <int W, class BD>
__device__ __forceinline__ void unmove_md(Pos<W, 0, BD> &s, int d, uint32_t nearer) {
    const int to = s.cell(), from = to - dir_cell<W>(d);
    s.bd.slide(s.bd.tile(from), from, to);
    s.hs.h += nearer ? 1 : -1;
    s.step(3 - d);
}

// Move the blank one cell in direction d: the tile at the neighbour slides
// into the blank.
template <int W, int HEUR, class BD>
__device__ __forceinline__ void do_move(Pos<W, HEUR, BD> &s, int d, const PdbCtx &P) {
    if constexpr (HEUR == 0) {
        move_md(s, d);
    } else {
        const int to = s.cell(), from = to + dir_cell<W>(d);
        const int t = s.bd.tile(from);
        pdb_apply(P, s.hs, t, from, to);
        s.bd.slide(t, from, to);
        s.step(d);
    }
}

// Blank move that updates the board and blank only.  Queue items of the PDB
// search need no heuristic state: a lane starting an item recomputes all
// eight lookups from the board (pdb_init), so the positions of donated and
// exported items are built with these moves and no table reads.
template <int W, int HEUR, class BD>
__device__ __forceinline__ void move_board_only(Pos<W, HEUR, BD> &s, int d) {
    const int to = s.cell(), from = to + dir_cell<W>(d);
    s.bd.slide(s.bd.tile(from), from, to);
    s.step(d);
}

// do_move for the PDB state with the moved tile's two lookup values known
template <int W, class BD>
__device__ __forceinline__ void do_move_known(Pos<W, 1, BD> &s, int d, const PdbCtx &P, uint32_t vals) {
    int nr = s.br + dir_dr(d);
    int nc = s.bc + dir_dc(d);
    int from = nr * W + nc, to = s.br * W + s.bc;
    int t = s.bd.tile(from);
    pdb_apply_known(P, s.hs, t, from, to, vals);
    s.bd.slide(t, from, to);
    s.br = nr;
    s.bc = nc;
}

// Blank move in direction d for the PDB search.  The moved tile's two new
// lookup values are either known (kv, packed regular | reflected << 8, from
// the parent's f-test) or read from the tables.  Returns the two values the
// move replaced, packed the same way, for the stack entry that undoes it.
template <int W, class BD>
__device__ __forceinline__ uint32_t move_pdb(Pos<W, 1, BD> &s, int d, const PdbCtx &P, bool known, uint32_t kv) {
    const int nr = s.br + dir_dr(d);
    const int nc = s.bc + dir_dc(d);
    const int from = nr * W + nc, to = s.br * W + s.bc;
    const int t = s.bd.tile(from);
    const PdbGeo &G = *P.geo;
    const uint2 wa = G.w[t][from], wb = G.w[t][to];
    const int kr = (int)(wa.x >> 28), kf = (int)(wa.y >> 28);
    HState<1> &h = s.hs;
    add4(h.i0, h.i1, h.i2, h.i3, kr, wb.x - wa.x);
    add4(h.j0, h.j1, h.j2, h.j3, kf, wb.y - wa.y);
    const int oldr = byte_of(h.v0, kr), oldf = byte_of(h.v1, kf);
    int vr, vf;
    if (known) {
        vr = (int)(kv & 0xffu);
        vf = (int)((kv >> 8) & 0xffu);
    } else {
        vr = pdb_read(P, sel4(h.i0, h.i1, h.i2, h.i3, kr));
        vf = pdb_read(P, sel4(h.j0, h.j1, h.j2, h.j3, kf));
    }
    h.A += vr - oldr;
    h.B += vf - oldf;
    h.v0 = set_byte(h.v0, kr, vr);
    h.v1 = set_byte(h.v1, kf, vf);
    h.h = h.A > h.B ? h.A : h.B;
    s.bd.slide(t, from, to);
    s.br = nr;
    s.bc = nc;
    return (uint32_t)oldr | ((uint32_t)oldf << 8);
}

// move_pdb in two halves, so that the table reads of a move and those of the
// f-test that follows it are in flight together: the first half moves the
// tile and updates the indices, and reads (or takes) the two new values; the
// second applies them to the sums once the f-test's reads are issued.
struct PendMove {
    int kr, kf;  // the lookups the move changed
    int vr, vf;  // their new values
};
template <int W, class BD>
__device__ __forceinline__ PendMove move_pdb_begin(Pos<W, 1, BD> &s, int d, const PdbCtx &P, bool known, uint32_t kv,
                                                   uint32_t &ov) {
    const int nr = s.br + dir_dr(d);
    const int nc = s.bc + dir_dc(d);
    const int from = nr * W + nc, to = s.br * W + s.bc;
    const int t = s.bd.tile(from);
    const PdbGeo &G = *P.geo;
    const uint2 wa = G.w[t][from], wb = G.w[t][to];
    PendMove m;
    m.kr = (int)(wa.x >> 28);
    m.kf = (int)(wa.y >> 28);
    HState<1> &h = s.hs;
    add4(h.i0, h.i1, h.i2, h.i3, m.kr, wb.x - wa.x);
    add4(h.j0, h.j1, h.j2, h.j3, m.kf, wb.y - wa.y);
    ov = (uint32_t)byte_of(h.v0, m.kr) | ((uint32_t)byte_of(h.v1, m.kf) << 8);
    if (known) {
        m.vr = (int)(kv & 0xffu);
        m.vf = (int)((kv >> 8) & 0xffu);
    } else {
        m.vr = pdb_read(P, sel4(h.i0, h.i1, h.i2, h.i3, m.kr));
        m.vf = pdb_read(P, sel4(h.j0, h.j1, h.j2, h.j3, m.kf));
    }
    s.bd.slide(t, from, to);
    s.br = nr;
    s.bc = nc;
    return m;
}
__device__ __forceinline__ void move_pdb_finish(HState<1> &h, const PendMove &m) {
    h.A += m.vr - byte_of(h.v0, m.kr);
    h.B += m.vf - byte_of(h.v1, m.kf);
    h.v0 = set_byte(h.v0, m.kr, m.vr);
    h.v1 = set_byte(h.v1, m.kf, m.vf);
    h.h = h.A > h.B ? h.A : h.B;
}

// values, before the move, of the two lookups containing the tile that a
// blank move in direction d slides (packed regular | reflected << 8)
template <int W, class BD>
__device__ __forceinline__ uint32_t pdb_moved_vals(const Pos<W, 1, BD> &s, int d, const PdbCtx &P) {
    int nr = s.br + dir_dr(d);
    int nc = s.bc + dir_dc(d);
    int t = s.bd.tile(nr * W + nc);
    const PdbGeo &G = *P.geo;
    int kr = G.reg_k[t], kf = G.ref_k[t] - 4;
    return (uint32_t)byte_of(s.hs.v0, kr) | ((uint32_t)byte_of(s.hs.v1, kf) << 8);
}

// One DFS stack level: the move d taken from it, the passing children R
// still to visit (always moves after d), and for the PDB search the values
// the move overwrote, so that undoing it needs no table reads.
template <int HEUR> struct StackEntry;
// Manhattan: bits 0-1 the move, 2-5 the children still to visit, 6 whether
// the move brought h down (move_md's result)
template <> struct StackEntry<0> {
    using T = uint8_t;
    static __device__ __forceinline__ T make(int d, uint32_t R, uint32_t v) {
        return (T)((v << 6) | (R << 2) | (uint32_t)d);
    }
    static __device__ __forceinline__ int dir(uint32_t e) { return (int)(e & 3u); }
    static __device__ __forceinline__ uint32_t rest(uint32_t e) { return (e >> 2) & 15u; }
    static __device__ __forceinline__ uint32_t vals(uint32_t e) { return e >> 6; }
};
template <> struct StackEntry<1> {
    using T = uint16_t;
    // bits 0-3: code of (d, R) = base(d) + (R >> (d+1)), base = 0, 8, 12, 14
    // bits 4-9 / 10-15: regular / reflected value before the move (< 64)
    static __device__ __forceinline__ uint32_t base(int d) { return d == 0 ? 0u : d == 1 ? 8u : d == 2 ? 12u : 14u; }
    static __device__ __forceinline__ T make(int d, uint32_t R, uint32_t v) {
        uint32_t code = base(d) + (R >> (d + 1));
        return (T)(code | ((v & 63u) << 4) | (((v >> 8) & 63u) << 10));
    }
    static __device__ __forceinline__ int dir(uint32_t e) {
        uint32_t c = e & 15u;
        return c >= 14u ? 3 : (c >= 12u ? 2 : (c >= 8u ? 1 : 0));
    }
    static __device__ __forceinline__ uint32_t rest(uint32_t e) {
        int d = dir(e);
        return ((e & 15u) - base(d)) << (d + 1);
    }
    static __device__ __forceinline__ uint32_t vals(uint32_t e) { return ((e >> 4) & 63u) | (((e >> 10) & 63u) << 8); }
};

template <int W, int HEUR, class BD>
__device__ __forceinline__ void init_heur(Pos<W, HEUR, BD> &s, const PdbCtx &P, uint32_t meta) {
    if constexpr (HEUR == 0) {
        s.hs.h = meta_h(meta);
    } else {
        pdb_init(P, s.hs, s.bd);
    }
}

} // namespace spz
