// spz_api.hip -- host implementation of the libspz C ABI (include/spz.h).
#include "../../include/spz.h"
#include "spz_board.h"
#include "spz_dedup.h"
#include "spz_mlp.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace spz {
hipError_t launch_round(int width, int heur, const KParams &p, const PdbGeo *geo, int wg, hipStream_t s,
                        hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
size_t round_lds_bytes(int width, int maxd, int wg, int heur);
const void *round_kernel_ptr(int width, int heur);
hipError_t launch_restart(int width, int heur, const KParams &p, const PdbGeo *geo, hipStream_t s);
hipError_t launch_spec(const KParams &p, hipStream_t s);
hipError_t launch_control(const KParams &p, hipStream_t s);
hipError_t launch_split_level(int width, int heur, const KParams &p, const PdbGeo *geo, const Item *lvl_in,
                              const unsigned *n_in, Item *lvl_out, unsigned *n_out, unsigned *nfront,
                              unsigned *topgen, int *top_goal, int *frozen, int depth, int max_depth,
                              unsigned target, const unsigned long long *lp_in, unsigned long long *lp_out,
                              unsigned long long *fp_out, hipStream_t s);
hipError_t launch_eval_h(int width, int heur, const KParams &p, const PdbGeo *geo, Item *items, int n,
                         hipStream_t s);
hipError_t launch_scatter_roots(const Item *roots, Item *lvl, int capl, int n, hipStream_t s);
hipError_t launch_eval_h_bytes(int width, int heur, const uint8_t *bytes, int n, const PdbGeo *geo,
                               const uint8_t *pdb, int *h, hipStream_t s);
hipError_t pdb_bfs_build(int pat, uint8_t *compact, uint32_t *work, unsigned long long *d_cnt, hipStream_t s,
                         int *levels_out);
hipError_t pdb_spread(const uint8_t *compact, uint8_t *direct, hipStream_t s);
hipError_t launch_gather_probe(const uint8_t *t, uint64_t span, int blocks, int per_thread, uint64_t seed, uint32_t *out,
                               hipStream_t s);
constexpr uint64_t kPlacementsHost = 127512000ull;
uint32_t pdb_goal_rank(int pattern);
} // namespace spz

using namespace spz;

namespace {

// device buffer that grows on demand; freed when it goes out of scope
struct DBuf {
    void *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p); // the old contents are not needed; a failed free leaks at worst
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T> T *as() const { return (T *)p; }
};

// pinned host buffer that grows on demand; freed when it goes out of scope
struct HBuf {
    void *p = nullptr;
    size_t n = 0;
    HBuf() = default;
    HBuf(const HBuf &) = delete;
    HBuf &operator=(const HBuf &) = delete;
    ~HBuf() { release(); }
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

} // namespace

// Scratch of spz_eval_h_noalloc: the boards (kNoAllocMaxBoards x 25 bytes),
// then their heuristic values (kNoAllocMaxBoards int32): 1.9 MB on every GPU
// the library runs on, held until the process ends.  Host code never uses
// the array's name as an address: spz_create looks up each device's own
// copy once (scratch_for_device) and keeps that pointer.
constexpr int kNoAllocMaxBoards = 65536;
constexpr size_t kNoAllocHOff = (size_t)kNoAllocMaxBoards * 25;
static __device__ __align__(256) uint8_t scratch[kNoAllocHOff + (size_t)kNoAllocMaxBoards * 4];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// the device's copy of `scratch`; the device must be current
static hipError_t scratch_for_device(int dev, void **out) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

struct spz_ctx {
    spz_config cfg;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    int n_lanes = 0;
    int wg = 256;
    double wall_ticks_per_us = 0; // device wall clock (wall_clock64) rate
    std::string err;
    double stats[SPZ_STATS_LEN] = {0};
    DBuf prof;             // lane-state counters (env SPZ_PROFILE=1)
    DBuf wtime;            // per-wave clocks of one round (env SPZ_WAVE_TIMES)
    double timeout_s = 0;

    PdbGeo geo_h;
    DBuf geo_d;
    DBuf pdb_compact[2]; // file-order tables
    DBuf pdb_direct;     // both tables in 25^6 addressing, end to end (the search's)
    bool have_pdb = false;

    // lane contexts
    DBuf c_st, c_rows, c_root, c_meta, c_rmeta, c_inst, c_cnt, c_pdb, c_stack;
    // queue and control
    DBuf q, q_ctl, ctl;
    size_t qcap = 0;
    // per instance
    DBuf roots, L, it, cnt, pend, found, done, opt, iter_counts, restart_list;
    DBuf itn, drained, dead, nlive, Lbase, listed, want_next; // iteration slots (lookahead)
    DBuf frontier, nfront, topgen, top_goal, frozen, lvl[2], nlvl[2];
    // load balancing
    DBuf c_origin, load, samples, nsamp, pred;
    // solution paths (record_paths)
    DBuf qpath, c_ipath, fpath, lpath[2], sol, sol_len;
    int paths_n = 0;              // instances of the last solve with recorded paths
    std::vector<int> paths_slot; // their final iteration's slot
    MlpState mlp;
    bool mlp_alloced = false;
    HBuf dd_items, dd_paths; // staging of the duplicate-pruning pass
    uint8_t *scratch_d = nullptr; // this device's copy of `scratch` (spz_eval_h_noalloc)
};

static int fail(spz_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

// No exception may leave a C entry point: a host-side failure (out of host
// memory, a thread that cannot start) returns SPZ_E_HOST with its text.
static int host_failure(spz_ctx *c, const char *what) noexcept {
    if (c) {
        try {
            c->err = std::string("host: ") + what;
        } catch (...) {
        }
    }
    return SPZ_E_HOST;
}
#define SPZ_CATCH(c)                                                                                       \
    catch (const std::exception &ex) {                                                                     \
        return host_failure((c), ex.what());                                                              \
    }                                                                                                      \
    catch (...) {                                                                                          \
        return host_failure((c), "unknown exception");                                                    \
    }

#define HIPCHK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(ctx, SPZ_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                             \
    } while (0)

static bool is_solvable(int w, const uint8_t *b) {
    int n2 = w * w;
    bool seen[25] = {false};
    int inv = 0, blank = -1;
    for (int i = 0; i < n2; ++i) {
        if (b[i] >= n2 || seen[b[i]]) return false;
        seen[b[i]] = true;
        if (b[i] == 0) blank = i;
        for (int j = i + 1; j < n2; ++j)
            if (b[i] > b[j]) ++inv;
    }
    // every move flips both the permutation parity and the blank's row+col parity
    return (inv & 1) == ((blank / w + blank % w) & 1);
}

static void pack_root(int w, const uint8_t *b, int inst, Item *out) {
    int blank = 0;
    for (int c = 0; c < w * w; ++c)
        if (b[c] == 0) blank = c;
    if (w == 4) {
        Board<4> bd;
        board_from_bytes<4>(b, bd);
        bd.pack(out->b);
    } else {
        Board<5> bd;
        board_from_bytes<5>(b, bd);
        bd.pack(out->b);
    }
    out->meta = meta_pack(blank, 0, 0, DIR_NONE);
    out->inst = (uint32_t)inst;
    out->x0 = out->x1 = 0;
}

extern "C" {

void spz_default_config(spz_config *cfg, int32_t width, int32_t heuristic) {
    memset(cfg, 0, sizeof(*cfg));
    cfg->width = width;
    cfg->heuristic = heuristic;
    cfg->mode = SPZ_MODE_ALL;
    cfg->frontier_target = 2048;
    cfg->max_split_depth = 24;
    cfg->workgroup_size = 256;
    cfg->lanes = 0;
    cfg->round_budget = 0;
    cfg->lds_stack_depth = 0;
    cfg->lb_policy = SPZ_LB_UNIFORM;
    cfg->device = 0;
    cfg->time_kernels = 0;
    cfg->log_samples = 0;
    cfg->use_graph = 1;
    cfg->queue_items = 0;
    cfg->lookahead = 1;
}

// Steps per lane and round when the configuration leaves it open: shorter
// rounds refill idle lanes sooner, longer ones spend less on the context save
// and restore.  `full`: the solve has work for every lane of the context (its
// automatic lane count reached the capacity).  Measured on MI355X: 1024 beats
// 256 and 512 for the PDB search, where 2048 is level with it; the Manhattan
// search gains from longer rounds only when the batch fills every lane (4096
// mixed 15-puzzles: 2048 steps / 3.5 ms 829 instances/s, 1024 / 2.5 ms 773),
// and loses with one instance on a part of them (44.8 -> 39.6).
static int default_round_budget(int heur, bool full) { return (heur != SPZ_H_PDB666 && full) ? 2048 : 1024; }

// Time limit of a round when the configuration leaves it open.  With a step
// budget alone, waves that step faster (fewer lanes on divergent paths) end
// early and wait for the slowest -- in the 24-puzzle PDB search under load,
// waves end 4.75 ms (10th percentile) to 5.5 ms (last) into a 1024-step
// round.  The limit cuts that tail; lightly loaded rounds, which step faster,
// still end at their step budget (measured on MI355X: PDB 4500 us beat 3200,
// 4000, 4800 and 5600; Manhattan batches with work for every lane, 3500 us
// beat 2000 and 3000 at 2048 steps).  A solve with fewer items than lanes
// (one random 15-puzzle per call, 2048 frontier nodes on 65536 lanes) keeps
// lanes fed only by the donations made at each round's end, so its rounds
// are short: 700 us, 48.1 instances/s against 47.9 at 500, 47.3 at 1000 and
// 45.9 at 2500 (1024 steps).
static int default_round_time_us(int heur, bool full) {
    return heur == SPZ_H_PDB666 ? 4500 : (full ? 3500 : 700);
}

// Lanes per compute unit of the 24-puzzle searches: more resident lanes than
// this lose more to L2 sharing and finer work splitting than they gain in
// latency hiding (measured on MI355X: 768 beat 1024 and 512 for both
// heuristics).
constexpr int kLanesPerCu5 = 768;
// Lanes per initial work item (root-split frontier node) when the lane count
// is automatic: a batch with little work runs on fewer lanes, which split it
// less finely (a single 15-puzzle instance, 2048 frontier nodes: 65536 lanes
// ran 2.4x as fast as 524288, and faster than 32768 or 131072).
constexpr size_t kLanesPerItem = 32;
// step budget of the rounds that run while the queue holds little work
constexpr int kShortRoundSteps = 128;

// Lane capacity (from occupancy unless given) and lane-context arrays for
// the context's workgroup size and LDS stack depth.
static int configure_lanes(spz_ctx *ctx) {
    const spz_config *cfg = &ctx->cfg;
    const int wg = ctx->wg;
    hipError_t e;
    int lanes = cfg->lanes;
    if (lanes <= 0) {
        int blocks = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, round_kernel_ptr(cfg->width, cfg->heuristic), wg,
                                                         round_lds_bytes(cfg->width, cfg->lds_stack_depth, wg, cfg->heuristic));
        if (e != hipSuccess || blocks <= 0) blocks = 2;
        if (cfg->width == 5) blocks = std::max(1, std::min(blocks, kLanesPerCu5 / wg));
        lanes = ctx->n_cu * blocks * wg;
    }
    lanes = (lanes + wg - 1) / wg * wg;
    ctx->n_lanes = lanes;
    size_t L = (size_t)lanes;
    int maxd = cfg->lds_stack_depth;
    if ((e = ctx->c_st.ensure(L * 4)) != hipSuccess || (e = ctx->c_rows.ensure(L * 4 * 5)) != hipSuccess ||
        (e = ctx->c_root.ensure(L * 4 * 4)) != hipSuccess || (e = ctx->c_meta.ensure(L * 4)) != hipSuccess ||
        (e = ctx->c_rmeta.ensure(L * 4)) != hipSuccess || (e = ctx->c_inst.ensure(L * 4)) != hipSuccess ||
        (e = ctx->c_cnt.ensure(L * 4)) != hipSuccess || (e = ctx->c_pdb.ensure(L * 4 * 10)) != hipSuccess ||
        (e = ctx->c_origin.ensure(L * 4)) != hipSuccess || (e = ctx->c_stack.ensure(L * (size_t)maxd * 2)) != hipSuccess)
        return fail(ctx, SPZ_E_HIP, "lane context allocation: %s", hipGetErrorString(e));
    return SPZ_OK;
}

static_assert(kMaxIters == SPZ_MAX_ITERS, "the header's iteration limit is the library's");

// Largest root-split target: a level holds at most 4 x target items per
// instance, so this keeps that capacity far inside 32-bit counts and the
// frontier buffer at 128 MB per instance.
constexpr int kMaxFrontierTarget = 1 << 20;

static int normalise_config(const spz_config *in, spz_config *c, int *wg) {
    *c = *in;
    if (c->mode != SPZ_MODE_ALL && c->mode != SPZ_MODE_FIRST) return SPZ_E_ARG;
    if (c->frontier_target > kMaxFrontierTarget) return SPZ_E_ARG;
    // 2^30 queue items (32 GiB a ring) keep the buffer sizes far from
    // wrapping; 2^24 lanes also keep the search kernel's int offsets into
    // the lane contexts (at most 10 x lanes) in range
    if (c->queue_items > (1ll << 30) || c->lanes > (1 << 24)) return SPZ_E_ARG;
    // a round's step budget bounds one launch when its time limit is off
    if (c->round_budget > (1 << 20)) return SPZ_E_ARG;
    *wg = c->workgroup_size ? c->workgroup_size : 256;
    if (*wg != 64 && *wg != 128 && *wg != 256) return SPZ_E_ARG;
    c->workgroup_size = *wg;
    if (c->frontier_target < 1) c->frontier_target = 1;
    if (c->max_split_depth <= 0) c->max_split_depth = 24;
    if (c->lds_stack_depth <= 0)
        c->lds_stack_depth = c->width == 4 ? Geo<4>::DEFD : (c->heuristic == SPZ_H_PDB666 ? kDefDepthPdb : Geo<5>::DEFD);
    if (c->lds_stack_depth < 8 || c->lds_stack_depth > Geo<5>::MAXD) return SPZ_E_ARG;
    if (c->lb_policy < SPZ_LB_UNIFORM || c->lb_policy > SPZ_LB_PREVLOAD) return SPZ_E_ARG;
    if (c->resplit < 0 || c->resplit > 1024) return SPZ_E_ARG;
    return SPZ_OK;
}

int spz_create(spz_ctx **out, const spz_config *cfg) try {
    spz_ctx *ctx = nullptr;
    if (!out || !cfg) return SPZ_E_ARG;
    *out = nullptr;
    if (cfg->width != 4 && cfg->width != 5) return SPZ_E_ARG;
    if (cfg->heuristic == SPZ_H_PDB666 && cfg->width != 5) return SPZ_E_ARG;
    if (cfg->heuristic != SPZ_H_MANHATTAN && cfg->heuristic != SPZ_H_PDB666) return SPZ_E_ARG;
    spz_config ncfg;
    int wg = 256;
    if (normalise_config(cfg, &ncfg, &wg) != SPZ_OK) return SPZ_E_ARG;
    ctx = new spz_ctx();
    ctx->cfg = ncfg;
    ctx->wg = wg;
    {
        const char *t = getenv("SPZ_TIMEOUT_S");
        ctx->timeout_s = t ? atof(t) : 900.0;
    }
    int rc = SPZ_OK;
    do {
        hipError_t e = hipSetDevice(cfg->device);
        if (e != hipSuccess) {
            rc = fail(ctx, SPZ_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
            break;
        }
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, cfg->device);
        if (e != hipSuccess) {
            rc = fail(ctx, SPZ_E_HIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
            break;
        }
        ctx->n_cu = prop.multiProcessorCount;
        {
            int khz = 0;
            if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, cfg->device) == hipSuccess && khz > 0)
                ctx->wall_ticks_per_us = khz / 1000.0;
            else
                (void)hipGetLastError();
        }
        e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            rc = fail(ctx, SPZ_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
            break;
        }
        {
            void *sp = nullptr;
            e = scratch_for_device(cfg->device, &sp);
            if (e != hipSuccess) {
                rc = fail(ctx, SPZ_E_HIP, "scratch address: %s", hipGetErrorString(e));
                break;
            }
            ctx->scratch_d = (uint8_t *)sp;
        }
        pdb_geo_build(&ctx->geo_h);
        if ((e = ctx->geo_d.ensure(sizeof(PdbGeo))) != hipSuccess ||
            (e = hipMemcpy(ctx->geo_d.p, &ctx->geo_h, sizeof(PdbGeo), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = ctx->q_ctl.ensure(2 * kQCtl * 8)) != hipSuccess || (e = ctx->ctl.ensure(64 * 4)) != hipSuccess) {
            rc = fail(ctx, SPZ_E_HIP, "context allocation: %s", hipGetErrorString(e));
            break;
        }
        rc = configure_lanes(ctx);
    } while (0);
    if (rc != SPZ_OK) {
        // keep the context so the caller can read the error, but report failure
        *out = ctx;
        return rc;
    }
    *out = ctx;
    return SPZ_OK;
} SPZ_CATCH(nullptr)

int spz_reconfigure(spz_ctx *ctx, const spz_config *cfg) try {
    if (!ctx || !cfg) return SPZ_E_ARG;
    if (cfg->width != ctx->cfg.width || cfg->heuristic != ctx->cfg.heuristic || cfg->device != ctx->cfg.device)
        return fail(ctx, SPZ_E_ARG, "width, heuristic and device are fixed for a context");
    spz_config ncfg;
    int wg = 256;
    if (normalise_config(cfg, &ncfg, &wg) != SPZ_OK) return fail(ctx, SPZ_E_ARG, "invalid configuration");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    ctx->cfg = ncfg;
    ctx->wg = wg;
    return configure_lanes(ctx);
} SPZ_CATCH(ctx)

void spz_get_config(spz_ctx *ctx, spz_config *cfg) {
    if (ctx && cfg) *cfg = ctx->cfg;
}

// Coordinate descent over workgroup size, root-split target, LDS stack depth,
// round budget, lookahead on/off, the round time limit on/off and (under a
// policy with estimates) the heavy-node re-split factor on a sample batch;
// every candidate must give the same result.  The context is left configured
// with the winner.
//
// What "the same result" is depends on the configuration:
//  - lengths and thresholds always;
//  - all-solutions mode: every per-iteration count;
//  - first-solution mode: every count but the last iteration's, which stops
//    at a goal and so depends on the schedule;
//  - duplicate pruning: the pruned counts depend on the root-split depth, so
//    the frontier target is not varied (the other dimensions keep counts).
int spz_autotune(spz_ctx *ctx, const uint8_t *boards, int32_t n, int32_t max_iters, int32_t passes,
                 spz_config *best, double *best_ms, int32_t *tried) try {
    if (!ctx || !boards || n <= 0 || !best) return SPZ_E_ARG;
    if (max_iters < 1 || max_iters > kMaxIters) return fail(ctx, SPZ_E_ARG, "max_iters must be in 1..%d", kMaxIters);
    if (passes <= 0) passes = 1;
    std::vector<spz_result> ref((size_t)n), cur((size_t)n);
    std::vector<uint64_t> ic_ref((size_t)n * max_iters), ic((size_t)n * max_iters);
    spz_config base = ctx->cfg;
    const int T = base.frontier_target;
    const int wgs[3] = {64, 128, 256};
    const int tgts[5] = {std::max(1, T / 16), std::max(1, T / 4), T, std::min(T * 2, kMaxFrontierTarget),
                         std::min(T * 4, kMaxFrontierTarget)};
    const int deps_w4[4] = {48, 64, 80, 112}, deps_w5[4] = {64, 80, 96, 128};
    const int *deps = base.width == 4 ? deps_w4 : deps_w5;
    const int buds[4] = {256, 512, 1024, 2048};
    const int splits[4] = {0, 1, 2, 4};
    const bool first = base.mode == SPZ_MODE_FIRST;
    auto run = [&](const spz_config &c, double *ms, bool check) -> int {
        int rc = spz_reconfigure(ctx, &c);
        if (rc) return rc;
        auto t = std::chrono::steady_clock::now();
        rc = spz_solve_batch(ctx, boards, n, max_iters, check ? cur.data() : ref.data(),
                             check ? ic.data() : ic_ref.data());
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
        if (rc) return rc;
        if (check) {
            for (int i = 0; i < n; ++i) {
                const spz_result &a = cur[i], &b = ref[i];
                if (a.opt_len != b.opt_len || a.n_iters != b.n_iters || a.h0 != b.h0)
                    return fail(ctx, SPZ_E_MISMATCH, "autotune: candidate changed the length of board %d", i);
                // per-iteration counts; in first-solution mode the final one
                // is partial and schedule-dependent
                const int cmp = first ? a.n_iters - 1 : max_iters;
                for (int k = 0; k < cmp; ++k)
                    if (ic[(size_t)i * max_iters + k] != ic_ref[(size_t)i * max_iters + k])
                        return fail(ctx, SPZ_E_MISMATCH,
                                    "autotune: candidate changed count %d of board %d", k, i);
            }
        }
        return SPZ_OK;
    };
    spz_config bc = base;
    if (bc.round_budget <= 0)
        bc.round_budget = default_round_budget(bc.heuristic, kLanesPerItem * (size_t)n * (size_t)std::max(1, bc.frontier_target) >=
                                                                 (size_t)ctx->n_lanes);
    double bms = 0;
    int count = 0;
    int rc = run(bc, &bms, false); // reference result (also warms the buffers)
    if (rc) return rc;
    rc = run(bc, &bms, true);
    if (rc) return rc;
    count = 2;
    for (int pass = 0; pass < passes; ++pass) {
        for (int dim = 0; dim < 7; ++dim) {
            if (dim == 1 && base.dedup) continue;
            if (dim == 6 && base.lb_policy == SPZ_LB_UNIFORM) continue;
            int nv = dim == 0 ? 3 : dim == 1 ? 5 : (dim == 4 || dim == 5) ? 2 : 4;
            spz_config winner = bc;
            for (int v = 0; v < nv; ++v) {
                spz_config c = bc;
                if (dim == 0) c.workgroup_size = wgs[v];
                if (dim == 1) c.frontier_target = tgts[v];
                if (dim == 2) c.lds_stack_depth = deps[v];
                if (dim == 3) c.round_budget = buds[v];
                if (dim == 4) c.lookahead = v;
                if (dim == 5) c.round_time_us = v ? -1 : 0; // time limit (auto) or step budget only
                if (dim == 6) c.resplit = splits[v];
                if (memcmp(&c, &bc, sizeof c) == 0) continue;
                c.lanes = 0;
                double ms = 0;
                rc = run(c, &ms, true);
                ++count;
                if (rc) return rc;
                if (ms < bms) {
                    bms = ms;
                    winner = c;
                }
            }
            bc = winner;
        }
    }
    rc = spz_reconfigure(ctx, &bc);
    if (rc) return rc;
    *best = ctx->cfg;
    if (best_ms) *best_ms = bms;
    if (tried) *tried = count;
    return SPZ_OK;
} SPZ_CATCH(ctx)

const char *spz_last_error(spz_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void spz_destroy(spz_ctx *ctx) {
    if (!ctx) return;
    DBuf *all[] = {&ctx->geo_d,       &ctx->pdb_compact[0], &ctx->pdb_compact[1], &ctx->pdb_direct,
                   &ctx->c_st,         &ctx->c_rows,         &ctx->c_root,
                   &ctx->c_meta,      &ctx->c_rmeta,        &ctx->c_inst,         &ctx->c_cnt,
                   &ctx->c_pdb,       &ctx->c_stack,        &ctx->q,              &ctx->q_ctl,
                   &ctx->ctl,         &ctx->roots,          &ctx->L,              &ctx->it,
                   &ctx->cnt,         &ctx->pend,           &ctx->found,          &ctx->done,
                   &ctx->opt,         &ctx->iter_counts,    &ctx->restart_list,   &ctx->frontier,
                   &ctx->nfront,      &ctx->topgen,         &ctx->top_goal,       &ctx->frozen,
                   &ctx->lvl[0],      &ctx->lvl[1],         &ctx->nlvl[0],        &ctx->nlvl[1],
                   &ctx->c_origin,    &ctx->load,           &ctx->samples,        &ctx->nsamp,
                   &ctx->pred,        &ctx->qpath,          &ctx->c_ipath,        &ctx->fpath,
                   &ctx->lpath[0],    &ctx->lpath[1],       &ctx->sol,            &ctx->sol_len,
                   &ctx->itn,         &ctx->drained,        &ctx->nlive,          &ctx->Lbase,
                   &ctx->listed,      &ctx->want_next,      &ctx->dead,           &ctx->prof,
                   &ctx->wtime};
    for (DBuf *b : all) b->release();
    ctx->dd_items.release();
    ctx->dd_paths.release();
    if (ctx->mlp_alloced) mlp_free(ctx->mlp);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int spz_last_stats(spz_ctx *ctx, double *stats) try {
    if (!ctx || !stats) return SPZ_E_ARG;
    memcpy(stats, ctx->stats, sizeof ctx->stats);
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_set_lb(spz_ctx *ctx, int32_t lb_policy) try {
    if (!ctx) return SPZ_E_ARG;
    if (lb_policy != SPZ_LB_UNIFORM && lb_policy != SPZ_LB_LEARNED && lb_policy != SPZ_LB_PREVLOAD)
        return fail(ctx, SPZ_E_ARG, "unknown lb policy %d", lb_policy);
    ctx->cfg.lb_policy = lb_policy;
    return SPZ_OK;
} SPZ_CATCH(ctx)

// ---------------------------------------------------------------- PDB ---
static int upload_direct(spz_ctx *ctx) {
    HIPCHK(ctx->pdb_direct.ensure(2 * kPdbDirect));
    for (int t = 0; t < 2; ++t)
        HIPCHK(pdb_spread(ctx->pdb_compact[t].as<uint8_t>(), ctx->pdb_direct.as<uint8_t>() + t * kPdbDirect,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_pdb = true;
    return SPZ_OK;
}

// PDB cache file: a 64-byte header, then the two canonical tables (file
// order).  Header: magic "SPZPDB01", u64 bytes per table, u64 table count
// (2), u64 checksum of each table, zero padding.  A file that is not exactly
// that size, or whose magic, sizes or checksums do not match, is rejected.
constexpr char kCacheMagic[8] = {'S', 'P', 'Z', 'P', 'D', 'B', '0', '1'};
constexpr size_t kCacheHeader = 64;

// checksum of a table (64-bit multiply-xor over 8-byte words; detects
// truncation and corruption, not tampering)
static uint64_t table_checksum(const uint8_t *p, size_t n) {
    uint64_t h = 0x9e3779b97f4a7c15ull ^ n;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, p + i, 8);
        h = (h ^ w) * 0xff51afd7ed558ccdull;
        h ^= h >> 29;
    }
    for (; i < n; ++i) h = (h ^ p[i]) * 0xc4ceb9fe1a85ec53ull;
    return h ^ (h >> 32);
}

// the two canonical tables from host memory to the device, then the direct
// tables the search reads
static int upload_compact(spz_ctx *ctx, const uint8_t *t0, const uint8_t *t1) {
    const uint8_t *t[2] = {t0, t1};
    for (int k = 0; k < 2; ++k)
        HIPCHK(hipMemcpy(ctx->pdb_compact[k].p, t[k], kPlacementsHost, hipMemcpyHostToDevice));
    return upload_direct(ctx);
}

static int read_cache(spz_ctx *ctx, const char *path, FILE *f) {
    std::vector<uint8_t> buf(2 * kPlacementsHost);
    uint8_t hdr[kCacheHeader];
    const size_t got = fread(hdr, 1, kCacheHeader, f);
    const size_t got2 = got == kCacheHeader ? fread(buf.data(), 1, buf.size(), f) : 0;
    const bool at_end = got2 == buf.size() && fgetc(f) == EOF;
    fclose(f);
    uint64_t v[4];
    memcpy(v, hdr + 8, sizeof v);
    if (got != kCacheHeader || memcmp(hdr, kCacheMagic, 8) != 0 || v[0] != kPlacementsHost || v[1] != 2)
        return fail(ctx, SPZ_E_IO, "%s is not a PDB cache file of this library", path);
    if (!at_end) return fail(ctx, SPZ_E_IO, "PDB cache %s has the wrong size", path);
    for (int k = 0; k < 2; ++k)
        if (table_checksum(buf.data() + k * kPlacementsHost, kPlacementsHost) != v[2 + k])
            return fail(ctx, SPZ_E_IO, "PDB cache %s: table %d fails its checksum", path, k);
    return upload_compact(ctx, buf.data(), buf.data() + kPlacementsHost);
}

// written to a temporary file beside the target, then renamed over it, so a
// reader never sees a partial cache
static int write_cache(spz_ctx *ctx, const char *path) {
    std::vector<uint8_t> buf(2 * kPlacementsHost);
    for (int k = 0; k < 2; ++k)
        HIPCHK(hipMemcpy(buf.data() + k * kPlacementsHost, ctx->pdb_compact[k].p, kPlacementsHost,
                         hipMemcpyDeviceToHost));
    uint8_t hdr[kCacheHeader];
    memset(hdr, 0, sizeof hdr);
    memcpy(hdr, kCacheMagic, 8);
    const uint64_t v[4] = {kPlacementsHost, 2, table_checksum(buf.data(), kPlacementsHost),
                           table_checksum(buf.data() + kPlacementsHost, kPlacementsHost)};
    memcpy(hdr + 8, v, sizeof v);
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return fail(ctx, SPZ_E_IO, "cannot write PDB cache %s", tmp.c_str());
    const bool ok = fwrite(hdr, 1, sizeof hdr, f) == sizeof hdr && fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    const bool closed = fclose(f) == 0;
    if (!ok || !closed) {
        remove(tmp.c_str());
        return fail(ctx, SPZ_E_IO, "short write to %s", tmp.c_str());
    }
    if (rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return fail(ctx, SPZ_E_IO, "cannot rename %s to %s", tmp.c_str(), path);
    }
    return SPZ_OK;
}

int spz_build_pdb(spz_ctx *ctx, const char *cache_path) try {
    if (!ctx) return SPZ_E_ARG;
    if (ctx->cfg.width != 5) return fail(ctx, SPZ_E_ARG, "PDBs exist for the 24-puzzle only");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    for (int t = 0; t < 2; ++t) HIPCHK(ctx->pdb_compact[t].ensure(kPlacementsHost));
    if (cache_path) {
        FILE *f = fopen(cache_path, "rb");
        if (f) return read_cache(ctx, cache_path, f);
    }
    DBuf work, cntb;
    hipError_t e = work.ensure(3ull * kPlacementsHost * 4);
    if (e == hipSuccess) e = cntb.ensure(8);
    if (e != hipSuccess) {
        work.release();
        cntb.release();
        return fail(ctx, SPZ_E_HIP, "PDB scratch allocation: %s", hipGetErrorString(e));
    }
    auto t0 = std::chrono::steady_clock::now();
    for (int t = 0; t < 2; ++t) {
        int levels = 0;
        e = pdb_bfs_build(t, ctx->pdb_compact[t].as<uint8_t>(), work.as<uint32_t>(),
                          cntb.as<unsigned long long>(), ctx->stream, &levels);
        if (e != hipSuccess) break;
        ctx->stats[4 + t] = levels;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->stats[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    work.release();
    cntb.release();
    if (e != hipSuccess) return fail(ctx, SPZ_E_HIP, "PDB BFS: %s", hipGetErrorString(e));
    int rc = upload_direct(ctx);
    if (rc != SPZ_OK) return rc;
    if (cache_path) return write_cache(ctx, cache_path);
    return SPZ_OK;
} SPZ_CATCH(ctx)

// The reference's table files (src/24puzzle_with_pdb.cc:77-112): each is
// exactly 127,512,000 bytes, one byte per placement of the pattern's six
// tiles in nested-loop order.  Besides the size, the entry of the goal
// placement (every pattern tile home) must be 0.
int spz_load_pdb_tab(spz_ctx *ctx, const char *path_1256712, const char *path_34891314) try {
    if (!ctx || !path_1256712 || !path_34891314) return SPZ_E_ARG;
    if (ctx->cfg.width != 5) return fail(ctx, SPZ_E_ARG, "PDBs exist for the 24-puzzle only");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    for (int t = 0; t < 2; ++t) HIPCHK(ctx->pdb_compact[t].ensure(kPlacementsHost));
    std::vector<uint8_t> buf(2 * kPlacementsHost);
    const char *paths[2] = {path_1256712, path_34891314};
    for (int t = 0; t < 2; ++t) {
        FILE *f = fopen(paths[t], "rb");
        if (!f) return fail(ctx, SPZ_E_IO, "cannot open %s", paths[t]);
        const size_t got = fread(buf.data() + t * kPlacementsHost, 1, kPlacementsHost, f);
        const bool at_end = got == kPlacementsHost && fgetc(f) == EOF;
        fclose(f);
        if (!at_end) return fail(ctx, SPZ_E_IO, "%s is not %llu bytes", paths[t], (unsigned long long)kPlacementsHost);
        const uint32_t goal = pdb_goal_rank(t);
        if (buf[t * kPlacementsHost + goal] != 0)
            return fail(ctx, SPZ_E_IO, "%s: the goal placement's entry is not 0", paths[t]);
    }
    ctx->have_pdb = false;
    return upload_compact(ctx, buf.data(), buf.data() + kPlacementsHost);
} SPZ_CATCH(ctx)

int spz_get_pdb(spz_ctx *ctx, int32_t pattern, uint8_t *out) try {
    if (!ctx || !out || pattern < 0 || pattern > 1) return SPZ_E_ARG;
    if (!ctx->have_pdb) return fail(ctx, SPZ_E_NOPDB, "no PDB built");
    HIPCHK(hipMemcpy(out, ctx->pdb_compact[pattern].p, kPlacementsHost, hipMemcpyDeviceToHost));
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_export_pdb_tab(spz_ctx *ctx, const char *p0, const char *p1) try {
    if (!ctx || !p0 || !p1) return SPZ_E_ARG;
    if (!ctx->have_pdb) return fail(ctx, SPZ_E_NOPDB, "no PDB built");
    std::vector<uint8_t> buf(kPlacementsHost);
    const char *paths[2] = {p0, p1};
    for (int t = 0; t < 2; ++t) {
        HIPCHK(hipMemcpy(buf.data(), ctx->pdb_compact[t].p, buf.size(), hipMemcpyDeviceToHost));
        FILE *f = fopen(paths[t], "wb");
        if (!f) return fail(ctx, SPZ_E_IO, "cannot write %s", paths[t]);
        size_t w = fwrite(buf.data(), 1, buf.size(), f);
        fclose(f);
        if (w != buf.size()) return fail(ctx, SPZ_E_IO, "short write to %s", paths[t]);
    }
    return SPZ_OK;
} SPZ_CATCH(ctx)

// -------------------------------------------------------------- solve ---
// Duplicate pruning in the root split (config dedup).  Before a level is
// expanded, a node is dropped when a kept node of its instance's top tree --
// the root, a node of a shallower level, or a copy in the same level --
// holds the same board with a path-max f no larger than its own.  The kept
// copy's g is no larger (levels run in depth order), so below it every path
// of the dropped copy has a counterpart whose every f, and so path-max f, is
// no larger: it is expanded in every iteration in which the dropped copy's
// would be, for any admissible heuristic, consistent or not.  Optimal
// lengths stay; the dropped node's subtree is no longer counted.  Host side,
// between split levels: it changes the set-up only, never the search
// kernels.  The per-instance pass is dedup_level_host (spz_dedup.h).
static BoardKey board_key(int width, const Item &it) {
    return BoardKey{{it.b[0], it.b[1], width == 4 ? 0u : it.b[2], width == 4 ? 0u : it.b[3]}};
}

static hipError_t dedup_level(spz_ctx *ctx, int n, int capl, int width, int cur, bool paths,
                              std::vector<SeenBoards> &seen, hipStream_t s);

static KParams base_params(spz_ctx *ctx, int n, int max_iters, int capl) {
    KParams p;
    memset(&p, 0, sizeof p);
    p.q = ctx->q.as<Item>();
    p.qmask = ctx->qcap - 1;
    p.q_ctl = ctx->q_ctl.as<unsigned long long>();
    p.n_inst = n;
    p.max_iters = max_iters;
    p.mode = ctx->cfg.mode;
    p.roots = ctx->roots.as<Item>();
    p.L = ctx->L.as<int>();
    p.it = ctx->it.as<int>();
    p.itn = ctx->itn.as<int>();
    p.drained = ctx->drained.as<int>();
    p.dead = ctx->dead.as<int>();
    p.nlive = ctx->nlive.as<int>();
    p.Lbase = ctx->Lbase.as<int>();
    p.listed = ctx->listed.as<int>();
    p.want_next = ctx->want_next.as<int>();
    p.cnt = ctx->cnt.as<unsigned long long>();
    p.pend = ctx->pend.as<unsigned>();
    p.found = ctx->found.as<int>();
    p.done = ctx->done.as<int>();
    p.opt = ctx->opt.as<int>();
    p.iter_counts = ctx->iter_counts.as<unsigned long long>();
    p.restart_list = ctx->restart_list.as<int>();
    p.ctl = ctx->ctl.as<unsigned>();
    p.frontier = ctx->frontier.as<Item>();
    p.nfront = ctx->nfront.as<unsigned>();
    p.topgen = ctx->topgen.as<unsigned>();
    p.top_goal = ctx->top_goal.as<int>();
    p.capl = capl;
    p.n_lanes = ctx->n_lanes;
    p.maxd = ctx->cfg.lds_stack_depth;
    p.prof = nullptr; // a solve turns the counters on (env SPZ_PROFILE)
    p.c_st = ctx->c_st.as<uint32_t>();
    p.c_rows = ctx->c_rows.as<uint32_t>();
    p.c_root = ctx->c_root.as<uint32_t>();
    p.c_meta = ctx->c_meta.as<uint32_t>();
    p.c_rmeta = ctx->c_rmeta.as<uint32_t>();
    p.c_inst = ctx->c_inst.as<uint32_t>();
    p.c_cnt = ctx->c_cnt.as<uint32_t>();
    p.c_origin = ctx->c_origin.as<uint32_t>();
    p.load = ctx->load.as<unsigned long long>();
    p.lb = ctx->cfg.lb_policy;
    p.resplit = ctx->cfg.resplit;
    p.pred = ctx->pred.as<float>();
    p.samples = ctx->samples.as<Sample>();
    p.n_samples = ctx->nsamp.as<unsigned>();
    p.max_samples = ctx->cfg.log_samples > 0 && ctx->samples.p ? (unsigned)ctx->cfg.log_samples : 0u;
    p.c_pdb = ctx->c_pdb.as<uint32_t>();
    p.c_stack = ctx->c_stack.as<uint8_t>();
    p.pdb = ctx->pdb_direct.as<uint8_t>();
    return p;
}

static size_t next_pow2(size_t x) {
    size_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

int spz_eval_h(spz_ctx *ctx, const uint8_t *boards, int32_t n, int32_t *h) try {
    if (!ctx || !boards || !h || n < 0) return SPZ_E_ARG;
    if (n == 0) return SPZ_OK;
    const int w = ctx->cfg.width, heur = ctx->cfg.heuristic;
    if (heur == SPZ_H_PDB666 && !ctx->have_pdb) return fail(ctx, SPZ_E_NOPDB, "call spz_build_pdb first");
    for (int i = 0; i < n; ++i)
        if (!is_solvable(w, boards + (size_t)i * w * w))
            return fail(ctx, SPZ_E_ARG, "board %d is not a solvable permutation", i);
    HIPCHK(hipSetDevice(ctx->cfg.device));
    std::vector<Item> items((size_t)n);
    for (int i = 0; i < n; ++i) pack_root(w, boards + (size_t)i * w * w, i, &items[i]);
    DBuf d;
    HIPCHK(d.ensure(sizeof(Item) * (size_t)n));
    KParams p = base_params(ctx, 0, 1, 1);
    hipError_t e = hipMemcpyAsync(d.p, items.data(), sizeof(Item) * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_eval_h(w, heur, p, ctx->geo_d.as<PdbGeo>(), d.as<Item>(), n, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(items.data(), d.p, sizeof(Item) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    d.release();
    if (e != hipSuccess) return fail(ctx, SPZ_E_HIP, "eval_h: %s", hipGetErrorString(e));
    for (int i = 0; i < n; ++i) h[i] = meta_h(items[i].meta);
    return SPZ_OK;
} SPZ_CATCH(ctx)

// spz_eval_h without any allocation: boards and values pass through the
// device's `scratch` array (every call on a GPU shares it; calls that
// overlap on one GPU race), the kernel reads the tables the context holds.
int spz_eval_h_noalloc(spz_ctx *ctx, const uint8_t *boards, int32_t n, int32_t *h) try {
    if (!ctx || !boards || !h || n < 0) return SPZ_E_ARG;
    if (n == 0) return SPZ_OK;
    if (n > kNoAllocMaxBoards)
        return fail(ctx, SPZ_E_ARG, "spz_eval_h_noalloc takes at most %d boards (got %d)", kNoAllocMaxBoards, n);
    if (!ctx->scratch_d) return fail(ctx, SPZ_E_HIP, "no scratch on this device");
    const int w = ctx->cfg.width, heur = ctx->cfg.heuristic, n2 = w * w;
    if (heur == SPZ_H_PDB666 && !ctx->have_pdb) return fail(ctx, SPZ_E_NOPDB, "call spz_build_pdb first");
    for (int i = 0; i < n; ++i)
        if (!is_solvable(w, boards + (size_t)i * n2))
            return fail(ctx, SPZ_E_ARG, "board %d is not a solvable permutation", i);
    HIPCHK(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->stream;
    int *hd = (int *)(ctx->scratch_d + kNoAllocHOff);
    HIPCHK(hipMemcpyAsync(ctx->scratch_d, boards, (size_t)n * n2, hipMemcpyHostToDevice, s));
    HIPCHK(launch_eval_h_bytes(w, heur, ctx->scratch_d, n, ctx->geo_d.as<PdbGeo>(), ctx->pdb_direct.as<uint8_t>(),
                               hd, s));
    HIPCHK(hipMemcpyAsync(h, hd, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_solve_batch(spz_ctx *ctx, const uint8_t *boards, int32_t n, int32_t max_iters, spz_result *out,
                    uint64_t *iter_counts) try {
    if (!ctx || !boards || !out || n < 0) return SPZ_E_ARG;
    if (max_iters < 1 || max_iters > kMaxIters)
        return fail(ctx, SPZ_E_ARG, "max_iters must be in 1..%d", kMaxIters);
    if (n == 0) return SPZ_OK;
    if (n > (1 << 24)) return fail(ctx, SPZ_E_ARG, "batch too large");
    const int w = ctx->cfg.width, heur = ctx->cfg.heuristic;
    if (heur == SPZ_H_PDB666 && !ctx->have_pdb) return fail(ctx, SPZ_E_NOPDB, "call spz_build_pdb first");
    for (int i = 0; i < n; ++i)
        if (!is_solvable(w, boards + (size_t)i * w * w))
            return fail(ctx, SPZ_E_ARG, "board %d is not a solvable permutation", i);
    HIPCHK(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->stream;
    auto t_start = std::chrono::steady_clock::now();

    const unsigned target = (unsigned)ctx->cfg.frontier_target;
    const int capl = (int)std::max<unsigned>(4u, 4u * target);
    const size_t N = (size_t)n;
    const size_t front_items = N * (size_t)capl;
    // lanes of this solve: all of the context's, or (automatic count) as many
    // as the batch's initial work keeps useful, at least one workgroup per CU
    size_t lanes = (size_t)ctx->n_lanes;
    if (ctx->cfg.lanes <= 0) {
        const size_t wg = (size_t)ctx->wg;
        size_t want = kLanesPerItem * N * (size_t)std::max(1u, target);
        want = std::max(want, (size_t)ctx->n_cu * wg);
        want = (want + wg - 1) / wg * wg;
        lanes = std::min(lanes, want);
    }
    size_t qcap = ctx->cfg.queue_items > 0 ? (size_t)ctx->cfg.queue_items
                                           : std::max<size_t>(1u << 22, 4 * front_items + 8 * lanes);
    qcap = next_pow2(qcap);
    ctx->qcap = qcap;

    HIPCHK(ctx->q.ensure(2 * qcap * sizeof(Item))); // ring 0: current iterations, ring 1: lookahead
    HIPCHK(ctx->roots.ensure(N * sizeof(Item)));
    HIPCHK(ctx->L.ensure(2 * N * 4));
    HIPCHK(ctx->itn.ensure(2 * N * 4));
    HIPCHK(ctx->cnt.ensure(2 * N * 8));
    HIPCHK(ctx->pend.ensure(2 * N * 4));
    HIPCHK(ctx->found.ensure(2 * N * 4));
    HIPCHK(ctx->drained.ensure(2 * N * 4));
    HIPCHK(ctx->dead.ensure(2 * N * 4));
    HIPCHK(ctx->it.ensure(N * 4));
    HIPCHK(ctx->nlive.ensure(N * 4));
    HIPCHK(ctx->Lbase.ensure(N * 4));
    HIPCHK(ctx->listed.ensure(N * 4));
    HIPCHK(ctx->want_next.ensure(N * 4));
    HIPCHK(ctx->done.ensure(N * 4));
    HIPCHK(ctx->opt.ensure(N * 4));
    HIPCHK(ctx->iter_counts.ensure(N * (size_t)max_iters * 8));
    HIPCHK(ctx->restart_list.ensure(N * 4));
    HIPCHK(ctx->frontier.ensure(front_items * sizeof(Item)));
    HIPCHK(ctx->nfront.ensure(N * 4));
    HIPCHK(ctx->topgen.ensure(N * kFBins * 4));
    HIPCHK(ctx->top_goal.ensure(N * 4));
    HIPCHK(ctx->frozen.ensure(N * 4));
    HIPCHK(ctx->load.ensure(front_items * 8));
    HIPCHK(ctx->nsamp.ensure(64));
    if (ctx->cfg.log_samples > 0) HIPCHK(ctx->samples.ensure((size_t)ctx->cfg.log_samples * sizeof(Sample)));
    if (ctx->cfg.lb_policy == SPZ_LB_LEARNED) {
        if (!ctx->mlp_alloced || !ctx->mlp.ready)
            return fail(ctx, SPZ_E_ARG, "learned load balancing needs MLP parameters (spz_mlp_set_params/train)");
        HIPCHK(ctx->pred.ensure(front_items * kPredK * 4));
    }
    for (int k = 0; k < 2; ++k) {
        HIPCHK(ctx->lvl[k].ensure(front_items * sizeof(Item)));
        HIPCHK(ctx->nlvl[k].ensure(N * 4));
    }

    KParams p = base_params(ctx, n, max_iters, capl);
    p.n_lanes = (int)lanes;
    // lookahead runs two iterations of an instance at once; the per-node
    // loads (load-based LB, training samples) describe one iteration at a time
    p.lookahead = ctx->cfg.lookahead != 0 && ctx->cfg.lb_policy != SPZ_LB_PREVLOAD && p.max_samples == 0 ? 1 : 0;

    // solution paths: every queued item carries its moves from the root
    const bool paths = ctx->cfg.record_paths != 0;
    ctx->paths_n = 0;
    if (paths) {
        if (ctx->cfg.max_split_depth > 32)
            return fail(ctx, SPZ_E_ARG, "record_paths needs max_split_depth <= 32");
        // the kernels address queued paths in 32-bit words
        if (2 * qcap * (size_t)kPathWords > (1ull << 32))
            return fail(ctx, SPZ_E_ARG, "record_paths needs at most 2^27 queue items a ring (smaller batch)");
        HIPCHK(ctx->qpath.ensure(2 * qcap * kPathWords * 4));
        HIPCHK(ctx->c_ipath.ensure((size_t)kPathWords * lanes * 4));
        HIPCHK(ctx->fpath.ensure(front_items * 8));
        for (int k = 0; k < 2; ++k) HIPCHK(ctx->lpath[k].ensure(front_items * 8));
        HIPCHK(ctx->sol.ensure(2 * N * kPathWords * 4)); // per iteration slot
        HIPCHK(ctx->sol_len.ensure(2 * N * 4));
        HIPCHK(hipMemsetAsync(ctx->lpath[0].p, 0, front_items * 8, s)); // the roots' empty paths
        HIPCHK(hipMemsetAsync(ctx->sol.p, 0, 2 * N * kPathWords * 4, s));
        HIPCHK(hipMemsetAsync(ctx->sol_len.p, 0xff, 2 * N * 4, s)); // -1: none yet
        p.qpath = ctx->qpath.as<uint32_t>();
        p.c_ipath = ctx->c_ipath.as<uint32_t>();
        p.fpath = ctx->fpath.as<unsigned long long>();
        p.sol = ctx->sol.as<uint32_t>();
        p.sol_len = ctx->sol_len.as<int>();
    }

    // roots and their heuristic
    std::vector<Item> roots(N);
    for (int i = 0; i < n; ++i) pack_root(w, boards + (size_t)i * w * w, i, &roots[i]);
    HIPCHK(hipMemcpyAsync(ctx->roots.p, roots.data(), N * sizeof(Item), hipMemcpyHostToDevice, s));
    HIPCHK(launch_eval_h(w, heur, p, ctx->geo_d.as<PdbGeo>(), ctx->roots.as<Item>(), n, s));
    HIPCHK(hipMemcpyAsync(roots.data(), ctx->roots.p, N * sizeof(Item), hipMemcpyDeviceToHost, s));

    // root split: level 0 holds the root of each instance at slot 0
    HIPCHK(hipMemsetAsync(ctx->topgen.p, 0, N * kFBins * 4, s));
    HIPCHK(hipMemsetAsync(ctx->frozen.p, 0, N * 4, s));
    HIPCHK(hipMemsetAsync(ctx->nfront.p, 0, N * 4, s));
    HIPCHK(hipStreamSynchronize(s));
    {
        std::vector<int> tg(N, kNoGoal);
        HIPCHK(hipMemcpyAsync(ctx->top_goal.p, tg.data(), N * 4, hipMemcpyHostToDevice, s));
        std::vector<unsigned> ones(N, 1u);
        HIPCHK(hipMemcpyAsync(ctx->nlvl[0].p, ones.data(), N * 4, hipMemcpyHostToDevice, s));
        HIPCHK(launch_scatter_roots(ctx->roots.as<Item>(), ctx->lvl[0].as<Item>(), capl, n, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    const int max_depth = ctx->cfg.max_split_depth;
    auto t_split = std::chrono::steady_clock::now();
    int cur = 0;
    std::vector<SeenBoards> seen(ctx->cfg.dedup ? N : 0);
    // the closed list starts with the root (src_horie/idas_smem.cu:1068)
    for (size_t i = 0; i < seen.size(); ++i)
        seen[i].emplace_back(board_key(w, roots[i]), meta_h(roots[i].meta));
    for (int depth = 0; depth <= max_depth; ++depth) {
        if (ctx->cfg.dedup && depth > 0) HIPCHK(dedup_level(ctx, n, capl, w, cur, paths, seen, s));
        HIPCHK(hipMemsetAsync(ctx->nlvl[cur ^ 1].p, 0, N * 4, s));
        HIPCHK(launch_split_level(w, heur, p, ctx->geo_d.as<PdbGeo>(), ctx->lvl[cur].as<Item>(),
                                  ctx->nlvl[cur].as<unsigned>(), ctx->lvl[cur ^ 1].as<Item>(),
                                  ctx->nlvl[cur ^ 1].as<unsigned>(), ctx->nfront.as<unsigned>(),
                                  ctx->topgen.as<unsigned>(), ctx->top_goal.as<int>(), ctx->frozen.as<int>(),
                                  depth, max_depth, target,
                                  paths ? ctx->lpath[cur].as<unsigned long long>() : nullptr,
                                  paths ? ctx->lpath[cur ^ 1].as<unsigned long long>() : nullptr,
                                  paths ? ctx->fpath.as<unsigned long long>() : nullptr, s));
        cur ^= 1;
    }

    if (getenv("SPZ_PROFILE") && atoi(getenv("SPZ_PROFILE")) > 0) {
        HIPCHK(ctx->prof.ensure(64));
        HIPCHK(hipMemsetAsync(ctx->prof.p, 0, 64, s));
        p.prof = ctx->prof.as<unsigned long long>();
        // with SPZ_WAVE_TIMES=r: every wave's start and end clock in round r
        if (getenv("SPZ_WAVE_TIMES")) {
            HIPCHK(ctx->wtime.ensure(lanes / 64 * 16));
            HIPCHK(hipMemsetAsync(ctx->wtime.p, 0, lanes / 64 * 16, s));
            p.wtime = ctx->wtime.as<unsigned long long>();
            p.wtime_round = (unsigned)atoi(getenv("SPZ_WAVE_TIMES"));
        }
    }
    HIPCHK(hipMemsetAsync(ctx->load.p, 0, front_items * 8, s));
    HIPCHK(hipMemsetAsync(ctx->nsamp.p, 0, 64, s));
    if (ctx->cfg.lb_policy == SPZ_LB_LEARNED)
        HIPCHK(mlp_predict_frontier(ctx->mlp, w, ctx->frontier.as<Item>(), ctx->nfront.as<unsigned>(),
                                    ctx->roots.as<Item>(), n, capl, ctx->pred.as<float>(), s));
    HIPCHK(hipStreamSynchronize(s));
    ctx->stats[10] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_split).count();

    // per-instance iteration state: every instance on the restart list with
    // no iteration started (k_restart starts iteration 0)
    {
        std::vector<int> Lh(N), nog(2 * N, kNoGoal), neg(N, -1), list(N), ones(N, 1);
        for (int i = 0; i < n; ++i) {
            Lh[i] = meta_h(roots[i].meta);
            list[i] = i;
        }
        HIPCHK(hipMemcpyAsync(ctx->Lbase.p, Lh.data(), N * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(ctx->L.p, 0, 2 * N * 4, s));
        HIPCHK(hipMemsetAsync(ctx->itn.p, 0, 2 * N * 4, s));
        HIPCHK(hipMemsetAsync(ctx->cnt.p, 0, 2 * N * 8, s));
        HIPCHK(hipMemsetAsync(ctx->pend.p, 0, 2 * N * 4, s));
        HIPCHK(hipMemsetAsync(ctx->drained.p, 0, 2 * N * 4, s));
        HIPCHK(hipMemsetAsync(ctx->dead.p, 0, 2 * N * 4, s));
        HIPCHK(hipMemcpyAsync(ctx->found.p, nog.data(), 2 * N * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(ctx->it.p, 0, N * 4, s));
        HIPCHK(hipMemsetAsync(ctx->nlive.p, 0, N * 4, s));
        HIPCHK(hipMemcpyAsync(ctx->listed.p, ones.data(), N * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(ctx->want_next.p, 0, N * 4, s));
        HIPCHK(hipMemsetAsync(ctx->done.p, 0, N * 4, s));
        HIPCHK(hipMemcpyAsync(ctx->opt.p, neg.data(), N * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(ctx->iter_counts.p, 0, N * (size_t)max_iters * 8, s));
        HIPCHK(hipMemcpyAsync(ctx->restart_list.p, list.data(), N * 4, hipMemcpyHostToDevice, s));
        unsigned ctl[64];
        memset(ctl, 0, sizeof ctl);
        ctl[0] = (unsigned)n; // restart list length
        ctl[1] = (unsigned)n; // active instances
        const bool full = lanes >= (size_t)ctx->n_lanes;
        int big = ctx->cfg.round_budget > 0 ? ctx->cfg.round_budget : default_round_budget(heur, full);
        // a round ends after its step budget or its time limit, whichever
        // comes first
        const int rt = ctx->cfg.round_time_us == 0 ? default_round_time_us(heur, full) : ctx->cfg.round_time_us;
        p.round_ticks = (rt > 0 && ctx->wall_ticks_per_us > 0)
                            ? (unsigned long long)(rt * ctx->wall_ticks_per_us) : 0ull;
        ctl[9] = (unsigned)big;
        // rounds while the queue cannot keep half the lanes busy: short, so
        // that donations refill idle lanes soon (measured on MI355X, 32 / 64
        // / 128 / 256 / 512 steps: 100 40-move 15-puzzles 5270 / 6300 /
        // 7270 / 7220 / 6040 instances/s, one random 15-puzzle per launch
        // 44.7 / 45.2 / 46.5 / 45.1 / 39.4, the PDB search level)
        ctl[10] = (unsigned)std::min(big, kShortRoundSteps);
        ctl[5] = ctl[10];
        HIPCHK(hipMemcpyAsync(ctx->ctl.p, ctl, sizeof ctl, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(ctx->q_ctl.p, 0, 2 * kQCtl * 8, s));
        HIPCHK(hipMemsetAsync(ctx->c_st.p, 0, lanes * 4, s));
        HIPCHK(hipStreamSynchronize(s));
    }

    HIPCHK(launch_restart(w, heur, p, ctx->geo_d.as<PdbGeo>(), s));
    HIPCHK(launch_control(p, s));
    auto t_rounds = std::chrono::steady_clock::now();
    ctx->stats[13] = std::chrono::duration<double, std::milli>(t_rounds - t_start).count();

    // rounds until every instance is done; poll every few rounds.  Each
    // round is a bounded kernel; the host stops on a wall-clock limit or if
    // work vanishes while instances are still active (a bookkeeping bug).
    const int per_poll = 4;
    const double limit_s = ctx->timeout_s;
    const bool trace = getenv("SPZ_TRACE") && atoi(getenv("SPZ_TRACE")) > 0; // per-poll progress on stderr
    int stalled = 0;
    const bool timing = ctx->cfg.time_kernels != 0;
    hipEvent_t ev[2 * per_poll];
    if (timing)
        for (auto &x : ev) HIPCHK(hipEventCreate(&x));
    double round_ms = 0;
    long timed = 0;
    unsigned *h_ctl = nullptr; // pinned: 64 control words, then the two rings' queue counters
    HIPCHK(hipHostMalloc((void **)&h_ctl, 64 * 4 + 2 * kQCtl * 8, hipHostMallocDefault));
    unsigned long long *h_qc = (unsigned long long *)(h_ctl + 64);
    int rc = SPZ_OK;
    long rounds = 0;

    // one poll period: per_poll rounds, then the control words to the host
    auto enqueue_period = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        for (int k = 0; k < per_poll && e == hipSuccess; ++k) {
            // timing: the runtime records the events with the dispatch
            if (e == hipSuccess)
                e = launch_round(w, heur, p, ctx->geo_d.as<PdbGeo>(), ctx->wg, s, timing ? ev[2 * k] : nullptr,
                                 timing ? ev[2 * k + 1] : nullptr);
            if (e == hipSuccess && p.lookahead) e = launch_spec(p, s);
            if (e == hipSuccess) e = launch_restart(w, heur, p, ctx->geo_d.as<PdbGeo>(), s);
            if (e == hipSuccess) e = launch_control(p, s);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(h_ctl, ctx->ctl.p, 64 * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(h_qc, ctx->q_ctl.p, 2 * kQCtl * 8, hipMemcpyDeviceToHost, s);
        return e;
    };
    // hipGraph: the period is captured once and replayed; the kernels read
    // all their state from device memory, so replays need no new arguments
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // kernel timing runs the rounds eagerly: HIP events recorded by a captured
    // graph cannot be timed against each other
    bool use_graph = ctx->cfg.use_graph != 0 && !timing;
    if (use_graph) {
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
        hipError_t e2 = e == hipSuccess ? enqueue_period() : e;
        hipError_t e3 = e == hipSuccess ? hipStreamEndCapture(s, &graph) : e;
        if (e2 == hipSuccess && e3 == hipSuccess) e3 = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            // capture unsupported here: run the same sequence eagerly
            if (gexec) (void)hipGraphExecDestroy(gexec);
            if (graph) (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
            use_graph = false;
            (void)hipGetLastError();
        }
    }
    ctx->stats[15] = use_graph ? 1.0 : 0.0;
    for (;;) {
        hipError_t e = use_graph ? hipGraphLaunch(gexec, s) : enqueue_period();
        rounds += per_poll;
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (timing && e == hipSuccess) {
            for (int k = 0; k < per_poll && e == hipSuccess; ++k) {
                float ms = 0;
                e = hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]);
                round_ms += ms;
                ++timed;
            }
        }
        if (e != hipSuccess) {
            rc = fail(ctx, SPZ_E_HIP, "search rounds: %s", hipGetErrorString(e));
            break;
        }
        if (trace) {
            fprintf(stderr, "spz poll %ld: %.1f ms, busy lanes %u, active instances %u, queued %llu", rounds,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_rounds).count(),
                    h_ctl[12], h_ctl[1],
                    (unsigned long long)(h_qc[2] - h_qc[0] + h_qc[kQCtl + 2] - h_qc[kQCtl]));
            unsigned long long pf[4];
            if (p.prof && hipMemcpy(pf, p.prof, sizeof pf, hipMemcpyDeviceToHost) == hipSuccess)
                fprintf(stderr, ", lane-steps idle %llu expanding %llu popping %llu", pf[0], pf[1], pf[2]);
            fputc('\n', stderr);
        }
        if (h_ctl[3]) {
            rc = fail(ctx, SPZ_E_CAPACITY, "work queue overflow: raise queue_items (now %zu items)", ctx->qcap);
            break;
        }
        if (h_ctl[1] == 0) break;
        // no lane busy at the end of the last round and nothing queued (a
        // finished iteration has queued its next one by now): no work can
        // appear any more
        stalled = (h_qc[2] == h_qc[0] && h_qc[kQCtl + 2] == h_qc[kQCtl] && h_ctl[12] == 0) ? stalled + 1 : 0;
        if (stalled > 2) {
            rc = fail(ctx, SPZ_E_STALL, "search stalled with %u active instances", h_ctl[1]);
            break;
        }
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        if (limit_s > 0 && el > limit_s) {
            rc = fail(ctx, SPZ_E_TIMEOUT, "search exceeded %.0f s", limit_s);
            break;
        }
    }
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (timing)
        for (auto &x : ev) (void)hipEventDestroy(x);
    if (rc != SPZ_OK) {
        (void)hipHostFree(h_ctl);
        return rc;
    }
    ctx->stats[14] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_rounds).count();
    ctx->stats[8] = round_ms;
    ctx->stats[9] = (double)timed;
    ctx->stats[0] = (double)h_ctl[6];
    ctx->stats[2] = (double)h_ctl[7];
    ctx->stats[3] = (double)lanes;
    ctx->stats[7] = (double)h_ctl[8];
    ctx->stats[23] = (double)h_ctl[13];
    ctx->stats[24] = (double)h_ctl[14];
    ctx->stats[25] = (double)h_ctl[15];
    (void)hipHostFree(h_ctl);

    std::vector<int> opt(N), its(N);
    std::vector<unsigned long long> ic(N * (size_t)max_iters);
    HIPCHK(hipMemcpyAsync(opt.data(), ctx->opt.p, N * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(its.data(), ctx->it.p, N * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ic.data(), ctx->iter_counts.p, ic.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    ctx->stats[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    double all_nodes = 0;
    for (int i = 0; i < n; ++i) {
        spz_result &r = out[i];
        r.opt_len = opt[i];
        r.h0 = meta_h(roots[i].meta);
        r.n_iters = its[i] + 1;
        r.status = opt[i] >= 0 ? SPZ_ST_SOLVED : SPZ_ST_ITER_LIMIT;
        uint64_t tot = 0;
        for (int k = 0; k < r.n_iters; ++k) tot += ic[(size_t)i * max_iters + k];
        r.total_nodes = tot;
        all_nodes += (double)tot;
        if (iter_counts)
            for (int k = 0; k < max_iters; ++k) iter_counts[(size_t)i * max_iters + k] = ic[(size_t)i * max_iters + k];
    }
    ctx->stats[11] = (double)n;
    ctx->paths_n = paths ? n : 0;
    if (paths) {
        ctx->paths_slot.resize(N);
        for (int i = 0; i < n; ++i) ctx->paths_slot[i] = its[i] & 1;
    }
    if (p.prof) {
        unsigned long long pf[7];
        HIPCHK(hipMemcpy(pf, p.prof, sizeof pf, hipMemcpyDeviceToHost));
        for (int k = 0; k < 7; ++k) ctx->stats[16 + k] = (double)pf[k];
    }
    if (p.wtime) {
        // end times of the recorded round's waves after its first start, on
        // stderr: the spread is the round's tail
        const size_t nw = lanes / 64;
        std::vector<unsigned long long> wt(2 * nw);
        HIPCHK(hipMemcpy(wt.data(), p.wtime, wt.size() * 8, hipMemcpyDeviceToHost));
        std::vector<double> ends;
        unsigned long long t0 = ~0ull;
        for (size_t k = 0; k < nw; ++k)
            if (wt[2 * k + 1]) t0 = std::min(t0, wt[2 * k]);
        for (size_t k = 0; k < nw; ++k)
            if (wt[2 * k + 1]) ends.push_back((double)(wt[2 * k + 1] - t0));
        std::sort(ends.begin(), ends.end());
        if (!ends.empty()) {
            auto q = [&](double f) { return ends[(size_t)(f * (ends.size() - 1))]; };
            fprintf(stderr, "spz wave end ticks in round %u (%zu waves): p10 %.0f p50 %.0f p90 %.0f p99 %.0f max %.0f\n",
                    p.wtime_round, ends.size(), q(0.1), q(0.5), q(0.9), q(0.99), ends.back());
        }
    }
    ctx->stats[12] = all_nodes;
    return SPZ_OK;
} SPZ_CATCH(ctx)

static hipError_t dedup_level(spz_ctx *ctx, int n, int capl, int width, int cur, bool paths,
                              std::vector<SeenBoards> &seen, hipStream_t s) {
    const size_t N = (size_t)n, C = (size_t)capl;
    std::vector<unsigned> nl(N);
    hipError_t e = hipMemcpyAsync(nl.data(), ctx->nlvl[cur].p, N * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    // only the first max_i nl[i] items of each instance's region travel, in
    // one strided copy each way, through pinned staging buffers
    size_t width_items = 0;
    for (size_t i = 0; i < N; ++i) width_items = std::max<size_t>(width_items, std::min<size_t>(nl[i], C));
    if (width_items == 0) return hipSuccess;
    if ((e = ctx->dd_items.ensure(N * width_items * sizeof(Item))) != hipSuccess) return e;
    if (paths && (e = ctx->dd_paths.ensure(N * width_items * 8)) != hipSuccess) return e;
    Item *items = (Item *)ctx->dd_items.p;
    unsigned long long *lp = (unsigned long long *)ctx->dd_paths.p;
    const size_t W = width_items;
    e = hipMemcpy2DAsync(items, W * sizeof(Item), ctx->lvl[cur].p, C * sizeof(Item), W * sizeof(Item), N,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && paths)
        e = hipMemcpy2DAsync(lp, W * 8, ctx->lpath[cur].p, C * 8, W * 8, N, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    dedup_level_host(
        items, paths ? lp : nullptr, N, W, nl.data(), seen, [&](const Item &it) { return board_key(width, it); },
        [](const Item &it) { return std::max((int)it.x0, meta_g(it.meta) + meta_h(it.meta)); },
        [](const Item &it) { return meta_prev(it.meta); },
        std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    e = hipMemcpyAsync(ctx->nlvl[cur].p, nl.data(), N * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(ctx->lvl[cur].p, C * sizeof(Item), items, W * sizeof(Item), W * sizeof(Item), N,
                             hipMemcpyHostToDevice, s);
    if (e == hipSuccess && paths)
        e = hipMemcpy2DAsync(ctx->lpath[cur].p, C * 8, lp, W * 8, W * 8, N, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

int spz_get_paths(spz_ctx *ctx, uint8_t *moves, int32_t max_len, int32_t *lens) try {
    if (!ctx || !lens || max_len < 0 || (max_len > 0 && !moves)) return SPZ_E_ARG;
    if (ctx->paths_n <= 0) return fail(ctx, SPZ_E_ARG, "the last solve recorded no paths (config record_paths)");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    const size_t n = (size_t)ctx->paths_n;
    // one path per iteration slot; the instance's answer is in the slot of
    // its last iteration
    std::vector<int> len(2 * n);
    std::vector<uint32_t> w(2 * n * kPathWords);
    HIPCHK(hipMemcpy(len.data(), ctx->sol_len.p, 2 * n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(w.data(), ctx->sol.p, 2 * n * kPathWords * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        const size_t sl = 2 * i + (size_t)ctx->paths_slot[i];
        lens[i] = len[sl];
        for (int j = 0; j < max_len; ++j)
            moves[i * max_len + j] = j < len[sl] ? (uint8_t)((w[sl * kPathWords + j / 16] >> (2 * (j % 16))) & 3u) : 0;
    }
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_gather_probe(spz_ctx *ctx, int64_t n_gathers, int64_t span_bytes, double *ms, double *gathers_per_s) try {
    if (!ctx || !ms || !gathers_per_s || n_gathers <= 0) return SPZ_E_ARG;
    // the grid is n_gathers / 2^16 workgroups of 256 threads: keep it an int
    if (n_gathers > (1ll << 36)) return fail(ctx, SPZ_E_ARG, "n_gathers above 2^36");
    if (!ctx->have_pdb) return fail(ctx, SPZ_E_NOPDB, "call spz_build_pdb first");
    const uint64_t both = 2ull * kPdbDirect; // the two direct tables, end to end
    const uint64_t span = (span_bytes <= 0 || (uint64_t)span_bytes > both) ? both : (uint64_t)span_bytes;
    HIPCHK(hipSetDevice(ctx->cfg.device));
    const int per_thread = 256;
    long long threads = (n_gathers + per_thread - 1) / per_thread;
    int blocks = (int)std::max<long long>(1, (threads + 255) / 256);
    DBuf out;
    HIPCHK(out.ensure((size_t)blocks * 256 * 4));
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    hipError_t e = launch_gather_probe(ctx->pdb_direct.as<uint8_t>(), span, blocks, per_thread, 12345,
                                       out.as<uint32_t>(), ctx->stream); // warm
    if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
    if (e == hipSuccess)
        e = launch_gather_probe(ctx->pdb_direct.as<uint8_t>(), span, blocks, per_thread, 777, out.as<uint32_t>(),
                                ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    out.release();
    if (e != hipSuccess) return fail(ctx, SPZ_E_HIP, "gather probe: %s", hipGetErrorString(e));
    *ms = t;
    *gathers_per_s = (double)blocks * 256.0 * per_thread / (t * 1e-3);
    return SPZ_OK;
} SPZ_CATCH(ctx)

// ------------------------------------------------------------ learned LB ---
int32_t spz_mlp_param_count(void) { return (int32_t)kMlpParams; }

static int mlp_ready_alloc(spz_ctx *ctx) {
    if (!ctx->mlp_alloced) {
        HIPCHK(mlp_alloc(ctx->mlp));
        ctx->mlp_alloced = true;
    }
    return SPZ_OK;
}

int spz_mlp_init(spz_ctx *ctx, uint64_t seed) try {
    if (!ctx) return SPZ_E_ARG;
    HIPCHK(hipSetDevice(ctx->cfg.device));
    int rc = mlp_ready_alloc(ctx);
    if (rc) return rc;
    HIPCHK(mlp_init_random(ctx->mlp, seed, ctx->stream));
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_mlp_set_params(spz_ctx *ctx, const float *params) try {
    if (!ctx || !params) return SPZ_E_ARG;
    HIPCHK(hipSetDevice(ctx->cfg.device));
    int rc = mlp_ready_alloc(ctx);
    if (rc) return rc;
    HIPCHK(mlp_set_params(ctx->mlp, params, ctx->stream));
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_mlp_get_params(spz_ctx *ctx, float *params) try {
    if (!ctx || !params) return SPZ_E_ARG;
    if (!ctx->mlp_alloced || !ctx->mlp.ready) return fail(ctx, SPZ_E_ARG, "no MLP parameters");
    HIPCHK(mlp_get_params(ctx->mlp, params, ctx->stream));
    return SPZ_OK;
} SPZ_CATCH(ctx)

// boards -> Items with the context's heuristic in meta -> feature rows
static int mlp_load_rows(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr, int32_t n, DBuf &items,
                         DBuf &thr_d) {
    const int w = ctx->cfg.width;
    if (ctx->cfg.heuristic == SPZ_H_PDB666 && !ctx->have_pdb) return fail(ctx, SPZ_E_NOPDB, "call spz_build_pdb first");
    for (int i = 0; i < n; ++i)
        if (!is_solvable(w, boards + (size_t)i * w * w))
            return fail(ctx, SPZ_E_ARG, "board %d is not a solvable permutation", i);
    std::vector<Item> it((size_t)n);
    for (int i = 0; i < n; ++i) pack_root(w, boards + (size_t)i * w * w, i, &it[i]);
    HIPCHK(items.ensure((size_t)n * sizeof(Item)));
    HIPCHK(thr_d.ensure((size_t)n * 4));
    HIPCHK(hipMemcpyAsync(items.p, it.data(), (size_t)n * sizeof(Item), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(thr_d.p, thr, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    KParams p = base_params(ctx, 0, 1, 1);
    HIPCHK(launch_eval_h(w, ctx->cfg.heuristic, p, ctx->geo_d.as<PdbGeo>(), items.as<Item>(), n, ctx->stream));
    HIPCHK(mlp_reserve(ctx->mlp, (size_t)n));
    HIPCHK(mlp_features_items(ctx->mlp, w, items.as<Item>(), thr_d.as<int>(), n, ctx->stream));
    return SPZ_OK;
}

int spz_mlp_forward(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr, int32_t n, float *out) try {
    if (!ctx || !boards || !thr || !out || n <= 0) return SPZ_E_ARG;
    if (!ctx->mlp_alloced || !ctx->mlp.ready) return fail(ctx, SPZ_E_ARG, "no MLP parameters");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    DBuf items, thr_d;
    int rc = mlp_load_rows(ctx, boards, thr, n, items, thr_d);
    if (rc) return rc;
    HIPCHK(mlp_forward(ctx->mlp, n, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, ctx->mlp.y, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SPZ_OK;
} SPZ_CATCH(ctx)

static_assert(kPredK == kPredThr, "one prediction per threshold slot of the restart");

int spz_mlp_predict(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr0, int32_t n, float *out) try {
    if (!ctx || !boards || !thr0 || !out || n <= 0) return SPZ_E_ARG;
    if (!ctx->mlp_alloced || !ctx->mlp.ready) return fail(ctx, SPZ_E_ARG, "no MLP parameters");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    DBuf items, thr_d, y;
    int rc = mlp_load_rows(ctx, boards, thr0, n, items, thr_d);
    if (rc) return rc;
    HIPCHK(y.ensure((size_t)n * kPredThr * 4));
    HIPCHK(mlp_predict_items(ctx->mlp, ctx->cfg.width, items.as<Item>(), thr_d.as<int>(), n, y.as<float>(),
                             ctx->stream));
    HIPCHK(hipMemcpyAsync(out, y.p, (size_t)n * kPredThr * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_mlp_grad(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr, const float *target, int32_t n,
                 float *loss, float *grad) try {
    if (!ctx || !boards || !thr || !target || !loss || !grad || n <= 0) return SPZ_E_ARG;
    if (!ctx->mlp_alloced || !ctx->mlp.ready) return fail(ctx, SPZ_E_ARG, "no MLP parameters");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    DBuf items, thr_d, tgt;
    int rc = mlp_load_rows(ctx, boards, thr, n, items, thr_d);
    if (rc) return rc;
    HIPCHK(tgt.ensure((size_t)n * 4));
    HIPCHK(hipMemcpyAsync(tgt.p, target, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(mlp_forward(ctx->mlp, n, ctx->stream));
    HIPCHK(mlp_backward(ctx->mlp, tgt.as<float>(), n, ctx->stream));
    HIPCHK(hipMemcpyAsync(loss, ctx->mlp.loss, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(grad, ctx->mlp.grad, kMlpParams * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_mlp_train(spz_ctx *ctx, const uint8_t *boards, const int32_t *thr, const float *target, int32_t n,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss) try {
    if (!ctx || !boards || !thr || !target || n <= 0 || epochs <= 0 || batch <= 0) return SPZ_E_ARG;
    if (!ctx->mlp_alloced || !ctx->mlp.ready) return fail(ctx, SPZ_E_ARG, "initialise the MLP first");
    HIPCHK(hipSetDevice(ctx->cfg.device));
    const int w = ctx->cfg.width;
    hipStream_t s = ctx->stream;
    // all rows, thresholds and targets to the device once; each epoch uploads
    // one permutation and its batches are gathered on the device, so the
    // host waits once per epoch (for the epoch's summed batch losses)
    std::vector<Item> it((size_t)n);
    for (int i = 0; i < n; ++i) {
        if (!is_solvable(w, boards + (size_t)i * w * w))
            return fail(ctx, SPZ_E_ARG, "board %d is not a solvable permutation", i);
        pack_root(w, boards + (size_t)i * w * w, i, &it[i]);
    }
    DBuf items, thr_d, tgt_d, perm_d, btgt, eloss;
    HIPCHK(items.ensure((size_t)n * sizeof(Item)));
    HIPCHK(thr_d.ensure((size_t)n * 4));
    HIPCHK(tgt_d.ensure((size_t)n * 4));
    HIPCHK(perm_d.ensure((size_t)n * 4));
    HIPCHK(btgt.ensure((size_t)batch * 4));
    HIPCHK(eloss.ensure((size_t)epochs * 4));
    HIPCHK(hipMemcpyAsync(items.p, it.data(), (size_t)n * sizeof(Item), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(thr_d.p, thr, (size_t)n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(tgt_d.p, target, (size_t)n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(eloss.p, 0, (size_t)epochs * 4, s));
    KParams p = base_params(ctx, 0, 1, 1);
    HIPCHK(launch_eval_h(w, ctx->cfg.heuristic, p, ctx->geo_d.as<PdbGeo>(), items.as<Item>(), n, s));
    HIPCHK(mlp_reserve(ctx->mlp, (size_t)batch));
    std::vector<int> perm((size_t)n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    uint64_t x = seed;
    std::vector<int> nbatches((size_t)epochs, 0);
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = n - 1; i > 0; --i) { // Fisher-Yates with splitmix64
            uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            z ^= z >> 31;
            std::swap(perm[i], perm[(int)(z % (uint64_t)(i + 1))]);
        }
        // the upload is ordered before this epoch's kernels on the stream; the
        // host vector is reshuffled only after the epoch's synchronisation
        HIPCHK(hipMemcpyAsync(perm_d.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
        float *el = eloss.as<float>() + ep;
        for (int b0 = 0; b0 + batch <= n || (b0 == 0 && n < batch); b0 += batch) {
            const int nbat = std::min(batch, n - b0);
            HIPCHK(mlp_features_gather(ctx->mlp, w, items.as<Item>(), thr_d.as<int>(), tgt_d.as<float>(),
                                       perm_d.as<int>() + b0, nbat, btgt.as<float>(), s));
            HIPCHK(mlp_forward(ctx->mlp, nbat, s));
            HIPCHK(mlp_backward(ctx->mlp, btgt.as<float>(), nbat, s, el));
            HIPCHK(mlp_adam_step(ctx->mlp, lr, s));
            ++nbatches[ep];
        }
        HIPCHK(hipStreamSynchronize(s));
    }
    std::vector<float> sums((size_t)epochs);
    HIPCHK(hipMemcpy(sums.data(), eloss.p, (size_t)epochs * 4, hipMemcpyDeviceToHost));
    if (epoch_loss)
        for (int ep = 0; ep < epochs; ++ep) epoch_loss[ep] = nbatches[ep] ? sums[ep] / (float)nbatches[ep] : 0.f;
    return SPZ_OK;
} SPZ_CATCH(ctx)

int spz_get_samples(spz_ctx *ctx, uint8_t *boards, int32_t *thr, uint64_t *loads, int32_t max, int32_t *n_out) try {
    if (!ctx || !n_out) return SPZ_E_ARG;
    *n_out = 0;
    if (!ctx->samples.p || !ctx->nsamp.p) return SPZ_OK;
    unsigned cnt = 0;
    HIPCHK(hipMemcpy(&cnt, ctx->nsamp.p, 4, hipMemcpyDeviceToHost));
    unsigned have = std::min<unsigned>(cnt, (unsigned)std::max(ctx->cfg.log_samples, 0));
    unsigned n = std::min<unsigned>(have, (unsigned)std::max(max, 0));
    if (n == 0 || !boards) {
        *n_out = (int32_t)have;
        return SPZ_OK;
    }
    std::vector<Sample> sm(n);
    HIPCHK(hipMemcpy(sm.data(), ctx->samples.p, (size_t)n * sizeof(Sample), hipMemcpyDeviceToHost));
    const int w = ctx->cfg.width, n2 = w * w;
    for (unsigned i = 0; i < n; ++i) {
        if (w == 4) {
            Board<4> bd;
            bd.unpack(sm[i].item.b);
            for (int c = 0; c < n2; ++c) boards[(size_t)i * n2 + c] = (uint8_t)bd.tile(c);
        } else {
            Board<5> bd;
            bd.unpack(sm[i].item.b);
            for (int c = 0; c < n2; ++c) boards[(size_t)i * n2 + c] = (uint8_t)bd.tile(c);
        }
        if (thr) thr[i] = sm[i].thr;
        if (loads) loads[i] = sm[i].load;
    }
    *n_out = (int32_t)n;
    return SPZ_OK;
} SPZ_CATCH(ctx)

void spz_pdb_geometry(uint8_t *tiles48, uint8_t *maps200) {
    PdbGeo g;
    pdb_geo_build(&g);
    if (tiles48) memcpy(tiles48, g.tiles, 48);
    if (maps200) memcpy(maps200, g.map, 200);
}

// ----------------------------------------------------------- generators ---
static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix64(&x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    int below(int k) { return (int)(((next() >> 32) * (uint64_t)k) >> 32); }
};

int32_t spz_gen_walk(int32_t width, int32_t steps, uint64_t seed, uint8_t *out) {
    if (!out || width < 2 || width > 5 || steps < 0) return SPZ_E_ARG;
    int n2 = width * width, blank = 0, prev = -1;
    Xoshiro r(seed);
    for (int i = 0; i < n2; ++i) out[i] = (uint8_t)i;
    for (int s = 0; s < steps; ++s) {
        int cand[4], nc = 0;
        int row = blank / width, col = blank % width;
        for (int d = 0; d < 4; ++d) {
            if (prev >= 0 && d == 3 - prev) continue;
            bool ok = (d == DIR_UP && row > 0) || (d == DIR_RIGHT && col < width - 1) ||
                      (d == DIR_LEFT && col > 0) || (d == DIR_DOWN && row < width - 1);
            if (ok) cand[nc++] = d;
        }
        int d = cand[r.below(nc)];
        int nb = blank + (d == DIR_UP ? -width : d == DIR_DOWN ? width : d == DIR_LEFT ? -1 : 1);
        out[blank] = out[nb];
        out[nb] = 0;
        blank = nb;
        prev = d;
    }
    return SPZ_OK;
}

int32_t spz_gen_random(int32_t width, uint64_t seed, uint8_t *out) {
    if (!out || width < 2 || width > 5) return SPZ_E_ARG;
    int n2 = width * width;
    Xoshiro r(seed);
    for (;;) {
        for (int i = 0; i < n2; ++i) out[i] = (uint8_t)i;
        for (int i = n2 - 1; i > 0; --i) {
            int j = r.below(i + 1);
            std::swap(out[i], out[j]);
        }
        if (is_solvable(width, out)) return SPZ_OK;
    }
}

} // extern "C"
