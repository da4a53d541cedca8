// spz_dedup.h -- host pass of the optional duplicate pruning of the root
// split (config `dedup`): one split level of every instance at a time, on
// host threads.  Plain C++ (no HIP): the library calls it on the level it
// copied back from the device, and the CPU tests build it on its own under
// the address, undefined-behaviour and thread sanitizers.
//
// A node is dropped when a kept node of its instance's top tree -- the root,
// a node of a shallower level, or a copy in the same level -- holds the same
// board with a path-max f no larger than its own.  Within a level the copy
// kept is the one with the smallest (path-max f, previous move): the level's
// order comes from the split kernel's atomics, and copies equal in both are
// the same node, so the result does not depend on that order (the closed
// list of src_horie/idas_smem.cu:1057-1149 keeps the first copy it meets).
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <utility>
#include <vector>

namespace spz {

struct BoardKey {
    uint32_t w[4];
    bool operator<(const BoardKey &o) const {
        for (int i = 0; i < 4; ++i)
            if (w[i] != o.w[i]) return w[i] < o.w[i];
        return false;
    }
    bool operator==(const BoardKey &o) const {
        return w[0] == o.w[0] && w[1] == o.w[1] && w[2] == o.w[2] && w[3] == o.w[3];
    }
};
// an instance's kept top-tree boards with the smallest path-max f kept for
// each, sorted by board
using SeenBoards = std::vector<std::pair<BoardKey, int>>;

// One level of N instances: instance i holds nl[i] items (at most W) at
// items[i * W ...], with their paths at lp[i * W ...] when lp is not null.
// key(item), pf(item), prev(item): board, path-max f, previous move.  Kept
// items move to the front of their instance's region in their original
// order, nl[i] becomes their number, and seen[i] (sorted by board) takes
// the level's kept boards.  Instances are independent: thread t of
// n_threads handles i = t, t + n_threads, ...
template <class ItemT, class KeyF, class FF, class PrevF>
void dedup_level_host(ItemT *items, unsigned long long *lp, size_t N, size_t W, unsigned *nl,
                      std::vector<SeenBoards> &seen, KeyF key, FF pf, PrevF prev, unsigned n_threads) {
    struct Entry {
        BoardKey k;
        int f, prev;
        uint32_t j;
        bool operator<(const Entry &o) const {
            if (!(k == o.k)) return k < o.k;
            if (f != o.f) return f < o.f;
            return prev < o.prev;
        }
    };
    auto one = [&](size_t i) {
        const size_t cnt = std::min<size_t>(nl[i], W);
        std::vector<Entry> v(cnt);
        for (size_t j = 0; j < cnt; ++j) {
            const ItemT &it = items[i * W + j];
            v[j].k = key(it);
            v[j].f = pf(it);
            v[j].prev = prev(it);
            v[j].j = (uint32_t)j;
        }
        std::sort(v.begin(), v.end());
        std::vector<char> keep(cnt, 0);
        SeenBoards &sn = seen[i];
        SeenBoards merged;
        merged.reserve(sn.size() + cnt);
        size_t a = 0;
        for (size_t x = 0; x < cnt; ++x) {
            if (x > 0 && v[x].k == v[x - 1].k) continue; // not the board's best copy
            while (a < sn.size() && sn[a].first < v[x].k) merged.push_back(sn[a++]);
            if (a < sn.size() && sn[a].first == v[x].k) {
                if (sn[a].second <= v[x].f) continue; // a shallower kept copy covers it
                merged.emplace_back(v[x].k, v[x].f);
                ++a;
            } else {
                merged.emplace_back(v[x].k, v[x].f);
            }
            keep[v[x].j] = 1;
        }
        while (a < sn.size()) merged.push_back(sn[a++]);
        sn.swap(merged);
        size_t out = 0;
        for (size_t j = 0; j < cnt; ++j) {
            if (!keep[j]) continue;
            if (out != j) {
                items[i * W + out] = items[i * W + j];
                if (lp) lp[i * W + out] = lp[i * W + j];
            }
            ++out;
        }
        nl[i] = (unsigned)out;
    };
    const size_t nth = std::min<size_t>(N, std::max(1u, n_threads));
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nth; ++t)
        pool.emplace_back([&, t]() {
            for (size_t i = t; i < N; i += nth) one(i);
        });
    for (auto &th : pool) th.join();
}

} // namespace spz
