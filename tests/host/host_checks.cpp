// Host-side checks of libspz's plain C++ parts, built by
// tests/test_host_sanitizers.py under the address + undefined-behaviour
// sanitizers and, separately, the thread sanitizer:
//  - dedup_level_host (root-split duplicate pruning on host threads) against
//    a direct restatement of its rule, on random levels with many duplicates;
//  - read_problem_file (the CLI's problem-file reader) on valid, short,
//    malformed and out-of-range files.
// Prints "OK" and exits 0 when every check passes.
#include "../../sliding-puzzle_amd/csrc/spz_dedup.h"
#include "../../sliding-puzzle_amd/host/spz_problem.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <string>
#include <tuple>
#include <vector>

using spz::BoardKey;
using spz::SeenBoards;

static int failures = 0;
#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            ++failures;                                                           \
        }                                                                         \
    } while (0)

// a level item: a board from a small set (so that copies repeat), its
// path-max f and previous move; the payload is a function of the three, as
// copies of one node are identical
struct TItem {
    uint32_t board;
    int f, prev;
    uint32_t payload;
};
static BoardKey key_of(const TItem &t) { return BoardKey{{t.board, t.board * 2654435761u, 0u, 0u}}; }
static uint32_t payload_of(uint32_t board, int f, int prev) { return board * 1000003u + (uint32_t)f * 17u + (uint32_t)prev; }

// the rule restated: a level's copy survives iff it is the first copy of its
// board with the smallest (f, prev) in the level and no kept shallower copy
// of the board has a path-max f no larger
static std::vector<TItem> reference_level(const std::vector<TItem> &level, std::map<uint32_t, int> &seen) {
    std::map<uint32_t, std::pair<int, int>> best;
    for (const TItem &t : level) {
        auto it = best.find(t.board);
        if (it == best.end() || std::make_pair(t.f, t.prev) < it->second) best[t.board] = {t.f, t.prev};
    }
    std::vector<TItem> kept;
    std::map<uint32_t, int> taken;
    for (const TItem &t : level) {
        if (std::make_pair(t.f, t.prev) != best[t.board] || taken.count(t.board)) continue;
        taken[t.board] = 1;
        auto s = seen.find(t.board);
        if (s != seen.end() && s->second <= t.f) continue;
        kept.push_back(t);
    }
    for (const TItem &t : kept) seen[t.board] = t.f;
    return kept;
}

static void check_dedup(unsigned seed, size_t N, size_t W, unsigned n_threads, int levels) {
    std::mt19937 rng(seed);
    std::vector<SeenBoards> seen(N);
    std::vector<std::map<uint32_t, int>> ref_seen(N);
    for (int lv = 0; lv < levels; ++lv) {
        std::vector<TItem> items(N * W);
        std::vector<unsigned> nl(N);
        std::vector<unsigned long long> lp(N * W);
        std::vector<std::vector<TItem>> expect(N);
        for (size_t i = 0; i < N; ++i) {
            nl[i] = (unsigned)(rng() % (W + 1));
            std::vector<TItem> level;
            const uint32_t boards = 1 + rng() % 40;
            for (unsigned j = 0; j < nl[i]; ++j) {
                TItem t;
                t.board = (uint32_t)(rng() % boards);
                t.f = 20 + lv + (int)(rng() % 4);
                t.prev = (int)(rng() % 4);
                t.payload = payload_of(t.board, t.f, t.prev);
                items[i * W + j] = t;
                lp[i * W + j] = t.payload;
                level.push_back(t);
            }
            expect[i] = reference_level(level, ref_seen[i]);
        }
        spz::dedup_level_host(
            items.data(), lp.data(), N, W, nl.data(), seen, [](const TItem &t) { return key_of(t); },
            [](const TItem &t) { return t.f; }, [](const TItem &t) { return t.prev; }, n_threads);
        for (size_t i = 0; i < N; ++i) {
            CHECK(nl[i] == expect[i].size());
            // kept copies of a board are identical, so compare as multisets
            std::vector<std::tuple<uint32_t, int, int, uint32_t>> got, want;
            for (unsigned j = 0; j < nl[i] && j < W; ++j) {
                const TItem &t = items[i * W + j];
                CHECK(lp[i * W + j] == t.payload);
                CHECK(t.payload == payload_of(t.board, t.f, t.prev));
                got.emplace_back(t.board, t.f, t.prev, t.payload);
            }
            for (const TItem &t : expect[i]) want.emplace_back(t.board, t.f, t.prev, t.payload);
            std::sort(got.begin(), got.end());
            std::sort(want.begin(), want.end());
            CHECK(got == want);
            // seen: sorted by board, one entry per board, the reference's f
            CHECK(seen[i].size() == ref_seen[i].size());
            for (size_t a = 1; a < seen[i].size(); ++a) CHECK(seen[i][a - 1].first < seen[i][a].first);
            for (const auto &e : seen[i]) {
                bool found = false;
                for (const auto &r : ref_seen[i])
                    if (key_of(TItem{r.first, 0, 0, 0}) == e.first) {
                        found = true;
                        CHECK(r.second == e.second);
                    }
                CHECK(found);
            }
        }
    }
}

static std::string write_file(const std::string &dir, const std::string &name, const std::string &text) {
    const std::string p = dir + "/" + name;
    FILE *f = std::fopen(p.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", p.c_str());
        std::exit(2);
    }
    std::fputs(text.c_str(), f);
    std::fclose(f);
    return p;
}

static void check_problem_files(const std::string &dir) {
    std::string b16, b25;
    for (int i = 0; i < 16; ++i) b16 += std::to_string((i + 1) % 16) + (i % 4 == 3 ? "\n" : " ");
    for (int i = 0; i < 25; ++i) b25 += std::to_string((i + 3) % 25) + (i % 5 == 4 ? "\n" : " ");
    std::vector<uint8_t> b;
    int w = 0;
    CHECK(read_problem_file(write_file(dir, "p16", b16), b, w) && w == 4 && b.size() == 16 && b[0] == 1 && b[15] == 0);
    CHECK(read_problem_file(write_file(dir, "p16b", b16 + " 7 7 7\n"), b, w) && b.size() == 32); // trailing ignored
    CHECK(!read_problem_file(write_file(dir, "p25", b25), b, w) && b.size() == 32);              // width differs
    w = 0;
    b.clear();
    CHECK(read_problem_file(write_file(dir, "p25b", b25), b, w) && w == 5 && b.size() == 25 && b[22] == 0);
    w = 0;
    CHECK(!read_problem_file(write_file(dir, "empty", ""), b, w) && w == 0);
    CHECK(!read_problem_file(write_file(dir, "short", "1 2 3\n4 0 5\n"), b, w));
    CHECK(!read_problem_file(write_file(dir, "text", "a b c d e f g h i j k l m n o p q\n"), b, w));
    CHECK(!read_problem_file(write_file(dir, "neg", "-1 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15\n"), b, w));
    CHECK(!read_problem_file(write_file(dir, "big", "99 1 2 3 4 5 6 7 8 9 10 11 12 13 14 0\n"), b, w));
    CHECK(!read_problem_file(write_file(dir, "huge", "99999999999999999999 1 2 3 4 5 6 7 8 9 10 11 12 13 14 0\n"), b, w));
    CHECK(!read_problem_file(dir + "/does-not-exist", b, w));
    CHECK(b.size() == 25 && w == 0);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: host_checks TMPDIR\n");
        return 2;
    }
    check_dedup(1, 64, 300, 8, 4);
    check_dedup(2, 5, 2000, 16, 3);
    check_dedup(3, 1, 50, 1, 6);
    check_problem_files(argv[1]);
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
