// spz_solve -- command-line front end of libspz.
//
// Takes over from the reference's solver processes: read problem files
// (N*N whitespace-separated integers, 0 = blank; src/24puzzle.cc:72-87,
// src_horie/idas_cpu.c:336-360), solve, and print one line per file in the
// reference's format "<path> <len> <nodes>" (cumulative all-solutions count,
// src/24puzzle_expand.cc:154).  --iters adds the per-iteration lines of
// src_horie/idas_cpu_expand.c:288-304 ("f_limit=%3d, expanded nodes = %lld"),
// --csv writes per-instance seconds like the reference's result/*.csv files.
#include "spz.h"
#include "spz_problem.h"

#include <cerrno>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

static void usage() {
    std::fprintf(stderr,
                 "usage: spz_solve [options] FILE...\n"
                 "  -H manhattan|pdb       heuristic (default manhattan)\n"
                 "  -m all|first           all solutions of the last iteration (default) or stop at a goal\n"
                 "  --batch                solve all files in one batch (default: one per launch)\n"
                 "  --iters                print per-iteration counts\n"
                 "  --path                 print an optimal solution (\"len=N: d d ...\", blank moves\n"
                 "                         0 up, 1 right, 2 left, 3 down)\n"
                 "  --csv FILE             write per-instance seconds\n"
                 "  --pdb-cache FILE       read the PDB tables from FILE, or write them there\n"
                 "  --pdb-tab DIR          load the reference's pat24.1256712.tab and\n"
                 "                         pat24.34891314.tab from DIR instead of building them\n"
                 "  --export-tab DIR       write pat24.1256712.tab and pat24.34891314.tab\n"
                 "  --lb uniform|prevload|learned  dispatch policy of frontier nodes (learned\n"
                 "                         needs --mlp-params)\n"
                 "  --mlp-params FILE      predictor parameters: spz_mlp_param_count() float32\n"
                 "  --resplit K            re-split frontier nodes above K x the mean estimate\n"
                 "  --frontier N --wg N --lds-depth N --budget N --max-iters N (1..%d)\n"
                 "  --autotune             tune the launch configuration on the files first\n"
                 "  --no-lookahead         run each instance's iterations strictly one at a time\n"
                 "  --dedup                drop duplicate boards from the root split (fewer nodes;\n"
                 "                         not the reference's counts)\n"
                 "  --round-time US        a search round also ends after US microseconds (-1: off;\n"
                 "                         default 4500 for the PDB search; Manhattan 3500 when the\n"
                 "                         batch has work for every lane, else 700)\n",
                 SPZ_MAX_ITERS);
}

int main(int argc, char **argv) {
    spz_config cfg;
    int heur = SPZ_H_MANHATTAN, mode = SPZ_MODE_ALL, lb = SPZ_LB_UNIFORM;
    int frontier = 2048, wg = 256, depth = 0, budget = 0, max_iters = 64, round_us = 0;
    bool batch = false, iters = false, autotune = false, path = false, lookahead = true, dedup = false;
    std::string csv, cache, tabdir, tabin, mlpfile;
    int resplit = 0;
    std::vector<std::string> files;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        // a whole decimal int, else the usage text and exit code 2
        auto num = [&]() -> int {
            const char *v = next();
            char *end = nullptr;
            errno = 0;
            const long x = std::strtol(v, &end, 10);
            if (end == v || *end != '\0' || errno == ERANGE || x < INT_MIN || x > INT_MAX) {
                usage();
                std::exit(2);
            }
            return (int)x;
        };
        if (a == "-H") {
            std::string v = next();
            if (v == "pdb")
                heur = SPZ_H_PDB666;
            else if (v == "manhattan")
                heur = SPZ_H_MANHATTAN;
            else {
                usage();
                return 2;
            }
        } else if (a == "-m") {
            std::string v = next();
            if (v == "first")
                mode = SPZ_MODE_FIRST;
            else if (v == "all")
                mode = SPZ_MODE_ALL;
            else {
                usage();
                return 2;
            }
        } else if (a == "--batch") {
            batch = true;
        } else if (a == "--iters") {
            iters = true;
        } else if (a == "--path") {
            path = true;
        } else if (a == "--csv") {
            csv = next();
        } else if (a == "--pdb-cache") {
            cache = next();
        } else if (a == "--export-tab") {
            tabdir = next();
        } else if (a == "--pdb-tab") {
            tabin = next();
        } else if (a == "--mlp-params") {
            mlpfile = next();
        } else if (a == "--resplit") {
            resplit = num();
        } else if (a == "--lb") {
            std::string v = next();
            if (v == "prevload")
                lb = SPZ_LB_PREVLOAD;
            else if (v == "learned")
                lb = SPZ_LB_LEARNED;
            else if (v == "uniform")
                lb = SPZ_LB_UNIFORM;
            else {
                usage();
                return 2;
            }
        } else if (a == "--frontier") {
            frontier = num();
        } else if (a == "--wg") {
            wg = num();
        } else if (a == "--lds-depth") {
            depth = num();
        } else if (a == "--budget") {
            budget = num();
        } else if (a == "--round-time") {
            round_us = num();
        } else if (a == "--max-iters") {
            max_iters = num();
        } else if (a == "--dedup") {
            dedup = true;
        } else if (a == "--no-lookahead") {
            lookahead = false;
        } else if (a == "--autotune") {
            autotune = true;
        } else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else if (!a.empty() && a[0] == '-') {
            usage();
            return 2;
        } else {
            files.push_back(a);
        }
    }
    if ((files.empty() && tabdir.empty()) || max_iters < 1 || max_iters > SPZ_MAX_ITERS) {
        usage();
        return 2;
    }
    int width = heur == SPZ_H_PDB666 ? 5 : 0;
    std::vector<uint8_t> boards;
    for (auto &f : files)
        if (!read_problem_file(f, boards, width)) {
            std::fprintf(stderr, "spz_solve: %s: not a %s board (N*N integers 0..N*N-1, 0 = blank)\n", f.c_str(),
                         width == 5 ? "5x5" : width == 4 ? "4x4" : "4x4 or 5x5");
            return 1;
        }
    if (width == 0) width = 5;
    spz_default_config(&cfg, width, heur);
    cfg.mode = mode;
    cfg.frontier_target = frontier;
    cfg.workgroup_size = wg;
    cfg.lds_stack_depth = depth;
    cfg.round_budget = budget;
    cfg.lb_policy = lb;
    cfg.record_paths = path ? 1 : 0;
    cfg.lookahead = lookahead ? 1 : 0;
    cfg.round_time_us = round_us;
    cfg.dedup = dedup ? 1 : 0;
    cfg.resplit = resplit;
    if (lb == SPZ_LB_LEARNED && mlpfile.empty()) {
        std::fprintf(stderr, "spz_solve: --lb learned needs --mlp-params FILE\n");
        return 2;
    }
    spz_ctx *ctx = nullptr;
    int rc = spz_create(&ctx, &cfg);
    if (rc != SPZ_OK) {
        std::fprintf(stderr, "spz_create: %s\n", ctx ? spz_last_error(ctx) : "bad configuration");
        spz_destroy(ctx);
        return 1;
    }
    if (!mlpfile.empty()) {
        std::vector<float> prm((size_t)spz_mlp_param_count());
        FILE *f = std::fopen(mlpfile.c_str(), "rb");
        const size_t got = f ? std::fread(prm.data(), sizeof(float), prm.size(), f) : 0;
        if (f) std::fclose(f);
        if (got != prm.size()) {
            std::fprintf(stderr, "spz_solve: %s does not hold %zu float32 parameters\n", mlpfile.c_str(),
                         prm.size());
            spz_destroy(ctx);
            return 1;
        }
        rc = spz_mlp_set_params(ctx, prm.data());
        if (rc != SPZ_OK) {
            std::fprintf(stderr, "spz_solve: %s\n", spz_last_error(ctx));
            spz_destroy(ctx);
            return 1;
        }
    }
    if (heur == SPZ_H_PDB666 || !tabdir.empty()) {
        if (!tabin.empty()) {
            std::string a = tabin + "/pat24.1256712.tab", b = tabin + "/pat24.34891314.tab";
            rc = spz_load_pdb_tab(ctx, a.c_str(), b.c_str());
        } else {
            rc = spz_build_pdb(ctx, cache.empty() ? nullptr : cache.c_str());
        }
        if (rc == SPZ_OK && !tabdir.empty()) {
            std::string a = tabdir + "/pat24.1256712.tab", b = tabdir + "/pat24.34891314.tab";
            rc = spz_export_pdb_tab(ctx, a.c_str(), b.c_str());
        }
        if (rc != SPZ_OK) {
            std::fprintf(stderr, "spz_solve: %s\n", spz_last_error(ctx));
            spz_destroy(ctx);
            return 1;
        }
    }
    const int n = (int)files.size(), n2 = width * width;
    if (autotune && n > 0) {
        spz_config best;
        double ms = 0;
        int tried = 0;
        rc = spz_autotune(ctx, boards.data(), n, max_iters, 1, &best, &ms, &tried);
        if (rc != SPZ_OK) {
            std::fprintf(stderr, "spz_solve: autotune: %s\n", spz_last_error(ctx));
            spz_destroy(ctx);
            return 1;
        }
        std::fprintf(stderr, "autotune: wg=%d frontier=%d lds_depth=%d budget=%d (%.1f ms, %d candidates)\n",
                     best.workgroup_size, best.frontier_target, best.lds_stack_depth, best.round_budget, ms, tried);
    }
    std::vector<spz_result> res((size_t)n);
    std::vector<uint64_t> ic((size_t)n * max_iters);
    std::vector<double> secs((size_t)n);
    const int max_len = 256;
    std::vector<uint8_t> moves((size_t)n * max_len);
    std::vector<int32_t> lens((size_t)n, -1);
    if (batch) {
        auto t0 = std::chrono::steady_clock::now();
        rc = spz_solve_batch(ctx, boards.data(), n, max_iters, res.data(), ic.data());
        double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (auto &x : secs) x = s / n;
        if (rc == SPZ_OK && path) rc = spz_get_paths(ctx, moves.data(), max_len, lens.data());
    } else {
        for (int i = 0; i < n && rc == SPZ_OK; ++i) {
            auto t0 = std::chrono::steady_clock::now();
            rc = spz_solve_batch(ctx, boards.data() + (size_t)i * n2, 1, max_iters, &res[i],
                                 ic.data() + (size_t)i * max_iters);
            secs[i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (rc == SPZ_OK && path) rc = spz_get_paths(ctx, moves.data() + (size_t)i * max_len, max_len, &lens[i]);
        }
    }
    if (rc != SPZ_OK) {
        std::fprintf(stderr, "spz_solve: %s\n", spz_last_error(ctx));
        spz_destroy(ctx);
        return 1;
    }
    FILE *cf = csv.empty() ? nullptr : std::fopen(csv.c_str(), "w");
    for (int i = 0; i < n; ++i) {
        if (iters)
            for (int k = 0; k < res[i].n_iters; ++k)
                std::printf("f_limit=%3d, expanded nodes = %llu\n", res[i].h0 + 2 * k,
                            (unsigned long long)ic[(size_t)i * max_iters + k]);
        std::printf("%s %d %llu\n", files[i].c_str(), res[i].opt_len, (unsigned long long)res[i].total_nodes);
        if (path && lens[i] >= 0) {
            // the reference CPU solver's stack_dump line (src_horie/idas_cpu.c:50-57)
            std::printf("len=%d: ", lens[i]);
            for (int j = 0; j < lens[i] && j < max_len; ++j) std::printf("%d ", moves[(size_t)i * max_len + j]);
            std::printf("\n");
        }
        if (cf) std::fprintf(cf, "%f\n", secs[i]);
    }
    if (cf) std::fclose(cf);
    spz_destroy(ctx);
    return 0;
}
