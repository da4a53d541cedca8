/* The CPU oracle (oracle/spz_oracle.c) run on small inputs under the
 * address + undefined-behaviour sanitizers, built by
 * tests/test_host_sanitizers.py: the checker every parity test relies on
 * must itself stay inside its buffers.  Covers the generators (widths
 * 2..5), solvability and Manhattan, IDA* in both modes, the root split with
 * and without duplicate pruning, PDB rank/unrank, the small-pattern
 * builder, and the PDB lookups and search over tables of full size (all
 * zeros, then all ones), with a few invariants checked on the way.  Prints
 * "OK" and exits 0 when all hold. */
#include "../../oracle/spz_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int failures = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            ++failures;                                                       \
        }                                                                     \
    } while (0)

static int same_run(int n, const uint64_t *a, const int *ta, const uint64_t *b, const int *tb)
{
    for (int i = 0; i < n; ++i)
        if (a[i] != b[i] || ta[i] != tb[i])
            return 0;
    return 1;
}

int main(void)
{
    uint8_t b[25];
    uint64_t c0[40], c1[40];
    int t0[40], t1[40], n0 = 0, n1 = 0;

    for (int w = 2; w <= 5; ++w)
        for (int i = 0; i < 40; ++i) {
            orc_walk(w, 3 * i, 100 + (uint64_t)i, b);
            CHECK(orc_is_permutation(w, b) && orc_is_solvable(w, b));
            CHECK(orc_manhattan(w, b) <= 3 * i);
            orc_random_solvable(w, 200 + (uint64_t)i, b);
            CHECK(orc_is_permutation(w, b) && orc_is_solvable(w, b));
        }

    /* IDA*, both modes, and the split against it (8- and 15-puzzle) */
    for (int i = 0; i < 12; ++i) {
        const int w = i < 6 ? 3 : 4;
        orc_walk(w, 10 + 2 * i, 300 + (uint64_t)i, b);
        const int r = orc_ida(w, b, ORC_H_MANHATTAN, ORC_MODE_ALL, 40, c0, t0, &n0);
        CHECK(r >= 0 && n0 > 0);
        const int rf = orc_ida(w, b, ORC_H_MANHATTAN, ORC_MODE_FIRST, 40, c1, t1, &n1);
        CHECK(rf == r && n1 == n0);
        for (int target = 1; target <= 4096; target *= 8) {
            CHECK(orc_ida_split(w, b, ORC_H_MANHATTAN, target, 24, 0, 40, c1, t1, &n1) == r);
            CHECK(n1 == n0 && same_run(n0, c0, t0, c1, t1));
            CHECK(orc_ida_split(w, b, ORC_H_MANHATTAN, target, 24, 1, 40, c1, t1, &n1) == r);
            CHECK(n1 == n0);
            for (int k = 0; k < n0 && k < n1; ++k)
                CHECK(c1[k] <= c0[k] && t1[k] == t0[k]);
        }
        CHECK(orc_ida_split(w, b, ORC_H_MANHATTAN, 1 << 20, 3, 1, 40, c1, t1, &n1) == r);
    }
    /* bad input is refused, not read */
    memset(b, 0, sizeof b);
    CHECK(orc_ida(4, b, ORC_H_MANHATTAN, ORC_MODE_ALL, 40, c0, t0, &n0) == -2);
    CHECK(orc_ida_split(4, b, ORC_H_MANHATTAN, 16, 24, 1, 40, c0, t0, &n0) == -2);

    /* PDB rank/unrank over random 6-cell placements */
    for (int i = 0; i < 2000; ++i) {
        uint8_t cells[6], back[6];
        uint64_t x = orc_rng_u64(7, i);
        int used[25] = {0};
        for (int j = 0; j < 6; ++j) {
            int cell = (int)(x % 25);
            x = x / 25 + (uint64_t)j * 2654435761u;
            while (used[cell])
                cell = (cell + 1) % 25;
            used[cell] = 1;
            cells[j] = (uint8_t)cell;
        }
        const uint32_t rk = orc_pdb_rank(cells);
        CHECK(rk < ORC_PDB_ENTRIES);
        orc_pdb_unrank(rk, back);
        CHECK(memcmp(cells, back, 6) == 0);
    }

    /* small-pattern builder: 8-puzzle, 3 tiles, every goal kind */
    {
        const uint8_t tiles[3] = {1, 2, 3};
        const size_t n = 9 * 8 * 7;
        uint8_t *tab = (uint8_t *)malloc(n);
        for (int g = 0; g <= 1; ++g) {
            CHECK(orc_pdb_build_small(3, 3, tiles, g, tab) >= 0);
            const uint8_t home[3] = {1, 2, 3};
            CHECK(tab[orc_rank_small(9, 3, home)] == 0);
        }
        free(tab);
    }

    /* PDB lookups over zero tables of the full size, then a tiny search */
    {
        uint8_t *z0 = (uint8_t *)calloc(ORC_PDB_ENTRIES, 1), *z1 = (uint8_t *)calloc(ORC_PDB_ENTRIES, 1);
        orc_pdb_set_tables(z0, z1);
        int out[8];
        for (int i = 0; i < 200; ++i) {
            orc_random_solvable(5, 400 + (uint64_t)i, b);
            CHECK(orc_pdb_h(b) == 0);
            orc_pdb_lookups(b, out);
            for (int k = 0; k < 8; ++k)
                CHECK(out[k] == 0);
        }
        for (int k = 0; k < 8; ++k) {
            int tab;
            uint8_t map[25], tl[6];
            orc_pdb_lookup_spec(k, &tab, map, tl);
            CHECK(tab == 0 || tab == 1);
        }
        /* tables of ones: h is 4 everywhere, so no board is a goal (the
         * goal test is h == 0) and three iterations search depth <= 4 */
        memset(z0, 1, ORC_PDB_ENTRIES);
        memset(z1, 1, ORC_PDB_ENTRIES);
        orc_walk(5, 6, 500, b);
        CHECK(orc_pdb_h(b) == 4);
        const int r = orc_ida(5, b, ORC_H_PDB, ORC_MODE_ALL, 3, c0, t0, &n0);
        CHECK(r == -1 && n0 == 3 && t0[2] == 8);
        CHECK(orc_ida_split(5, b, ORC_H_PDB, 64, 24, 0, 3, c1, t1, &n1) == -1);
        CHECK(n1 == 3 && same_run(3, c0, t0, c1, t1));
        CHECK(orc_ida_split(5, b, ORC_H_PDB, 64, 24, 1, 3, c1, t1, &n1) == -1);
        free(z0);
        free(z1);
    }

    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("OK\n");
    return 0;
}
