// Prints what read_problem_file (the CLI's problem-file reader) makes of
// each file named on the command line: "w c0 c1 ..." or "REJECT".  Built by
// tests/test_problem_reader.py to compare the reader with the Python
// mirror's read_problem.
#include "../../sliding-puzzle_amd/host/spz_problem.h"

#include <cstdio>

int main(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
        std::vector<uint8_t> b;
        int w = 0;
        if (!read_problem_file(argv[i], b, w)) {
            std::printf("REJECT\n");
            continue;
        }
        std::printf("%d", w);
        for (uint8_t c : b) std::printf(" %d", (int)c);
        std::printf("\n");
    }
    return 0;
}
