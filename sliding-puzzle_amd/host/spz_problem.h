// spz_problem.h -- reading a problem file of the reference's solvers: N*N
// whitespace-separated integers, row-major, 0 = blank (src/24puzzle.cc:72-87,
// src_horie/idas_cpu.c:336-360).  Plain C++: the CLI uses it, and the CPU
// tests build it on its own under the address and undefined-behaviour
// sanitizers.
#pragma once
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

// Appends the board of `path` to b.  The width comes from the number of
// integers (25 or more: 5x5, 16 or more: 4x4; integers past the board are
// ignored, as the reference reads only the board) and must equal `width`
// unless that is 0, which the first file sets.  Cells outside 0..24 fail
// here; whether the cells form a solvable permutation is the library's test.
inline bool read_problem_file(const std::string &path, std::vector<uint8_t> &b, int &width) {
    std::ifstream f(path);
    if (!f) return false;
    std::vector<long long> v;
    long long x;
    while (v.size() < 25 && f >> x) v.push_back(x);
    const int w = v.size() >= 25 ? 5 : (v.size() >= 16 ? 4 : 0);
    if (w == 0) return false;
    if (width != 0 && w != width) return false;
    for (int i = 0; i < w * w; ++i)
        if (v[(size_t)i] < 0 || v[(size_t)i] >= w * w) return false;
    width = w;
    for (int i = 0; i < w * w; ++i) b.push_back((uint8_t)v[(size_t)i]);
    return true;
}
