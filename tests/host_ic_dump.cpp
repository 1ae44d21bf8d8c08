// Writes the driver's primitive rows for one case file and resolution
// (test helper: the host half of the initial state, checked on the CPU
// against the reference's initial state through the oracle's prim_to_cons).
//
//     host_ic_dump <case file> <resolution index> <output file>
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "config.hpp"
#include "initial_conditions.hpp"

int main(int argc, char** argv)
{
    if (argc != 4) return 2;
    try {
        const auto c = dimhost::load_case<DIM_>(argv[1]);
        const auto N = c.N_list.at(std::atoi(argv[2]));
        const dimhost::InitialState<DIM_> ic(c, N, (int)c.materials.size());
        const auto rows = ic.rows(3);
        FILE* f = std::fopen(argv[3], "wb");
        if (!f) return 3;
        const size_t n = std::fwrite(rows.data(), sizeof(double), rows.size(), f);
        std::fclose(f);
        return n == rows.size() ? 0 : 4;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return 1;
    }
}
