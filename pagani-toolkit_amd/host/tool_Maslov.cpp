// Maslov -- one degree-preserving random rewiring of a network (drop-in for
// W/Maslov/Maslov/Maslov_main.cpp:11-81).
//
//   Maslov network.csr
// Writes network_Maslov.csr: the same row array, the rewired columns and, for
// a weighted file, each edge's weight carried with it (Maslov_weighted_1).
// Seeded with PAGANI_SEED (wall clock when unset).
#include <iostream>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::cerr << "1 argument is required.\nInput format example: Maslov /data/net_0.06.csr" << std::endl;
        return 1;
    }
    const std::string in = argv[1];
    if (in.size() < 5 || in.compare(in.size() - 4, 4, ".csr") != 0) {
        std::cerr << "expected a .csr file: " << in << std::endl;
        return 1;
    }
    const std::string out = in.substr(0, in.size() - 4) + "_Maslov.csr";
    const bool weighted = pgtool::csr_has_values(in);
    std::cout << "Reading CSR file..." << std::endl;
    pgio::Csr g;
    std::string err;
    if (!pgio::read_csr(in, weighted, g, err)) {
        std::cerr << err << std::endl;
        return 2;
    }
    pgtool::Ctx ctx;
    const uint64_t seed = pgtool::random_seed();
    std::cout << "Rewiring... (seed " << seed << ")" << std::endl;
    pgtool::Timer tm;
    const pgio::Csr r = pgtool::rewired(ctx, g, seed);
    std::cout << "Elapsed time:\t" << tm.seconds() << "s" << std::endl;
    if (!pgio::write_csr(out, r.row.data(), r.n, r.col.data(), (int64_t)r.col.size(),
                         r.val.empty() ? nullptr : r.val.data())) {
        std::cerr << "Can't open\t" << out << std::endl;
        return 2;
    }
    std::cout << "===========================================================" << std::endl;
    return 0;
}
