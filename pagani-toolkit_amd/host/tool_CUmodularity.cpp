// CUmodularity -- Newman leading-eigenvector modules of every .csr in a
// directory (drop-in for W/Newman_Modularity_GPU/CUmodularity/
// modularity_GPU.cu:527-720).  Output: <file>_modu.nm (float module labels
// 1..K) and <file>_modu.txt (module count and Q).
//
//   CUmodularity [--strength] dir_for_csr num_of_random_networks
// Default degrees are binary with weighted adjacency, as the reference tool;
// --strength uses strengths and the total weight instead.
// num_of_random_networks must be 0 (no Maslov rewiring in this build).
#include <iomanip>
#include <iostream>
#include <set>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    int a0 = 1;
    int kmode = PG_MOD_BINARY_K;
    if (a0 < argc && !std::strcmp(argv[a0], "--strength")) {
        kmode = PG_MOD_STRENGTH;
        a0++;
    }
    if (argc - a0 != 2) {
        std::cerr << "Input format: CUmodularity [--strength] dir_for_csr num_of_random_networks \n"
                     "For example: CUmodularity /data 0" << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[a0]);
    if (std::atoi(argv[a0 + 1]) != 0) {
        std::cerr << "random-network comparison (num_of_random_networks > 0) is not available\n";
        return 1;
    }
    const auto files = pgio::list_dir(dir, ".csr");
    std::cout << files.size() << " files to be processed." << std::endl;
    pgtool::Ctx ctx;
    pgtool::Timer total;
    for (const auto &f : files) {
        const std::string a = pgio::join(dir, f);
        std::cout << "\nModular analysis for " << a << " ..." << std::endl;
        const bool weighted = pgtool::csr_has_values(a);
        pgio::Csr g;
        std::string err;
        if (!pgio::read_csr(a, weighted, g, err)) {
            std::cerr << err << "\n";
            return 2;
        }
        std::vector<int32_t> lab(g.n);
        double q = 0.0;
        int32_t rounds = 0;
        pgtool::Timer tm;
        ctx.check(pg_modularity(ctx.c, g.row.data(), g.col.data(), weighted ? g.val.data() : nullptr, g.n,
                                (int64_t)g.col.size(), kmode, lab.data(), &q, nullptr, nullptr, &rounds),
                  "pg_modularity");
        const double secs = tm.seconds();
        const std::set<int32_t> mods(lab.begin(), lab.end());
        std::cout << "Number of Modules: " << mods.size() << ",\tQ=" << q << "\nElapsed time:   " << secs << "s"
                  << std::endl;
        {
            std::ofstream fo(pgio::strip_ext(a) + "_modu.txt");
            fo << "Number of Modules: " << mods.size() << ",\tQ=" << q << "\nBisection rounds: " << rounds
               << "\nElapsed time:   " << secs << "s" << std::endl;
        }
        std::vector<float> out(g.n);
        for (int i = 0; i < g.n; i++) out[i] = (float)lab[i];
        const std::string x = pgio::strip_ext(a) + "_modu.nm";
        std::cout << "Save partition results as " << x << std::endl;
        if (!pgio::write_nm(x, out.data(), g.n)) return 2;
        pgio::time_log(std::string("Modularity\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
    }
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
