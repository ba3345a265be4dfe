// CUBFW_Lp -- all-pairs shortest paths by blocked Floyd-Warshall, then nodal
// efficiency and characteristic path length (drop-in for
// W/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp_main.cpp:18-270).  Distance d = 1/w for
// weighted files (W FormBlock.cpp:64), hop count for binary ones.
//
//   CUBFW_Lp dir_for_csr num_of_random_networks {g|n|l|gn|gl|nl|gnl}
// Outputs: <file>_eff.nm ('n'), <file>_Lp.txt ('g': the network's Lp, then
// one line per random network), <file>_lambda.txt ('l': Lp over the mean Lp of
// the random networks), as W CUBFW_Lp_main.cpp:187-249.  Random networks are
// degree-preserving rewirings (pg_rewire) with each weight travelling with
// its edge (Maslov_weighted_1); network l is seeded with PAGANI_SEED + l (wall
// clock when unset).
#include <iomanip>
#include <iostream>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::cerr << "Input format: CUBFW_Lp dir_for_csr num_of_random_networks parameter_type \n"
                     "For example: CUBFW_Lp /data 0 gn" << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[1]);
    const int nrand = std::atoi(argv[2]);
    const std::string sel = argv[3];
    if (nrand < 0) {
        std::cerr << "num_of_random_networks must be >= 0\n";
        return 1;
    }
    const bool global = sel.find('g') != std::string::npos;
    const bool nodal = sel.find('n') != std::string::npos;
    const bool lambda = sel.find('l') != std::string::npos;
    if (lambda && nrand == 0) std::cerr << "lambda needs random networks (num_of_random_networks > 0): writing nan\n";
    const uint64_t seed = pgtool::random_seed();
    if (nrand > 0) std::cout << "random network seed " << seed << std::endl;
    const auto files = pgio::list_dir(dir, ".csr");
    std::cout << files.size() << " files to be processed." << std::endl;
    pgtool::Ctx ctx;
    pgtool::Timer total;
    for (const auto &f : files) {
        const std::string a = pgio::join(dir, f);
        std::cout << "\ncalculating Lp for " << a << " ..." << std::endl;
        const bool weighted = pgtool::csr_has_values(a);
        pgio::Csr g;
        std::string err;
        if (!pgio::read_csr(a, weighted, g, err)) {
            std::cerr << err << "\n";
            return 2;
        }
        std::vector<float> d((size_t)g.n * g.n), eff(g.n);
        // Lp of a network (nodal efficiency into eff)
        auto fw_lp = [&](const pgio::Csr &h, int *isolated) {
            double lp = 0.0;
            ctx.check(pg_fw_init(ctx.c, h.row.data(), h.col.data(), weighted ? h.val.data() : nullptr, h.n,
                                 (int64_t)h.col.size(), weighted ? 0 : 2, d.data()),
                      "pg_fw_init");
            ctx.check(pg_fw(ctx.c, d.data(), h.n, 0), "pg_fw");
            ctx.check(pg_fw_efficiency(ctx.c, d.data(), h.n, eff.data(), &lp, isolated), "pg_fw_efficiency");
            return lp;
        };
        int isolated = 0;
        pgtool::Timer tm;
        const double lp = fw_lp(g, &isolated);
        const double secs = tm.seconds();
        if (isolated > 0) std::cout << "\nisolated vertices number : " << isolated << std::endl;
        std::cout << "average Lp:\t" << std::setprecision(6) << lp << "\nElapsed time: " << secs << " s." << std::endl;
        pgio::time_log(std::string("BFW_Lp\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
        if (nodal) {
            const std::string x = pgio::strip_ext(a) + "_eff.nm";
            std::cout << "Save nodal efficiency for each node as " << x << std::endl;
            if (!pgio::write_nm(x, eff.data(), g.n)) return 2;
        }
        std::vector<double> lps{lp};
        if (nrand > 0) {
            std::cout << "Calculating Lp for random networks..." << std::endl;
            pgtool::Timer tr;
            for (int l = 0; l < nrand; l++) lps.push_back(fw_lp(pgtool::rewired(ctx, g, seed + (uint64_t)l), nullptr));
            const double rs = tr.seconds();
            std::cout << "Elapsed time: " << rs << " s." << std::endl;
            pgio::time_log(std::string("BFW_Lp\tRandomGPU\t(Maslov+kernel) time\t") + std::to_string(rs) + "s");
        }
        if (global) {
            const std::string x = pgio::strip_ext(a) + "_Lp.txt";
            std::ofstream fo(x);
            for (double v : lps) fo << std::setprecision(6) << v << std::endl;
            std::cout << "Save Lp for the brain network and random networks as " << x << std::endl;
        }
        if (lambda) {
            double rmean = 0.0;
            for (int l = 1; l <= nrand; l++) rmean += lps[l];
            rmean /= nrand;  // nan without random networks
            const std::string x = pgio::strip_ext(a) + "_lambda.txt";
            std::cout << "Save corresponding lambda for the brain network and random networks as " << x << std::endl;
            std::ofstream fo(x);
            fo << std::setprecision(6) << lp / rmean << std::endl;
        }
    }
    std::cout << "Total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
