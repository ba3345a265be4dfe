// CUBFS_Lp -- binary nodal efficiency and characteristic path length from BFS
// hop counts (drop-in for U/BFS_MulCPU/BFS_MulCPU/MUBFS_LP_main.cpp:16-169).
//
//   CUBFS_Lp dir_for_csr num_of_random_networks {g|n|b}[l]
// Reads every <dir>/*.csr (row and column arrays; weights, if present, are
// ignored as in the reference).  Outputs: <file>_eff.nm ('n' or 'b'),
// <file>_Lp.txt ('g' or 'b': the network's Lp, then one line per random
// network, U MUBFS_LP_main.cpp:125-149) and, an addition to the reference's
// binary tool, <file>_lambda.txt ('l': Lp over the mean Lp of the random
// networks, as W CUBFW_Lp_main.cpp:242-249).  Random networks are degree-
// preserving rewirings (pg_rewire); network l is seeded with PAGANI_SEED + l
// (wall clock when unset).
#include <iomanip>
#include <iostream>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::cerr << "Input format: CUBFS_Lp dir_for_csr num_of_random_networks parameter_type \n"
                     "For example: CUBFS_Lp /data 0 b" << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[1]);
    const int nrand = std::atoi(argv[2]);
    if (nrand < 0) {
        std::cerr << "num_of_random_networks must be >= 0\n";
        return 1;
    }
    const char sel = argv[3][0];
    const bool nodal = sel == 'n' || sel == 'b';
    const bool global = sel == 'g' || sel == 'b';
    const bool lambda = std::strchr(argv[3], 'l') != nullptr;
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
        pgio::Csr g;
        std::string err;
        if (!pgio::read_csr(a, false, g, err)) {
            std::cerr << err << "\n";
            return 2;
        }
        std::vector<float> eff(g.n);
        double lp = 0.0;
        int isolated = 0, unconnected = 0;
        pgtool::Timer tm;
        ctx.check(pg_bfs_efficiency(ctx.c, g.row.data(), g.col.data(), g.n, (int64_t)g.col.size(), eff.data(),
                                    &lp, &isolated, &unconnected),
                  "pg_bfs_efficiency");
        const double secs = tm.seconds();
        if (isolated > 0) std::cout << "\nisolated vertices number : " << isolated << std::endl;
        else if (unconnected) std::cout << "\nThe network is unconnected. " << std::endl;
        pgio::time_log(std::string("BFS_Lp\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
        if (nodal) {
            const std::string x = pgio::strip_ext(a) + "_eff.nm";
            std::cout << "Save nodal efficiency for each node as " << x << std::endl;
            if (!pgio::write_nm(x, eff.data(), g.n)) return 2;
        }
        std::vector<double> lps{lp};
        if (nrand > 0) {
            std::cout << "Calculating Lp for random networks..." << std::endl;
            pgtool::Timer tr;
            std::vector<float> reff(g.n);
            for (int l = 0; l < nrand; l++) {
                const pgio::Csr h = pgtool::rewired(ctx, g, seed + (uint64_t)l);
                double rl = 0.0;
                ctx.check(pg_bfs_efficiency(ctx.c, h.row.data(), h.col.data(), h.n, (int64_t)h.col.size(),
                                            reff.data(), &rl, nullptr, nullptr),
                          "pg_bfs_efficiency");
                lps.push_back(rl);
            }
            const double rs = tr.seconds();
            std::cout << "Elapsed time: " << rs << " s." << std::endl;
            pgio::time_log(std::string("BFS_Lp\tRandomGPU\t(Maslov+kernel) time\t") + std::to_string(rs) + "s");
        }
        if (global) {
            std::cout << "average Lp:\t" << std::setprecision(6) << lp << "\nElapsed time: " << secs << " s."
                      << std::endl;
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
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
