// Cp -- clustering coefficient of every .csr in a directory (drop-in for
// U/Cp/Cp/Cp_main.cpp:28-215 and W/Cp/Cp/Cp_main.cpp).
//
//   Cp dir_for_csr num_of_random_networks {g|n|k|gn|gk|nk|gnk}
//   Cp --weighted dir_for_csr num_of_random_networks [1|2] {g|n|...}
//     1 = Barrat, 2 = Onnela (default, weights divided by the largest weight)
// Outputs: <file>_cp.nm (nodal, 'n'), <file>_cp.txt (the network's mean Cp,
// then one line per random network, 'g'), <file>_gamma.txt (mean Cp over the
// mean of the random networks' mean Cp, 'k').  Random networks are degree-
// preserving rewirings (pg_rewire) of the network, weights travelling with
// their edge (U Cp_main.cpp:150-195, W Cp_main.cpp:172-213); network l is
// seeded with PAGANI_SEED + l (wall clock when unset).
#include <iomanip>
#include <iostream>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    int a0 = 1;
    bool weighted = false;
    if (a0 < argc && !std::strcmp(argv[a0], "--weighted")) {
        weighted = true;
        a0++;
    }
    if (argc - a0 < 3) {
        std::cerr << "Input format: Cp [--weighted] dir_for_csr num_of_random_networks [1|2] parameter_type "
                     "\nFor example: Cp /data 0 n" << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[a0]);
    const int nrand = std::atoi(argv[a0 + 1]);
    int kind = PG_CP_ONNELA;
    int q = a0 + 2;
    if (weighted && argc - q >= 2) {
        if (argv[q][0] == '1') kind = PG_CP_BARRAT;
        else if (argv[q][0] == '2') kind = PG_CP_ONNELA;
        else {
            std::cerr << "Cp algorithm type invalid, just 1 or 2 can be handled\n";
            return 1;
        }
        q++;
    }
    const std::string sel = argv[q];
    const bool global = sel.find('g') != std::string::npos;
    const bool nodal = sel.find('n') != std::string::npos;
    const bool gamma = sel.find('k') != std::string::npos;
    if (nrand < 0) {
        std::cerr << "num_of_random_networks must be >= 0\n";
        return 1;
    }
    if (gamma && nrand == 0) std::cerr << "gamma needs random networks (num_of_random_networks > 0): writing nan\n";
    const uint64_t seed = pgtool::random_seed();
    if (nrand > 0) std::cout << "random network seed " << seed << std::endl;
    const auto files = pgio::list_dir(dir, ".csr");
    std::cout << files.size() << " files to be processed." << std::endl;
    pgtool::Ctx ctx;
    pgtool::Timer total;
    for (const auto &f : files) {
        const std::string a = pgio::join(dir, f);
        std::cout << "\ncalculating Cp for " << a << " ..." << std::endl;
        pgio::Csr g;
        std::string err;
        if (!pgio::read_csr(a, weighted, g, err)) {
            std::cerr << err << "\n";
            return 2;
        }
        std::vector<float> out(g.n);
        // mean Cp of a network; the per-node values into out
        auto mean_cp = [&](const pgio::Csr &h) {
            double mean = 0.0;
            if (weighted) {
                std::vector<double> cp(h.n);
                ctx.check(pg_cp_weighted(ctx.c, h.row.data(), h.col.data(), h.val.data(), h.n,
                                         (int64_t)h.col.size(), kind, 0, cp.data(), &mean),
                          "pg_cp_weighted");
                for (int i = 0; i < h.n; i++) out[i] = (float)cp[i];
            } else {
                ctx.check(pg_cp_binary(ctx.c, h.row.data(), h.col.data(), h.n, (int64_t)h.col.size(), out.data(),
                                       &mean),
                          "pg_cp_binary");
            }
            return mean;
        };
        pgtool::Timer tm;
        const double mean = mean_cp(g);
        const double secs = tm.seconds();
        std::cout << "average Cp:\t" << std::setprecision(6) << mean << std::endl;
        std::cout << "Elapsed time: " << secs << " s." << std::endl;
        pgio::time_log(std::string("Cp\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
        if (nodal) {
            const std::string x = pgio::strip_ext(a) + "_cp.nm";
            std::cout << "Save Clustering Coefficient for each node as " << x << std::endl;
            if (!pgio::write_nm(x, out.data(), g.n)) return 2;
        }
        std::vector<double> means{mean};
        if (nrand > 0) {
            std::cout << "Calculating Cp for random networks..." << std::endl;
            pgtool::Timer tr;
            for (int l = 0; l < nrand; l++) means.push_back(mean_cp(pgtool::rewired(ctx, g, seed + (uint64_t)l)));
            const double rs = tr.seconds();
            std::cout << "Elapsed time: " << rs << " s." << std::endl;
            pgio::time_log(std::string("Cp\tRandomGPU\t(Maslov+kernel) time\t") + std::to_string(rs) + "s");
        }
        double rmean = 0.0;
        for (int l = 1; l <= nrand; l++) rmean += means[l];
        rmean /= nrand;  // nan when there are no random networks, as the reference
        if (global) {
            const std::string x = pgio::strip_ext(a) + "_cp.txt";
            std::cout << "Save average Cp for the brain network and random networks as " << x << std::endl;
            std::ofstream fo(x);
            for (double v : means) fo << std::setprecision(6) << v << std::endl;
        }
        if (gamma) {
            const std::string x = pgio::strip_ext(a) + "_gamma.txt";
            std::cout << "Save corresponding gamma for the brain network and random networks as " << x << std::endl;
            std::ofstream fo(x);
            fo << std::setprecision(6) << mean / rmean << std::endl;
        }
    }
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
