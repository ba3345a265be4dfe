// CUmodularity -- Newman leading-eigenvector modules of every .csr in a
// directory (drop-in for W/Newman_Modularity_GPU/CUmodularity/
// modularity_GPU.cu:527-720).  Output: <file>_modu.nm (float module labels
// 1..K) and <file>_modu.txt (module count and Q).
//
//   CUmodularity [--strength] [--rayleigh[=tol]] dir_for_csr num_of_random_networks
// Default degrees are binary with weighted adjacency, as the reference tool;
// --strength uses strengths and the total weight instead.  --rayleigh stops
// each power iteration when the Rayleigh quotient changes by at most tol
// (default 1e-8) relative, instead of the reference's vector test.  Networks
// of 20000 nodes or more print one progress line per wave of rounds.
// With num_of_random_networks > 0 each network's degree-preserving rewirings
// (pg_rewire, weights carried with their edges; network l seeded with
// PAGANI_SEED + l, wall clock when unset) are partitioned too, as the
// reference does (modularity_GPU.cu:665-684); their module counts and Q are
// printed and appended to _modu.txt (the reference only prints them).
#include <iomanip>
#include <iostream>
#include <set>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    int a0 = 1;
    pg_mod_opts mo;
    std::memset(&mo, 0, sizeof mo);
    mo.kmode = PG_MOD_BINARY_K;
    mo.stop = PG_MOD_STOP_VECTOR;
    for (; a0 < argc && argv[a0][0] == '-' && argv[a0][1] == '-'; a0++) {
        if (!std::strcmp(argv[a0], "--strength")) {
            mo.kmode = PG_MOD_STRENGTH;
        } else if (!std::strncmp(argv[a0], "--rayleigh", 10) && (argv[a0][10] == 0 || argv[a0][10] == '=')) {
            mo.stop = PG_MOD_STOP_RAYLEIGH;
            if (argv[a0][10] == '=') mo.tol = std::atof(argv[a0] + 11);
        } else {
            std::cerr << "unknown option " << argv[a0] << "\n";
            return 1;
        }
    }
    if (argc - a0 != 2) {
        std::cerr << "Input format: CUmodularity [--strength] [--rayleigh[=tol]] dir_for_csr num_of_random_networks \n"
                     "For example: CUmodularity /data 0" << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[a0]);
    const int nrand = std::atoi(argv[a0 + 1]);
    if (nrand < 0) {
        std::cerr << "num_of_random_networks must be >= 0\n";
        return 1;
    }
    const uint64_t seed = pgtool::random_seed();
    if (nrand > 0) std::cout << "random network seed " << seed << std::endl;
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
        int32_t rounds = 0;
        // modules of a network into lab; returns Q
        auto partition = [&](const pgio::Csr &h, int32_t *rnd) {
            double q = 0.0;
            pg_mod_opts o = mo;
            o.progress = h.n >= 20000;
            ctx.check(pg_modularity_ex(ctx.c, h.row.data(), h.col.data(), weighted ? h.val.data() : nullptr, h.n,
                                       (int64_t)h.col.size(), &o, lab.data(), &q, nullptr, nullptr, rnd),
                      "pg_modularity_ex");
            return q;
        };
        pgtool::Timer tm;
        const double q = partition(g, &rounds);
        const double secs = tm.seconds();
        const std::set<int32_t> mods(lab.begin(), lab.end());
        std::cout << "Number of Modules: " << mods.size() << ",\tQ=" << q << "\nElapsed time:   " << secs << "s"
                  << std::endl;
        std::vector<float> out(g.n);
        for (int i = 0; i < g.n; i++) out[i] = (float)lab[i];
        std::ofstream fo(pgio::strip_ext(a) + "_modu.txt");
        fo << "Number of Modules: " << mods.size() << ",\tQ=" << q << "\nBisection rounds: " << rounds
           << "\nElapsed time:   " << secs << "s" << std::endl;
        if (nrand > 0) {
            std::cout << "Modular analysis for random networks..." << std::endl;
            pgtool::Timer tr;
            for (int l = 0; l < nrand; l++) {
                const double rq = partition(pgtool::rewired(ctx, g, seed + (uint64_t)l), nullptr);
                const std::set<int32_t> rm(lab.begin(), lab.end());
                std::cout << "\nNumber of Modules: " << rm.size() << ",\tQ=" << rq << std::endl;
                fo << "Random network " << l + 1 << ": Number of Modules: " << rm.size() << ",\tQ=" << rq
                   << std::endl;
            }
            pgio::time_log(std::string("Modularity\tRandomGPU\t(Maslov+kernel) time\t") +
                           std::to_string(tr.seconds()) + "s");
        }
        const std::string x = pgio::strip_ext(a) + "_modu.nm";
        std::cout << "Save partition results as " << x << std::endl;
        if (!pgio::write_nm(x, out.data(), g.n)) return 2;
        pgio::time_log(std::string("Modularity\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
    }
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
