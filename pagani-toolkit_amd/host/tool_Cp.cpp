// Cp -- clustering coefficient of every .csr in a directory (drop-in for
// U/Cp/Cp/Cp_main.cpp:28-215 and W/Cp/Cp/Cp_main.cpp).
//
//   Cp dir_for_csr num_of_random_networks {g|n|k|gn|gk|nk|gnk}
//   Cp --weighted dir_for_csr num_of_random_networks [1|2] {g|n|...}
//     1 = Barrat, 2 = Onnela (default, weights divided by the largest weight)
// Outputs: <file>_cp.nm (nodal, 'n'), <file>_cp.txt (network mean, 'g').
// Random (Maslov-rewired) reference networks are not generated: the rewiring
// is seeded by wall-clock time in the reference and is outside this build, so
// num_of_random_networks must be 0 and 'k' (gamma) is rejected.
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
    if (sel.find('k') != std::string::npos || nrand != 0) {
        std::cerr << "random-network comparison (gamma, num_of_random_networks > 0) is not available\n";
        return 1;
    }
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
        double mean = 0.0;
        pgtool::Timer tm;
        if (weighted) {
            std::vector<double> cp(g.n);
            ctx.check(pg_cp_weighted(ctx.c, g.row.data(), g.col.data(), g.val.data(), g.n, (int64_t)g.col.size(),
                                     kind, 0, cp.data(), &mean),
                      "pg_cp_weighted");
            for (int i = 0; i < g.n; i++) out[i] = (float)cp[i];
        } else {
            ctx.check(pg_cp_binary(ctx.c, g.row.data(), g.col.data(), g.n, (int64_t)g.col.size(), out.data(), &mean),
                      "pg_cp_binary");
        }
        const double secs = tm.seconds();
        std::cout << "average Cp:\t" << std::setprecision(6) << mean << std::endl;
        std::cout << "Elapsed time: " << secs << " s." << std::endl;
        pgio::time_log(std::string("Cp\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
        if (nodal) {
            const std::string x = pgio::strip_ext(a) + "_cp.nm";
            std::cout << "Save Clustering Coefficient for each node as " << x << std::endl;
            if (!pgio::write_nm(x, out.data(), g.n)) return 2;
        }
        if (global) {
            std::ofstream fo(pgio::strip_ext(a) + "_cp.txt");
            fo << std::setprecision(6) << mean << std::endl;
        }
    }
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
