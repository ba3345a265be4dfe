// CUBFW_Lp -- all-pairs shortest paths by blocked Floyd-Warshall, then nodal
// efficiency and characteristic path length (drop-in for
// W/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp_main.cpp:18-270).  Distance d = 1/w for
// weighted files (W FormBlock.cpp:64), hop count for binary ones.
//
//   CUBFW_Lp dir_for_csr num_of_random_networks {g|n|l|gn|gl|nl|gnl}
// Outputs: <file>_eff.nm ('n'), <file>_Lp.txt ('g').  Random reference
// networks (lambda, 'l') are not generated: num_of_random_networks must be 0.
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
    if (sel.find('l') != std::string::npos || nrand != 0) {
        std::cerr << "random-network comparison (lambda, num_of_random_networks > 0) is not available\n";
        return 1;
    }
    const bool global = sel.find('g') != std::string::npos;
    const bool nodal = sel.find('n') != std::string::npos;
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
        double lp = 0.0;
        int isolated = 0;
        pgtool::Timer tm;
        ctx.check(pg_fw_init(ctx.c, g.row.data(), g.col.data(), weighted ? g.val.data() : nullptr, g.n,
                             (int64_t)g.col.size(), weighted ? 0 : 2, d.data()),
                  "pg_fw_init");
        ctx.check(pg_fw(ctx.c, d.data(), g.n, 0), "pg_fw");
        ctx.check(pg_fw_efficiency(ctx.c, d.data(), g.n, eff.data(), &lp, &isolated), "pg_fw_efficiency");
        const double secs = tm.seconds();
        if (isolated > 0) std::cout << "\nisolated vertices number : " << isolated << std::endl;
        std::cout << "average Lp:\t" << std::setprecision(6) << lp << "\nElapsed time: " << secs << " s." << std::endl;
        pgio::time_log(std::string("BFW_Lp\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
        if (nodal) {
            const std::string x = pgio::strip_ext(a) + "_eff.nm";
            std::cout << "Save nodal efficiency for each node as " << x << std::endl;
            if (!pgio::write_nm(x, eff.data(), g.n)) return 2;
        }
        if (global) {
            std::ofstream fo(pgio::strip_ext(a) + "_Lp.txt");
            fo << std::setprecision(6) << lp << std::endl;
        }
    }
    std::cout << "Total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
