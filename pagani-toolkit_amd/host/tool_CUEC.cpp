// CUEC -- eigenvector centrality of every network in a directory (drop-in for
// U/CUEC/CUEC/CUEC_unweighted.cu:198-312 and, with --weighted, W/CUEC/
// weighted-net.cu:205-331).
//
//   CUEC [--weighted] dir_for_csr
// Outputs per <file>.csr: <file>_ec.nm and <file>_evalue.txt.  Binary mode
// ignores stored weights and writes the normalised vector of the last power
// iteration and the eigenvalue as a float; weighted mode uses the stored
// weights and writes A times that vector and the eigenvalue as a double, as the
// two reference tools do.
#include <iomanip>
#include <iostream>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    int a0 = 1;
    bool weighted = false;
    if (argc > 1 && std::string(argv[1]) == "--weighted") {
        weighted = true;
        a0 = 2;
    }
    if (argc - a0 < 1) {
        std::cerr << "Input format: CUEC [--weighted] dir_for_csr\nFor example: CUEC /data" << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[a0]);
    const auto files = pgio::list_dir(dir, ".csr");
    std::cout << files.size() << " files to be processed." << std::endl;
    pgtool::Ctx ctx;
    for (const auto &f : files) {
        const std::string a = pgio::join(dir, f);
        std::cout << "\ncalculating eigenvalue centrality for " << a << " ..." << std::endl;
        pgio::Csr g;
        std::string err;
        if (!pgio::read_csr(a, weighted, g, err)) {
            std::cerr << err << "\n";
            return 2;
        }
        std::vector<double> vh(g.n), av(g.n);
        double lam = 0.0;
        int iters = 0;
        pgtool::Timer tm;
        ctx.check(pg_eigenvector_centrality(ctx.c, g.row.data(), g.col.data(), weighted ? g.val.data() : nullptr,
                                            g.n, (int64_t)g.col.size(), vh.data(), av.data(), &lam, &iters),
                  "pg_eigenvector_centrality");
        const double secs = tm.seconds();
        std::cout << "Iterations:\t" << iters << "\teigenvalue:\t" << lam << "\tcalculate time: " << secs << " s."
                  << std::endl;
        std::vector<float> out(g.n);
        const std::vector<double> &src = weighted ? av : vh;
        for (int k = 0; k < g.n; k++) out[k] = (float)src[k];
        const std::string base = pgio::strip_ext(a);
        std::cout << "Save eigenvector centrality for each node as " << base << "_ec.nm" << std::endl;
        if (!pgio::write_nm(base + "_ec.nm", out.data(), g.n)) return 2;
        std::ofstream fo(base + "_evalue.txt");
        if (weighted) fo << std::setprecision(6) << lam << std::endl;
        else fo << std::setprecision(6) << (float)lam << std::endl;
    }
    return 0;
}
