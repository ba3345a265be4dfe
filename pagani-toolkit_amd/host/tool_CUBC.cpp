// CUBC -- betweenness centrality of every .csr in a directory (drop-in for
// U/BC/CUBC/CUBC_main.cpp:15-111).  Output: <file>_bc.nm (int32 N, float32[N]),
// Brandes BC over ordered pairs, unnormalised, from every source.
//
//   CUBC dir_for_csr [--sources K]   (K > 0: the first K of a seeded sample)
#include <iostream>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "Input format: CUBC dir_for_csr \nFor example: CUBC /data " << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[1]);
    long long nsrc_opt = 0;
    for (int q = 2; q + 1 < argc; q++)
        if (!std::strcmp(argv[q], "--sources")) nsrc_opt = std::atoll(argv[q + 1]);
    const auto files = pgio::list_dir(dir, ".csr");
    std::cout << files.size() << " files to be processed." << std::endl;
    pgtool::Ctx ctx;
    pgtool::Timer total;
    for (const auto &f : files) {
        const std::string a = pgio::join(dir, f);
        std::cout << "\ncalculating Betweenness for " << a << " ..." << std::endl;
        pgio::Csr g;
        std::string err;
        if (!pgio::read_csr(a, false, g, err)) {
            std::cerr << err << "\n";
            return 2;
        }
        std::vector<double> bc(g.n);
        std::vector<int32_t> src;
        if (nsrc_opt > 0 && nsrc_opt < g.n) {
            // evenly spaced sources (deterministic sample)
            for (long long q = 0; q < nsrc_opt; q++) src.push_back((int32_t)(q * g.n / nsrc_opt));
        }
        pgtool::Timer tm;
        ctx.check(pg_bc(ctx.c, g.row.data(), g.col.data(), g.n, (int64_t)g.col.size(),
                        src.empty() ? nullptr : src.data(), (int64_t)src.size(), bc.data()),
                  "pg_bc");
        const double secs = tm.seconds();
        const double te = (double)(src.empty() ? g.n : (long long)src.size()) * (double)g.col.size();
        std::cout << secs * 1000.0 << " ms for GPU BC, " << te / secs / 1e9 << " GTEPS" << std::endl;
        std::vector<float> out(g.n);
        for (int i = 0; i < g.n; i++) out[i] = (float)bc[i];
        const std::string xbc = pgio::strip_ext(a) + "_bc.nm";
        std::cout << "Save betweenness data for each node as " << xbc << std::endl;
        if (!pgio::write_nm(xbc, out.data(), g.n)) return 2;
        pgio::time_log(std::string("BC\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
    }
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
