// Degree -- degree (binary) or strength (weighted) of every .csr in a
// directory (drop-in for U/Degree/Degree/Degree.cpp:12-117 and the weighted
// W/Degree/Degree/Degree.cpp).  Output: <file>_deg.nm; the binary tool also
// writes isolated_v_mark.txt (isolated vertex count per file).
//
//   Degree [--weighted] dir_for_csr
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
    if (argc - a0 < 1) {
        std::cerr << "Input format: Degree [--weighted] dir_for_csr \nFor example: Degree /data " << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[a0]);
    const auto files = pgio::list_dir(dir, ".csr");
    std::cout << files.size() << " files to be processed." << std::endl;
    pgtool::Ctx ctx;
    std::ofstream iso;
    if (!weighted) iso.open(pgio::join(dir, "isolated_v_mark.txt"));
    pgtool::Timer total;
    for (const auto &f : files) {
        const std::string a = pgio::join(dir, f);
        std::cout << "\ncalculating Degree for " << a << " ..." << std::endl;
        pgio::Csr g;
        std::string err;
        if (!pgio::read_csr(a, weighted, g, err)) {
            std::cerr << err << "\n";
            return 2;
        }
        std::vector<float> out(g.n);
        pgtool::Timer tm;
        if (weighted) {
            ctx.check(pg_strength(ctx.c, g.row.data(), g.val.data(), g.n, out.data()), "pg_strength");
        } else {
            std::vector<int32_t> deg(g.n);
            ctx.check(pg_degree(ctx.c, g.row.data(), g.n, deg.data(), out.data()), "pg_degree");
            int isolated = 0;
            for (int i = 0; i < g.n; i++) isolated += out[i] == 0.0f;
            iso << isolated << std::endl;
        }
        const double secs = tm.seconds();
        std::cout << "Elapsed time: " << secs * 1000.0 << " ms. " << std::endl;
        pgio::time_log(std::string("Degree\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
        const std::string x = pgio::strip_ext(a) + "_deg.nm";
        std::cout << "Save degree data for each node as " << x << std::endl;
        if (!pgio::write_nm(x, out.data(), g.n)) return 2;
    }
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
