// PC_CPU -- participation coefficient of every network in a directory from its
// module labels (drop-in for W/PC_CPU/PC_CPU/main.cpp:9-205; the computation
// runs on the GPU through pg_participation).
//
//   PC_CPU dir [x]
// For every directory entry whose name contains ".csr" (the reference's scan),
// with base = that name minus its first ".csr": reads <base>.csr (unit weights
// when the file stores none) and <base>_modu.nm (float module labels 1..K, as
// CUmodularity writes them; an entry without one is skipped), and writes
// <base>_pc.nm and <base>_pc.nm_Normalized.nm (pc / ((K-1)/K)).  The optional
// second argument is accepted and ignored, as in the reference.
#include <dirent.h>

#include <algorithm>
#include <iostream>

#include "tool_common.hpp"

int main(int argc, char **argv)
{
    if (argc != 2 && argc != 3) {
        std::cout << "No Directory Input" << std::endl;
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[1]);
    DIR *dp = opendir(dir.c_str());
    if (!dp) {
        std::cout << "can't open " << dir << std::endl;
        return 1;
    }
    std::vector<std::string> bases;
    while (dirent *e = readdir(dp)) {
        std::string nm = e->d_name;
        const size_t pos = nm.find(".csr");
        if (pos != std::string::npos) bases.push_back(nm.erase(pos, 4));
    }
    closedir(dp);
    std::sort(bases.begin(), bases.end());
    pgtool::Ctx ctx;
    for (const auto &b : bases) {
        const std::string base = pgio::join(dir, b);
        std::cout << "calculate participant coefficient of " << base << ".csr" << std::endl << std::endl;
        pgio::Csr g;
        std::string err;
        if (!pgio::read_csr(base + ".csr", pgtool::csr_has_values(base + ".csr"), g, err)) {
            std::cout << base << ".csr not found!" << std::endl;
            return 2;
        }
        std::vector<float> mod;
        if (!pgio::read_nm(base + "_modu.nm", mod)) {
            std::cout << base << "_modu.nm not found!" << std::endl;
            continue;
        }
        if ((int)mod.size() != g.n) {
            std::cerr << base << "_modu.nm holds " << mod.size() << " labels for " << g.n << " vertices\n";
            return 2;
        }
        std::vector<int32_t> lab(g.n);
        for (int i = 0; i < g.n; i++) lab[i] = (int)mod[i];
        std::vector<float> pc(g.n), pcn(g.n);
        int k = 0;
        ctx.check(pg_participation(ctx.c, g.row.data(), g.col.data(), g.val.empty() ? nullptr : g.val.data(), g.n,
                                   (int64_t)g.col.size(), lab.data(), pc.data(), pcn.data(), &k),
                  "pg_participation");
        if (!pgio::write_nm(base + "_pc.nm", pc.data(), g.n)) return 2;
        if (!pgio::write_nm(base + "_pc.nm_Normalized.nm", pcn.data(), g.n)) return 2;
    }
    return 0;
}
