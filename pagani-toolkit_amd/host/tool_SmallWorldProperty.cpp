// SmallWorldProperty -- small-world index sigma = gamma / lambda of every
// network whose _lambda.txt (CUBFW_Lp / CUBFS_Lp 'l') and _gamma.txt (Cp 'k')
// are in a directory (drop-in for W/SmallWorldProperty/SmallWorldProperty.cpp:
// 15-228).
//
//   SmallWorldProperty dir
// A lambda file is paired with the first unpaired gamma file whose name has
// the same characters 1-4 and the same 5 characters after "spa" (the
// reference's matching rule, file names taken in sorted order); each pair
// writes <dir>/<lambda file name up to its last '_'>_smallWorld.txt holding
// sigma with six significant digits.  Host-only: no GPU work is involved.
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "io.hpp"

namespace {

// name ends in <tag>.txt (the reference also requires a minimum length)
bool has_tag(const std::string &nm, const std::string &tag, size_t min_len)
{
    const size_t dot = nm.find_last_of('.');
    if (nm.size() <= min_len || dot == std::string::npos || dot < tag.size()) return false;
    return nm.compare(dot - tag.size(), tag.size(), tag) == 0 && nm.compare(dot, std::string::npos, ".txt") == 0;
}

// characters 1-4, and 5 characters after "spa" (a name without "spa" gives
// position 2, as the reference's size_t arithmetic does)
void features(const std::string &nm, std::string &f0, std::string &f1)
{
    f0 = nm.substr(1, 4);
    f1 = nm.substr(nm.find("spa") + 3, 5);
}

bool read_value(const std::string &path, double &x)
{
    std::ifstream f(path);
    return static_cast<bool>(f >> x);
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::cerr << "Input format: SmallWorldProperty dir_with_lambda_and_gamma_files\n"
                     "For example: SmallWorldProperty /data" << std::endl;
        return 1;
    }
    std::string dir = argv[1];
    while (dir.size() > 1 && (dir.back() == '/' || dir.back() == '\\')) dir.pop_back();
    std::vector<std::string> lam, gam;
    for (const auto &nm : pgio::list_dir(dir, ".txt")) {
        if (has_tag(nm, "_lambda", 11)) lam.push_back(nm);
        if (has_tag(nm, "_gamma", 10)) gam.push_back(nm);
    }
    std::cout << lam.size() << " files within lambda value to be processed." << std::endl;
    std::cout << gam.size() << " files within gamma value to be processed." << std::endl;
    if (lam.empty())
        std::cerr << "Lack lambda files. Please run lambda algorithm before executing small world algorithm."
                  << std::endl;
    if (gam.empty())
        std::cerr << "Lack gamma files. Please run gamma algorithm before executing small world algorithm."
                  << std::endl;
    std::cout << "Matching...." << std::endl;
    std::vector<bool> used(gam.size(), false);
    std::vector<std::pair<size_t, size_t>> pairs;
    std::vector<size_t> lone_lam;
    for (size_t i = 0; i < lam.size(); i++) {
        std::string a0, a1;
        features(lam[i], a0, a1);
        bool found = false;
        for (size_t j = 0; j < gam.size() && !found; j++) {
            if (used[j]) continue;
            std::string b0, b1;
            features(gam[j], b0, b1);
            if (a0 == b0 && a1 == b1) {
                used[j] = true;
                pairs.emplace_back(i, j);
                found = true;
            }
        }
        if (!found) lone_lam.push_back(i);
    }
    if (!lone_lam.empty()) {
        std::cout << lone_lam.size() << " files within lambda value have no matching gamma file:" << std::endl;
        for (size_t i : lone_lam) std::cout << "  " << lam[i] << std::endl;
    }
    for (size_t j = 0; j < gam.size(); j++)
        if (!used[j]) std::cout << "gamma file without a matching lambda file: " << gam[j] << std::endl;
    std::vector<double> sigma;
    for (const auto &p : pairs) {
        const std::string a = pgio::join(dir, lam[p.first]), b = pgio::join(dir, gam[p.second]);
        std::cout << "\ncalculating small world property for " << a << " and " << b << " ..." << std::endl;
        double lambda = 0.0, gamma = 0.0;
        if (!read_value(a, lambda)) {
            std::cerr << "Can't open\t" << a << std::endl;
            return 2;
        }
        if (!read_value(b, gamma)) {
            std::cerr << "Can't open\t" << b << std::endl;
            return 2;
        }
        sigma.push_back(gamma / lambda);
        std::cout << "Small world property equals " << std::setprecision(6) << sigma.back() << std::endl;
    }
    std::cout << "\nConclude each result as the table view:\n File   sparsity   small world property" << std::endl;
    for (size_t q = 0; q < pairs.size(); q++) {
        std::string f0, f1;
        features(lam[pairs[q].first], f0, f1);
        std::cout << " " << f0 << "    " << f1 << "%          " << std::setprecision(6) << sigma[q] << std::endl;
    }
    for (size_t q = 0; q < pairs.size(); q++) {
        const std::string &nm = lam[pairs[q].first];
        const std::string out = pgio::join(dir, nm.substr(0, nm.find_last_of('_'))) + "_smallWorld.txt";
        std::ofstream fo(out);
        fo << std::setprecision(6) << sigma[q] << std::endl;
        if (!fo.good()) {
            std::cerr << "Can't write\t" << out << std::endl;
            return 2;
        }
    }
    if (!pairs.empty()) std::cout << "Already save each result to the corresponding file with .txt format..." << std::endl;
    return 0;
}
