// io.cpp -- PAGANI file formats (see io.hpp for the layouts and references).
#include "io.hpp"

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>

namespace pgio {

namespace {
constexpr int HDR = 352;

bool read_header(std::ifstream &f, short h[HDR / 2])
{
    f.read(reinterpret_cast<char *>(h), HDR);
    return f.good();
}
}  // namespace

std::vector<std::string> list_dir(const std::string &dir, const std::string &ext, const std::string &exclude)
{
    std::vector<std::string> out;
    DIR *dp = opendir(dir.c_str());
    if (!dp) return out;
    while (dirent *e = readdir(dp)) {
        std::string nm = e->d_name;
        // the reference matches an extension of exactly the given length
        // after the last '.' (U CorMat.cpp:76-83)
        size_t dot = nm.find_last_of('.');
        if (dot == std::string::npos) continue;
        if (nm.substr(dot) != ext) continue;
        if (!exclude.empty() && nm == exclude) continue;
        out.push_back(nm);
    }
    closedir(dp);
    std::sort(out.begin(), out.end());
    return out;
}

std::string join(const std::string &dir, const std::string &name)
{
    if (dir.empty()) return name;
    char last = dir.back();
    if (last == '/' || last == '\\') return dir + name;
    return dir + "/" + name;
}

std::string strip_ext(const std::string &path)
{
    size_t dot = path.find_last_of('.');
    size_t sep = path.find_last_of("/\\");
    if (dot == std::string::npos || (sep != std::string::npos && dot < sep)) return path;
    return path.substr(0, dot);
}

bool make_dir(const std::string &dir)
{
    struct stat st;
    if (stat(dir.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
    return mkdir(dir.c_str(), 0755) == 0;
}

bool read_mask(const std::string &path, std::vector<float> &mask, int dims[3], std::string &err, short *hdr)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        err = "can't open " + path;
        return false;
    }
    short h[HDR / 2];
    if (!read_header(f, h)) {
        err = "short header in " + path;
        return false;
    }
    if (hdr) std::memcpy(hdr, h, HDR);
    dims[0] = h[21];
    dims[1] = h[22];
    dims[2] = h[23];
    const long long total = (long long)h[21] * h[22] * h[23];
    if (total <= 0) {
        err = "bad mask dimensions in " + path;
        return false;
    }
    mask.resize(total);
    if (h[35] == 2) {
        std::vector<unsigned char> u(total);
        f.read(reinterpret_cast<char *>(u.data()), total);
        for (long long k = 0; k < total; k++) mask[k] = (float)u[k];
    } else if (h[35] == 16) {
        f.read(reinterpret_cast<char *>(mask.data()), total * 4);
    } else {
        err = "mask data-type error, only unsigned char and float can be handled";
        return false;
    }
    if (!f.good()) {
        err = "truncated mask " + path;
        return false;
    }
    return true;
}

bool read_bold(const std::string &path, const std::vector<float> &mask, float thr, int &n, int &t,
               std::vector<float> &bold, std::string &err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        err = "can't open " + path;
        return false;
    }
    short h[HDR / 2];
    if (!read_header(f, h)) {
        err = "short header in " + path;
        return false;
    }
    const long long total = (long long)h[21] * h[22] * h[23];
    if (total != (long long)mask.size()) {
        err = "data and mask dimensions differ in " + path;
        return false;
    }
    t = h[24];
    n = 0;
    for (long long k = 0; k < total; k++) n += (mask[k] >= thr);
    bold.assign((size_t)n * t, 0.0f);
    if (h[36] == 64) {
        std::vector<double> in((size_t)total * t);
        f.read(reinterpret_cast<char *>(in.data()), in.size() * 8);
        if (!f.good()) {
            err = "truncated data " + path;
            return false;
        }
        for (long long k = 0, i = -1; k < total; k++)
            if (mask[k] >= thr) {
                i++;
                for (int l = 0; l < t; l++) bold[(size_t)l * n + i] = (float)in[(size_t)l * total + k];
            }
    } else if (h[36] == 32) {
        std::vector<float> in((size_t)total * t);
        f.read(reinterpret_cast<char *>(in.data()), in.size() * 4);
        if (!f.good()) {
            err = "truncated data " + path;
            return false;
        }
        for (long long k = 0, i = -1; k < total; k++)
            if (mask[k] >= thr) {
                i++;
                for (int l = 0; l < t; l++) bold[(size_t)l * n + i] = in[(size_t)l * total + k];
            }
    } else {
        err = "data type is neither float nor double";
        return false;
    }
    return true;
}

bool read_csr(const std::string &path, bool weighted, Csr &g, std::string &err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        err = "can't open " + path;
        return false;
    }
    int32_t rlen = 0, clen = 0, vlen = 0;
    f.read(reinterpret_cast<char *>(&rlen), 4);
    if (!f.good() || rlen < 1) {
        err = "csr damaged (row length) " + path;
        return false;
    }
    g.n = rlen - 1;
    g.row.resize(rlen);
    f.read(reinterpret_cast<char *>(g.row.data()), (size_t)rlen * 4);
    f.read(reinterpret_cast<char *>(&clen), 4);
    if (!f.good() || clen < 0 || g.row[0] != 0 || g.row[g.n] != clen) {
        err = "csr damaged (R[N] != Clength) " + path;
        return false;
    }
    g.col.resize(clen);
    f.read(reinterpret_cast<char *>(g.col.data()), (size_t)clen * 4);
    if (!f.good()) {
        err = "csr damaged (columns) " + path;
        return false;
    }
    g.val.clear();
    if (weighted) {
        f.read(reinterpret_cast<char *>(&vlen), 4);
        if (!f.good() || vlen != clen) {
            err = "csr damaged (Vlength != Clength) " + path;
            return false;
        }
        g.val.resize(vlen);
        f.read(reinterpret_cast<char *>(g.val.data()), (size_t)vlen * 4);
        if (!f.good()) {
            err = "csr damaged (values) " + path;
            return false;
        }
    }
    for (int32_t i = 0; i < g.n; i++)
        if (g.row[i + 1] < g.row[i]) {
            err = "csr damaged (row offsets decrease) " + path;
            return false;
        }
    for (int32_t c : g.col)
        if (c < 0 || c >= g.n) {
            err = "csr damaged (column out of range) " + path;
            return false;
        }
    return true;
}

bool write_csr(const std::string &path, const int32_t *row, int32_t n, const int32_t *col, int64_t nnz,
               const float *val)
{
    std::ofstream f(path, std::ios::binary);
    if (!f.good()) return false;
    int32_t rlen = n + 1;
    // the reference writes the low 32 bits of a 64-bit Clength
    int32_t clen = (int32_t)nnz;
    f.write(reinterpret_cast<const char *>(&rlen), 4);
    f.write(reinterpret_cast<const char *>(row), (size_t)rlen * 4);
    f.write(reinterpret_cast<const char *>(&clen), 4);
    f.write(reinterpret_cast<const char *>(col), (size_t)nnz * 4);
    if (val) {
        f.write(reinterpret_cast<const char *>(&clen), 4);
        f.write(reinterpret_cast<const char *>(val), (size_t)nnz * 4);
    }
    return f.good();
}

bool write_nm(const std::string &path, const float *x, int32_t n)
{
    std::ofstream f(path, std::ios::binary);
    if (!f.good()) return false;
    f.write(reinterpret_cast<const char *>(&n), 4);
    f.write(reinterpret_cast<const char *>(x), (size_t)n * 4);
    return f.good();
}

bool read_nm(const std::string &path, std::vector<float> &x)
{
    std::ifstream f(path, std::ios::binary);
    int32_t n = 0;
    f.read(reinterpret_cast<char *>(&n), 4);
    if (!f.good() || n < 0) return false;
    x.resize(n);
    f.read(reinterpret_cast<char *>(x.data()), (size_t)n * 4);
    return f.good();
}

void time_log(const std::string &line)
{
    std::ofstream f("BNA_time_log", std::ios::app);
    f << line << "\n";
}

}  // namespace pgio
