// tool_common.hpp -- shared helpers of the command-line tools.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "io.hpp"
#include "pagani.h"

namespace pgtool {

struct Ctx {
    pg_ctx *c = nullptr;
    Ctx()
    {
        pg_status s = pg_create(&c, 0);
        if (s != PG_OK) {
            std::fprintf(stderr, "pagani: no usable GPU (%s)\n", pg_status_string(s));
            std::exit(3);
        }
    }
    ~Ctx() { pg_destroy(c); }
    void check(pg_status s, const char *what) const
    {
        if (s != PG_OK) {
            std::fprintf(stderr, "pagani: %s failed: %s (%s)\n", what, pg_status_string(s), pg_last_error(c));
            std::exit(4);
        }
    }
};

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double seconds() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// A .csr file carries weights when bytes follow the column array.
inline bool csr_has_values(const std::string &path)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f.good()) return false;
    const long long size = (long long)f.tellg();
    f.seekg(0);
    int32_t rlen = 0, clen = 0;
    f.read(reinterpret_cast<char *>(&rlen), 4);
    if (rlen < 1) return false;
    f.seekg(4 + 4LL * rlen);
    f.read(reinterpret_cast<char *>(&clen), 4);
    return size > 8 + 4LL * rlen + 4LL * clen;
}

inline std::string dir_arg(const char *a)
{
    std::string d = a;
    while (d.size() > 1 && (d.back() == '/' || d.back() == '\\')) d.pop_back();
    return d;
}

}  // namespace pgtool
