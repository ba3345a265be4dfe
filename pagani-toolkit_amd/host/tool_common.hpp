// tool_common.hpp -- shared helpers of the command-line tools.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
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

// Seed of the random reference networks: PAGANI_SEED when set, else the wall
// clock (the reference calls srand(time(0)) inside every rewiring, so its
// networks made within one second repeat; here network l uses seed + l).
inline uint64_t random_seed()
{
    const char *e = std::getenv("PAGANI_SEED");
    if (e && *e) return std::strtoull(e, nullptr, 10);
    return (uint64_t)std::time(nullptr);
}

// Random reference network of g (Maslov / Maslov_weighted_1 of the reference):
// degree-preserving rewiring with 2M attempts; weights travel with their edge.
inline pgio::Csr rewired(const Ctx &ctx, const pgio::Csr &g, uint64_t seed)
{
    pgio::Csr r;
    r.n = g.n;
    r.row = g.row;
    r.col.resize(g.col.size());
    if (!g.val.empty()) r.val.resize(g.val.size());
    ctx.check(pg_rewire(ctx.c, g.row.data(), g.col.data(), g.val.empty() ? nullptr : g.val.data(), g.n,
                        (int64_t)g.col.size(), seed, 0, r.col.data(), r.val.empty() ? nullptr : r.val.data(),
                        nullptr),
              "pg_rewire");
    return r;
}

inline std::string dir_arg(const char *a)
{
    std::string d = a;
    while (d.size() > 1 && (d.back() == '/' || d.back() == '\\')) d.pop_back();
    return d;
}

}  // namespace pgtool
