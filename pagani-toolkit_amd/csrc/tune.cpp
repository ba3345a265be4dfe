// tune.cpp -- the autotuner's cache of launch shapes.
//
// A tuned value is keyed by the device architecture (gcnArchName), the knob
// and the half-octave buckets of up to three sizes (voxels or vertices, time
// points or nothing, edges or nothing).  Values live in memory for the process
// and in a text file, one "arch knob b0 b1 b2 value ms" line each, read once
// and appended to on every new result: $PAGANI_TUNE_CACHE, else
// $HOME/.cache/pagani/tune.txt.  PAGANI_TUNE=off ignores the cache.
#include "pg_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <tuple>

namespace {

struct Entry {
    int value;
    double ms;
};

std::mutex g_mu;
std::string g_path;  // the file the in-memory cache mirrors
bool g_loaded = false;
std::map<std::string, Entry> g_cache;

int bucket(long long x) { return x <= 0 ? -1 : (int)std::floor(2.0 * std::log2((double)x)); }

std::string cache_path()
{
    if (const char *p = std::getenv("PAGANI_TUNE_CACHE")) return p;
    const char *h = std::getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/pagani/tune.txt";
}

bool disabled()
{
    const char *e = std::getenv("PAGANI_TUNE");
    return e && !std::strcmp(e, "off");
}

std::string key(const std::string &arch, const char *knob, int b0, int b1, int b2)
{
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s %s %d %d %d", arch.c_str(), knob, b0, b1, b2);
    return buf;
}

void load_locked()
{
    const std::string path = cache_path();
    if (g_loaded && path == g_path) return;
    g_cache.clear();
    g_path = path;
    g_loaded = true;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return;
    char arch[128], knob[64];
    int b0, b1, b2, v;
    double ms;
    while (std::fscanf(f, "%127s %63s %d %d %d %d %lf", arch, knob, &b0, &b1, &b2, &v, &ms) == 7)
        g_cache[key(arch, knob, b0, b1, b2)] = Entry{v, ms};  // later lines win
    std::fclose(f);
}

void mkdirs(const std::string &path)
{
    for (size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1))
        (void)mkdir(path.substr(0, p).c_str(), 0755);
}

}  // namespace

std::string pg_arch(pg_ctx *ctx)
{
    if (!ctx->arch.empty()) return ctx->arch;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) {
        (void)hipGetLastError();
        return "unknown";
    }
    std::string a = prop.gcnArchName;
    const size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

int pg_tune_get(pg_ctx *ctx, const char *knob, long long a, long long b, long long c)
{
    if (disabled()) return 0;
    const std::string k = key(pg_arch(ctx), knob, bucket(a), bucket(b), bucket(c));
    std::lock_guard<std::mutex> lock(g_mu);
    load_locked();
    auto it = g_cache.find(k);
    return it == g_cache.end() ? 0 : it->second.value;
}

void pg_tune_put(pg_ctx *ctx, const char *knob, long long a, long long b, long long c, int value, double ms)
{
    const std::string arch = pg_arch(ctx);
    const int b0 = bucket(a), b1 = bucket(b), b2 = bucket(c);
    std::lock_guard<std::mutex> lock(g_mu);
    load_locked();
    g_cache[key(arch, knob, b0, b1, b2)] = Entry{value, ms};
    if (disabled()) return;
    const std::string path = cache_path();
    mkdirs(path);
    if (FILE *f = std::fopen(path.c_str(), "a")) {
        std::fprintf(f, "%s %s %d %d %d %d %.6f\n", arch.c_str(), knob, b0, b1, b2, value, ms);
        std::fclose(f);
    }
}
