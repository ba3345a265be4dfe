// autotune.cpp -- searches for the launch shapes the reference hard-codes:
// the Gram walk's supertile edge (the reference's Batch_size = 1024*3 tiles,
// U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:238), the number of 64-source BFS
// batches per launch of BC and BFS (its grid from free memory,
// U/BC/CUBC/CUBC.cpp:45-64) and the Floyd-Warshall tile (sizeBlock,
// W/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp.cpp:15).  Each search times the candidates
// on the caller's own input through the ordinary entry points, with the shape
// forced on the context, and keeps the fastest in the tuning cache
// (tune.cpp); later calls on inputs of the same size bucket use it.  With
// PAGANI_AUTOTUNE=1 the entry points search by themselves on first use.
#include "pg_internal.hpp"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

bool refresh()
{
    const char *e = std::getenv("PAGANI_TUNE");
    return e && !std::strcmp(e, "refresh");
}

// Runs fn once per candidate with the knob forced (two calls each, the second
// timed by `time`), restores the knob, and returns the fastest.
template <class Fn>
pg_status search(pg_ctx *ctx, int *knob, const std::vector<int> &cand, Fn fn, pg_tune_result *res)
{
    const int saved = *knob;
    const bool was = ctx->tuning;
    ctx->tuning = true;
    pg_status st = PG_OK;
    res->value = 0;
    res->ms = 1e300;
    res->candidates = 0;
    for (int v : cand) {
        *knob = v;
        double ms = 0.0;
        st = fn(&ms);  // warm-up (buffers sized, code loaded)
        if (st == PG_OK) st = fn(&ms);
        if (st != PG_OK) break;
        res->candidates++;
        if (v == cand[0]) res->ms_default = ms;
        if (ms < res->ms) {
            res->ms = ms;
            res->value = v;
        }
    }
    *knob = saved;
    ctx->tuning = was;
    return st;
}

double host_ms(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" pg_status pg_autotune_corr(pg_ctx *ctx, const float *bold, int32_t n, int32_t t,
                                      const pg_corr_opts *opts, pg_tune_result *res)
{
    if (!ctx) return PG_EINVAL;
    if (!bold || !opts || !res) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    std::memset(res, 0, sizeof *res);
    const int cached = pg_tune_get(ctx, "gram_super", n, t, 0);
    if (cached > 0 && !refresh()) {
        res->value = cached;
        return PG_OK;
    }
    PG_TRY(search(ctx, &ctx->force_gram_super, {8, 4, 16},
                  [&](double *ms) {
                      pg_corr_info info;
                      const pg_status s = pg_corr_build(ctx, bold, n, t, opts, &info);
                      *ms = info.ms_gram_emit;  // the pass the supertile walk shapes
                      return s;
                  },
                  res));
    pg_tune_put(ctx, "gram_super", n, t, 0, res->value, res->ms);
    return PG_OK;
}

extern "C" pg_status pg_autotune_bc(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n, int64_t nnz,
                                    const int32_t *sources, int64_t nsrc, pg_tune_result *res)
{
    if (!ctx) return PG_EINVAL;
    if (!res || !sources || nsrc <= 0) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    std::memset(res, 0, sizeof *res);
    const int cached = pg_tune_get(ctx, "bc_batches", n, nnz, 0);
    if (cached > 0 && !refresh()) {
        res->value = cached;
        return PG_OK;
    }
    double *bc = nullptr;
    PG_TRY(pg_bufT(ctx, "tune.bc", (size_t)n + 1, &bc));
    std::vector<int> cand = {0};  // 0: the size heuristic
    for (int nb : {1, 2, 4, 8, 16})
        if ((long long)nb * 64 <= nsrc + 63) cand.push_back(nb);
    PG_TRY(search(ctx, &ctx->force_bc_batches, cand,
                  [&](double *ms) {
                      const auto t0 = std::chrono::steady_clock::now();
                      const pg_status s = pg_bc(ctx, row, col, n, nnz, sources, nsrc, bc);
                      *ms = host_ms(t0);
                      return s;
                  },
                  res));
    if (res->value == 0) {
        // the heuristic won: store what it chose
        ctx->tuning = true;
        res->value = pg_bc_batches_for(ctx, n, nnz, nsrc);
        ctx->tuning = false;
    }
    pg_tune_put(ctx, "bc_batches", n, nnz, 0, res->value, res->ms);
    return PG_OK;
}

extern "C" pg_status pg_autotune_fw(pg_ctx *ctx, const float *d, int32_t n, pg_tune_result *res)
{
    if (!ctx) return PG_EINVAL;
    if (!d || !res || n < 1) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    std::memset(res, 0, sizeof *res);
    const int cached = pg_tune_get(ctx, "fw_block", n, 0, 0);
    if (cached > 0 && !refresh()) {
        res->value = cached;
        return PG_OK;
    }
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = (size_t)n * n * 4;
    float *work = nullptr;
    PG_TRY(pg_buf(ctx, "tune.fw", bytes, (void **)&work));
    PG_TRY(search(ctx, &ctx->force_fw_block, {64, 128},
                  [&](double *ms) {
                      PG_HIP(ctx, hipMemcpyAsync(work, d, bytes, pg_is_device_ptr(d) ? hipMemcpyDeviceToDevice
                                                                                      : hipMemcpyHostToDevice,
                                                 ctx->stream));
                      PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
                      const auto t0 = std::chrono::steady_clock::now();
                      const pg_status s = pg_fw(ctx, work, n, 0);
                      *ms = host_ms(t0);
                      return s;
                  },
                  res));
    pg_tune_put(ctx, "fw_block", n, 0, 0, res->value, res->ms);
    return PG_OK;
}

bool pg_autotune_on_first_use(pg_ctx *ctx)
{
    if (ctx->tuning) return false;
    const char *e = std::getenv("PAGANI_AUTOTUNE");
    return e && e[0] == '1';
}
