// pg_internal.hpp -- context, device arena and helpers shared by the
// libpagani translation units.  gfx950 (MI355X) only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>

#include "pagani.h"

struct PgBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct PgEmit;  // kept pairs of pg_corr_emit (corr.hip)

struct pg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    size_t lds_per_block = 65536;
    std::string err;
    std::map<std::string, PgBuf> bufs;  // named grow-only device buffers
    hipEvent_t ev[12] = {};
    // stage-timing events of pg_corr_build (pg_set_timing): 2 every stage,
    // 1 the emitting Gram kernel only (events 8, 9), 0 none
    int timing = 2;
    void *static_scratch = nullptr;  // this device's fixed scratch array (pg_static_scratch)
    PgEmit *emit = nullptr;          // pg_corr_emit state, owned
    // launch shapes forced by the autotuner while it times candidates (0: none)
    int force_gram_super = 0, force_bc_batches = 0, force_fw_block = 0;
    bool tuning = false;             // inside a search: no nested search
    std::string arch;                // gcnArchName without features, read once
    void *h_report = nullptr;        // pinned host buffer for small result read-backs
    // the Gram supertile list last copied to the "corr.stl" buffer and its key
    // (row tiles, column tiles, tile aspect, supertile edge, length)
    void *stl_dev = nullptr;
    long long stl_key[5] = {-1, -1, -1, -1, -1};

    // result of the last pg_corr_build
    int32_t cn = 0, ct = 0;
    int64_t cnnz = -1;
    bool cweighted = false;
    int32_t *c_row = nullptr, *c_col = nullptr;
    float *c_val = nullptr;
    double *c_z64 = nullptr;

    // dense packed correlation matrix (pg_cormat_*): n voxels, nsub subjects
    // added, fisher mode, finished flag; the buffer is the named "cm.acc"
    int32_t cm_n = -1, cm_nsub = 0, cm_fisher = 0;
    bool cm_finished = false;
    float *cm_acc = nullptr;
};

pg_status pg_fail(pg_ctx *ctx, pg_status st, const char *fmt, ...);

// size of pg_ctx::h_report
constexpr size_t PG_REPORT_BYTES = 4096;

#define PG_HIP(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return pg_fail((ctx), e_ == hipErrorOutOfMemory ? PG_ENOMEM : PG_EDEVICE,      \
                           "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,       \
                           __LINE__);                                                      \
    } while (0)

#define PG_TRY(expr)                 \
    do {                             \
        pg_status s_ = (expr);       \
        if (s_ != PG_OK) return s_;  \
    } while (0)

#define PG_KERNEL_CHECK(ctx) PG_HIP(ctx, hipGetLastError())

// Grow-only named device buffer; contents are not preserved on growth.
pg_status pg_buf(pg_ctx *ctx, const char *name, size_t bytes, void **out);
template <class T>
inline pg_status pg_bufT(pg_ctx *ctx, const char *name, size_t count, T **out)
{
    void *p = nullptr;
    pg_status s = pg_buf(ctx, name, count * sizeof(T) + 16, &p);
    *out = (T *)p;
    return s;
}

bool pg_is_device_ptr(const void *p);

// Device view of an input: device pointers pass through, host pointers are
// copied into the named buffer.
template <class T>
inline pg_status pg_in(pg_ctx *ctx, const T *p, size_t count, const char *name, const T **dev)
{
    if (p == nullptr || count == 0 || pg_is_device_ptr(p)) {
        *dev = p;
        return PG_OK;
    }
    T *d = nullptr;
    PG_TRY(pg_bufT(ctx, name, count, &d));
    PG_HIP(ctx, hipMemcpyAsync(d, p, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    *dev = d;
    return PG_OK;
}

// Device destination for an output: device pointers pass through, host ones
// get a staging buffer copied back by pg_out_finish.
template <class T>
inline pg_status pg_out(pg_ctx *ctx, T *p, size_t count, const char *name, T **dev)
{
    if (p == nullptr || count == 0 || pg_is_device_ptr(p)) {
        *dev = p;
        return PG_OK;
    }
    return pg_bufT(ctx, name, count, dev);
}

template <class T>
inline pg_status pg_out_finish(pg_ctx *ctx, T *host, const T *dev, size_t count)
{
    if (host == nullptr || count == 0 || host == dev) return PG_OK;
    PG_HIP(ctx, hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    return PG_OK;
}

pg_status pg_check_csr_args(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                            int64_t nnz);

inline int64_t pg_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// FP64 z-scores of a device time series into z64[i * t + l] (voxel-major), by
// the same kernels and arithmetic as pg_corr_build.
pg_status pg_zscore64(pg_ctx *ctx, const float *dbold, int layout, int n, int t, double *z64);

// Device check of a staged CSR (device pointers): row non-decreasing from 0 to
// nnz, every column in [0, n), and with `sorted` strictly ascending columns in
// each row.  PG_EFORMAT otherwise.  Synchronises the stream.
pg_status pg_validate_csr(pg_ctx *ctx, const int32_t *drow, const int32_t *dcol, int32_t n, int64_t nnz,
                          bool sorted);
// The same check with a caller-provided device flag (no allocation).
pg_status pg_validate_csr_flag(pg_ctx *ctx, const int32_t *drow, const int32_t *dcol, int32_t n, int64_t nnz,
                               bool sorted, int *bad);

// Device address of this device's fixed scratch array of the allocation-free
// entry points (graph.hip), looked up once per device; called by pg_create
// with the device current.
pg_status pg_static_scratch(pg_ctx *ctx, void **ptr);

// pg_corr_emit state: dropped by any build that reuses its buffers, freed
// with the context.
void pg_emit_invalidate(pg_ctx *ctx);
void pg_emit_free(PgEmit *e);

// Autotuner cache (tune.cpp): the tuned value of a knob for this device's
// architecture and the half-octave buckets of up to three sizes (0 or less
// where unused), or 0 when none is stored; pg_tune_put stores one (and the
// time it measured) in memory and in the cache file.
std::string pg_arch(pg_ctx *ctx);
int pg_tune_get(pg_ctx *ctx, const char *knob, long long a, long long b, long long c);
void pg_tune_put(pg_ctx *ctx, const char *knob, long long a, long long b, long long c, int value, double ms);

// Captured cohort (pg_cohort_run): the construction part (corr.hip) and the
// metrics part (graph.hip) plan their workspaces for the largest subject, then
// enqueue each subject's launches with no allocation, synchronisation or
// memset, so the whole cohort can be captured into one graph.
struct PgCohortCorr;
struct PgCohortMetrics;
pg_status pg_cohort_corr_plan(pg_ctx *ctx, int nsub, const int32_t *n, const int32_t *t, const float *spars,
                              PgCohortCorr **out);
bool pg_cohort_corr_in_graph(const PgCohortCorr *P, int s);
pg_status pg_cohort_corr_enqueue(PgCohortCorr *P, int s, const float *dbold, int32_t *row, int32_t *col,
                                 unsigned *status);
void pg_cohort_corr_free(PgCohortCorr *P);
pg_status pg_cohort_metrics_plan(pg_ctx *ctx, int max_n, long long max_nsrc, int levels, PgCohortMetrics **out);
int pg_cohort_metrics_group(const PgCohortMetrics *M);
pg_status pg_cohort_metrics_enqueue(PgCohortMetrics *M, const int32_t *row, const int32_t *col, int n,
                                    const int32_t *srcs, long long nsrc, int32_t *deg, float *cp, double *bc,
                                    unsigned *status);
void pg_cohort_metrics_free(PgCohortMetrics *M);
long long pg_sparsity_k(long long n, float s);

// Autotuner hooks: PAGANI_AUTOTUNE=1 and not inside a search; the BC batch
// count the heuristic (or a forced or tuned value) gives.
bool pg_autotune_on_first_use(pg_ctx *ctx);
int pg_bc_batches_for(pg_ctx *ctx, int n, long long nnz, long long nsrc);
