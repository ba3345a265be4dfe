// cohort.hip -- many subjects in captured HIP graphs (pg_cohort_run).
//
// The reference builds one subject's network after another in a host loop
// (U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:190-378) and runs each metric tool
// over the files afterwards.  Here the per-subject construction and metrics
// are enqueued back to back on the context's stream for every subject, with
// workspaces sized for the largest one, captured into graphs of consecutive
// subjects (about GRAPH_NODES kernel nodes each) that are launched back to
// back: no host round trip between kernels or subjects.  The
// decisions the ordinary path takes on the host (the single sparsity pass's
// checks, the BFS depth) are device flags here, read once after the launch;
// a flagged subject is rebuilt through the ordinary entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "pg_internal.hpp"

namespace {

constexpr int BFS_LEVELS = 8;       // level launches per BFS inside the graph
constexpr int GRAPH_NODES = 2048;   // kernel nodes per captured graph, at most about

// ts[s] = the device wall clock when the stream reaches it
__global__ void stamp(unsigned long long *ts, int s)
{
    if (threadIdx.x == 0) ts[s] = wall_clock64();
}

__global__ void clear_u32(unsigned *p, int count)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) p[k] = 0u;
}

__global__ void fill_i32(int *p, long long count, int v)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) p[k] = v;
}

// Kernel nodes one subject adds to a graph: the construction (24 launches)
// and its time stamp, degree, triangles and Cp, and per group of BFS batches
// init, seed, `levels` forward launches, the depth check, the coefficient
// init and `levels` - 1 backward launches.
long long subject_nodes(bool in_graph, long long nsrc, int group, int levels)
{
    if (!in_graph) return 1;
    long long k = 25 + 3;
    if (nsrc > 0) k += 2 + pg_cdiv(nsrc, group) * (2 * levels + 3);
    return k;
}

struct GraphGuard {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    GraphGuard() = default;
    GraphGuard(const GraphGuard &) = delete;
    GraphGuard &operator=(const GraphGuard &) = delete;
    GraphGuard(GraphGuard &&o) noexcept : g(o.g), e(o.e)
    {
        o.g = nullptr;
        o.e = nullptr;
    }
    ~GraphGuard()
    {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
    }
};

}  // namespace

extern "C" int64_t pg_sparsity_pairs(int32_t n, float sparsity)
{
    if (n < 2 || !(sparsity > 0.0f && sparsity <= 100.0f)) return 0;
    return pg_sparsity_k(n, sparsity);
}

extern "C" pg_status pg_cohort_run(pg_ctx *ctx, pg_subject *subj, int32_t nsub, int32_t launches,
                                   pg_cohort_info *info)
{
    if (!ctx) return PG_EINVAL;
    if (!subj || nsub < 0 || launches < 1) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const auto t0 = std::chrono::steady_clock::now();
    pg_cohort_info inf;
    std::memset(&inf, 0, sizeof inf);
    std::vector<int32_t> n((size_t)nsub), t((size_t)nsub);
    std::vector<float> sp((size_t)nsub);
    int max_n = 1;
    long long max_src = 0;
    for (int s = 0; s < nsub; s++) {
        const pg_subject &S = subj[s];
        if (S.n < 1 || S.t < 2 || !(S.sparsity > 0.0f && S.sparsity <= 100.0f))
            return pg_fail(ctx, PG_EINVAL, "subject %d: n=%d t=%d sparsity=%g", s, S.n, S.t, (double)S.sparsity);
        if (!pg_is_device_ptr(S.bold) || !pg_is_device_ptr(S.row) ||
            (pg_sparsity_pairs(S.n, S.sparsity) > 0 && !pg_is_device_ptr(S.col)) ||
            (S.deg && !pg_is_device_ptr(S.deg)) || (S.cp && !pg_is_device_ptr(S.cp)) ||
            (S.bc && !pg_is_device_ptr(S.bc)))
            return pg_fail(ctx, PG_EINVAL, "subject %d: series and outputs must be device memory", s);
        if (S.bc && S.nsrc > 0 && !S.sources) return pg_fail(ctx, PG_EINVAL, "subject %d: sources is NULL", s);
        if ((size_t)((S.n + 31) / 32) * 8 + 64 > ctx->lds_per_block)
            return pg_fail(ctx, PG_ENOTSUP, "subject %d: n=%d exceeds the row-sort LDS bitmap", s, S.n);
        n[s] = S.n;
        t[s] = S.t;
        sp[s] = S.sparsity;
        max_n = std::max(max_n, S.n);
        if (S.bc) max_src = std::max<long long>(max_src, S.nsrc);
    }

    // plans and workspaces (every allocation happens here, before capture)
    PgCohortCorr *pc = nullptr;
    PgCohortMetrics *pm = nullptr;
    PG_TRY(pg_cohort_corr_plan(ctx, nsub, n.data(), t.data(), sp.data(), &pc));
    std::unique_ptr<PgCohortCorr, void (*)(PgCohortCorr *)> pc_guard(pc, pg_cohort_corr_free);
    // PAGANI_COHORT_LEVELS: BFS level launches per batch group in the graph
    int levels = BFS_LEVELS;
    if (const char *e = std::getenv("PAGANI_COHORT_LEVELS")) levels = std::max(1, std::min(1024, std::atoi(e)));
    PG_TRY(pg_cohort_metrics_plan(ctx, max_n, std::max<long long>(max_src, 1), levels, &pm));
    std::unique_ptr<PgCohortMetrics, void (*)(PgCohortMetrics *)> pm_guard(pm, pg_cohort_metrics_free);
    const int grp = pg_cohort_metrics_group(pm);
    // sources of every subject, padded with -1 to whole groups of batches
    std::vector<long long> src_off((size_t)nsub, 0);
    long long src_total = 0;
    for (int s = 0; s < nsub; s++) {
        src_off[s] = src_total;
        if (subj[s].bc && subj[s].nsrc > 0) src_total += pg_cdiv(subj[s].nsrc, grp) * grp;
    }
    int32_t *dsrc = nullptr;
    unsigned *dstatus = nullptr;
    unsigned long long *dts = nullptr;
    PG_TRY(pg_bufT(ctx, "coh.src", (size_t)src_total + 1, &dsrc));
    PG_TRY(pg_bufT(ctx, "coh.status", (size_t)nsub + 1, &dstatus));
    PG_TRY(pg_bufT(ctx, "coh.ts", (size_t)nsub + 1, &dts));
    {
        std::vector<int32_t> hsrc((size_t)src_total, -1);
        for (int s = 0; s < nsub; s++) {
            const pg_subject &S = subj[s];
            if (!S.bc || S.nsrc <= 0) continue;
            int32_t *dst = hsrc.data() + src_off[s];
            if (pg_is_device_ptr(S.sources))
                PG_HIP(ctx, hipMemcpy(dst, S.sources, (size_t)S.nsrc * 4, hipMemcpyDeviceToHost));
            else std::memcpy(dst, S.sources, (size_t)S.nsrc * 4);
            for (long long q = 0; q < S.nsrc; q++)
                if (dst[q] < 0 || dst[q] >= S.n)
                    return pg_fail(ctx, PG_EINVAL, "subject %d: source %d out of range", s, dst[q]);
        }
        if (src_total > 0)
            PG_HIP(ctx, hipMemcpy(dsrc, hsrc.data(), (size_t)src_total * 4, hipMemcpyHostToDevice));
    }
    PG_HIP(ctx, hipStreamSynchronize(st));

    // capture: every subject's launches back to back, a time stamp between;
    // consecutive subjects share a graph up to about GRAPH_NODES kernel nodes
    // (PAGANI_COHORT_GRAPH_NODES), and the graphs are launched back to back
    long long budget = GRAPH_NODES;
    if (const char *e = std::getenv("PAGANI_COHORT_GRAPH_NODES")) budget = std::max(1, std::atoi(e));
    std::vector<GraphGuard> graphs;
    graphs.reserve((size_t)nsub + 1);
    int s0 = 0;
    while (s0 < nsub) {
        graphs.emplace_back();
        GraphGuard &G = graphs.back();
        PG_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        pg_status cap = PG_OK;
        if (s0 == 0) clear_u32<<<(unsigned)pg_cdiv(nsub + 1, 256), 256, 0, st>>>(dstatus, nsub + 1);
        long long nodes = 0;
        int s = s0;
        for (; s < nsub && cap == PG_OK; s++) {
            const long long est =
                subject_nodes(pg_cohort_corr_in_graph(pc, s), subj[s].bc ? subj[s].nsrc : 0, grp, levels);
            if (s > s0 && nodes + est > budget) break;  // the next graph takes it
            nodes += est;
            pg_subject &S = subj[s];
            stamp<<<1, 64, 0, st>>>(dts, s);
            if (!pg_cohort_corr_in_graph(pc, s)) continue;
            cap = pg_cohort_corr_enqueue(pc, s, S.bold, S.row, S.col, dstatus + s);
            if (cap == PG_OK)
                cap = pg_cohort_metrics_enqueue(pm, S.row, S.col, S.n, dsrc + src_off[s], S.bc ? S.nsrc : 0, S.deg,
                                                S.cp, S.bc, dstatus + s);
        }
        if (s == nsub) stamp<<<1, 64, 0, st>>>(dts, nsub);
        const hipError_t ec = hipStreamEndCapture(st, &G.g);
        if (cap != PG_OK) return cap;
        PG_HIP(ctx, ec);
        PG_HIP(ctx, hipGraphInstantiate(&G.e, G.g, nullptr, nullptr, 0));
        size_t gn = 0;
        PG_HIP(ctx, hipGraphGetNodes(G.g, nullptr, &gn));
        inf.graph_nodes += (int32_t)gn;
        inf.graphs++;
        s0 = s;
    }
    inf.ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    for (int l = 0; l < launches; l++)
        for (auto &G : graphs) PG_HIP(ctx, hipGraphLaunch(G.e, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    std::vector<unsigned> hstatus((size_t)nsub + 1);
    std::vector<unsigned long long> hts((size_t)nsub + 1);
    PG_HIP(ctx, hipMemcpy(hstatus.data(), dstatus, (size_t)(nsub + 1) * 4, hipMemcpyDeviceToHost));
    PG_HIP(ctx, hipMemcpy(hts.data(), dts, (size_t)(nsub + 1) * 8, hipMemcpyDeviceToHost));
    int rate_khz = 0;
    PG_HIP(ctx, hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
    const double tick_ms = rate_khz > 0 ? 1.0 / (double)rate_khz : 0.0;
    inf.ms_graph = (double)(hts[nsub] - hts[0]) * tick_ms;
    for (int s = 0; s < nsub; s++) {
        subj[s].ms = (float)((double)(hts[s + 1] - hts[s]) * tick_ms);
        subj[s].status = pg_cohort_corr_in_graph(pc, s) ? hstatus[s] : PG_COHORT_OUTSIDE;
    }

    // subjects the graph could not finish: the ordinary entry points
    const auto r0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsub; s++) {
        pg_subject &S = subj[s];
        if (S.status == 0) continue;
        inf.reruns++;
        pg_corr_opts o;
        o.thr_kind = PG_THR_SPARSITY;
        o.thr = S.sparsity;
        o.weighted = 0;
        o.prec = PG_PREC_BF16X3;
        o.layout = PG_BOLD_TIME_MAJOR;
        pg_corr_info ci;
        PG_TRY(pg_corr_build(ctx, S.bold, S.n, S.t, &o, &ci));
        if (ci.nnz > 0 || S.n > 0) PG_TRY(pg_corr_csr(ctx, S.row, ci.nnz > 0 ? S.col : nullptr, nullptr));
        if (S.deg) PG_TRY(pg_degree(ctx, S.row, S.n, S.deg, nullptr));
        if (S.cp) PG_TRY(pg_cp_binary(ctx, S.row, S.col, S.n, ci.nnz, S.cp, nullptr));
        if (S.bc) {
            if (S.nsrc > 0) {
                PG_TRY(pg_bc(ctx, S.row, S.col, S.n, ci.nnz, dsrc + src_off[s], S.nsrc, S.bc));
            } else {
                fill_i32<<<(unsigned)pg_cdiv(2ll * S.n, 256), 256, 0, st>>>((int *)S.bc, 2ll * S.n, 0);  // 0.0
                PG_KERNEL_CHECK(ctx);
            }
        }
    }
    PG_HIP(ctx, hipStreamSynchronize(st));
    inf.ms_reruns = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
    if (info) *info = inf;
    return PG_OK;
}
