// efficiency.hip -- binary nodal efficiency and characteristic path length
// from all-pairs hop counts (U/BFS_MulCPU/BFS_MulCPU/APSP_BFS.cpp:100-192),
// MI355X (gfx950).
//
// The reference fills a dense n x n float hop matrix with one queue BFS per
// source and then walks it.  Here the hop counts never exist as a matrix:
//
//  * Bit-parallel BFS.  A batch of S sources (S a multiple of 4096) gives every
//    vertex S/64 visited words, one bit per source.  Level d -> d+1 is one
//    pull launch: one wave per (vertex w, 64-word chunk), lane k owns word k of
//    the chunk and ORs word k of every neighbour's visited set (512 contiguous
//    bytes per neighbour per wave), stopping as soon as the bits w still lacks
//    are all found.  New bits are the sources at distance exactly d+1.
//  * Bit-sliced levels.  Bit b of the level at which (w, s) is reached is
//    OR-ed into plane b, so a batch keeps ceil(log2(depth+1)) planes of S/8
//    bytes per vertex instead of a byte or a float per pair.  Planes are
//    allocated as the depth grows.
//  * Ordered sums.  One lane per source walks j = 0..n-1 in order, rebuilds
//    d(s, j) from the planes (wave-uniform words) and adds 1.0f / d in FP32
//    -- the reference's precision and order, so e_s is bit-identical -- and
//    (double)(1.0f / d) for j > s into the source's FP64 share of the Lp sum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pg_internal.hpp"

namespace {

typedef unsigned long long u64;

constexpr int MAXPLANES = 31;  // levels < 2^31
constexpr int LEVEL_GROUP = 4;  // level launches per host progress check

struct Planes {
    u64 *p[MAXPLANES];
};

// visited words of a batch: vis[v * W + k]; sources base + 64k + b
__global__ void eff_init(int n, int W, int base, int nsrc, u64 *__restrict__ vis)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * W) return;
    const int v = (int)(t / W), k = (int)(t % W);
    const int lo = 64 * k;
    u64 x;
    if (lo >= nsrc) x = ~0ull;                // every bit an absent source
    else if (nsrc - lo >= 64) x = 0ull;
    else x = ~0ull << (nsrc - lo);
    const int own = v - base;
    if (own >= 0 && own < nsrc && (own >> 6) == k) x |= 1ull << (own & 63);
    vis[t] = x;
}

__global__ void eff_zero(long long count, u64 *__restrict__ p)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = 0ull;
}

// Level d -> d+1.  prog[d] counts waves that found new bits; a launch whose
// previous level found none returns at once (the BFS has ended).
__global__ __launch_bounds__(256) void eff_level(const int *__restrict__ row, const int *__restrict__ col, int n,
                                                 int W, int d, const u64 *__restrict__ visA,
                                                 u64 *__restrict__ visB, Planes planes, int nplanes,
                                                 unsigned *__restrict__ prog)
{
    if (d > 0 && prog[d - 1] == 0u) return;
    const int lane = threadIdx.x & 63;
    const int chunks = W >> 6;
    const long long wid = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= (long long)n * chunks) return;
    const int w = __builtin_amdgcn_readfirstlane((int)(wid / chunks));
    const int c = __builtin_amdgcn_readfirstlane((int)(wid % chunks));
    const int k = (c << 6) + lane;
    const u64 vis = visA[(size_t)w * W + k];
    const u64 need = ~vis;
    u64 acc = 0;
    if (need) {
        const int a = row[w], b = row[w + 1];
        int e = a;
        for (; e + 8 <= b; e += 8) {
            u64 x = 0;
#pragma unroll
            for (int q = 0; q < 8; q++) x |= visA[(size_t)col[e + q] * W + k];
            acc |= x;
            if ((acc & need) == need) break;
        }
        if ((acc & need) != need)
            for (; e < b; e++) acc |= visA[(size_t)col[e] * W + k];
    }
    const u64 nb = acc & need;
    visB[(size_t)w * W + k] = vis | nb;
    if (nb) {
        const unsigned lv = (unsigned)d + 1u;
        for (int b = 0; b < nplanes; b++)
            if ((lv >> b) & 1u) planes.p[b][(size_t)w * W + k] |= nb;
    }
    if (__ballot(nb != 0) != 0 && lane == 0) atomicAdd(&prog[d], 1u);
}

// Level 0 -> 1 from the column indices alone: the sources at distance 1 from
// w are its neighbours inside the batch.  One wave per (w, 64-word chunk); the
// chunk's 4096 sources are a contiguous column range of w's sorted row, found
// from a sampled ballot and read 64 entries at a time, each such neighbour
// setting its bit in the LDS word of the lane owning it -- no neighbour
// visited words are read.
__global__ __launch_bounds__(256) void eff_level0(const int *__restrict__ row, const int *__restrict__ col, int n,
                                                  int W, int base, const u64 *__restrict__ visA,
                                                  u64 *__restrict__ visB, Planes planes,
                                                  unsigned *__restrict__ prog)
{
    __shared__ u64 sbits[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int chunks = W >> 6;
    const long long wid = (long long)blockIdx.x * 4 + wv;
    if (wid >= (long long)n * chunks) return;
    const int w = __builtin_amdgcn_readfirstlane((int)(wid / chunks));
    const int c = __builtin_amdgcn_readfirstlane((int)(wid % chunks));
    const int k = (c << 6) + lane;
    const long long v0 = (long long)base + 4096ll * c, v1 = v0 + 4096;
    const int a = row[w], b = row[w + 1], dw = b - a;
    // a start at or before the first column >= v0: 64 evenly spaced samples
    // and a ballot (the loop below filters v >= v0 itself)
    int start = b;
    if (dw > 0) {
        const int ps = a + (int)(((long long)lane * dw) >> 6);  // < b
        const unsigned long long m = __ballot(col[ps] >= v0);
        const int f = m ? __builtin_ctzll(m) : 64;
        start = f == 0 ? a : a + (int)(((long long)(f - 1) * dw) >> 6);
    }
    // the chunk's neighbours, 64 at a time, each OR-ed into the LDS word of
    // the lane owning it
    sbits[wv][lane] = 0ull;
    __threadfence_block();
    for (int e0 = start; e0 < b; e0 += 64) {
        const int e = e0 + lane;
        const long long v = e < b ? (long long)col[e] : v1;
        if (v >= v0 && v < v1) {
            const int rel = (int)(v - v0);
            atomicOr(&sbits[wv][rel >> 6], 1ull << (rel & 63));
        }
        if (__ballot(e < b && v >= v1)) break;  // columns ascend: the chunk is done
    }
    __threadfence_block();
    const u64 acc = atomicOr(&sbits[wv][lane], 0ull);
    const u64 vis = visA[(size_t)w * W + k];
    const u64 nb = acc & ~vis;
    visB[(size_t)w * W + k] = vis | nb;
    if (nb) planes.p[0][(size_t)w * W + k] |= nb;
    if (__ballot(nb != 0) != 0 && lane == 0) atomicAdd(&prog[0], 1u);
}

// One wave per 64 sources (word k of the batch), one lane per source.  The
// plane words of 64 consecutive j are fetched by one vector load per plane
// (lane l holds j0 + l), then j walks them in order through readlane, so the
// FP32 sums keep the reference's order while the loads are batched.
template <int NP>
__global__ __launch_bounds__(64) void eff_sum(int n, int W, int base, int nsrc, Planes planes,
                                              float *__restrict__ eff, double *__restrict__ lpp,
                                              int *__restrict__ flags)
{
    const int lane = threadIdx.x;
    const int k = blockIdx.x;
    const int s = base + 64 * k + lane;
    float acc = 0.0f;
    double dacc = 0.0;
    int reached = 0;
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int jl = j0 + lane;
        uint32_t wlo[NP], whi[NP];
#pragma unroll
        for (int b = 0; b < NP; b++) {
            const u64 x = jl < n ? planes.p[b][(size_t)jl * W + k] : 0ull;
            wlo[b] = (uint32_t)x;
            whi[b] = (uint32_t)(x >> 32);
        }
        const int cnt = min(64, n - j0);
        for (int jj = 0; jj < cnt; jj++) {
            int lv = 0;
#pragma unroll
            for (int b = 0; b < NP; b++) {
                const uint32_t lo = __builtin_amdgcn_readlane(wlo[b], jj);
                const uint32_t hi = __builtin_amdgcn_readlane(whi[b], jj);
                const uint32_t word = lane < 32 ? lo : hi;
                lv |= (int)((word >> (lane & 31)) & 1u) << b;
            }
            if (lv != 0) {
                const float r = 1.0f / (float)lv;
                acc += r;
                reached++;
                if (j0 + jj > s) dacc += (double)r;
            }
        }
    }
    if (64 * k + lane >= nsrc) return;
    const float e = acc / (float)(n - 1);
    eff[s] = e;
    lpp[s] = dacc;
    if (e == 0.0f) atomicAdd(&flags[0], 1);
    if (reached < n - 1) atomicOr(&flags[1], 1);
}

template <int NP>
void launch_sum(hipStream_t st, int n, int W, int base, int nsrc, const Planes &pl, float *eff, double *lpp,
                int *flags)
{
    eff_sum<NP><<<(unsigned)(W), 64, 0, st>>>(n, W, base, nsrc, pl, eff, lpp, flags);
}

typedef void (*SumFn)(hipStream_t, int, int, int, int, const Planes &, float *, double *, int *);

SumFn sum_fn(int np)
{
    static const SumFn t[MAXPLANES + 1] = {
        launch_sum<1>,  launch_sum<1>,  launch_sum<2>,  launch_sum<3>,  launch_sum<4>,  launch_sum<5>,
        launch_sum<6>,  launch_sum<7>,  launch_sum<8>,  launch_sum<9>,  launch_sum<10>, launch_sum<11>,
        launch_sum<12>, launch_sum<13>, launch_sum<14>, launch_sum<15>, launch_sum<16>, launch_sum<17>,
        launch_sum<18>, launch_sum<19>, launch_sum<20>, launch_sum<21>, launch_sum<22>, launch_sum<23>,
        launch_sum<24>, launch_sum<25>, launch_sum<26>, launch_sum<27>, launch_sum<28>, launch_sum<29>,
        launch_sum<30>, launch_sum<31>};
    return t[np];
}

int bit_width(unsigned x)
{
    int b = 0;
    while (x) {
        b++;
        x >>= 1;
    }
    return b;
}

// Sources per batch: all of them when the visited words and four planes fit
// in the budget, else the largest multiple of 4096 that does.
int batch_sources(int n, size_t budget)
{
    const long long all = pg_cdiv(n, 4096) * 4096;
    const double per_source = (double)n / 8.0 * 6.0;
    long long fit = (long long)((double)budget / per_source) / 4096 * 4096;
    fit = std::max<long long>(fit, 4096);
    return (int)std::min(all, fit);
}

}  // namespace

extern "C" pg_status pg_bfs_efficiency(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                                       int64_t nnz, float *eff, double *lp, int32_t *isolated,
                                       int32_t *unconnected)
{
    if (!ctx) return PG_EINVAL;
    PG_HIP(ctx, hipSetDevice(ctx->device));
    PG_TRY(pg_check_csr_args(ctx, row, col, n, nnz));
    if (n < 1) return pg_fail(ctx, PG_EINVAL, "n must be >= 1");
    const int *drow, *dcol;
    PG_TRY(pg_in(ctx, row, (size_t)n + 1, "in.row", &drow));
    PG_TRY(pg_in(ctx, col, (size_t)nnz, "in.col", &dcol));
    hipStream_t st = ctx->stream;

    const char *env = std::getenv("PAGANI_EFF_BATCH");
    int S = batch_sources(n, (size_t)4 << 30);
    if (env && std::atoi(env) > 0) S = (int)std::min<long long>(pg_cdiv(std::atoi(env), 4096) * 4096, S);
    const int W = S / 64;
    const size_t words = (size_t)n * W;
    u64 *va, *vb;
    unsigned *prog;
    float *deff;
    double *lpp;
    int *flags;
    PG_TRY(pg_bufT(ctx, "eff.visA", words, &va));
    PG_TRY(pg_bufT(ctx, "eff.visB", words, &vb));
    PG_TRY(pg_bufT(ctx, "eff.prog", (size_t)n + LEVEL_GROUP + 2, &prog));
    PG_TRY(pg_bufT(ctx, "eff.lpp", (size_t)n, &lpp));
    PG_TRY(pg_bufT(ctx, "eff.flags", 2, &flags));
    PG_TRY(pg_out(ctx, eff, (size_t)n, "out.eff", &deff));
    if (!deff) PG_TRY(pg_bufT(ctx, "eff.scratch", (size_t)n, &deff));
    PG_HIP(ctx, hipMemsetAsync(flags, 0, 8, st));
    PG_TRY(pg_validate_csr(ctx, drow, dcol, n, nnz, true));

    Planes pl = {};
    int nplanes = 0;
    auto grow = [&](int want) -> pg_status {
        for (; nplanes < want; nplanes++) {
            char name[32];
            std::snprintf(name, sizeof name, "eff.plane%d", nplanes);
            PG_TRY(pg_bufT(ctx, name, words, &pl.p[nplanes]));
            eff_zero<<<(unsigned)pg_cdiv((long long)words, 256), 256, 0, st>>>((long long)words, pl.p[nplanes]);
        }
        return PG_OK;
    };

    for (int base = 0; base < n; base += S) {
        const int nsrc = std::min(S, n - base);
        for (int b = 0; b < nplanes; b++)
            eff_zero<<<(unsigned)pg_cdiv((long long)words, 256), 256, 0, st>>>((long long)words, pl.p[b]);
        PG_HIP(ctx, hipMemsetAsync(prog, 0, ((size_t)n + LEVEL_GROUP + 2) * 4, st));
        eff_init<<<(unsigned)pg_cdiv((long long)words, 256), 256, 0, st>>>(n, W, base, nsrc, va);
        PG_KERNEL_CHECK(ctx);
        const unsigned blocks = (unsigned)pg_cdiv((long long)n * (W / 64), 4);
        int depth = -1;
        for (int d = 0; depth < 0; d += LEVEL_GROUP) {
            if (d > n) return pg_fail(ctx, PG_EDEVICE, "BFS did not converge");
            const int top = std::min(d + LEVEL_GROUP, n + 1);  // levels reachable by this group
            if (bit_width((unsigned)top) > MAXPLANES) return pg_fail(ctx, PG_ENOTSUP, "graph too deep");
            PG_TRY(grow(bit_width((unsigned)top)));
            for (int q = 0; q < LEVEL_GROUP; q++) {
                if (d + q == 0)
                    eff_level0<<<blocks, 256, 0, st>>>(drow, dcol, n, W, base, va, vb, pl, prog);
                else
                    eff_level<<<blocks, 256, 0, st>>>(drow, dcol, n, W, d + q, va, vb, pl, nplanes, prog);
                std::swap(va, vb);
            }
            PG_KERNEL_CHECK(ctx);
            unsigned h[LEVEL_GROUP];
            PG_HIP(ctx, hipMemcpyAsync(h, prog + d, sizeof h, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            for (int q = 0; q < LEVEL_GROUP && depth < 0; q++)
                if (h[q] == 0) depth = d + q;  // deepest level present
        }
        const int np = std::max(1, bit_width((unsigned)depth));
        sum_fn(np)(st, n, W, base, nsrc, pl, deff, lpp, flags);
        PG_KERNEL_CHECK(ctx);
    }
    PG_TRY(pg_out_finish(ctx, eff, deff, (size_t)n));
    std::vector<double> h((size_t)n);
    int hf[2] = {0, 0};
    PG_HIP(ctx, hipMemcpyAsync(h.data(), lpp, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipMemcpyAsync(hf, flags, 8, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (lp) {
        double x = 0.0;
        for (int s = 0; s < n; s++) x += h[(size_t)s];
        x /= (double)n;
        x /= (double)(n - 1);
        *lp = 1 / x / 2;
    }
    if (isolated) *isolated = hf[0];
    if (unconnected) *unconnected = hf[1];
    return PG_OK;
}

namespace {
// bad = 1 unless row is non-decreasing from row[0] = 0 to row[n] = nnz and
// every col is in [0, n); sorted: also strictly ascending columns inside each
// row.  One wave per row; a row whose bounds fall outside [0, nnz] is flagged
// before any of its columns is read, so the check itself stays in bounds.
// col == nullptr checks the row offsets only.
__global__ __launch_bounds__(256) void csr_validate(const int *__restrict__ row, const int *__restrict__ col, int n,
                                                    long long nnz, int sorted, int *__restrict__ bad)
{
    const long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    bool ok = true;
    if (w < n) {
        const long long a = row[w], b = row[w + 1];
        ok = a >= 0 && a <= b && b <= nnz;
        if (ok && col)
            for (long long e = a + lane; e < b; e += 64) {
                const int c = col[e];
                ok = ok && c >= 0 && c < n && !(sorted && e > a && col[e - 1] >= c);
            }
    }
    if (w == 0 && lane == 0) ok = ok && row[0] == 0 && (long long)row[n] == nnz;
    if (!ok) atomicOr(bad, 1);
}
}  // namespace

namespace {
__global__ void flag_clear(int *f) { *f = 0; }
}  // namespace

pg_status pg_validate_csr_flag(pg_ctx *ctx, const int32_t *drow, const int32_t *dcol, int32_t n, int64_t nnz,
                               bool sorted, int *bad)
{
    hipStream_t st = ctx->stream;
    int hbad = 0;
    flag_clear<<<1, 1, 0, st>>>(bad);
    const long long waves = std::max<long long>(n, 1);
    csr_validate<<<(unsigned)pg_cdiv(waves, 4), 256, 0, st>>>(drow, dcol, n, nnz, sorted ? 1 : 0, bad);
    PG_KERNEL_CHECK(ctx);
    PG_HIP(ctx, hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (hbad)
        return pg_fail(ctx, PG_EFORMAT, "malformed CSR (row not 0..nnz non-decreasing, col out of range%s)",
                       sorted ? " or columns not ascending in a row" : "");
    return PG_OK;
}

pg_status pg_validate_csr(pg_ctx *ctx, const int32_t *drow, const int32_t *dcol, int32_t n, int64_t nnz,
                          bool sorted)
{
    int *bad;
    PG_TRY(pg_bufT(ctx, "csr.bad", 1, &bad));
    return pg_validate_csr_flag(ctx, drow, dcol, n, nnz, sorted, bad);
}
