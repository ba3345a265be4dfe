// modularity.hip -- Newman leading-eigenvector modularity on MI355X (gfx950).
//
// Replaces Partition_GPU / Sub_Partition_GPU / Lead_Vector_GPU
// (W/Newman_Modularity_GPU/CUmodularity/modularity_GPU.cu:225-525) and the CPU
// twins used for groups of up to 25000 nodes (Modularity_CPU.cpp:164-298).
// The power iteration on B^(g) + beta I is one CSR SpMV (FP32 weights, FP64
// accumulation) plus the rank-one degree term and the diagonal correction; all
// norms and dot products are FP64 on the device.  Every round whose members
// are known (a wave) iterates together: two launches per iteration cover all
// of the wave's groups, each group stops on the device at the reference's
// test, and the host checks the wave once per BATCH iterations.
//
// The bisection bookkeeping (group membership, sub-CSR, label order) is host
// code like the reference's Partition_GPU and runs in the reference's round
// order, so labels match it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pg_internal.hpp"

namespace {

constexpr int MAX_ITER = 10000;      // modularity_GPU.cu:25
constexpr double EPS = 1e-6;         // modularity_GPU.cu:28
constexpr double RAYLEIGH_TOL = 1e-8;  // default relative tolerance of the quotient stop
constexpr int TPB = 256;             // threads per block
constexpr int NPB_MAX = 1024;        // nodes per work block, upper limit
constexpr int GSUM_MAX = 512;        // target for block partials per group
constexpr int BATCH = 64;            // power iterations queued per host check

// One wave of independent groups (the rounds created but not yet processed),
// concatenated: group g owns nodes [no[g], no[g] + ng[g]) of the wave and
// blocks [fb[g], fb[g] + nb[g]); a block never spans two groups, so every
// per-group sum is a fixed-order sum of its blocks' partials.
struct Wave {
    int nnodes, nblk;
    int npb;               // nodes per work block (64 .. NPB_MAX)
    const int *bgrp;       // [nblk] group of each block
    const int *no, *ng, *fb, *nb;   // [G]
    const int *gid;        // [nnodes] global id (start vector)
    const int *srow;       // [nnodes + 1] into scol/sval (wave-global)
    const int *scol;       // wave-global node index
    const float *sval;
    const double *k, *sumbg;
    double m2;
    int stopmode;          // pg_mod_stop
    double tol;            // Rayleigh-quotient tolerance
    int max_iter;          // iterations per pass
    // per group state
    int ngroups;
    int *stop, *iters, *it0;
    const int *restart;    // [G] 1: (re)start the group's pass from the start vector
    double *beta, *vnorm, *lam;
    // vectors: iteration li of a group reads buf[li & 1] and leaves its
    // normalised result in buf[(li + 1) & 1]; per-block partials
    double *buf[2], *vres, *pdot, *pss, *pe1, *pe2;
    // Rayleigh quotient: per-block partials of x.y and x.x (phase S), and the
    // quotient of iteration li of group g in rho[2 g + (li & 1)] (phase N)
    double *pxy, *pxx, *rho;
};

__device__ double block_sum(double x, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = x;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// fixed-order sum of the group's block partials (every block gets the same value)
__device__ double group_sum(const double *part, int f, int m, double *red)
{
    double s = 0.0;
    for (int q = threadIdx.x; q < m; q += TPB) s += part[f + q];
    return block_sum(s, red);
}

// One power iteration is two launches.  (S) the stop test on the previous
// iteration's partials, then y = B^(g) x + beta x (CSR SpMV, FP32 weights,
// FP64 accumulation, one wave per row) into the other buffer, and the partial
// of |y|^2; x . k comes from the partials of the previous launch.  (N) y /= |y|
// in place, the partials of |y - x|^2, |y + x|^2 and y . k.  Work blocks b =
// blockIdx.x, blockIdx.x + gridDim.x, ... cover the wave; every skip below is
// uniform in a block.
struct PhaseSmem {
    double red[TPB];
    double rowv[NPB_MAX];
    int skip;
};

__device__ __forceinline__ void phase_s(const Wave &w, int it, PhaseSmem &sm)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int b = blockIdx.x; b < w.nblk; b += gridDim.x) {
        const int g = w.bgrp[b];
        const int li = it - w.it0[g];  // iterations completed in this pass
        // stop[g] may be set by this launch's first block of g; every block of
        // g reaches the same decision below, so either reading skips
        if (threadIdx.x == 0) sm.skip = w.stop[g] || li < 0;
        __syncthreads();
        const int skip = sm.skip;
        __syncthreads();
        if (skip) continue;
        if (li > 0) {
            bool done;
            if (w.stopmode == PG_MOD_STOP_VECTOR) {
                const double e1 = sqrt(group_sum(w.pe1, w.fb[g], w.nb[g], sm.red));
                const double e2 = sqrt(group_sum(w.pe2, w.fb[g], w.nb[g], sm.red));
                done = !(e1 > EPS && e2 > EPS);
            } else {
                // quotients of the last two completed iterations (written by
                // earlier launches)
                done = false;
                if (li >= 2) {
                    const double ra = w.rho[2 * g + ((li - 1) & 1)], rb = w.rho[2 * g + ((li - 2) & 1)];
                    done = fabs(ra - rb) <= w.tol * fabs(ra);
                }
            }
            if (done || li >= w.max_iter) {
                if (b == w.fb[g] && threadIdx.x == 0) {
                    w.iters[g] = li;
                    w.stop[g] = 1;
                }
                continue;
            }
        }
        const double *x = w.buf[li & 1];
        double *y = w.buf[(li + 1) & 1];
        const double vk = group_sum(w.pdot, w.fb[g], w.nb[g], sm.red);
        const double beta = w.beta[g];
        const int i0 = w.no[g] + (b - w.fb[g]) * w.npb, i1 = min(w.no[g] + w.ng[g], i0 + w.npb);
        for (int i = i0 + wv; i < i1; i += TPB / 64) {
            double acc = 0.0;
            for (int e = w.srow[i] + lane; e < w.srow[i + 1]; e += 64) acc += x[w.scol[e]] * (double)w.sval[e];
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
            if (lane == 0) sm.rowv[i - i0] = acc;
        }
        __syncthreads();
        double s = 0.0, sxy = 0.0, sxx = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += TPB) {
            double acc = sm.rowv[i - i0];
            acc -= vk / w.m2 * w.k[i] + (w.sumbg[i] - beta) * x[i];
            y[i] = acc;
            s += acc * acc;
            sxy += x[i] * acc;
            sxx += x[i] * x[i];
        }
        s = block_sum(s, sm.red);
        if (threadIdx.x == 0) w.pss[b] = s;
        if (w.stopmode == PG_MOD_STOP_RAYLEIGH) {
            sxy = block_sum(sxy, sm.red);
            sxx = block_sum(sxx, sm.red);
            if (threadIdx.x == 0) {
                w.pxy[b] = sxy;
                w.pxx[b] = sxx;
            }
        }
    }
}

__device__ __forceinline__ void phase_n(const Wave &w, int it, PhaseSmem &sm)
{
#pragma clang fp contract(off)
    for (int b = blockIdx.x; b < w.nblk; b += gridDim.x) {
        const int g = w.bgrp[b];
        const int li = it - w.it0[g];
        if (w.stop[g] || li < 0) continue;  // stop[] is not written during this launch
        const double *x = w.buf[li & 1];
        double *y = w.buf[(li + 1) & 1];
        const double vn = sqrt(group_sum(w.pss, w.fb[g], w.nb[g], sm.red));
        const int i0 = w.no[g] + (b - w.fb[g]) * w.npb, i1 = min(w.no[g] + w.ng[g], i0 + w.npb);
        double p1 = 0.0, p2 = 0.0, pk = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += TPB) {
            const double v = y[i] / vn;
            y[i] = v;
            const double a = v - x[i], c = v + x[i];
            p1 += a * a;
            p2 += c * c;
            pk += v * w.k[i];
        }
        p1 = block_sum(p1, sm.red);
        p2 = block_sum(p2, sm.red);
        pk = block_sum(pk, sm.red);
        if (threadIdx.x == 0) {
            w.pe1[b] = p1;
            w.pe2[b] = p2;
            w.pdot[b] = pk;
            if (b == w.fb[g]) w.vnorm[g] = vn;
        }
        if (w.stopmode == PG_MOD_STOP_RAYLEIGH && b == w.fb[g]) {
            // this iteration's quotient from phase S's partials (fixed order)
            const double xy = group_sum(w.pxy, w.fb[g], w.nb[g], sm.red);
            const double xx = group_sum(w.pxx, w.fb[g], w.nb[g], sm.red);
            if (threadIdx.x == 0) w.rho[2 * g + (li & 1)] = xy / xx;
        }
    }
}

// One phase of one iteration.  (A cooperative single-launch form with grid
// barriers between the phases measured slower: 80 us against 62 us per
// iteration of the three-launch form on a 1.1M-edge wave.)
template <int P>
__global__ __launch_bounds__(TPB) void mod_phase(Wave w, int it)
{
    __shared__ PhaseSmem sm;
    if (P == 0) phase_s(w, it, sm);
    else phase_n(w, it, sm);
}

// Groups flagged in restart[]: buf[0] = global ids (the start vector) and the
// partials of x . k for the first iteration.
__global__ __launch_bounds__(TPB) void mod_start(Wave w)
{
#pragma clang fp contract(off)
    __shared__ double red[TPB];
    for (int b = blockIdx.x; b < w.nblk; b += gridDim.x) {
        const int g = w.bgrp[b];
        if (!w.restart[g]) continue;
        const int i0 = w.no[g] + (b - w.fb[g]) * w.npb, i1 = min(w.no[g] + w.ng[g], i0 + w.npb);
        double pk = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += TPB) {
            const double x = (double)w.gid[i];
            w.buf[0][i] = x;
            pk += x * w.k[i];
        }
        pk = block_sum(pk, red);
        if (threadIdx.x == 0) w.pdot[b] = pk;
    }
}

// For the groups in list[]: lam[g] = +-|v| by the sign of v[mi] v0[mi] at the
// first largest |v[mi]| (the reference's strict '>' scan), v = the final
// vector, v0 = the one before (Rayleigh stop: the last quotient); vres[] = v.
__global__ __launch_bounds__(TPB) void mod_lambda(Wave w, const int *list)
{
    __shared__ double rv[TPB];
    __shared__ int ri[TPB];
    const int g = list[blockIdx.x], tid = threadIdx.x;
    const int a = w.no[g], z = a + w.ng[g];
    const int li = w.iters[g];
    const double *v = w.buf[li & 1], *v0 = w.buf[(li + 1) & 1];
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int i = a + tid; i < z; i += TPB) {
        const double x = fabs(v[i]);
        w.vres[i] = v[i];
        if (x > best) { best = x; bi = i; }
    }
    rv[tid] = best;
    ri[tid] = bi;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (tid < o) {
            const double x = rv[tid + o];
            const int ix = ri[tid + o];
            if (x > rv[tid] || (x == rv[tid] && ix < ri[tid])) { rv[tid] = x; ri[tid] = ix; }
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int mi = ri[0];
        const double vn = w.vnorm[g];
        if (w.stopmode == PG_MOD_STOP_RAYLEIGH) w.lam[g] = w.rho[2 * g + ((li - 1) & 1)];
        else w.lam[g] = (mi < z && v[mi] * v0[mi] > 0) ? vn : -vn;
    }
}

}  // namespace

extern "C" pg_status pg_modularity(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val,
                                   int32_t n, int64_t nnz, int32_t kmode, int32_t *labels, double *q,
                                   double *lambda0, double *eig0, int32_t *rounds)
{
    pg_mod_opts o;
    std::memset(&o, 0, sizeof o);
    o.kmode = kmode;
    o.stop = PG_MOD_STOP_VECTOR;
    return pg_modularity_ex(ctx, row, col, val, n, nnz, &o, labels, q, lambda0, eig0, rounds);
}

extern "C" pg_status pg_modularity_ex(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val,
                                      int32_t n, int64_t nnz, const pg_mod_opts *opts, int32_t *labels,
                                      double *q, double *lambda0, double *eig0, int32_t *rounds)
{
    if (!ctx) return PG_EINVAL;
    if (!labels || !opts) return pg_fail(ctx, PG_EINVAL, "labels/opts is NULL");
    const int kmode = opts->kmode;
    if (kmode != PG_MOD_STRENGTH && kmode != PG_MOD_BINARY_K) return pg_fail(ctx, PG_EINVAL, "bad kmode");
    if (opts->stop != PG_MOD_STOP_VECTOR && opts->stop != PG_MOD_STOP_RAYLEIGH)
        return pg_fail(ctx, PG_EINVAL, "bad stop rule %d", opts->stop);
    if (opts->tol < 0.0 || opts->max_iter < 0 || opts->max_depth < 0)
        return pg_fail(ctx, PG_EINVAL, "bad tol / max_iter / max_depth");
    const double tol = opts->tol > 0.0 ? opts->tol : RAYLEIGH_TOL;
    const int max_iter = opts->max_iter > 0 ? opts->max_iter : MAX_ITER;
    PG_TRY(pg_check_csr_args(ctx, row, col, n, nnz));
    PG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // host copies: the bisection driver builds each group's sub-CSR
    std::vector<int> R((size_t)n + 1), Cc((size_t)nnz);
    std::vector<float> V(val ? (size_t)nnz : 0);
    auto fetch = [&](void *dst, const void *src, size_t bytes) -> pg_status {
        if (bytes == 0) return PG_OK;
        if (pg_is_device_ptr(src)) PG_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        else std::memcpy(dst, src, bytes);
        return PG_OK;
    };
    PG_TRY(fetch(R.data(), row, ((size_t)n + 1) * 4));
    PG_TRY(fetch(Cc.data(), col, (size_t)nnz * 4));
    if (val) PG_TRY(fetch(V.data(), val, (size_t)nnz * 4));
    // the driver indexes host and device arrays with these offsets and columns
    if (R[0] != 0 || R[n] != nnz) return pg_fail(ctx, PG_EFORMAT, "row[0]=%d row[n]=%d, nnz=%lld", R[0], R[n], (long long)nnz);
    for (int i = 0; i < n; i++)
        if (R[i] > R[i + 1]) return pg_fail(ctx, PG_EFORMAT, "row offsets decrease at %d", i);
    for (int64_t e = 0; e < nnz; e++)
        if (Cc[(size_t)e] < 0 || Cc[(size_t)e] >= n)
            return pg_fail(ctx, PG_EFORMAT, "column %d at %lld outside [0, %d)", Cc[(size_t)e], (long long)e, n);

    std::vector<double> k((size_t)n);
    double m2 = 0.0;
    for (int i = 0; i < n; i++) {
        double ki = 0.0;
        if (kmode == PG_MOD_BINARY_K || !val) ki = (double)(R[i + 1] - R[i]);
        else for (int e = R[i]; e < R[i + 1]; e++) ki += (double)V[e];
        k[i] = ki;
        m2 += ki;
    }
    std::vector<int> res((size_t)n, 0), adj((size_t)2 * n + 2, 0);
    int maxr = 0, index = 1, nrounds = 0;
    if (lambda0) *lambda0 = 0.0;
    if (n == 0 || m2 <= 0.0) {
        for (int i = 0; i < n; i++) res[i] = 0;
        adj[0] = 1;
        maxr = -1;
    }

    // Rounds are processed in waves: every round created by the previous wave
    // has fixed membership, so the whole wave's power iterations run together;
    // the bookkeeping (splits, new round numbers, labels) then runs in round
    // order exactly as the sequential driver (modularity_GPU.cu:405-525).
    int lo = 0, nwave = 0;
    std::vector<int> rloc((size_t)n);
    // PAGANI_MOD_TRACE=1: one line per wave on stderr (groups, nodes, edges,
    // iterations, seconds)
    const char *trace_env = std::getenv("PAGANI_MOD_TRACE");
    const bool trace = opts->progress || (trace_env && trace_env[0] == '1');
    while (lo <= maxr) {
        const int hi = maxr + 1;
        const int G = hi - lo;
        if (opts->max_depth > 0 && nwave >= opts->max_depth) {
            // depth limit: this level's groups are final, labelled in round order
            std::vector<char> used((size_t)G, 0);
            for (int i = 0; i < n; i++)
                if (res[i] >= lo && res[i] < hi) used[res[i] - lo] = 1;
            for (int g = 0; g < G; g++)
                if (used[g]) adj[lo + g] = index++;
            break;
        }
        const auto wave_t0 = std::chrono::steady_clock::now();
        // members of each round, in increasing global id
        std::vector<int> gcount((size_t)G, 0);
        for (int i = 0; i < n; i++)
            if (res[i] >= lo && res[i] < hi) gcount[res[i] - lo]++;
        std::vector<int> no((size_t)G), ng((size_t)G), fb((size_t)G), nb((size_t)G);
        int nnodes = 0, nblk = 0;
        // small work blocks for parallelism, large enough that no group has
        // more than GSUM_MAX block partials to sum
        int gmax = 1;
        for (int g = 0; g < G; g++) gmax = std::max(gmax, gcount[g]);
        int npb = 64;
        while (npb < NPB_MAX && pg_cdiv(gmax, npb) > GSUM_MAX) npb *= 2;
        for (int g = 0; g < G; g++) {
            no[g] = nnodes;
            ng[g] = gcount[g];
            fb[g] = nblk;
            nb[g] = (int)std::max<long long>(1, pg_cdiv(ng[g], npb));
            nnodes += ng[g];
            nblk += nb[g];
        }
        std::vector<int> gid((size_t)nnodes), fill(no), bgrp((size_t)nblk);
        for (int g = 0; g < G; g++)
            for (int q = 0; q < nb[g]; q++) bgrp[fb[g] + q] = g;
        for (int i = 0; i < n; i++)
            if (res[i] >= lo && res[i] < hi) {
                const int g = res[i] - lo;
                rloc[i] = fill[g];
                gid[fill[g]++] = i;
            }
        // sub-CSRs: edges inside the same round, wave-global node indices
        std::vector<int> srow((size_t)nnodes + 1, 0), scol;
        std::vector<float> sval;
        std::vector<double> kw((size_t)nnodes), sumbg((size_t)nnodes), Kg((size_t)G, 0.0);
        for (int a = 0; a < nnodes; a++) Kg[res[gid[a]] - lo] += k[gid[a]];
        scol.reserve((size_t)nnz);
        sval.reserve((size_t)nnz);
        for (int a = 0; a < nnodes; a++) {
            const int i = gid[a], r = res[i];
            double kin = 0.0;
            for (int e = R[i]; e < R[i + 1]; e++) {
                const int j = Cc[e];
                if (res[j] != r) continue;
                const float wv = val ? V[e] : 1.0f;
                scol.push_back(rloc[j]);
                sval.push_back(wv);
                kin += (kmode == PG_MOD_BINARY_K && val) ? 1.0 : (double)wv;
            }
            srow[a + 1] = (int)scol.size();
            kw[a] = k[i];
            sumbg[a] = kin - k[i] * Kg[r - lo] / m2;
        }
        const size_t gn = scol.size();

        Wave w;
        int *d_bgrp, *d_no, *d_ng, *d_fb, *d_nb, *d_gid, *d_srow, *d_scol, *d_restart, *d_stop, *d_iters, *d_it0,
            *d_list;
        float *d_sval;
        double *d_k, *d_sumbg, *d_beta, *d_vnorm, *d_lam, *d_b0, *d_b1, *d_vres, *d_pdot, *d_pss, *d_pe1, *d_pe2,
            *d_pxy, *d_pxx, *d_rho;
        PG_TRY(pg_bufT(ctx, "mod.bgrp", (size_t)nblk, &d_bgrp));
        PG_TRY(pg_bufT(ctx, "mod.no", (size_t)G, &d_no));
        PG_TRY(pg_bufT(ctx, "mod.ng", (size_t)G, &d_ng));
        PG_TRY(pg_bufT(ctx, "mod.fb", (size_t)G, &d_fb));
        PG_TRY(pg_bufT(ctx, "mod.nb", (size_t)G, &d_nb));
        PG_TRY(pg_bufT(ctx, "mod.gid", (size_t)nnodes + 1, &d_gid));
        PG_TRY(pg_bufT(ctx, "mod.srow", (size_t)nnodes + 1, &d_srow));
        PG_TRY(pg_bufT(ctx, "mod.scol", gn + 1, &d_scol));
        PG_TRY(pg_bufT(ctx, "mod.sval", gn + 1, &d_sval));
        PG_TRY(pg_bufT(ctx, "mod.k", (size_t)nnodes + 1, &d_k));
        PG_TRY(pg_bufT(ctx, "mod.sumbg", (size_t)nnodes + 1, &d_sumbg));
        PG_TRY(pg_bufT(ctx, "mod.restart", (size_t)G, &d_restart));
        PG_TRY(pg_bufT(ctx, "mod.stop", (size_t)G, &d_stop));
        PG_TRY(pg_bufT(ctx, "mod.iters", (size_t)G, &d_iters));
        PG_TRY(pg_bufT(ctx, "mod.it0", (size_t)G, &d_it0));
        PG_TRY(pg_bufT(ctx, "mod.list", (size_t)G, &d_list));
        PG_TRY(pg_bufT(ctx, "mod.beta", (size_t)G, &d_beta));
        PG_TRY(pg_bufT(ctx, "mod.vnorm", (size_t)G, &d_vnorm));
        PG_TRY(pg_bufT(ctx, "mod.lam", (size_t)G, &d_lam));
        PG_TRY(pg_bufT(ctx, "mod.b0", (size_t)nnodes + 1, &d_b0));
        PG_TRY(pg_bufT(ctx, "mod.b1", (size_t)nnodes + 1, &d_b1));
        PG_TRY(pg_bufT(ctx, "mod.vres", (size_t)nnodes + 1, &d_vres));
        PG_TRY(pg_bufT(ctx, "mod.pdot", (size_t)nblk, &d_pdot));
        PG_TRY(pg_bufT(ctx, "mod.pss", (size_t)nblk, &d_pss));
        PG_TRY(pg_bufT(ctx, "mod.pe1", (size_t)nblk, &d_pe1));
        PG_TRY(pg_bufT(ctx, "mod.pe2", (size_t)nblk, &d_pe2));
        PG_TRY(pg_bufT(ctx, "mod.pxy", (size_t)nblk, &d_pxy));
        PG_TRY(pg_bufT(ctx, "mod.pxx", (size_t)nblk, &d_pxx));
        PG_TRY(pg_bufT(ctx, "mod.rho", (size_t)G * 2, &d_rho));
        auto h2d = [&](void *dst, const void *src, size_t bytes) -> pg_status {
            if (bytes) PG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
            return PG_OK;
        };
        std::vector<int> zeros((size_t)G, 0);
        std::vector<double> dzeros((size_t)G, 0.0);
        PG_TRY(h2d(d_bgrp, bgrp.data(), (size_t)nblk * 4));
        PG_TRY(h2d(d_no, no.data(), (size_t)G * 4));
        PG_TRY(h2d(d_ng, ng.data(), (size_t)G * 4));
        PG_TRY(h2d(d_fb, fb.data(), (size_t)G * 4));
        PG_TRY(h2d(d_nb, nb.data(), (size_t)G * 4));
        PG_TRY(h2d(d_gid, gid.data(), (size_t)nnodes * 4));
        PG_TRY(h2d(d_srow, srow.data(), ((size_t)nnodes + 1) * 4));
        PG_TRY(h2d(d_scol, scol.data(), gn * 4));
        PG_TRY(h2d(d_sval, sval.data(), gn * 4));
        PG_TRY(h2d(d_k, kw.data(), (size_t)nnodes * 8));
        PG_TRY(h2d(d_sumbg, sumbg.data(), (size_t)nnodes * 8));
        std::vector<int> ones((size_t)G, 1);
        PG_TRY(h2d(d_restart, ones.data(), (size_t)G * 4));
        PG_TRY(h2d(d_stop, zeros.data(), (size_t)G * 4));
        PG_TRY(h2d(d_it0, zeros.data(), (size_t)G * 4));
        PG_TRY(h2d(d_beta, dzeros.data(), (size_t)G * 8));
        w.nnodes = nnodes;
        w.nblk = nblk;
        w.npb = npb;
        w.bgrp = d_bgrp;
        w.no = d_no;
        w.ng = d_ng;
        w.fb = d_fb;
        w.nb = d_nb;
        w.gid = d_gid;
        w.srow = d_srow;
        w.scol = d_scol;
        w.sval = d_sval;
        w.k = d_k;
        w.sumbg = d_sumbg;
        w.m2 = m2;
        w.stopmode = opts->stop;
        w.tol = tol;
        w.max_iter = max_iter;
        w.ngroups = G;
        w.stop = d_stop;
        w.restart = d_restart;
        w.iters = d_iters;
        w.it0 = d_it0;
        w.beta = d_beta;
        w.vnorm = d_vnorm;
        w.lam = d_lam;
        w.buf[0] = d_b0;
        w.buf[1] = d_b1;
        w.vres = d_vres;
        w.pdot = d_pdot;
        w.pss = d_pss;
        w.pe1 = d_pe1;
        w.pe2 = d_pe2;
        w.pxy = d_pxy;
        w.pxx = d_pxx;
        w.rho = d_rho;

        // pass state per group: 0 first pass, 1 shifted pass, 2 finished
        std::vector<int> pass((size_t)G, 0), list, hit0((size_t)G, 0);
        std::vector<double> lam_total((size_t)G, 0.0), hbeta((size_t)G, 0.0);
        mod_start<<<(unsigned)nblk, TPB, 0, st>>>(w);
        PG_KERNEL_CHECK(ctx);
        int it = 0, running = G;
        std::vector<int> hdone((size_t)G);
        std::vector<double> hlam((size_t)G);
        while (running > 0) {
            for (int q = 0; q < BATCH; q++) {
                mod_phase<0><<<(unsigned)nblk, TPB, 0, st>>>(w, it + q);
                mod_phase<1><<<(unsigned)nblk, TPB, 0, st>>>(w, it + q);
            }
            it += BATCH;
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, hipMemcpyAsync(hdone.data(), d_stop, (size_t)G * 4, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            list.clear();
            for (int g = 0; g < G; g++)
                if (pass[g] < 2 && hdone[g]) list.push_back(g);
            if (list.empty()) continue;
            const int m = (int)list.size();
            PG_TRY(h2d(d_list, list.data(), (size_t)m * 4));
            mod_lambda<<<(unsigned)m, TPB, 0, st>>>(w, d_list);
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, hipMemcpyAsync(hlam.data(), d_lam, (size_t)G * 8, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            std::vector<int> restart;
            for (int g : list) {
                const double lam = hlam[g];
                if (pass[g] == 0) {
                    lam_total[g] = lam;
                    // shift rule: groups above 25000 nodes use lambda < -0.01
                    // (modularity_GPU.cu:369), smaller ones lambda < 0
                    // (Modularity_CPU.cpp:183)
                    const bool shift = ng[g] > 25000 ? (lam < -0.01) : (lam < 0.0);
                    if (shift) {
                        pass[g] = 1;
                        restart.push_back(g);
                        continue;
                    }
                } else {
                    lam_total[g] += lam;
                }
                pass[g] = 2;
                running--;
            }
            if (!restart.empty()) {
                // second pass on B + beta I, beta = -lambda, from the start vector
                // no iteration ran since hdone (stop flags) was read: it is
                // the device state
                std::vector<int> rflag((size_t)G, 0);
                for (int g : restart) {
                    hbeta[g] = -lam_total[g];
                    hit0[g] = it;
                    hdone[g] = 0;
                    rflag[g] = 1;
                }
                PG_TRY(h2d(d_restart, rflag.data(), (size_t)G * 4));
                mod_start<<<(unsigned)nblk, TPB, 0, st>>>(w);
                PG_KERNEL_CHECK(ctx);
                PG_TRY(h2d(d_beta, hbeta.data(), (size_t)G * 8));
                PG_TRY(h2d(d_it0, hit0.data(), (size_t)G * 4));
                PG_TRY(h2d(d_stop, hdone.data(), (size_t)G * 4));
                PG_HIP(ctx, hipStreamSynchronize(st));
            }
        }
        std::vector<double> hv((size_t)nnodes);
        PG_HIP(ctx, hipMemcpyAsync(hv.data(), d_vres, (size_t)nnodes * 8, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        if (trace) {
            std::vector<int> hits((size_t)G);
            PG_HIP(ctx, hipMemcpy(hits.data(), d_iters, (size_t)G * 4, hipMemcpyDeviceToHost));
            int mx = 0, big = 0;
            for (int g = 0; g < G; g++) {
                mx = std::max(mx, hits[g]);
                big = std::max(big, ng[g]);
            }
            const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - wave_t0).count();
            std::fprintf(stderr, "pg_modularity wave %d: rounds %d..%d, %d nodes (largest group %d), %zu edges, "
                         "%d launched iterations, max %d in one pass, %.3f s\n",
                         nwave, lo, hi - 1, nnodes, big, gn, it, mx, secs);
        }
        nwave++;

        // bookkeeping in round order
        for (int g = 0; g < G; g++) {
            const int round = lo + g;
            if (ng[g] == 0) continue;
            nrounds++;
            const double lam = lam_total[g];
            const double *gv = hv.data() + no[g];
            if (round == 0) {
                if (lambda0) *lambda0 = lam;
                if (eig0) {
                    std::vector<double> e0((size_t)n, 0.0);
                    for (int a = 0; a < ng[g]; a++) e0[gid[no[g] + a]] = gv[a];
                    if (pg_is_device_ptr(eig0)) PG_HIP(ctx, hipMemcpy(eig0, e0.data(), (size_t)n * 8, hipMemcpyHostToDevice));
                    else std::memcpy(eig0, e0.data(), (size_t)n * 8);
                }
            }
            long long sp = 0, sn = 0;
            for (int a = 0; a < ng[g]; a++) {
                sp += gv[a] > 0;
                sn += gv[a] <= 0;
            }
            const bool split = lam > 0.01 && sp > 1 && sn > 1;  // LAMBDA, MIN_GROUP
            if (split) {
                for (int a = 0; a < ng[g]; a++)
                    res[gid[no[g] + a]] = maxr + 1 + (gv[a] * ((double)(sp - sn) + 0.5) <= 0);
                maxr += 2;
            } else {
                adj[round] = index++;
            }
        }
        lo = hi;
    }
    std::vector<int> lab((size_t)n);
    for (int i = 0; i < n; i++) lab[i] = adj[res[i]];
    if (pg_is_device_ptr(labels)) PG_HIP(ctx, hipMemcpy(labels, lab.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    else std::memcpy(labels, lab.data(), (size_t)n * 4);
    if (q) {
        // W modularity_GPU.cu:500-509 with the same k and m convention
        int maxl = 0;
        for (int i = 0; i < n; i++) maxl = std::max(maxl, lab[i]);
        std::vector<double> kc((size_t)maxl + 2, 0.0);
        double in = 0.0;
        for (int i = 0; i < n; i++) {
            kc[lab[i]] += k[i];
            for (int e = R[i]; e < R[i + 1]; e++)
                if (lab[Cc[e]] == lab[i]) in += (kmode == PG_MOD_BINARY_K || !val) ? 1.0 : (double)V[e];
        }
        double s2 = 0.0;
        for (double x : kc) s2 += x * x;
        *q = m2 > 0.0 ? (in - s2 / m2) / m2 : 0.0;
    }
    if (rounds) *rounds = nrounds;
    return PG_OK;
}
