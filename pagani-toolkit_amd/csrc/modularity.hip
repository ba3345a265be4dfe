// modularity.hip -- Newman leading-eigenvector modularity on MI355X (gfx950).
//
// Replaces Partition_GPU / Sub_Partition_GPU / Lead_Vector_GPU
// (W/Newman_Modularity_GPU/CUmodularity/modularity_GPU.cu:225-525) and the CPU
// twins used for groups of up to 25000 nodes (Modularity_CPU.cpp:164-298).
// The power iteration on B^(g) + beta I is one CSR SpMV (FP32 weights, FP64
// accumulation) plus the rank-one degree term and the diagonal correction; all
// norms and dot products are FP64 on the device.  Groups up to SMALL_N nodes
// run the whole iteration inside one work-group (no host round trips); larger
// groups use one kernel per step with device-side reductions.
//
// The bisection driver (group membership, sub-CSR, label bookkeeping) is host
// code like the reference's Partition_GPU; groups are processed in the
// reference's Round order so labels match it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pg_internal.hpp"

namespace {

constexpr int SMALL_N = 16384;
constexpr int MAX_ITER = 10000;      // modularity_GPU.cu:25
constexpr double EPS = 1e-6;         // modularity_GPU.cu:28
constexpr int TPB = 1024;

struct Group {
    int ng;
    const int *gid;     // global ids (for the starting vector)
    const int *srow;    // [ng+1]
    const int *scol;    // local ids
    const double *sval;
    const double *kg;
    const double *sumbg;
};

__device__ double block_sum1024(double v, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int off = TPB / 2; off > 0; off >>= 1) {
        if (tid < off) red[tid] += red[tid + off];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// First index of the largest |v| (ties to the lowest index, as the
// reference's strict '>' scan).
__device__ int block_argmax_abs(const double *v, int ng, double *redv, int *redi)
{
    const int tid = threadIdx.x;
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int i = tid; i < ng; i += TPB) {
        double a = fabs(v[i]);
        if (a > best) { best = a; bi = i; }
    }
    redv[tid] = best;
    redi[tid] = bi;
    __syncthreads();
    for (int off = TPB / 2; off > 0; off >>= 1) {
        if (tid < off) {
            double a = redv[tid + off];
            int ia = redi[tid + off];
            if (a > redv[tid] || (a == redv[tid] && ia < redi[tid])) { redv[tid] = a; redi[tid] = ia; }
        }
        __syncthreads();
    }
    int r = redi[0];
    __syncthreads();
    return r;
}

// Whole power iteration (both the plain and the shifted pass) of one group.
__global__ __launch_bounds__(TPB) void lead_small(Group g, double m2, double *__restrict__ v,
                                                  double *__restrict__ v0, double *__restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double red[TPB];
    __shared__ int redi[TPB];
    const int tid = threadIdx.x, ng = g.ng;
    double lam_total = 0.0;
    int iters = 0;
    for (int pass = 0; pass < 2; pass++) {
        const double beta = pass == 0 ? 0.0 : -lam_total;
        for (int i = tid; i < ng; i += TPB) v[i] = (double)g.gid[i];
        __syncthreads();
        double err1 = 1.0, err2 = 1.0, vnorm = 0.0;
        int it = 0;
        while (err1 > EPS && err2 > EPS && it < MAX_ITER) {
            double pk = 0.0;
            for (int i = tid; i < ng; i += TPB) {
                double x = v[i];
                v0[i] = x;
                pk += x * g.kg[i];
            }
            const double vk = block_sum1024(pk, red);  // includes a barrier after v0 writes
            double py = 0.0;
            for (int i = tid; i < ng; i += TPB) {
                double acc = 0.0;
                for (int e = g.srow[i]; e < g.srow[i + 1]; e++) acc += v0[g.scol[e]] * g.sval[e];
                acc -= vk / m2 * g.kg[i] + (g.sumbg[i] - beta) * v0[i];
                v[i] = acc;
                py += acc * acc;
            }
            vnorm = sqrt(block_sum1024(py, red));
            double p1 = 0.0, p2 = 0.0;
            for (int i = tid; i < ng; i += TPB) {
                double x = v[i] / vnorm;
                v[i] = x;
                double a = x - v0[i], b = x + v0[i];
                p1 += a * a;
                p2 += b * b;
            }
            err1 = sqrt(block_sum1024(p1, red));
            err2 = sqrt(block_sum1024(p2, red));
            it++;
        }
        iters += it;
        const int mi = block_argmax_abs(v, ng, red, redi);
        const double lam = (v[mi] * v0[mi] > 0) ? vnorm : -vnorm;
        if (pass == 0) {
            lam_total = lam;
            // shift rule: groups above 25000 nodes use lambda < -0.01
            // (modularity_GPU.cu:369), smaller ones lambda < 0 (Modularity_CPU.cpp:183)
            bool shift = ng > 25000 ? (lam < -0.01) : (lam < 0.0);
            if (!shift) break;
        } else {
            lam_total += lam;
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = lam_total;
        out[1] = (double)iters;
    }
}

// ---- large groups: one kernel per step, FP64 reductions through partials

__global__ void big_start(Group g, double *__restrict__ v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g.ng) v[i] = (double)g.gid[i];
}

// v0 = v, partial sums of v0 . k
__global__ __launch_bounds__(TPB) void big_copy_dot(Group g, const double *__restrict__ v, double *__restrict__ v0,
                                                    double *__restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double red[TPB];
    double s = 0.0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < g.ng; i += gridDim.x * TPB) {
        double x = v[i];
        v0[i] = x;
        s += x * g.kg[i];
    }
    s = block_sum1024(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// scalars[slot] = sqrt?(sum of part[0..np))
__global__ __launch_bounds__(TPB) void big_reduce(const double *__restrict__ part, int np, double *__restrict__ scal,
                                                  int slot, int take_sqrt)
{
    __shared__ double red[TPB];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += TPB) s += part[i];
    s = block_sum1024(s, red);
    if (threadIdx.x == 0) scal[slot] = take_sqrt ? sqrt(s) : s;
}

__global__ __launch_bounds__(TPB) void big_spmv(Group g, double m2, double beta, const double *__restrict__ v0,
                                                double *__restrict__ v, const double *__restrict__ scal,
                                                double *__restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double red[TPB];
    const double vk = scal[0];
    double s = 0.0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < g.ng; i += gridDim.x * TPB) {
        double acc = 0.0;
        for (int e = g.srow[i]; e < g.srow[i + 1]; e++) acc += v0[g.scol[e]] * g.sval[e];
        acc -= vk / m2 * g.kg[i] + (g.sumbg[i] - beta) * v0[i];
        v[i] = acc;
        s += acc * acc;
    }
    s = block_sum1024(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(TPB) void big_normalize(int ng, double *__restrict__ v, const double *__restrict__ v0,
                                                     const double *__restrict__ scal, double *__restrict__ part1,
                                                     double *__restrict__ part2)
{
#pragma clang fp contract(off)
    __shared__ double red[TPB];
    const double vn = scal[1];
    double p1 = 0.0, p2 = 0.0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < ng; i += gridDim.x * TPB) {
        double x = v[i] / vn;
        v[i] = x;
        double a = x - v0[i], b = x + v0[i];
        p1 += a * a;
        p2 += b * b;
    }
    p1 = block_sum1024(p1, red);
    p2 = block_sum1024(p2, red);
    if (threadIdx.x == 0) {
        part1[blockIdx.x] = p1;
        part2[blockIdx.x] = p2;
    }
}

__global__ __launch_bounds__(TPB) void big_lambda(int ng, const double *__restrict__ v, const double *__restrict__ v0,
                                                  const double *__restrict__ scal, double *__restrict__ out)
{
    __shared__ double red[TPB];
    __shared__ int redi[TPB];
    int mi = block_argmax_abs(v, ng, red, redi);
    if (threadIdx.x == 0) out[0] = (v[mi] * v0[mi] > 0) ? scal[1] : -scal[1];
}

struct Work {
    double *v, *v0, *part, *part2, *scal, *out;
    int *gid, *srow, *scol;
    double *sval, *kg, *sumbg;
};

pg_status lead_big(pg_ctx *ctx, const Group &g, double m2, Work &W, double *lam_out, int *iters)
{
    hipStream_t st = ctx->stream;
    const int nblk = (int)std::min<long long>(pg_cdiv(g.ng, TPB), 1024);
    double lam_total = 0.0;
    for (int pass = 0; pass < 2; pass++) {
        const double beta = pass == 0 ? 0.0 : -lam_total;
        big_start<<<(unsigned)pg_cdiv(g.ng, 256), 256, 0, st>>>(g, W.v);
        double err[2] = {1.0, 1.0};
        int it = 0;
        while (err[0] > EPS && err[1] > EPS && it < MAX_ITER) {
            big_copy_dot<<<nblk, TPB, 0, st>>>(g, W.v, W.v0, W.part);
            big_reduce<<<1, TPB, 0, st>>>(W.part, nblk, W.scal, 0, 0);
            big_spmv<<<nblk, TPB, 0, st>>>(g, m2, beta, W.v0, W.v, W.scal, W.part);
            big_reduce<<<1, TPB, 0, st>>>(W.part, nblk, W.scal, 1, 1);
            big_normalize<<<nblk, TPB, 0, st>>>(g.ng, W.v, W.v0, W.scal, W.part, W.part2);
            big_reduce<<<1, TPB, 0, st>>>(W.part, nblk, W.scal, 2, 1);
            big_reduce<<<1, TPB, 0, st>>>(W.part2, nblk, W.scal, 3, 1);
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, hipMemcpyAsync(err, W.scal + 2, 16, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            it++;
        }
        *iters += it;
        big_lambda<<<1, TPB, 0, st>>>(g.ng, W.v, W.v0, W.scal, W.out);
        double lam = 0.0;
        PG_HIP(ctx, hipMemcpyAsync(&lam, W.out, 8, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        if (pass == 0) {
            lam_total = lam;
            bool shift = g.ng > 25000 ? (lam < -0.01) : (lam < 0.0);
            if (!shift) break;
        } else {
            lam_total += lam;
        }
    }
    *lam_out = lam_total;
    return PG_OK;
}

}  // namespace

extern "C" pg_status pg_modularity(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val,
                                   int32_t n, int64_t nnz, int32_t kmode, int32_t *labels, double *q,
                                   double *lambda0, double *eig0, int32_t *rounds)
{
    if (!ctx) return PG_EINVAL;
    if (!labels) return pg_fail(ctx, PG_EINVAL, "labels is NULL");
    if (kmode != PG_MOD_STRENGTH && kmode != PG_MOD_BINARY_K) return pg_fail(ctx, PG_EINVAL, "bad kmode");
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
    if (R[n] != nnz) return pg_fail(ctx, PG_EFORMAT, "row[n]=%d != nnz=%lld", R[n], (long long)nnz);

    std::vector<double> k((size_t)n);
    double m2 = 0.0;
    for (int i = 0; i < n; i++) {
        double ki = 0.0;
        if (kmode == PG_MOD_BINARY_K || !val) ki = (double)(R[i + 1] - R[i]);
        else for (int e = R[i]; e < R[i + 1]; e++) ki += (double)V[e];
        k[i] = ki;
        m2 += ki;
    }
    std::vector<int> res((size_t)n, 0), adj((size_t)2 * n + 2, 0), loc((size_t)n, -1);
    std::vector<int> gid, srow, scol;
    std::vector<double> sval, kg, sumbg, hv;
    int round = 0, maxr = 0, index = 1, nrounds = 0;
    if (lambda0) *lambda0 = 0.0;
    if (n == 0 || m2 <= 0.0) {
        for (int i = 0; i < n; i++) res[i] = 0;
        adj[0] = 1;
        maxr = -1;
    }
    Work W;
    PG_TRY(pg_bufT(ctx, "mod.v", (size_t)n + 1, &W.v));
    PG_TRY(pg_bufT(ctx, "mod.v0", (size_t)n + 1, &W.v0));
    PG_TRY(pg_bufT(ctx, "mod.part", 1024, &W.part));
    PG_TRY(pg_bufT(ctx, "mod.part2", 1024, &W.part2));
    PG_TRY(pg_bufT(ctx, "mod.scal", 8, &W.scal));
    PG_TRY(pg_bufT(ctx, "mod.out", 4, &W.out));
    PG_TRY(pg_bufT(ctx, "mod.gid", (size_t)n + 1, &W.gid));
    PG_TRY(pg_bufT(ctx, "mod.srow", (size_t)n + 2, &W.srow));
    PG_TRY(pg_bufT(ctx, "mod.scol", (size_t)nnz + 1, &W.scol));
    PG_TRY(pg_bufT(ctx, "mod.sval", (size_t)nnz + 1, &W.sval));
    PG_TRY(pg_bufT(ctx, "mod.kg", (size_t)n + 1, &W.kg));
    PG_TRY(pg_bufT(ctx, "mod.sumbg", (size_t)n + 1, &W.sumbg));
    while (round <= maxr) {
        gid.clear();
        double Kg = 0.0;
        for (int i = 0; i < n; i++)
            if (res[i] == round) {
                loc[i] = (int)gid.size();
                gid.push_back(i);
                Kg += k[i];
            }
        const int ng = (int)gid.size();
        if (ng > 0) {
            nrounds++;
            srow.assign((size_t)ng + 1, 0);
            scol.clear();
            sval.clear();
            kg.resize(ng);
            sumbg.resize(ng);
            for (int a = 0; a < ng; a++) {
                const int i = gid[a];
                double kin = 0.0;
                for (int e = R[i]; e < R[i + 1]; e++) {
                    const int j = Cc[e];
                    if (res[j] != round) continue;
                    const double w = val ? (double)V[e] : 1.0;
                    scol.push_back(loc[j]);
                    sval.push_back(w);
                    kin += (kmode == PG_MOD_BINARY_K && val) ? 1.0 : w;
                }
                srow[a + 1] = (int)scol.size();
                kg[a] = k[i];
                sumbg[a] = kin - k[i] * Kg / m2;
            }
            const size_t gn = scol.size();
            PG_HIP(ctx, hipMemcpyAsync(W.gid, gid.data(), (size_t)ng * 4, hipMemcpyHostToDevice, st));
            PG_HIP(ctx, hipMemcpyAsync(W.srow, srow.data(), ((size_t)ng + 1) * 4, hipMemcpyHostToDevice, st));
            if (gn) {
                PG_HIP(ctx, hipMemcpyAsync(W.scol, scol.data(), gn * 4, hipMemcpyHostToDevice, st));
                PG_HIP(ctx, hipMemcpyAsync(W.sval, sval.data(), gn * 8, hipMemcpyHostToDevice, st));
            }
            PG_HIP(ctx, hipMemcpyAsync(W.kg, kg.data(), (size_t)ng * 8, hipMemcpyHostToDevice, st));
            PG_HIP(ctx, hipMemcpyAsync(W.sumbg, sumbg.data(), (size_t)ng * 8, hipMemcpyHostToDevice, st));
            Group g{ng, W.gid, W.srow, W.scol, W.sval, W.kg, W.sumbg};
            double lam = 0.0;
            int its = 0;
            if (ng <= SMALL_N) {
                lead_small<<<1, TPB, 0, st>>>(g, m2, W.v, W.v0, W.out);
                PG_KERNEL_CHECK(ctx);
                double o[2];
                PG_HIP(ctx, hipMemcpyAsync(o, W.out, 16, hipMemcpyDeviceToHost, st));
                PG_HIP(ctx, hipStreamSynchronize(st));
                lam = o[0];
                its = (int)o[1];
            } else {
                PG_TRY(lead_big(ctx, g, m2, W, &lam, &its));
            }
            hv.resize(ng);
            PG_HIP(ctx, hipMemcpyAsync(hv.data(), W.v, (size_t)ng * 8, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            if (round == 0) {
                if (lambda0) *lambda0 = lam;
                if (eig0) {
                    std::vector<double> e0((size_t)n, 0.0);
                    for (int a = 0; a < ng; a++) e0[gid[a]] = hv[a];
                    if (pg_is_device_ptr(eig0)) PG_HIP(ctx, hipMemcpy(eig0, e0.data(), (size_t)n * 8, hipMemcpyHostToDevice));
                    else std::memcpy(eig0, e0.data(), (size_t)n * 8);
                }
            }
            long long sp = 0, sn = 0;
            for (int a = 0; a < ng; a++) {
                sp += hv[a] > 0;
                sn += hv[a] <= 0;
            }
            const bool split = lam > 0.01 && sp > 1 && sn > 1;  // LAMBDA, MIN_GROUP
            if (split) {
                for (int a = 0; a < ng; a++) res[gid[a]] = maxr + 1 + (hv[a] * ((double)(sp - sn) + 0.5) <= 0);
                maxr += 2;
            } else {
                adj[round] = index++;
            }
            for (int a = 0; a < ng; a++) loc[gid[a]] = -1;
        }
        round++;
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
