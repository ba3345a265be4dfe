// graph.hip -- per-node metrics on a CSR graph, MI355X (gfx950).
//
//  * degree / strength            (U/Degree/Degree/Degree.cpp, W Degree.cpp)
//  * triangles, binary clustering (U/Cp/Cp/Cp.cpp:21-115) by sorted-adjacency
//    intersection against an LDS bitmap of the vertex's neighbourhood
//  * weighted clustering, Barrat and Onnela (W/Cp/Cp/Cp_1.cpp, Cp_2.cpp,
//    W/Cp_GPU/CUCP/Cp.cu)
//  * Brandes betweenness and BFS levels (U/BC/BC_CPU/BC_CPU/BC_CPU.cpp,
//    U/BC/CUBC/BC_GPU_node_pred.cu, U/BFS_MulCPU/BFS_MulCPU/APSP_BFS.cpp):
//    64 sources per batch, one per lane of a wave64; one wave per vertex pulls
//    from its neighbours, whose per-source rows (uint64 path counts, FP64
//    coefficients) are read coalesced.  Visited sets and hop levels mod 4 are
//    kept as 64-bit per-vertex masks so the neighbour tests are scalar.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "pg_internal.hpp"

namespace {

typedef unsigned long long u64;

// ------------------------------------------------------------------ degree

__global__ void degree_kernel(const int *__restrict__ row, int n, int *__restrict__ deg,
                              float *__restrict__ deg_ref)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = row[i], b = row[i + 1];
    if (deg) deg[i] = b - a;
    // U Degree.cpp:87: (float) R[k+1] - R[k], both offsets rounded to float
    if (deg_ref) deg_ref[i] = __fsub_rn((float)b, (float)a);
}

// W Degree.cpp:84-88: sequential FP32 sum in CSR order, one thread per row.
__global__ void strength_kernel(const int *__restrict__ row, const float *__restrict__ val, int n,
                                float *__restrict__ s)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int e = row[i]; e < row[i + 1]; e++) acc += val[e];
    s[i] = acc;
}

// -------------------------------------------------------------- clustering

// A scan start for the columns of row [rv, vb) above key, at or before the
// first of them: 64 evenly spaced samples (one load per lane) and a ballot,
// in place of a binary search's chain of dependent loads.  The scan filters
// col > key itself.  Called by a whole wave.
__device__ __forceinline__ int tail_start(const int *__restrict__ col, int rv, int vb, int key, int lane)
{
    const int dv = vb - rv;
    if (dv <= 0) return vb;
    const int ps = rv + (int)(((long long)lane * dv) >> 6);  // < vb
    const unsigned long long m = __ballot(col[ps] > key);
    if (m & 1ull) return rv;
    const int f = m ? __builtin_ctzll(m) : 64;
    return rv + (int)(((long long)(f - 1) * dv) >> 6);
}

template <class T>
__device__ T block_sum(T v, T *scratch)
{
    // fixed-order tree over 256 threads
    const int tid = threadIdx.x;
    scratch[tid] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) scratch[tid] += scratch[tid + off];
        __syncthreads();
    }
    T r = scratch[0];
    __syncthreads();
    return r;
}

// One block per vertex u: T_u = #{(v, w) : v < w, v, w in N(u), (v, w) in E}.
__global__ __launch_bounds__(256) void triangles_kernel(const int *__restrict__ row, const int *__restrict__ col,
                                                        int n, long long *__restrict__ tri)
{
    extern __shared__ uint32_t bits[];
    __shared__ long long red[256];
    const int u = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int a = row[u], b = row[u + 1], d = b - a;
    if (d < 2) {
        if (tid == 0) tri[u] = 0;
        return;
    }
    const int nw = (n + 31) >> 5;
    for (int q = tid; q < nw; q += 256) bits[q] = 0u;
    __syncthreads();
    for (int e = a + tid; e < b; e += 256) {
        int c = col[e];
        atomicOr(&bits[c >> 5], 1u << (c & 31));
    }
    __syncthreads();
    long long cnt = 0;
    for (int k = wv; k < d; k += 4) {
        const int v = col[a + k];
        const int vb = row[v + 1];
        const int start = tail_start(col, row[v], vb, v, lane);
        for (int e = start + lane; e < vb; e += 64) {
            const int w = col[e];
            cnt += (w > v) & ((bits[w >> 5] >> (w & 31)) & 1u);
        }
    }
    long long s = block_sum<long long>(cnt, red);
    if (tid == 0) tri[u] = s;
}

__global__ void cp_from_tri(const int *__restrict__ row, const long long *__restrict__ tri, int n,
                            float *__restrict__ cp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long k = row[i + 1] - row[i];
    float c = 0.0f;
    if (k >= 2) c = (float)((double)tri[i] / (double)k / (double)(k - 1) * 2.0);
    cp[i] = c;
}

template <class T>
__global__ __launch_bounds__(256) void mean_kernel(const T *__restrict__ x, int n, double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += (double)x[i];
    s = block_sum<double>(s, red);
    if (threadIdx.x == 0) *out = n > 0 ? s / (double)n : 0.0;
}

__global__ __launch_bounds__(256) void max_kernel(const float *__restrict__ x, long long m, float *__restrict__ out)
{
    __shared__ float red[256];
    float v = -INFINITY;
    for (long long i = threadIdx.x; i < m; i += 256) v = fmaxf(v, x[i]);
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Weighted clustering of vertex u (one block): bitmap + word prefix of N(u)
// give the position of w in N(u), hence w_uw.
__global__ __launch_bounds__(256) void cpw_kernel(const int *__restrict__ row, const int *__restrict__ col,
                                                  const float *__restrict__ val, int n, int kind,
                                                  const float *__restrict__ wmaxp, double *__restrict__ cp)
{
    extern __shared__ uint32_t sm[];
    __shared__ double red[256];
    __shared__ int part[256];
    const int nw = (n + 31) >> 5;
    uint32_t *bits = sm;
    int *pre = (int *)(sm + nw);
    const int u = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int a = row[u], b = row[u + 1], d = b - a;
    if (d < 2) {
        if (tid == 0) cp[u] = 0.0;
        return;
    }
    for (int q = tid; q < nw; q += 256) bits[q] = 0u;
    __syncthreads();
    for (int e = a + tid; e < b; e += 256) {
        int c = col[e];
        atomicOr(&bits[c >> 5], 1u << (c & 31));
    }
    __syncthreads();
    const int per = (nw + 255) / 256;
    const int w0 = tid * per, w1 = min(nw, w0 + per);
    int s = 0;
    for (int q = w0; q < w1; q++) s += __popc(bits[q]);
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (int q = w0; q < w1; q++) {
        pre[q] = run;
        run += __popc(bits[q]);
    }
    __syncthreads();
    const double wmax = (double)*wmaxp;
    double acc = 0.0, sk = 0.0;
    for (int e = a + tid; e < b; e += 256) sk += (double)val[e];
    for (int k = wv; k < d; k += 4) {
        const int v = col[a + k];
        const double wuv = (double)val[a + k];
        const int vb = row[v + 1];
        const int start = tail_start(col, row[v], vb, v, lane);
        for (int e = start + lane; e < vb; e += 64) {
            const int w = col[e];
            const uint32_t word = bits[w >> 5];
            if (w > v && ((word >> (w & 31)) & 1u)) {
                const double wvw = (double)val[e];
                if (wvw != 0.0) {
                    const int pos = pre[w >> 5] + __popc(word & ((1u << (w & 31)) - 1u));
                    const double wuw = (double)val[a + pos];
                    if (kind == PG_CP_BARRAT) acc += wuv + wuw;
                    else acc += pow((wvw / wmax) * (wuv / wmax) * (wuw / wmax), 1.0 / 3.0);
                }
            }
        }
    }
    acc = block_sum<double>(acc, red);
    sk = block_sum<double>(sk, red);
    if (tid == 0) {
        double c;
        if (kind == PG_CP_BARRAT) c = acc / (double)(d - 1) / sk;
        else c = 2.0 * acc / (double)d / (double)(d - 1);
        cp[u] = c;
    }
}

// ------------------------------------------------------- betweenness / BFS

constexpr int S = 64;  // sources per batch = lanes of a wave64
constexpr int GATHER = 8;
constexpr int MAX_BATCHES = 16;  // batches per launch (blockIdx.y)

// Per-launch state of up to MAX_BATCHES batches of S sources: batch b's
// arrays start at b * n * S (per-source rows), b * n (per-vertex masks),
// b * S (sources, -1 for an absent lane), b * lcap (level counters) and
// b * n (its share of the BC sums).
struct BatchBufs {
    int nb = 1;            // batches per launch
    int lcap = 0;          // level counters per batch
    u64 *sigma, *visA, *visB, *p0, *p1;
    int *lev, *srcs;
    double *coef, *bcpart;
    unsigned *prog;        // [nb][lcap]: waves that found new vertices at level d + 1
    unsigned *tot;         // [lcap]: the same summed over the batches
};

__global__ void bfs_init(int n, BatchBufs B)
{
    const long long ns = (long long)n * S;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;  // ns is a multiple of 64: whole waves return
    const int bb = blockIdx.y, lane = threadIdx.x & 63;
    const long long o = (long long)bb * ns + k;
    B.sigma[o] = 0;
    B.lev[o] = -1;
    // absent lanes count as visited
    const u64 absent = __ballot(B.srcs[bb * S + lane] < 0);
    if (lane == 0) {
        const long long v = (long long)bb * n + k / S;
        B.visA[v] = absent;
        B.p0[v] = 0;
        B.p1[v] = 0;
    }
    // the level counters restart with the group (lcap <= n + 2 <= n * S)
    if (k < B.lcap) {
        B.prog[(long long)bb * B.lcap + k] = 0u;
        if (bb == 0) B.tot[k] = 0u;
    }
}

__global__ void bfs_seed(int n, BatchBufs B)
{
    const int s = threadIdx.x, bb = blockIdx.y;
    const int v = B.srcs[bb * S + s];
    if (v < 0) return;
    const long long o = (long long)bb * n * S + (long long)v * S + s;
    B.sigma[o] = 1;
    B.lev[o] = 0;
    atomicOr(&B.visA[(long long)bb * n + v], 1ull << s);
}

// Level d -> d+1 of every batch.  A vertex w not yet reached from source s
// collects sigma[w][s] = sum of sigma[v][s] over neighbours v reached from s;
// in an undirected graph those are exactly its level-d neighbours.  A batch
// whose level d is empty (no progress at d - 1) returns at once.
__global__ __launch_bounds__(256) void bfs_forward(const int *__restrict__ row, const int *__restrict__ col, int n,
                                                   int d, int flip, BatchBufs B)
{
    const int bb = blockIdx.y;
    unsigned *prog = B.prog + (long long)bb * B.lcap;
    if (d > 0 && prog[d - 1] == 0u) return;
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const long long vb0 = (long long)bb * n;
    const u64 *__restrict__ visA = (flip ? B.visB : B.visA) + vb0;
    u64 *__restrict__ visB = (flip ? B.visA : B.visB) + vb0;
    u64 *__restrict__ sigma = B.sigma + vb0 * S;
    const u64 vis = visA[w];
    const u64 need = ~vis;
    if (need == 0) {
        if (lane == 0) visB[w] = vis;
        return;
    }
    const int a = row[w], b = row[w + 1];
    u64 acc = 0;
    bool hit = false;
    for (int base = a; base < b; base += 64) {
        const int e = base + lane;
        int v = 0;
        u64 m = 0;
        if (e < b) {
            v = col[e];
            m = visA[v] & need;
        }
        u64 any = __ballot(m != 0);
        while (any) {
            u64 g[GATHER];
            bool gh[GATHER];
#pragma unroll
            for (int q = 0; q < GATHER; q++) {
                g[q] = 0;
                gh[q] = false;
                if (any) {
                    const int sl = __ffsll((long long)any) - 1;
                    any &= any - 1;
                    const int vv = __shfl(v, sl);
                    const u64 mm = __shfl(m, sl);
                    if ((mm >> lane) & 1ull) {
                        g[q] = sigma[(long long)vv * S + lane];
                        gh[q] = true;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < GATHER; q++) {
                acc += g[q];
                hit |= gh[q];
            }
        }
    }
    const bool fresh = ((need >> lane) & 1ull) && hit;
    const u64 nbits = __ballot(fresh);
    if (fresh) {
        sigma[(long long)w * S + lane] = acc;
        B.lev[(vb0 + w) * S + lane] = d + 1;
    }
    if (lane == 0) {
        visB[w] = vis | nbits;
        if (nbits) {
            if ((d + 1) & 1) B.p0[vb0 + w] |= nbits;
            if ((d + 1) & 2) B.p1[vb0 + w] |= nbits;
            atomicAdd(&prog[d], 1u);
            atomicAdd(&B.tot[d], 1u);
        }
    }
}

__global__ void bc_coef_init(long long total, const u64 *__restrict__ sigma, double *__restrict__ coef)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    u64 s = sigma[k];
    coef[k] = s ? 1.0 / (double)s : 0.0;
}

// Level d of every batch: delta[v][s] = sigma[v][s] * sum over children w
// (level d+1) of (1 + delta[w][s]) / sigma[w][s]; coef[v][s] = (1 + delta[v][s]) /
// sigma[v][s].  The batch's BC share bcpart[v] += sum over sources s != v of
// delta[v][s] (fixed lane order).  A batch without level d + 1 returns.  vis:
// the final visited masks (either ping-pong buffer holds them once the
// batch's BFS has ended).
__global__ __launch_bounds__(256) void bc_backward(const int *__restrict__ row, const int *__restrict__ col, int n,
                                                   int d, BatchBufs B)
{
    const int bb = blockIdx.y;
    if (B.prog[(long long)bb * B.lcap + d] == 0u) return;
    const int lane = threadIdx.x & 63;
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= n) return;
    const long long vb0 = (long long)bb * n;
    const u64 *__restrict__ vis = B.visA + vb0;
    const u64 *__restrict__ p0 = B.p0 + vb0;
    const u64 *__restrict__ p1 = B.p1 + vb0;
    const u64 *__restrict__ sigma = B.sigma + vb0 * S;
    double *__restrict__ coef = B.coef + vb0 * S;
    const bool own = B.lev[(vb0 + v) * S + lane] == d;
    const u64 ownm = __ballot(own);
    if (ownm == 0) return;
    const u64 x0 = ((d + 1) & 1) ? ~0ull : 0ull;
    const u64 x1 = ((d + 1) & 2) ? ~0ull : 0ull;
    const int a = row[v], b = row[v + 1];
    double acc = 0.0;
    for (int base = a; base < b; base += 64) {
        const int e = base + lane;
        int w = 0;
        u64 m = 0;
        if (e < b) {
            w = col[e];
            m = ~(p0[w] ^ x0) & ~(p1[w] ^ x1) & vis[w] & ownm;
        }
        u64 any = __ballot(m != 0);
        while (any) {
            double g[GATHER];
#pragma unroll
            for (int q = 0; q < GATHER; q++) {
                g[q] = 0.0;
                if (any) {
                    const int sl = __ffsll((long long)any) - 1;
                    any &= any - 1;
                    const int ww = __shfl(w, sl);
                    const u64 mm = __shfl(m, sl);
                    if ((mm >> lane) & 1ull) g[q] = coef[(long long)ww * S + lane];
                }
            }
#pragma unroll
            for (int q = 0; q < GATHER; q++) acc += g[q];
        }
    }
    double contrib = 0.0;
    if (own) {
        const double sg = (double)sigma[(long long)v * S + lane];
        const double dl = sg * acc;
        coef[(long long)v * S + lane] = (1.0 + dl) / sg;
        if (B.srcs[bb * S + lane] != v) contrib = dl;
    }
    // fixed-order butterfly over the wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
    if (lane == 0) B.bcpart[vb0 + v] += contrib;
}

// bc[v] = sum over batch slots of their shares, in slot order
__global__ void bc_reduce(const double *__restrict__ part, int nb, int n, double *__restrict__ bc)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[(long long)b * n + v];
    bc[v] = s;
}

__global__ void fill_f64(double *__restrict__ p, long long count)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) p[k] = 0.0;
}

// Levels of batch bb's sources into out[(bb*S + s) * n + v] for its first
// nsrc[bb] lanes.
__global__ void bfs_export(int n, int nvalid, const int *__restrict__ lev, int *__restrict__ out)
{
    const int bb = blockIdx.y;
    const int nb = min(S, nvalid - bb * S);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nb <= 0 || k >= (long long)n * nb) return;
    const long long s = k / n, v = k % n;
    out[((long long)bb * S + s) * n + v] = lev[((long long)bb * n + v) * S + s];
}

pg_status batch_bufs(pg_ctx *ctx, int n, int nb, int lcap, BatchBufs &B, bool need_coef)
{
    const size_t ns = (size_t)n * S * nb, nv = (size_t)n * nb + 1;
    B.nb = nb;
    B.lcap = lcap;
    PG_TRY(pg_bufT(ctx, "bc.sigma", ns, &B.sigma));
    PG_TRY(pg_bufT(ctx, "bc.lev", ns, &B.lev));
    PG_TRY(pg_bufT(ctx, "bc.visA", nv, &B.visA));
    PG_TRY(pg_bufT(ctx, "bc.visB", nv, &B.visB));
    PG_TRY(pg_bufT(ctx, "bc.p0", nv, &B.p0));
    PG_TRY(pg_bufT(ctx, "bc.p1", nv, &B.p1));
    PG_TRY(pg_bufT(ctx, "bc.srcs", (size_t)S * nb, &B.srcs));
    PG_TRY(pg_bufT(ctx, "bc.prog", (size_t)nb * lcap, &B.prog));
    PG_TRY(pg_bufT(ctx, "bc.tot", (size_t)lcap, &B.tot));
    PG_TRY(pg_bufT(ctx, "bc.part", nv, &B.bcpart));
    B.coef = nullptr;
    if (need_coef) PG_TRY(pg_bufT(ctx, "bc.coef", ns, &B.coef));
    return PG_OK;
}

// Forward sweep of one group of batches (sources already in B.srcs), host
// driven: a level is launched while some batch progressed at the previous
// one.  Returns the number of levels launched after level 0; the final
// visited masks are in B.visA (and B.visB).
pg_status run_forward(pg_ctx *ctx, const int *row, const int *col, int n, BatchBufs &B, int *maxlev)
{
    hipStream_t st = ctx->stream;
    const long long ns = (long long)n * S;
    bfs_init<<<dim3((unsigned)pg_cdiv(ns, 256), (unsigned)B.nb), 256, 0, st>>>(n, B);
    bfs_seed<<<dim3(1, (unsigned)B.nb), S, 0, st>>>(n, B);
    PG_KERNEL_CHECK(ctx);
    int d = 0;
    for (;; d++) {
        if (d + 1 >= B.lcap) return pg_fail(ctx, PG_ENOTSUP, "BFS deeper than %d levels", B.lcap - 1);
        bfs_forward<<<dim3((unsigned)pg_cdiv(n, 4), (unsigned)B.nb), 256, 0, st>>>(row, col, n, d, d & 1, B);
        PG_KERNEL_CHECK(ctx);
        unsigned prog = 0;
        PG_HIP(ctx, hipMemcpyAsync(&prog, B.tot + d, 4, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        if (prog == 0) break;
    }
    *maxlev = d;  // levels 0..d hold vertices
    return PG_OK;
}

// Batches per launch: PAGANI_BC_BATCHES (1..16), else a value forced by the
// autotuner's search, else the tuned value for (arch, n, nnz), else enough
// 64-source batches for about 16384 work-groups of four vertex waves; capped
// by the sources and by 1 GiB of per-source state.
int bc_batches(pg_ctx *ctx, int n, long long nnz, long long nsrc)
{
    int nb = 0;
    if (const char *e = std::getenv("PAGANI_BC_BATCHES")) nb = std::atoi(e);
    if (nb <= 0) nb = ctx->force_bc_batches;
    if (nb <= 0) nb = pg_tune_get(ctx, "bc_batches", n, nnz, 0);
    if (nb <= 0) {
        nb = (int)std::max<long long>(1, std::min<long long>(8, pg_cdiv(65536, std::max(n, 1))));
        while (nb > 1 && (long long)nb * n * S * 20 > (1ll << 30)) nb--;
    }
    nb = std::max(1, std::min(MAX_BATCHES, nb));
    return (int)std::max<long long>(1, std::min<long long>(nb, pg_cdiv(std::max<long long>(nsrc, 1), S)));
}

}  // namespace

extern "C" pg_status pg_degree(pg_ctx *ctx, const int32_t *row, int32_t n, int32_t *deg, float *deg_ref)
{
    if (!ctx) return PG_EINVAL;
    if (!row || n < 0) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const int *drow;
    int *ddeg;
    float *dref;
    PG_TRY(pg_in(ctx, row, (size_t)n + 1, "in.row", &drow));
    PG_TRY(pg_out(ctx, deg, (size_t)n, "out.deg", &ddeg));
    PG_TRY(pg_out(ctx, deg_ref, (size_t)n, "out.degref", &dref));
    if (n > 0) degree_kernel<<<(unsigned)pg_cdiv(n, 256), 256, 0, ctx->stream>>>(drow, n, ddeg, dref);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, deg, ddeg, (size_t)n));
    PG_TRY(pg_out_finish(ctx, deg_ref, dref, (size_t)n));
    PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PG_OK;
}

extern "C" pg_status pg_strength(pg_ctx *ctx, const int32_t *row, const float *val, int32_t n, float *s)
{
    if (!ctx) return PG_EINVAL;
    if (!row || !val || !s || n < 0) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    int nnz = 0;
    if (pg_is_device_ptr(row)) PG_HIP(ctx, hipMemcpy(&nnz, row + n, 4, hipMemcpyDeviceToHost));
    else nnz = row[n];
    if (nnz < 0) return pg_fail(ctx, PG_EFORMAT, "row[n]=%d", nnz);
    const int *drow;
    const float *dval;
    float *ds;
    PG_TRY(pg_in(ctx, row, (size_t)n + 1, "in.row", &drow));
    PG_TRY(pg_validate_csr(ctx, drow, nullptr, n, nnz, false));  // offsets only
    PG_TRY(pg_in(ctx, val, (size_t)nnz, "in.val", &dval));
    PG_TRY(pg_out(ctx, s, (size_t)n, "out.s", &ds));
    if (n > 0) strength_kernel<<<(unsigned)pg_cdiv(n, 256), 256, 0, ctx->stream>>>(drow, dval, n, ds);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, s, ds, (size_t)n));
    PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PG_OK;
}

namespace {
// Device views of a caller's CSR, checked on the device before any kernel
// indexes with it (PG_EFORMAT for offsets outside [0, nnz], columns outside
// [0, n), and with `sorted` columns not strictly ascending in a row).
pg_status stage_csr(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val, int32_t n,
                    int64_t nnz, bool sorted, const int **drow, const int **dcol, const float **dval)
{
    PG_TRY(pg_check_csr_args(ctx, row, col, n, nnz));
    PG_TRY(pg_in(ctx, row, (size_t)n + 1, "in.row", drow));
    PG_TRY(pg_in(ctx, col, (size_t)nnz, "in.col", dcol));
    if (dval) PG_TRY(pg_in(ctx, val, (size_t)nnz, "in.val", dval));
    return pg_validate_csr(ctx, *drow, *dcol, n, nnz, sorted);
}
}  // namespace

extern "C" pg_status pg_triangles(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                                  int64_t nnz, int64_t *tri)
{
    if (!ctx) return PG_EINVAL;
    if (!tri) return pg_fail(ctx, PG_EINVAL, "tri is NULL");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const int *drow, *dcol;
    PG_TRY(stage_csr(ctx, row, col, nullptr, n, nnz, true, &drow, &dcol, nullptr));
    const int nw = (n + 31) / 32;
    if ((size_t)nw * 4 + 4096 > ctx->lds_per_block) return pg_fail(ctx, PG_ENOTSUP, "n too large for LDS bitmap");
    long long *dtri;
    PG_TRY(pg_out(ctx, (long long *)tri, (size_t)n, "out.tri", &dtri));
    if (n > 0) triangles_kernel<<<(unsigned)n, 256, (size_t)nw * 4, ctx->stream>>>(drow, dcol, n, dtri);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, (long long *)tri, dtri, (size_t)n));
    PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PG_OK;
}

extern "C" pg_status pg_cp_binary(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                                  int64_t nnz, float *cp, double *mean)
{
    if (!ctx) return PG_EINVAL;
    if (!cp) return pg_fail(ctx, PG_EINVAL, "cp is NULL");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const int *drow, *dcol;
    PG_TRY(stage_csr(ctx, row, col, nullptr, n, nnz, true, &drow, &dcol, nullptr));
    const int nw = (n + 31) / 32;
    if ((size_t)nw * 4 + 4096 > ctx->lds_per_block) return pg_fail(ctx, PG_ENOTSUP, "n too large for LDS bitmap");
    long long *dtri;
    float *dcp;
    double *dmean;
    PG_TRY(pg_bufT(ctx, "cp.tri", (size_t)n + 1, &dtri));
    PG_TRY(pg_bufT(ctx, "cp.mean", 1, &dmean));
    PG_TRY(pg_out(ctx, cp, (size_t)n, "out.cp", &dcp));
    hipStream_t st = ctx->stream;
    if (n > 0) {
        triangles_kernel<<<(unsigned)n, 256, (size_t)nw * 4, st>>>(drow, dcol, n, dtri);
        cp_from_tri<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(drow, dtri, n, dcp);
    }
    mean_kernel<float><<<1, 256, 0, st>>>(dcp, n, dmean);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, cp, dcp, (size_t)n));
    if (mean) PG_HIP(ctx, hipMemcpyAsync(mean, dmean, 8, pg_is_device_ptr(mean) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    return PG_OK;
}

int pg_bc_batches_for(pg_ctx *ctx, int n, long long nnz, long long nsrc) { return bc_batches(ctx, n, nnz, nsrc); }

// ------------------------------------------------ captured cohort: metrics

namespace {
// status bit when level `cap` of some batch still found vertices: the BFS may
// go deeper than the captured launches
__global__ void bc_depth_check(const unsigned *__restrict__ tot, int cap, unsigned *__restrict__ status)
{
    if (threadIdx.x == 0 && tot[cap - 1] != 0u) atomicOr(status, PG_COHORT_DEPTH);
}
}  // namespace

struct PgCohortMetrics {
    pg_ctx *ctx = nullptr;
    int levels = 0;    // forward launches per group (BFS levels 1..levels)
    int nb = 1;        // batches of 64 sources per launch
    BatchBufs B;
    long long *tri = nullptr;
};

pg_status pg_cohort_metrics_plan(pg_ctx *ctx, int max_n, long long max_nsrc, int levels, PgCohortMetrics **out)
{
    PgCohortMetrics *M = new PgCohortMetrics();
    std::unique_ptr<PgCohortMetrics> guard(M);
    M->ctx = ctx;
    M->levels = levels;
    // as many batches per launch as 2 GiB of per-source state allows (fewer
    // graph nodes: each batch group costs 2 * levels + 3 launches)
    M->nb = MAX_BATCHES;
    while (M->nb > 1 && (long long)M->nb * max_n * S * 20 > (2ll << 30)) M->nb--;
    M->nb = (int)std::max<long long>(1, std::min<long long>(M->nb, pg_cdiv(std::max<long long>(max_nsrc, 1), S)));
    PG_TRY(batch_bufs(ctx, max_n, M->nb, levels, M->B, true));
    PG_TRY(pg_bufT(ctx, "coh.tri", (size_t)max_n + 1, &M->tri));
    *out = guard.release();
    return PG_OK;
}

int pg_cohort_metrics_group(const PgCohortMetrics *M) { return M->nb * S; }

// Degree, binary clustering and BC (sources: nsrc device entries padded with
// -1 to a multiple of pg_cohort_metrics_group) of one subject's CSR, launched
// without allocation, synchronisation or memset; every BFS gets `levels`
// level launches, each returning at once once its batch has ended.
pg_status pg_cohort_metrics_enqueue(PgCohortMetrics *M, const int32_t *row, const int32_t *col, int n,
                                    const int32_t *srcs, long long nsrc, int32_t *deg, float *cp, double *bc,
                                    unsigned *status)
{
    pg_ctx *ctx = M->ctx;
    hipStream_t st = ctx->stream;
    if (n <= 0) return PG_OK;
    if (deg) degree_kernel<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(row, n, deg, nullptr);
    if (cp) {
        const int nw = (n + 31) / 32;
        triangles_kernel<<<(unsigned)n, 256, (size_t)nw * 4, st>>>(row, col, n, M->tri);
        cp_from_tri<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(row, M->tri, n, cp);
    }
    if (bc && nsrc <= 0) fill_f64<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(bc, n);
    if (bc && nsrc > 0) {
        BatchBufs B = M->B;
        const int nb = M->nb, L = M->levels;
        const long long ns = (long long)n * S;
        fill_f64<<<(unsigned)pg_cdiv((long long)nb * n, 256), 256, 0, st>>>(B.bcpart, (long long)nb * n);
        for (long long g0 = 0; g0 < nsrc; g0 += (long long)nb * S) {
            B.srcs = const_cast<int32_t *>(srcs) + g0;
            bfs_init<<<dim3((unsigned)pg_cdiv(ns, 256), (unsigned)nb), 256, 0, st>>>(n, B);
            bfs_seed<<<dim3(1, (unsigned)nb), S, 0, st>>>(n, B);
            for (int d = 0; d < L; d++)
                bfs_forward<<<dim3((unsigned)pg_cdiv(n, 4), (unsigned)nb), 256, 0, st>>>(row, col, n, d, d & 1, B);
            bc_depth_check<<<1, 64, 0, st>>>(B.tot, L, status);
            bc_coef_init<<<(unsigned)pg_cdiv(ns * nb, 256), 256, 0, st>>>(ns * nb, B.sigma, B.coef);
            for (int d = L - 1; d >= 1; d--)
                bc_backward<<<dim3((unsigned)pg_cdiv(n, 4), (unsigned)nb), 256, 0, st>>>(row, col, n, d, B);
        }
        bc_reduce<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(B.bcpart, nb, n, bc);
    }
    PG_KERNEL_CHECK(ctx);
    return PG_OK;
}

void pg_cohort_metrics_free(PgCohortMetrics *M) { delete M; }

// ------------------------------------------- allocation-free binary Cp

namespace {
// Fixed scratch of pg_cp_binary_static: triangle counts of up to
// PG_CP_STATIC_MAX_N vertices, then the FP64 mean, then the CSR check flag.
// One copy per GPU, held while the library is loaded; host code reaches it
// only through the address pg_static_scratch looks up per device.
constexpr size_t STATIC_TRI_BYTES = (size_t)PG_CP_STATIC_MAX_N * 8;
constexpr size_t STATIC_SCRATCH_BYTES = STATIC_TRI_BYTES + 256;
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

pg_status pg_static_scratch(pg_ctx *ctx, void **ptr)
{
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES)
        return pg_fail(ctx, PG_ENOTSUP, "device %d beyond the scratch table", ctx->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        PG_HIP(ctx, hipSetDevice(ctx->device));
        void *p = nullptr;
        PG_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[ctx->device] = p;
    }
    *ptr = g_scratch_ptr[ctx->device];
    return PG_OK;
}

extern "C" pg_status pg_cp_binary_static(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                                         int64_t nnz, float *cp, double *mean)
{
    if (!ctx) return PG_EINVAL;
    PG_TRY(pg_check_csr_args(ctx, row, col, n, nnz));
    if (!cp) return pg_fail(ctx, PG_EINVAL, "cp is NULL");
    if (n > PG_CP_STATIC_MAX_N)
        return pg_fail(ctx, PG_ENOTSUP, "n=%d exceeds the fixed scratch of %d vertices", n, PG_CP_STATIC_MAX_N);
    // staging a host buffer would need device memory
    if (!pg_is_device_ptr(row) || (nnz > 0 && !pg_is_device_ptr(col)) || (n > 0 && !pg_is_device_ptr(cp)))
        return pg_fail(ctx, PG_EINVAL, "row, col and cp must be device memory");
    if (!ctx->static_scratch) return pg_fail(ctx, PG_ESTATE, "no scratch address for device %d", ctx->device);
    PG_HIP(ctx, hipSetDevice(ctx->device));
    unsigned char *base = (unsigned char *)ctx->static_scratch;
    long long *dtri = (long long *)base;
    double *dmean = (double *)(base + STATIC_TRI_BYTES);
    int *bad = (int *)(base + STATIC_TRI_BYTES + 64);
    PG_TRY(pg_validate_csr_flag(ctx, row, col, n, nnz, true, bad));
    const int nw = (n + 31) / 32;
    if ((size_t)nw * 4 + 4096 > ctx->lds_per_block) return pg_fail(ctx, PG_ENOTSUP, "n too large for LDS bitmap");
    hipStream_t st = ctx->stream;
    if (n > 0) {
        triangles_kernel<<<(unsigned)n, 256, (size_t)nw * 4, st>>>(row, col, n, dtri);
        cp_from_tri<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(row, dtri, n, cp);
    }
    mean_kernel<float><<<1, 256, 0, st>>>(cp, n, dmean);
    PG_KERNEL_CHECK(ctx);
    if (mean) PG_HIP(ctx, hipMemcpyAsync(mean, dmean, 8, pg_is_device_ptr(mean) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    return PG_OK;
}

extern "C" pg_status pg_cp_weighted(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val,
                                    int32_t n, int64_t nnz, int32_t kind, int32_t wmax_mode, double *cp,
                                    double *mean)
{
    if (!ctx) return PG_EINVAL;
    if (!cp || !val) return pg_fail(ctx, PG_EINVAL, "cp/val is NULL");
    if (kind != PG_CP_BARRAT && kind != PG_CP_ONNELA) return pg_fail(ctx, PG_EINVAL, "unknown kind %d", kind);
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const int *drow, *dcol;
    const float *dval = nullptr;
    PG_TRY(stage_csr(ctx, row, col, val, n, nnz, true, &drow, &dcol, &dval));
    const int nw = (n + 31) / 32;
    if ((size_t)nw * 8 + 4096 > ctx->lds_per_block) return pg_fail(ctx, PG_ENOTSUP, "n too large for LDS bitmap");
    double *dcp, *dmean;
    float *dwmax;
    PG_TRY(pg_bufT(ctx, "cpw.wmax", 1, &dwmax));
    PG_TRY(pg_bufT(ctx, "cpw.mean", 1, &dmean));
    PG_TRY(pg_out(ctx, cp, (size_t)n, "out.cpw", &dcp));
    hipStream_t st = ctx->stream;
    long long lim = wmax_mode == 1 ? std::min<long long>(n, nnz) : nnz;
    max_kernel<<<1, 256, 0, st>>>(dval, lim, dwmax);
    if (n > 0) cpw_kernel<<<(unsigned)n, 256, (size_t)nw * 8, st>>>(drow, dcol, dval, n, kind, dwmax, dcp);
    mean_kernel<double><<<1, 256, 0, st>>>(dcp, n, dmean);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, cp, dcp, (size_t)n));
    if (mean) PG_HIP(ctx, hipMemcpyAsync(mean, dmean, 8, pg_is_device_ptr(mean) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    return PG_OK;
}

extern "C" pg_status pg_bc(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n, int64_t nnz,
                           const int32_t *sources, int64_t nsrc, double *bc)
{
    if (!ctx) return PG_EINVAL;
    if (!bc) return pg_fail(ctx, PG_EINVAL, "bc is NULL");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const int *drow, *dcol;
    PG_TRY(stage_csr(ctx, row, col, nullptr, n, nnz, false, &drow, &dcol, nullptr));
    std::vector<int32_t> hsrc;
    if (sources) {
        if (nsrc < 0) return pg_fail(ctx, PG_EINVAL, "nsrc < 0");
        hsrc.resize((size_t)nsrc);
        if (pg_is_device_ptr(sources)) PG_HIP(ctx, hipMemcpy(hsrc.data(), sources, nsrc * 4, hipMemcpyDeviceToHost));
        else std::memcpy(hsrc.data(), sources, nsrc * 4);
    } else {
        nsrc = n;
        hsrc.resize((size_t)n);
        for (int i = 0; i < n; i++) hsrc[i] = i;
    }
    for (auto s : hsrc)
        if (s < 0 || s >= n) return pg_fail(ctx, PG_EINVAL, "source %d out of range", s);
    if (n > 0 && nsrc > 0 && pg_autotune_on_first_use(ctx) && pg_tune_get(ctx, "bc_batches", n, nnz, 0) == 0) {
        // search on up to 1024 of the sources, then run with the result
        pg_tune_result tr;
        const long long m = std::min<long long>(nsrc, 1024);
        PG_TRY(pg_autotune_bc(ctx, drow, dcol, n, nnz, hsrc.data(), m, &tr));
    }
    double *dbc;
    PG_TRY(pg_out(ctx, bc, (size_t)n, "out.bc", &dbc));
    hipStream_t st = ctx->stream;
    if (n == 0) return PG_OK;
    const int nb = bc_batches(ctx, n, nnz, nsrc);
    BatchBufs B;
    PG_TRY(batch_bufs(ctx, n, nb, n + 2, B, true));
    fill_f64<<<(unsigned)pg_cdiv((long long)nb * n, 256), 256, 0, st>>>(B.bcpart, (long long)nb * n);
    const long long ns = (long long)n * S;
    std::vector<int32_t> pad((size_t)nb * S);
    for (long long g0 = 0; g0 < nsrc; g0 += (long long)nb * S) {
        for (int s = 0; s < nb * S; s++) pad[s] = g0 + s < nsrc ? hsrc[g0 + s] : -1;
        PG_HIP(ctx, hipMemcpyAsync(B.srcs, pad.data(), pad.size() * 4, hipMemcpyHostToDevice, st));
        int maxlev = 0;
        PG_TRY(run_forward(ctx, drow, dcol, n, B, &maxlev));
        bc_coef_init<<<(unsigned)pg_cdiv(ns * nb, 256), 256, 0, st>>>(ns * nb, B.sigma, B.coef);
        for (int d = maxlev - 1; d >= 1; d--)
            bc_backward<<<dim3((unsigned)pg_cdiv(n, 4), (unsigned)nb), 256, 0, st>>>(drow, dcol, n, d, B);
        PG_KERNEL_CHECK(ctx);
    }
    bc_reduce<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(B.bcpart, nb, n, dbc);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, bc, dbc, (size_t)n));
    PG_HIP(ctx, hipStreamSynchronize(st));
    return PG_OK;
}

namespace {
__global__ void to_float(const double *__restrict__ x, int n, float *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}
}  // namespace

extern "C" pg_status pg_bc_float(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n, int64_t nnz,
                                 const int32_t *sources, int64_t nsrc, float *bc)
{
    if (!ctx) return PG_EINVAL;
    if (!bc || n < 0) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    double *acc;
    PG_TRY(pg_bufT(ctx, "bcf.acc", (size_t)n + 1, &acc));
    PG_TRY(pg_bc(ctx, row, col, n, nnz, sources, nsrc, acc));
    float *dbc;
    PG_TRY(pg_out(ctx, bc, (size_t)n, "out.bcf", &dbc));
    if (n > 0) to_float<<<(unsigned)pg_cdiv(n, 256), 256, 0, ctx->stream>>>(acc, n, dbc);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, bc, dbc, (size_t)n));
    PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PG_OK;
}

extern "C" pg_status pg_bfs(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n, int64_t nnz,
                            const int32_t *sources, int64_t nsrc, int32_t *level)
{
    if (!ctx) return PG_EINVAL;
    if (!level || !sources || nsrc < 0) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const int *drow, *dcol;
    PG_TRY(stage_csr(ctx, row, col, nullptr, n, nnz, false, &drow, &dcol, nullptr));
    std::vector<int32_t> hsrc((size_t)nsrc);
    if (pg_is_device_ptr(sources)) PG_HIP(ctx, hipMemcpy(hsrc.data(), sources, nsrc * 4, hipMemcpyDeviceToHost));
    else std::memcpy(hsrc.data(), sources, nsrc * 4);
    for (auto s : hsrc)
        if (s < 0 || s >= n) return pg_fail(ctx, PG_EINVAL, "source %d out of range", s);
    int *dlev;
    PG_TRY(pg_out(ctx, level, (size_t)n * nsrc, "out.level", &dlev));
    hipStream_t st = ctx->stream;
    if (n > 0 && nsrc > 0) {
        const int nb = bc_batches(ctx, n, nnz, nsrc);
        BatchBufs B;
        PG_TRY(batch_bufs(ctx, n, nb, n + 2, B, false));
        std::vector<int32_t> pad((size_t)nb * S);
        for (long long g0 = 0; g0 < nsrc; g0 += (long long)nb * S) {
            for (int s = 0; s < nb * S; s++) pad[s] = g0 + s < nsrc ? hsrc[g0 + s] : -1;
            PG_HIP(ctx, hipMemcpyAsync(B.srcs, pad.data(), pad.size() * 4, hipMemcpyHostToDevice, st));
            int maxlev = 0;
            PG_TRY(run_forward(ctx, drow, dcol, n, B, &maxlev));
            const int nvalid = (int)std::min<long long>((long long)nb * S, nsrc - g0);
            bfs_export<<<dim3((unsigned)pg_cdiv((long long)n * S, 256), (unsigned)nb), 256, 0, st>>>(n, nvalid, B.lev,
                                                                                                 dlev + g0 * n);
            PG_KERNEL_CHECK(ctx);
        }
    }
    PG_TRY(pg_out_finish(ctx, level, dlev, (size_t)n * nsrc));
    PG_HIP(ctx, hipStreamSynchronize(st));
    return PG_OK;
}
