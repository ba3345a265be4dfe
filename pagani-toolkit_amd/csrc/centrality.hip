// centrality.hip -- eigenvector centrality and participation coefficient,
// MI355X (gfx950).
//
//  * Eigenvector centrality (U/CUEC/CUEC/CUEC_unweighted.cu:61-196, W/CUEC/
//    weighted-net.cu:55-203): the reference's power iteration -- start vector
//    0..n-1; per iteration u0 = u, |u|, u *= 1/|u|, u = A u, err1 = |u - u0|,
//    err2 = |u + u0|; stop when either is <= 1e-6 or after 10000 iterations --
//    as one launch per iteration.  Rows are split over a fixed set of blocks;
//    every block first reduces the previous launch's per-block partials (|u|^2,
//    |u - u0|^2, |u + u0|^2) in a fixed order, so all blocks take the same stop
//    decision on the device and launches past the stop return at once.  The
//    host looks once per batch of launches.  FP64 throughout, FP32 weights.
//    The SpMV is column-segmented: each block stages 12,288-entry segments of
//    the scaled vector in LDS and every wave gathers from there, one row at a
//    time, over the row's entries in that segment (rows are column-sorted, so
//    each is a contiguous range found once per call); an FP64 gather straight
//    from L2 would move a whole cache line per 8-byte element.
//  * Participation coefficient (W/PC_CPU/PC_CPU/main.cpp:96-160): one wave per
//    vertex, FP32 module sums accumulated in LDS in CSR order by one lane
//    (entries staged 64 at a time), then the FP32 sum of squared shares over
//    modules 1..K in order -- the reference's arithmetic, so bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pg_internal.hpp"

namespace {

constexpr int EC_MAX_ITER = 10000;  // CUEC_unweighted.cu:25
constexpr double EC_EPS = 1e-6;     // CUEC_unweighted.cu:28
constexpr int EC_TPB = 1024;
constexpr int EC_BATCH = 32;        // launches per host check
constexpr int EC_SEG = 12288;       // vector entries per LDS segment (96 KB)

struct EcState {
    int stopped, iters;
    double lam, err1, err2;
};

// fixed-order sum over the block (every thread gets the result)
__device__ double ec_block_sum(double x, double *red)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (lane == 0) red[wv] = x;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < EC_TPB / 64; q++) s += red[q];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(EC_TPB) void ec_init(int n, int R, double *__restrict__ u, double *__restrict__ part,
                                                  EcState *__restrict__ st)
{
    __shared__ double red[EC_TPB / 64];
    const int r0 = blockIdx.x * R, r1 = min(n, r0 + R);
    double s = 0.0;
    for (int i = r0 + threadIdx.x; i < r1; i += EC_TPB) {
        const double x = (double)i;
        u[i] = x;
        s += x * x;
    }
    s = ec_block_sum(s, red);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = 0.0;
        part[3 * blockIdx.x + 1] = 0.0;
        part[3 * blockIdx.x + 2] = s;
        if (blockIdx.x == 0) *st = EcState{0, 0, 0.0, 1.0, 1.0};
    }
}

// segoff[i (nseg + 1) + s] = first entry of row i with column >= s * EC_SEG
__global__ void ec_segments(const int *__restrict__ row, const int *__restrict__ col, int n, int nseg,
                            int *__restrict__ segoff)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = row[i], b = row[i + 1];
    for (int s = 0; s <= nseg; s++) {
        const long long x = (long long)s * EC_SEG;
        int lo = a, hi = b;  // first e in [a, b) with col[e] >= x
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (col[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        segoff[(long long)i * (nseg + 1) + s] = lo;
    }
}

// Iteration k: block b owns rows [b R, (b+1) R); per segment 16 lanes per row.
// Dynamic LDS: [EC_SEG] scaled vector segment, then [R] row partials.
__global__ __launch_bounds__(EC_TPB) void ec_iter(const int *__restrict__ col, const float *__restrict__ val,
                                                  int n, int R, int G, int k, int nseg,
                                                  const int *__restrict__ segoff, const double *__restrict__ uin,
                                                  double *__restrict__ uout, double *__restrict__ vhat,
                                                  const double *__restrict__ pin, double *__restrict__ pout,
                                                  EcState *__restrict__ st)
{
#pragma clang fp contract(off)
    extern __shared__ double xs[];
    double *yp = xs + EC_SEG;
    __shared__ double red[EC_TPB / 64];
    if (st->stopped) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int q = tid; q < G; q += EC_TPB) {
        a0 += pin[3 * q + 0];
        a1 += pin[3 * q + 1];
        a2 += pin[3 * q + 2];
    }
    a0 = ec_block_sum(a0, red);
    a1 = ec_block_sum(a1, red);
    a2 = ec_block_sum(a2, red);
    if (k > 0) {
        const double e1 = sqrt(a0), e2 = sqrt(a1);
        if (!(e1 > EC_EPS && e2 > EC_EPS) || k >= EC_MAX_ITER) {
            if (blockIdx.x == 0 && tid == 0) {
                st->err1 = e1;
                st->err2 = e2;
                st->iters = k;
                st->stopped = 1;
            }
            return;
        }
    }
    const double vnorm = sqrt(a2);
    const double t = 1 / vnorm;
    if (blockIdx.x == 0 && tid == 0) st->lam = vnorm;
    const int r0 = blockIdx.x * R, r1 = min(n, r0 + R);
    for (int q = tid; q < r1 - r0; q += EC_TPB) yp[q] = 0.0;
    for (int s = 0; s < nseg; s++) {
        const int x0 = s * EC_SEG, xn = min(EC_SEG, n - x0);
        __syncthreads();  // the previous segment's readers are done
        for (int x = tid; x < xn; x += EC_TPB) xs[x] = uin[x0 + x] * t;
        __syncthreads();
        // four rows per wave, 16 lanes each, two accumulators per lane: the
        // rows' index loads are in flight together
        const int grp = lane >> 4, gl = lane & 15;
        for (int base = r0 + wv * 4; base < r1; base += EC_TPB / 16) {
            const int i = base + grp;
            int a = 0, b = 0;
            if (i < r1) {
                const long long so = (long long)i * (nseg + 1) + s;
                a = segoff[so];
                b = segoff[so + 1];
            }
            double y0 = 0.0, y1 = 0.0;
            if (val) {
                for (int e = a + gl; e < b; e += 32) {
                    y0 = fma((double)val[e], xs[col[e] - x0], y0);
                    if (e + 16 < b) y1 = fma((double)val[e + 16], xs[col[e + 16] - x0], y1);
                }
            } else {
                for (int e = a + gl; e < b; e += 32) {
                    y0 += xs[col[e] - x0];
                    if (e + 16 < b) y1 += xs[col[e + 16] - x0];
                }
            }
            double y = y0 + y1;
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) y += __shfl_xor(y, o);
            if (i < r1 && gl == 0) yp[i - r0] += y;
        }
    }
    __syncthreads();
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = r0 + tid; i < r1; i += EC_TPB) {
        const double y = yp[i - r0], ui = uin[i];
        vhat[i] = ui * t;
        uout[i] = y;
        const double d1 = y - ui, d2 = y + ui;
        s0 += d1 * d1;
        s1 += d2 * d2;
        s2 += y * y;
    }
    s0 = ec_block_sum(s0, red);
    s1 = ec_block_sum(s1, red);
    s2 = ec_block_sum(s2, red);
    if (tid == 0) {
        pout[3 * blockIdx.x + 0] = s0;
        pout[3 * blockIdx.x + 1] = s1;
        pout[3 * blockIdx.x + 2] = s2;
    }
}

// ------------------------------------------------ participation coefficient

// flags[0] |= label < 1; flags[1] = max label
__global__ void label_scan(const int *__restrict__ lab, int n, int *__restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int l = lab[i];
    if (l < 1) atomicOr(&flags[0], 1);
    atomicMax(&flags[1], l);
}

__global__ __launch_bounds__(64) void pc_kernel(const int *__restrict__ row, const int *__restrict__ col,
                                                const float *__restrict__ val, const int *__restrict__ lab, int n,
                                                int K, float *__restrict__ pc, float *__restrict__ pcn)
{
#pragma clang fp contract(off)
    extern __shared__ float wm[];  // K module sums
    __shared__ int sm_m[64];
    __shared__ float sm_v[64];
    const int i = blockIdx.x, lane = threadIdx.x;
    const int a = row[i], b = row[i + 1];
    for (int j = lane; j < K; j += 64) wm[j] = 0.0f;
    float deg = 0.0f;  // lane 0: the reference's sequential FP32 sum
    for (int base = a; base < b; base += 64) {
        const int e = base + lane;
        if (e < b) {
            sm_m[lane] = lab[col[e]] - 1;
            sm_v[lane] = val ? val[e] : 1.0f;
        }
        __syncthreads();
        if (lane == 0) {
            const int cnt = min(64, b - base);
            for (int q = 0; q < cnt; q++) {
                deg += sm_v[q];
                wm[sm_m[q]] += sm_v[q];
            }
        }
        __syncthreads();
    }
    if (lane != 0) return;
    float r;
    if (deg == 0) {
        r = 0.0f;
    } else {
        float sum = 0.0f;
        for (int j = 0; j < K; j++) {
            const float t = wm[j] / deg;
            sum += t * t;
        }
        r = 1.0f - sum;
    }
    pc[i] = r;
    if (pcn) pcn[i] = K != 1 ? r / ((float)(K - 1) / (float)K) : r;
}

}  // namespace

extern "C" pg_status pg_eigenvector_centrality(pg_ctx *ctx, const int32_t *row, const int32_t *col,
                                               const float *val, int32_t n, int64_t nnz, double *vhat,
                                               double *avec, double *lambda, int32_t *iters)
{
    if (!ctx) return PG_EINVAL;
    PG_HIP(ctx, hipSetDevice(ctx->device));
    PG_TRY(pg_check_csr_args(ctx, row, col, n, nnz));
    if (n < 1) return pg_fail(ctx, PG_EINVAL, "n must be >= 1");
    const int *drow, *dcol;
    const float *dval = nullptr;
    PG_TRY(pg_in(ctx, row, (size_t)n + 1, "in.row", &drow));
    PG_TRY(pg_in(ctx, col, (size_t)nnz, "in.col", &dcol));
    if (val) PG_TRY(pg_in(ctx, val, (size_t)nnz, "in.val", &dval));
    PG_TRY(pg_validate_csr(ctx, drow, dcol, n, nnz, true));
    hipStream_t st = ctx->stream;
    // one block per CU (96 KB segment + row partials in LDS)
    const int G = (int)std::max<int64_t>(1, std::min<int64_t>(ctx->num_cu, pg_cdiv(n, 64)));
    const int R = (int)pg_cdiv(n, G);
    const int nseg = (int)pg_cdiv(n, EC_SEG);
    const size_t lds = ((size_t)EC_SEG + (size_t)R) * sizeof(double);
    if (lds > 160 * 1024) return pg_fail(ctx, PG_ENOTSUP, "n = %d needs %zu bytes of LDS per block", n, lds);
    PG_HIP(ctx, hipFuncSetAttribute((const void *)ec_iter, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int *segoff;
    PG_TRY(pg_bufT(ctx, "ec.segoff", (size_t)n * (nseg + 1), &segoff));
    ec_segments<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(drow, dcol, n, nseg, segoff);
    double *u[2], *part[2], *dvh, *dav;
    EcState *dst;
    PG_TRY(pg_bufT(ctx, "ec.u0", (size_t)n, &u[0]));
    PG_TRY(pg_bufT(ctx, "ec.u1", (size_t)n, &u[1]));
    PG_TRY(pg_bufT(ctx, "ec.p0", 3 * (size_t)G, &part[0]));
    PG_TRY(pg_bufT(ctx, "ec.p1", 3 * (size_t)G, &part[1]));
    PG_TRY(pg_bufT(ctx, "ec.state", 1, &dst));
    PG_TRY(pg_out(ctx, vhat, (size_t)n, "out.ecvhat", &dvh));
    if (!dvh) PG_TRY(pg_bufT(ctx, "ec.vhat", (size_t)n, &dvh));
    // iteration k reads u[k & 1] and part[(k + 1) & 1], writes the other two;
    // ec_init fills u[0] and part[1]
    ec_init<<<(unsigned)G, EC_TPB, 0, st>>>(n, R, u[0], part[1], dst);
    PG_KERNEL_CHECK(ctx);
    EcState h{};
    for (int k0 = 0; k0 <= EC_MAX_ITER; k0 += EC_BATCH) {
        for (int k = k0; k < k0 + EC_BATCH && k <= EC_MAX_ITER; k++)
            ec_iter<<<(unsigned)G, EC_TPB, lds, st>>>(dcol, dval, n, R, G, k, nseg, segoff, u[k & 1],
                                                      u[(k + 1) & 1], dvh, part[(k + 1) & 1], part[k & 1], dst);
        PG_KERNEL_CHECK(ctx);
        PG_HIP(ctx, hipMemcpyAsync(&h, dst, sizeof h, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        if (h.stopped) break;
    }
    if (!h.stopped) return pg_fail(ctx, PG_EDEVICE, "power iteration did not stop");
    // the last executed iteration (iters - 1) wrote y into u[iters & 1]
    PG_TRY(pg_out_finish(ctx, vhat, dvh, (size_t)n));
    if (avec) {
        PG_TRY(pg_out(ctx, avec, (size_t)n, "out.ecavec", &dav));
        PG_HIP(ctx, hipMemcpyAsync(dav, u[h.iters & 1], (size_t)n * 8, hipMemcpyDeviceToDevice, st));
        PG_TRY(pg_out_finish(ctx, avec, dav, (size_t)n));
    }
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (lambda) *lambda = h.lam;
    if (iters) *iters = h.iters;
    return PG_OK;
}

extern "C" pg_status pg_participation(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val,
                                      int32_t n, int64_t nnz, const int32_t *labels, float *pc,
                                      float *pc_norm, int32_t *nmod)
{
    if (!ctx) return PG_EINVAL;
    PG_HIP(ctx, hipSetDevice(ctx->device));
    PG_TRY(pg_check_csr_args(ctx, row, col, n, nnz));
    if (n < 1 || !labels || !pc) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    const int *drow, *dcol, *dlab;
    const float *dval = nullptr;
    PG_TRY(pg_in(ctx, row, (size_t)n + 1, "in.row", &drow));
    PG_TRY(pg_in(ctx, col, (size_t)nnz, "in.col", &dcol));
    if (val) PG_TRY(pg_in(ctx, val, (size_t)nnz, "in.val", &dval));
    PG_TRY(pg_in(ctx, labels, (size_t)n, "in.labels", &dlab));
    PG_TRY(pg_validate_csr(ctx, drow, dcol, n, nnz, false));
    hipStream_t st = ctx->stream;
    int *flags, hf[2] = {0, 0};
    PG_TRY(pg_bufT(ctx, "pc.flags", 2, &flags));
    PG_HIP(ctx, hipMemsetAsync(flags, 0, 8, st));
    label_scan<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(dlab, n, flags);
    PG_KERNEL_CHECK(ctx);
    PG_HIP(ctx, hipMemcpyAsync(hf, flags, 8, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (hf[0]) return pg_fail(ctx, PG_EINVAL, "module labels must be >= 1");
    const int K = hf[1];
    const size_t lds = (size_t)K * sizeof(float);
    if (lds > ctx->lds_per_block - 1024)
        return pg_fail(ctx, PG_ENOTSUP, "%d modules exceed the LDS budget of one wave", K);
    float *dpc, *dpn;
    PG_TRY(pg_out(ctx, pc, (size_t)n, "out.pc", &dpc));
    PG_TRY(pg_out(ctx, pc_norm, (size_t)n, "out.pcn", &dpn));
    pc_kernel<<<(unsigned)n, 64, lds, st>>>(drow, dcol, dval, dlab, n, K, dpc, dpn);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, pc, dpc, (size_t)n));
    PG_TRY(pg_out_finish(ctx, pc_norm, dpn, (size_t)n));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (nmod) *nmod = K;
    return PG_OK;
}
