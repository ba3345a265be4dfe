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
//    The SpMV is column-segmented and sliced: each block stages 12,288-entry
//    segments of the scaled vector in LDS (an FP64 gather straight from L2
//    would move a whole cache line per 8-byte element); the matrix is
//    re-laid once per call as slices of 64 rows, one per wave, whose entries
//    in each segment are stored in groups of 8 positions: group g holds, for
//    each of the 64 rows in turn, its entries 8g .. 8g+7 as 16-bit
//    segment-relative indices (one 16-byte load per lane, 1 KB per wave),
//    padded to the slice's longest row with the index of a zero LDS slot.
//    Each lane sums its own row in column order, y = y + w x, as the
//    reference's loop.
//  * Participation coefficient (W/PC_CPU/PC_CPU/main.cpp:96-160): one wave per
//    vertex, FP32 module sums accumulated in LDS in CSR order by one lane
//    (entries staged 64 at a time), then the FP32 sum of squared shares over
//    modules 1..K in order -- the reference's arithmetic, so bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pg_internal.hpp"
#include "pg_scan.hpp"

namespace {

constexpr int EC_MAX_ITER = 10000;  // CUEC_unweighted.cu:25
constexpr double EC_EPS = 1e-6;     // CUEC_unweighted.cu:28
constexpr int EC_TPB = 1024;
constexpr int EC_BATCH = 32;        // launches per host check
constexpr int EC_SEG = 12288;       // vector entries per LDS segment (96 KB)
constexpr int EC_RB = 256;          // rows per block of the sliced SpMV: 4 slices of 64
constexpr int EC_PG = 8;            // positions per lane load (16 bytes of indices)

struct EcState {
    int stopped, iters;
    double lam, err1, err2;
};

// fixed-order sum over a block of TPB threads (every thread gets the result)
template <int TPB = EC_TPB>
__device__ double ec_block_sum(double x, double *red)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (lane == 0) red[wv] = x;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < TPB / 64; q++) s += red[q];
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

// Slice q (rows 64q .. 64q+63), segment s: len = the longest row part.
__global__ void ec_slice_len(const int *__restrict__ segoff, int n, int nslice, int nseg, int *__restrict__ len)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nslice * nseg) return;
    const int q = x / nseg, s = x % nseg;
    int m = 0;
    for (int i = q * 64; i < min(n, q * 64 + 64); i++) {
        const long long so = (long long)i * (nseg + 1) + s;
        m = max(m, segoff[so + 1] - segoff[so]);
    }
    len[x] = (m + EC_PG - 1) / EC_PG * EC_PG;
}

// off[x] = 64 * (len[0] + ... + len[x - 1]), x = 0 .. m, by one work-group.
__global__ __launch_bounds__(1024) void ec_len_scan(const int *__restrict__ len, int m, long long *__restrict__ off)
{
    pgdev::block_scan_1024(len, m, off, 64);
}

// Row slot i (0 .. 64 nslice), segment s: entry p at off + (p / 8) 512 +
// (i % 64) 8 + p % 8, then the padding index EC_SEG (a zero in LDS) with
// weight 0.
__global__ void ec_fill(const int *__restrict__ col, const float *__restrict__ val, const int *__restrict__ segoff,
                        const long long *__restrict__ off, int n, int nslice, int nseg,
                        uint16_t *__restrict__ ecol, float *__restrict__ eval)
{
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (long long)nslice * 64 * nseg) return;
    const int s = (int)(x % nseg), i = (int)(x / nseg);
    const int q = i >> 6, l = i & 63;
    const long long base = off[(long long)q * nseg + s], len = (off[(long long)q * nseg + s + 1] - base) / 64;
    int a = 0, cnt = 0;
    if (i < n) {
        const long long so = (long long)i * (nseg + 1) + s;
        a = segoff[so];
        cnt = segoff[so + 1] - a;
    }
    const int x0 = s * EC_SEG;
    for (long long p = 0; p < len; p++) {
        const long long d = base + (p / EC_PG) * (64 * EC_PG) + l * EC_PG + p % EC_PG;
        if (p < cnt) {
            ecol[d] = (uint16_t)(col[a + p] - x0);
            if (eval) eval[d] = val[a + p];
        } else {
            ecol[d] = (uint16_t)EC_SEG;
            if (eval) eval[d] = 0.0f;
        }
    }
}

// Iteration k: block b owns rows [256 b, 256 b + 256), lane = row.
// Dynamic LDS: [EC_SEG + 1] scaled vector segment, the last entry 0.
__global__ __launch_bounds__(EC_RB) void ec_iter(const uint16_t *__restrict__ ecol, const float *__restrict__ eval,
                                                 const long long *__restrict__ off, int n, int nslice, int G, int k,
                                                 int nseg,
                                                 const double *__restrict__ uin, double *__restrict__ uout,
                                                 double *__restrict__ vhat, const double *__restrict__ pin,
                                                 double *__restrict__ pout, EcState *__restrict__ st)
{
#pragma clang fp contract(off)
    extern __shared__ double xs[];
    __shared__ double red[EC_RB / 64];
    if (st->stopped) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int q = tid; q < G; q += EC_RB) {
        a0 += pin[3 * q + 0];
        a1 += pin[3 * q + 1];
        a2 += pin[3 * q + 2];
    }
    a0 = ec_block_sum<EC_RB>(a0, red);
    a1 = ec_block_sum<EC_RB>(a1, red);
    a2 = ec_block_sum<EC_RB>(a2, red);
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
    if (tid == 0) xs[EC_SEG] = 0.0;  // the padding slot; staging never writes it
    const int i = blockIdx.x * EC_RB + tid;
    const int qi = blockIdx.x * (EC_RB / 64) + wv;  // this wave's slice (none past the last row)
    const long long qs = (long long)qi * nseg;
    double y = 0.0;
    for (int s = 0; s < nseg; s++) {
        const int x0 = s * EC_SEG, xn = min(EC_SEG, n - x0);
        __syncthreads();  // the previous segment's readers are done
        {
            // 16-byte loads, eight in flight per thread; an odd tail reads the
            // buffer's padding word into a slot no column refers to
            const double2 *src = (const double2 *)(uin + x0);
            double2 *dst = (double2 *)xs;
            const int n2 = (xn + 1) / 2;
            for (int x = tid; x < n2; x += EC_RB * 8) {
                double2 v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int q = x + u * EC_RB;
                    v[u] = q < n2 ? src[q] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int q = x + u * EC_RB;
                    if (q < n2) dst[q] = make_double2(v[u].x * t, v[u].y * t);
                }
            }
        }
        __syncthreads();
        const long long a = qi < nslice ? off[qs + s] : 0;
        const int ng = qi < nslice ? (int)((off[qs + s + 1] - a) / (64 * EC_PG)) : 0;
        const uint4 *cp = (const uint4 *)(ecol + a) + lane;  // 8 indices per lane and group
        if (eval) {
            const float4 *vp = (const float4 *)(eval + a) + 2 * lane;
#pragma unroll 4
            for (int g = 0; g < ng; g++) {
                const uint4 c = cp[(long long)g * 64];
                const float4 w0 = vp[(long long)g * 128], w1 = vp[(long long)g * 128 + 1];
                const unsigned cc[4] = {c.x, c.y, c.z, c.w};
                const float ww[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
                for (int h = 0; h < EC_PG; h++)
                    y = y + (double)ww[h] * xs[(cc[h >> 1] >> (16 * (h & 1))) & 0xffffu];
            }
        } else {
#pragma unroll 4
            for (int g = 0; g < ng; g++) {
                const uint4 c = cp[(long long)g * 64];
                const unsigned cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
                for (int h = 0; h < EC_PG; h++) y = y + xs[(cc[h >> 1] >> (16 * (h & 1))) & 0xffffu];
            }
        }
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (i < n) {
        const double ui = uin[i];
        vhat[i] = ui * t;
        uout[i] = y;
        const double d1 = y - ui, d2 = y + ui;
        s0 = d1 * d1;
        s1 = d2 * d2;
        s2 = y * y;
    }
    s0 = ec_block_sum<EC_RB>(s0, red);
    s1 = ec_block_sum<EC_RB>(s1, red);
    s2 = ec_block_sum<EC_RB>(s2, red);
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
    // 256 rows per block (4 slices of 64), the scaled vector staged in LDS
    // segment by segment
    const int G = (int)pg_cdiv(n, EC_RB);
    const int R = EC_RB;
    const int nseg = (int)pg_cdiv(n, EC_SEG);
    const int nslice = (int)pg_cdiv(n, 64);
    const size_t lds = ((size_t)EC_SEG + 1) * sizeof(double);
    PG_HIP(ctx, hipFuncSetAttribute((const void *)ec_iter, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int *segoff, *slen;
    long long *soff;
    PG_TRY(pg_bufT(ctx, "ec.segoff", (size_t)n * (nseg + 1), &segoff));
    PG_TRY(pg_bufT(ctx, "ec.slen", (size_t)nslice * nseg, &slen));
    PG_TRY(pg_bufT(ctx, "ec.soff", (size_t)nslice * nseg + 1, &soff));
    ec_segments<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(drow, dcol, n, nseg, segoff);
    ec_slice_len<<<(unsigned)pg_cdiv((long long)nslice * nseg, 256), 256, 0, st>>>(segoff, n, nslice, nseg, slen);
    ec_len_scan<<<1, 1024, 0, st>>>(slen, nslice * nseg, soff);
    PG_KERNEL_CHECK(ctx);
    long long npad = 0;
    PG_HIP(ctx, hipMemcpyAsync(&npad, soff + (long long)nslice * nseg, 8, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    uint16_t *ecol;
    float *eval = nullptr;
    PG_TRY(pg_bufT(ctx, "ec.ecol", (size_t)std::max<long long>(npad, 1), &ecol));
    if (dval) PG_TRY(pg_bufT(ctx, "ec.eval", (size_t)std::max<long long>(npad, 1), &eval));
    ec_fill<<<(unsigned)pg_cdiv((long long)nslice * 64 * nseg, 256), 256, 0, st>>>(dcol, dval, segoff, soff, n, nslice,
                                                                                  nseg, ecol, eval);
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
            ec_iter<<<(unsigned)G, EC_RB, lds, st>>>(ecol, eval, soff, n, nslice, G, k, nseg, u[k & 1], u[(k + 1) & 1],
                                                     dvh, part[(k + 1) & 1], part[k & 1], dst);
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
