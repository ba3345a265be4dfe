// cormat.hip -- the dense correlation matrix, for subject averaging and the
// .cormat output of CUCorMat (U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:236-262,
// 380-620, 775-838; W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:386-600, 853-885),
// MI355X (gfx950).
//
// The matrix is kept as the reference keeps it: the packed upper triangle,
// row-major over i < j, M1 = n(n-1)/2 FP32 values (6.7 GB at n = 58,000, well
// inside one GPU's HBM).  Per subject:
//
//  * FP64 Gram on MFMA (v_mfma_f64_16x16x4f64), 128 x 128 tiles of the upper
//    triangle, z staged through LDS 16 time points at a time.  Every pair is
//    rounded to FP32 exactly as the FP64 oracle rounds it: the MFMA sum and the
//    oracle's sequential sum differ by at most delta = 2.5 t 2^-53 (both are
//    FP64 dot products of unit-norm rows), so when r - delta and r + delta
//    round to the same float that float is the oracle's; otherwise (mostly
//    pairs with r near 0, where FP32 steps are tiny) the pair goes to a list
//    that a second kernel recomputes in the oracle's sequential order, one
//    thread per pair (a lane stalling its wave for a 240-step dependent sum
//    would idle the matrix cores).
//  * The epilogue writes every pair's FP32 r into a per-subject packed buffer
//    straight from the MFMA result registers (16 consecutive columns of a row
//    per 16 lanes); a streaming kernel then adds r, or its Fisher z, to the
//    packed accumulator (the first subject without Fisher is written in place).
//
// After the last subject: divide by the subject count, inverse Fisher.  Then
// thresholds as the reference applies them to the averaged matrix: r-mode
// keeps a > thr; s-mode keeps the k largest with the earliest (row-major) ties
// at the k-th value, found by a 3-pass radix select over the M1 values.  The
// CSR is built in order without sorting: a per-pair keep byte in packed order,
// then one work-group per 64 CSR rows that reads the lower part of its rows as
// 64-byte slices of the packed rows above it (column order for free) and the
// upper part from the rows' own packed segments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pg_internal.hpp"
#include "pg_scan.hpp"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef unsigned long long u64;

constexpr int GT = 128;       // Gram tile edge
constexpr int GKC = 16;       // time points per LDS stage
constexpr int GLD = GKC + 1;  // LDS row stride (doubles)

__host__ __device__ __forceinline__ long long pidx(long long i, long long j, long long n)
{
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

// U CorMat.cpp:35-42: FP32 arithmetic, the logarithm in FP64
__device__ float fisher_f(float r)
{
#pragma clang fp contract(off)
    if (r == 1.0f) r = (float)((double)r - 1e-6);
    const float q = (1.0f + r) / (1.0f - r);
    return (float)(0.5 * log((double)q));
}

// U CorMat.cpp:44-49: FP32 arithmetic, the exponential in FP64
__device__ float inv_fisher_f(float z)
{
#pragma clang fp contract(off)
    float r = (float)exp((double)(2.0f * z));
    r = (r - 1.0f) / (r + 1.0f);
    return r;
}

// The oracle's FP32 value of the pair, (float) of its clamped sequential FP64
// sum, recomputed in that order.
__device__ float seq_r(const double *__restrict__ zi, const double *__restrict__ zj, int t)
{
#pragma clang fp contract(off)
    double s = 0.0;
    for (int l = 0; l < t; l++) {
        const double p = zi[l] * zj[l];
        s += p;
    }
    if (s > 1.0) s = 1.0;
    if (s < -1.0) s = -1.0;
    return (float)s;
}

// (float) of the clamped MFMA r when every value within delta of it rounds to
// the same float (then it is the oracle's); *amb set otherwise.
__device__ __forceinline__ float round_r(double r, double delta, bool *amb)
{
    const double lo = fmax(-1.0, fmin(1.0, r - delta));
    const double hi = fmax(-1.0, fmin(1.0, r + delta));
    const float flo = (float)lo;
    *amb = flo != (float)hi;
    return flo;
}

// Tile (blockIdx.y, blockIdx.x) of rows x columns; 8 waves as 4 x 2, each
// 32 x 64 as 2 x 4 MFMA blocks of 16 x 16 (two waves per SIMD hide each
// other's prologue, barrier and epilogue).  MFMA f64 16x16x4 operands: lane l
// holds A[l % 16][l / 16] and B[l / 16][l % 16]; result element v of lane l is
// D[l / 16 + 4 v][l % 16].  Software pipeline: the next 16 time points of both
// panels are loaded into registers while the MFMAs run on the current LDS
// buffer, then stored into the other buffer; one barrier per step.
constexpr int GTHREADS = 512;
constexpr int GLDS = 2 * 2 * GT * GLD;      // doubles: 2 buffers x (A, B)
constexpr int GNLD = GT * GKC / GTHREADS;   // doubles per thread per panel and step

__global__ __launch_bounds__(GTHREADS) __attribute__((amdgpu_waves_per_eu(4))) void gram64_r(const double *__restrict__ z, int n, int t, int tp,
                                                     double delta, float *__restrict__ out, int2 *__restrict__ amb,
                                                     unsigned *__restrict__ namb, unsigned cap)
{
    extern __shared__ double sm[];  // [buf][A|B][GT][GLD]
    const int bi = blockIdx.y, bj = blockIdx.x;
    if (bj < bi) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;  // rows [32 wr, +32), columns [64 wc, +64)
    const int i0 = bi * GT, j0 = bj * GT;
    v4d c[2][4];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int q = 0; q < 4; q++) c[m][q] = v4d{0.0, 0.0, 0.0, 0.0};
    double ra[GNLD], rb[GNLD];
    auto gload = [&](int k0) {
#pragma unroll
        for (int x = 0; x < GNLD; x++) {
            const int e = tid + GTHREADS * x, r = e / GKC, k = e % GKC;
            ra[x] = z[(long long)(i0 + r) * tp + k0 + k];
            rb[x] = z[(long long)(j0 + r) * tp + k0 + k];
        }
    };
    auto sstore = [&](int buf) {
        double *A = sm + (size_t)(2 * buf) * GT * GLD, *B = A + GT * GLD;
#pragma unroll
        for (int x = 0; x < GNLD; x++) {
            const int e = tid + GTHREADS * x, r = e / GKC, k = e % GKC;
            A[r * GLD + k] = ra[x];
            B[r * GLD + k] = rb[x];
        }
    };
    const int nsteps = tp / GKC;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int s = 0; s < nsteps; s++) {
        const int cur = s & 1;
        if (s + 1 < nsteps) gload((s + 1) * GKC);
        const double *A = sm + (size_t)(2 * cur) * GT * GLD, *B = A + GT * GLD;
#pragma unroll
        for (int ks = 0; ks < GKC / 4; ks++) {
            double a[2], b[4];
#pragma unroll
            for (int m = 0; m < 2; m++) a[m] = A[(wr * 32 + m * 16 + (lane & 15)) * GLD + ks * 4 + (lane >> 4)];
#pragma unroll
            for (int q = 0; q < 4; q++) b[q] = B[(wc * 64 + q * 16 + (lane & 15)) * GLD + ks * 4 + (lane >> 4)];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int q = 0; q < 4; q++) c[m][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[q], c[m][q], 0, 0, 0);
        }
        if (s + 1 < nsteps) sstore(cur ^ 1);
        __syncthreads();
    }
    // every pair's FP32 r, straight from the MFMA registers (16 consecutive
    // columns of a row per 16 lanes)
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + wr * 32 + m * 16 + (lane >> 4) + 4 * v;
                const int j = j0 + wc * 64 + q * 16 + (lane & 15);
                if (i < j && j < n) {
                    bool a;
                    float f = round_r(c[m][q][v], delta, &a);
                    if (a) {
                        // left for gram64_recheck; inline when the list is full
                        const unsigned slot = atomicAdd(namb, 1u);
                        if (slot < cap) {
                            amb[slot] = make_int2(i, j);
                            continue;
                        }
                        f = seq_r(z + (long long)i * tp, z + (long long)j * tp, t);
                    }
                    out[pidx(i, j, n)] = f;
                }
            }
}

// The pairs whose rounding the Gram left open: one thread each, the oracle's
// sequential sum.
__global__ void gram64_recheck(const double *__restrict__ z, int n, int t, int tp, float *__restrict__ out,
                               const int2 *__restrict__ amb, const unsigned *__restrict__ namb, unsigned cap)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= min(*namb, cap)) return;
    const int2 p = amb[k];
    out[pidx(p.x, p.y, n)] = seq_r(z + (long long)p.x * tp, z + (long long)p.y * tp, t);
}

// post_block (U CorMat.cpp:806-838): acc += r, or its Fisher z
// 16-byte accesses (4 pairs per thread and step), the last m % 4 pairs one by
// one; the buffers are device allocations, so 16-byte aligned.
__global__ void cm_accumulate(long long m, const float *__restrict__ r, int fisher, float *__restrict__ acc)
{
#pragma clang fp contract(off)
    const long long m4 = m >> 2, stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const float4 *r4 = (const float4 *)r;
    float4 *a4 = (float4 *)acc;
    for (long long q = t0; q < m4; q += stride) {
        const float4 x = r4[q];
        float4 y = a4[q];
        y.x = y.x + (fisher ? fisher_f(x.x) : x.x);
        y.y = y.y + (fisher ? fisher_f(x.y) : x.y);
        y.z = y.z + (fisher ? fisher_f(x.z) : x.z);
        y.w = y.w + (fisher ? fisher_f(x.w) : x.w);
        a4[q] = y;
    }
    for (long long q = 4 * m4 + t0; q < m; q += stride) {
        const float x = r[q];
        acc[q] = acc[q] + (fisher ? fisher_f(x) : x);
    }
}

// z rows padded with zeros to [npad][tp]: exact zeros leave every FP64 sum
// unchanged and let the Gram load without bounds tests
__global__ void cm_pad(const double *__restrict__ z, int n, int t, int npad, int tp, double *__restrict__ zp)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)npad * tp) return;
    const int i = (int)(q / tp), l = (int)(q % tp);
    zp[q] = (i < n && l < t) ? z[(long long)i * t + l] : 0.0;
}

__global__ void cm_zero(long long m, float *__restrict__ a)
{
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (long long)gridDim.x * blockDim.x)
        a[q] = 0.0f;
}

// U CorMat.cpp:507-517
__global__ void cm_finish(long long m, float *__restrict__ a, int nsub, int fisher)
{
#pragma clang fp contract(off)
    const float s = (float)nsub;
    const long long m4 = m >> 2, stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float4 *a4 = (float4 *)a;
    for (long long q = t0; q < m4; q += stride) {  // 16-byte accesses, as cm_accumulate
        float4 v = a4[q];
        v.x = v.x / s;
        v.y = v.y / s;
        v.z = v.z / s;
        v.w = v.w / s;
        if (fisher) {
            v.x = inv_fisher_f(v.x);
            v.y = inv_fisher_f(v.y);
            v.z = inv_fisher_f(v.z);
            v.w = inv_fisher_f(v.w);
        }
        a4[q] = v;
    }
    for (long long q = 4 * m4 + t0; q < m; q += stride) {
        float x = a[q] / s;
        if (fisher) x = inv_fisher_f(x);
        a[q] = x;
    }
}

// ----------------------------------------------------- k-th largest value

__device__ __forceinline__ uint32_t fkey(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__host__ __device__ __forceinline__ float key_float(uint32_t k)
{
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct SelState {
    u64 k;            // rank still wanted inside the current prefix
    uint32_t prefix;  // key bits fixed so far
};

// pass 0, 1, 2 histogram key bits [21, 32), [10, 21), [0, 10)
__global__ __launch_bounds__(256) void sel_hist(const float *__restrict__ a, long long m, int pass,
                                                const SelState *__restrict__ ss, u64 *__restrict__ hist)
{
    __shared__ unsigned h[2048];
    for (int b = threadIdx.x; b < 2048; b += 256) h[b] = 0u;
    __syncthreads();
    const int sh = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
    const int hi = pass == 0 ? 32 : (pass == 1 ? 21 : 10);  // bits at and above hi are fixed
    const uint32_t want = hi >= 32 ? 0u : (ss->prefix >> hi);
    const uint32_t bmask = pass == 2 ? 1023u : 2047u;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < m;
         q += (long long)gridDim.x * blockDim.x) {
        const uint32_t k = fkey(a[q]);
        if (hi < 32 && (k >> hi) != want) continue;
        atomicAdd(&h[(k >> sh) & bmask], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 2048; b += 256)
        if (h[b]) atomicAdd(&hist[b], (u64)h[b]);
}

// one thread: the bin holding the k-th largest, scanning from the top
__global__ void sel_pick(int pass, SelState *__restrict__ ss, const u64 *__restrict__ hist)
{
    const int nb = pass == 2 ? 1024 : 2048;
    const int sh = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
    u64 k = ss->k, above = 0;
    int b = nb - 1;
    for (; b > 0; b--) {
        if (above + hist[b] >= k) break;
        above += hist[b];
    }
    ss->k = k - above;
    ss->prefix |= (uint32_t)b << sh;
}

// ------------------------------------------------------ CSR from the matrix

// Per packed row i (one work-group): values above the bound and equal to it.
__global__ __launch_bounds__(256) void cm_row_counts(const float *__restrict__ a, int n, float bound,
                                                     int *__restrict__ cgt, int *__restrict__ ctie)
{
    __shared__ int red[2][4];
    const int i = blockIdx.x;
    const long long base = pidx(i, i + 1, n);
    const int len = n - 1 - i;
    int g = 0, e = 0;
#pragma unroll 4
    for (int q = threadIdx.x; q < len; q += 256) {
        const float x = a[base + q];
        g += x > bound;
        e += x == bound;
    }
    for (int o = 32; o > 0; o >>= 1) {
        g += __shfl_xor(g, o);
        e += __shfl_xor(e, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = g;
        red[1][threadIdx.x >> 6] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        cgt[i] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        ctie[i] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// Exclusive scan of n ints into long long (one work-group), total in out[n].
__global__ __launch_bounds__(1024) void cm_scan(const int *__restrict__ in, int n, long long *__restrict__ out)
{
    pgdev::block_scan_1024(in, n, out, 1);
}

// Keep byte per pair in packed order; ties at the bound are kept while their
// row-major rank (tbase[i] + rank inside the row) is below ntie.  cu[i] = kept
// pairs of packed row i.
__global__ __launch_bounds__(256) void cm_keep(const float *__restrict__ a, int n, float bound, int ties,
                                               const long long *__restrict__ tbase, long long ntie,
                                               uint8_t *__restrict__ keep, int *__restrict__ cu)
{
    // one packed row per wave, 64 pairs per step and four steps in flight; tie
    // ranks advance in column order by ballot, with no block barriers
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wv;
    if (i >= n) return;
    const long long base = pidx(i, i + 1, n);
    const int len = n - 1 - i;
    long long rank = ties ? tbase[i] : 0;
    int nk = 0;
    for (int q0 = 0; q0 < len; q0 += 4 * 64) {
        float x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int q = q0 + 64 * u + lane;
            x[u] = q < len ? a[base + q] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int q = q0 + 64 * u + lane;
            const bool in = q < len;
            const bool gt = in && x[u] > bound, tie = in && ties && x[u] == bound;
            const u64 tm = __ballot(tie);
            const long long r = rank + __popcll(tm & ((1ull << lane) - 1ull));
            const bool k = gt || (tie && r < ntie);
            if (in) keep[base + q] = k ? 1 : 0;
            nk += __popcll(__ballot(k));
            rank += __popcll(tm);
        }
    }
    if (lane == 0) cu[i] = nk;
}

// One work-group per 64 CSR rows [i0, i0 + 64).  Count pass (rowptr == NULL):
// cl[i] = kept pairs (j, i) with j < i.  Fill pass: writes row i as its lower
// part (ascending j < i) followed by its upper part (ascending j > i).
__global__ __launch_bounds__(256) void cm_csr_group(const uint8_t *__restrict__ keep, const float *__restrict__ a,
                                                    int n, const long long *__restrict__ rowptr,
                                                    int *__restrict__ cl, int *__restrict__ col,
                                                    float *__restrict__ val)
{
    __shared__ int run[64];
    __shared__ int wcnt[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int i0 = blockIdx.x * 64;
    const int nr = min(64, n - i0);
    if (tid < 64) run[tid] = 0;
    __syncthreads();
    // lower parts: packed rows j = 0 .. i0 + nr - 2
    const int jend = i0 + nr - 1;
    for (int j0 = 0; j0 < jend; j0 += 256) {
        const int j = j0 + tid;
        u64 mask = 0;
        if (j < jend) {
            // rows r with i0 + r > j; their bytes are contiguous in packed row j:
            // fetched as at most five aligned 16-byte words (only those that
            // overlap [b, b + cnt), so no read passes b + cnt + 15), then
            // unpacked by compile-time byte position
            const int r0 = j >= i0 ? j - i0 + 1 : 0;
            const long long b = pidx(j, i0 + r0, n);
            const long long b16 = b & ~15ll;
            const int off = (int)(b - b16), cnt = nr - r0;
            uint4 wd[5];
#pragma unroll
            for (int q = 0; q < 5; q++)
                wd[q] = 16 * q < off + cnt ? *(const uint4 *)(keep + b16 + 16 * q) : make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const unsigned w4[4] = {wd[q].x, wd[q].y, wd[q].z, wd[q].w};
#pragma unroll
                for (int p = 0; p < 16; p++) {
                    const int r = 16 * q + p - off + r0;  // row of byte 16q + p
                    const bool set = ((w4[p >> 2] >> (8 * (p & 3))) & 0xffu) != 0u;
                    if (set && r >= r0 && r < nr) mask |= 1ull << r;
                }
            }
        }
        for (int r = 0; r < nr; r++) {
            const u64 bm = __ballot((mask >> r) & 1ull);
            if (lane == 0) wcnt[wv][r] = __popcll(bm);
        }
        __syncthreads();
        if (rowptr) {
            for (int r = 0; r < nr; r++) {
                const u64 bm = __ballot((mask >> r) & 1ull);
                if ((mask >> r) & 1ull) {
                    int pos = run[r] + __popcll(bm & ((1ull << lane) - 1ull));
                    for (int w = 0; w < wv; w++) pos += wcnt[w][r];
                    const long long at = rowptr[i0 + r] + pos;
                    col[at] = j;
                    if (val) val[at] = a[pidx(j, i0 + r, n)];
                }
            }
        }
        __syncthreads();
        if (tid < nr) run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        __syncthreads();
    }
    if (!rowptr) {
        if (tid < nr) cl[i0 + tid] = run[tid];
        return;
    }
    // upper parts: packed segment of row i, j = i+1 .. n-1, one row per wave
    // (no block barriers), 64 bytes per step and four steps in flight, each
    // kept column placed by its rank in the step's ballot
    for (int r = wv; r < nr; r += 4) {
        const int i = i0 + r;
        const long long b = pidx(i, i + 1, n);
        const int len = n - 1 - i;
        long long pos = rowptr[i] + cl[i];
        for (int q0 = 0; q0 < len; q0 += 4 * 64) {
            bool k[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + 64 * u + lane;
                k[u] = q < len && keep[b + q];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const u64 bm = __ballot(k[u]);
                if (k[u]) {
                    const int q = q0 + 64 * u + lane;
                    const long long at = pos + __popcll(bm & ((1ull << lane) - 1ull));
                    col[at] = i + 1 + q;
                    if (val) val[at] = a[b + q];
                }
                pos += __popcll(bm);
            }
        }
    }
}

__global__ void cm_degrees(int n, const int *__restrict__ cl, const int *__restrict__ cu, int *__restrict__ deg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) deg[i] = cl[i] + cu[i];
}

__global__ void cm_rowptr(int n, const long long *__restrict__ p, int *__restrict__ row)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) row[i] = (int)p[i];
}

pg_status need_matrix(pg_ctx *ctx)
{
    if (ctx->cm_n < 0 || !ctx->cm_acc) return pg_fail(ctx, PG_ESTATE, "no pg_cormat_begin");
    return PG_OK;
}

}  // namespace

extern "C" pg_status pg_cormat_begin(pg_ctx *ctx, int32_t n, int32_t fisher)
{
    if (!ctx) return PG_EINVAL;
    if (n < 2) return pg_fail(ctx, PG_EINVAL, "n must be >= 2");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const long long m1 = (long long)n * (n - 1) / 2;
    PG_TRY(pg_bufT(ctx, "cm.acc", (size_t)m1, &ctx->cm_acc));
    cm_zero<<<(unsigned)std::min<long long>(pg_cdiv(m1, 256), 65536), 256, 0, ctx->stream>>>(m1, ctx->cm_acc);
    PG_KERNEL_CHECK(ctx);
    ctx->cm_n = n;
    ctx->cm_nsub = 0;
    ctx->cm_fisher = fisher ? 1 : 0;
    ctx->cm_finished = false;
    return PG_OK;
}

extern "C" pg_status pg_cormat_add(pg_ctx *ctx, const float *bold, int32_t t, int32_t layout)
{
    if (!ctx) return PG_EINVAL;
    PG_TRY(need_matrix(ctx));
    if (ctx->cm_finished) return pg_fail(ctx, PG_ESTATE, "matrix already finished");
    if (!bold || t < 2 || (layout != PG_BOLD_TIME_MAJOR && layout != PG_BOLD_VOXEL_MAJOR))
        return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const int n = ctx->cm_n;
    hipStream_t st = ctx->stream;
    const float *dbold;
    double *z64;
    PG_TRY(pg_in(ctx, bold, (size_t)n * t, "cm.bold", &dbold));
    PG_TRY(pg_bufT(ctx, "cm.z64", (size_t)n * t + 1, &z64));
    PG_TRY(pg_zscore64(ctx, dbold, layout, n, t, z64));
    const int nt = (int)pg_cdiv(n, GT);
    const int npad = nt * GT, tp = (int)pg_cdiv(t, GKC) * GKC;
    double *zp;
    PG_TRY(pg_bufT(ctx, "cm.zp", (size_t)npad * tp, &zp));
    cm_pad<<<(unsigned)pg_cdiv((long long)npad * tp, 256), 256, 0, st>>>(z64, n, t, npad, tp, zp);
    const double delta = 2.5 * (double)t * std::ldexp(1.0, -53);
    const size_t shm = (size_t)GLDS * sizeof(double);
    PG_HIP(ctx, hipFuncSetAttribute((const void *)gram64_r, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    // ambiguous-rounding list: about one pair in 10^4 is expected
    const long long m1 = (long long)n * (n - 1) / 2;
    unsigned cap = (unsigned)std::min<long long>(std::max<long long>(m1 / 256, 1 << 16), 1 << 26);
    if (const char *e = std::getenv("PAGANI_CM_AMBCAP")) cap = (unsigned)std::max(0, std::atoi(e));
    int2 *amb;
    unsigned *namb;
    PG_TRY(pg_bufT(ctx, "cm.amb", std::max(cap, 1u), &amb));
    PG_TRY(pg_bufT(ctx, "cm.namb", 1, &namb));
    PG_HIP(ctx, hipMemsetAsync(namb, 0, 4, st));
    // the first subject without Fisher is its own sum (0 + r = r): written in
    // place; otherwise into the subject buffer, then accumulated
    const bool direct = ctx->cm_nsub == 0 && !ctx->cm_fisher;
    float *r = ctx->cm_acc;
    if (!direct) PG_TRY(pg_bufT(ctx, "cm.r", (size_t)m1, &r));
    gram64_r<<<dim3((unsigned)nt, (unsigned)nt), GTHREADS, shm, st>>>(zp, n, t, tp, delta, r, amb, namb, cap);
    if (cap > 0)
        gram64_recheck<<<(unsigned)pg_cdiv(cap, 256), 256, 0, st>>>(zp, n, t, tp, r, amb, namb, cap);
    if (!direct)
        cm_accumulate<<<(unsigned)std::min<long long>(pg_cdiv(m1, 256), 65536), 256, 0, st>>>(
            m1, r, ctx->cm_fisher, ctx->cm_acc);
    PG_KERNEL_CHECK(ctx);
    PG_HIP(ctx, hipStreamSynchronize(st));
    ctx->cm_nsub++;
    return PG_OK;
}

extern "C" pg_status pg_cormat_finish(pg_ctx *ctx)
{
    if (!ctx) return PG_EINVAL;
    PG_TRY(need_matrix(ctx));
    if (ctx->cm_finished) return PG_OK;
    if (ctx->cm_nsub < 1) return pg_fail(ctx, PG_ESTATE, "no subject added");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const long long m1 = (long long)ctx->cm_n * (ctx->cm_n - 1) / 2;
    cm_finish<<<(unsigned)std::min<long long>(pg_cdiv(m1, 256), 65536), 256, 0, ctx->stream>>>(
        m1, ctx->cm_acc, ctx->cm_nsub, ctx->cm_fisher);
    PG_KERNEL_CHECK(ctx);
    PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->cm_finished = true;
    return PG_OK;
}

extern "C" pg_status pg_cormat_read(pg_ctx *ctx, int64_t offset, int64_t count, float *dst)
{
    if (!ctx) return PG_EINVAL;
    PG_TRY(need_matrix(ctx));
    const long long m1 = (long long)ctx->cm_n * (ctx->cm_n - 1) / 2;
    if (!dst || offset < 0 || count < 0 || offset + count > m1) return pg_fail(ctx, PG_EINVAL, "bad range");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const hipMemcpyKind kind = pg_is_device_ptr(dst) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    PG_HIP(ctx, hipMemcpyAsync(dst, ctx->cm_acc + offset, (size_t)count * 4, kind, ctx->stream));
    PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PG_OK;
}

extern "C" pg_status pg_cormat_threshold(pg_ctx *ctx, int32_t thr_kind, float thr, int32_t weighted,
                                         pg_corr_info *info)
{
    if (!ctx) return PG_EINVAL;
    PG_TRY(need_matrix(ctx));
    if (!ctx->cm_finished) return pg_fail(ctx, PG_ESTATE, "pg_cormat_finish first");
    if (thr_kind != PG_THR_R && thr_kind != PG_THR_SPARSITY) return pg_fail(ctx, PG_EINVAL, "bad threshold kind");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int n = ctx->cm_n;
    const long long m1 = (long long)n * (n - 1) / 2;
    const float *a = ctx->cm_acc;
    pg_corr_info inf;
    std::memset(&inf, 0, sizeof inf);
    inf.n = n;
    float bound = thr;
    long long kwant = 0, ntie = 0;
    int *cgt, *ctie, *cu, *cl, *deg;
    long long *tbase, *rp;
    PG_TRY(pg_bufT(ctx, "cm.cgt", (size_t)n, &cgt));
    PG_TRY(pg_bufT(ctx, "cm.ctie", (size_t)n, &ctie));
    PG_TRY(pg_bufT(ctx, "cm.cu", (size_t)n, &cu));
    PG_TRY(pg_bufT(ctx, "cm.cl", (size_t)n, &cl));
    PG_TRY(pg_bufT(ctx, "cm.deg", (size_t)n, &deg));
    PG_TRY(pg_bufT(ctx, "cm.tbase", (size_t)n + 1, &tbase));
    PG_TRY(pg_bufT(ctx, "cm.rp", (size_t)n + 1, &rp));
    if (thr_kind == PG_THR_SPARSITY) {
        // U CorMat.cpp:311-313 (the same arithmetic as the per-subject path)
        const float prod = (float)(2 * m1) * thr;
        long long clen = (long long)((double)prod / 100.0);
        clen += clen % 2;
        kwant = clen / 2;
        if (kwant < 1 || kwant > m1) return pg_fail(ctx, PG_ETHRESH, "sparsity %g%% selects %lld pairs", thr, kwant);
        SelState *ss;
        u64 *hist;
        PG_TRY(pg_bufT(ctx, "cm.sel", 1, &ss));
        PG_TRY(pg_bufT(ctx, "cm.hist", 2048, &hist));
        SelState h0{(u64)kwant, 0u};
        PG_HIP(ctx, hipMemcpyAsync(ss, &h0, sizeof h0, hipMemcpyHostToDevice, st));
        const unsigned g = (unsigned)std::min<long long>(pg_cdiv(m1, 256 * 16), (long long)ctx->num_cu * 16);
        for (int pass = 0; pass < 3; pass++) {
            PG_HIP(ctx, hipMemsetAsync(hist, 0, 2048 * 8, st));
            sel_hist<<<std::max(g, 1u), 256, 0, st>>>(a, m1, pass, ss, hist);
            sel_pick<<<1, 1, 0, st>>>(pass, ss, hist);
        }
        PG_KERNEL_CHECK(ctx);
        SelState hs;
        PG_HIP(ctx, hipMemcpyAsync(&hs, ss, sizeof hs, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        bound = key_float(hs.prefix);
        cm_row_counts<<<(unsigned)n, 256, 0, st>>>(a, n, bound, cgt, ctie);
        cm_scan<<<1, 1024, 0, st>>>(cgt, n, rp);
        cm_scan<<<1, 1024, 0, st>>>(ctie, n, tbase);
        PG_KERNEL_CHECK(ctx);
        long long totgt = 0, tottie = 0;
        PG_HIP(ctx, hipMemcpyAsync(&totgt, rp + n, 8, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipMemcpyAsync(&tottie, tbase + n, 8, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        ntie = kwant - totgt;
        if (ntie < 1 || ntie > tottie)
            return pg_fail(ctx, PG_ETHRESH, "selection mismatch: %lld above, %lld ties, %lld wanted", totgt, tottie,
                           kwant);
        inf.s_thresh = thr;
        inf.r_thresh = bound;
        inf.t_star = (double)bound;
    } else {
        inf.r_thresh = thr;
        inf.t_star = (double)thr;
    }
    uint8_t *keep;
    PG_TRY(pg_bufT(ctx, "cm.keep", (size_t)m1, &keep));
    cm_keep<<<(unsigned)pg_cdiv(n, 4), 256, 0, st>>>(a, n, bound, thr_kind == PG_THR_SPARSITY, tbase, ntie, keep,
                                                      cu);
    const int ng = (int)pg_cdiv(n, 64);
    cm_csr_group<<<(unsigned)ng, 256, 0, st>>>(keep, a, n, nullptr, cl, nullptr, nullptr);
    cm_degrees<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(n, cl, cu, deg);
    cm_scan<<<1, 1024, 0, st>>>(deg, n, rp);
    PG_KERNEL_CHECK(ctx);
    long long nnz = 0;
    PG_HIP(ctx, hipMemcpyAsync(&nnz, rp + n, 8, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (nnz > INT32_MAX) return pg_fail(ctx, PG_ENOTSUP, "nnz %lld exceeds int32 CSR", nnz);
    if (thr_kind == PG_THR_SPARSITY && nnz != 2 * kwant)
        return pg_fail(ctx, PG_ETHRESH, "CSR holds %lld entries, %lld wanted", nnz, 2 * kwant);
    int *row, *col;
    float *val = nullptr;
    PG_TRY(pg_bufT(ctx, "cm.row", (size_t)n + 1, &row));
    PG_TRY(pg_bufT(ctx, "cm.col", (size_t)std::max<long long>(nnz, 1), &col));
    if (weighted) PG_TRY(pg_bufT(ctx, "cm.val", (size_t)std::max<long long>(nnz, 1), &val));
    cm_rowptr<<<(unsigned)pg_cdiv(n + 1, 256), 256, 0, st>>>(n, rp, row);
    cm_csr_group<<<(unsigned)ng, 256, 0, st>>>(keep, a, n, rp, cl, col, val);
    PG_KERNEL_CHECK(ctx);
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (thr_kind == PG_THR_R) inf.s_thresh = (float)(100.0 * nnz / m1 / 2);
    inf.nnz = nnz;
    ctx->cn = n;
    ctx->ct = 0;
    ctx->cnnz = nnz;
    ctx->cweighted = weighted != 0;
    ctx->c_row = row;
    ctx->c_col = col;
    ctx->c_val = val;
    ctx->c_z64 = nullptr;
    if (info) *info = inf;
    return PG_OK;
}
