// corr.hip -- voxel series -> thresholded CSR adjacency on MI355X (gfx950).
//
// Replaces, in one device pipeline, the reference's CorMat_gpu (host z-score +
// per-block cuBLAS Sgemm + 37.7 MB D2H per block, U/CorMat/CUCorMat/CUCORMAT/
// CorMat.cu:18-160), post_block (U CorMat.cpp:775-838), FormFullGraph /
// FormFullGraph_s (U CorMat.cpp:633-773, W CorMat.cpp:647-667) and
// FormCSRGraph (U CorMat.cpp:839-857, W CorMat.cpp:853-885).  The N x N
// correlation matrix is never stored: each 128 x 128 Gram tile is thresholded
// in registers and its edges compacted (wavefront ballot + mbcnt prefix) into
// the work-group's private edge segment.
//
// Exactness: the edge rule is evaluated on the FP64 correlation r64 (FP64
// z-scores, sequential sum, no FMA).  The MFMA pass yields r~ with a proven
// bound |r~ - r64| <= band; only pairs with |r~ - thr| <= band are recomputed in
// FP64, so the edge set equals the FP64 one exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pg_internal.hpp"

namespace {

constexpr int TILE = 128;      // output tile edge (rows and columns)
constexpr int KC = 32;         // K (time) chunk staged in LDS
constexpr int THREADS = 256;   // 4 wave64s, each owns a 64 x 64 quarter
constexpr int NBIN = 8192;     // r histogram bins over [-1, 1] (sparsity mode)
constexpr int LDB = KC + 8;    // bf16 LDS row stride (elements), 16-byte aligned rows
constexpr int LDF = KC + 4;    // fp32 LDS row stride (elements)

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

enum { PREC_F32 = 0, PREC_BF16X3 = 1 };
enum { MODE_EMIT = 0, MODE_HIST = 1 };

struct GramArgs {
    const void *p0;        // fp32 panel, or bf16 hi panel   [npad][ldp]
    const void *p1;        // bf16 lo panel (unused for fp32)
    int ldp;               // panel row stride = padded T
    int kchunks;           // ldp / KC
    int n;                 // voxels
    int nb;                // row blocks = npad / TILE
    long long ntiles;      // nb (nb + 1) / 2 upper-triangle tiles
    long long per;         // tiles per work-group
    float upper;           // r~ > upper: edge without recheck
    float lower;           // lower <= r~ <= upper: FP64 recheck
    int2 *seg;             // [gridDim.x][seg_cap] private edge segments
    long long seg_cap;
    unsigned long long *seg_cnt;   // [gridDim.x]
    int2 *band;            // recheck list
    float *band_r;         // r~ of each recheck entry
    long long band_cap;
    unsigned long long *band_cnt;  // [1]
    uint32_t *hist;        // [gridDim.x][NBIN]
};

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------- z-score
// U CorMat.cu:34-53: mean and sqrt(sum (x-mu)^2) in FP64 with each product
// rounded (no FMA), z = (x-mu)/norm.  Zero-variance voxels get z = 0 (the
// reference's NaN r is zeroed by post_block, U CorMat.cpp:834).
__global__ void zscore_stats(const float *__restrict__ x, int layout, int n, int t,
                             double *__restrict__ mu, double *__restrict__ sd)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s1 = 0.0;
    if (layout == 0)
        for (int l = 0; l < t; l++) s1 += (double)x[(long long)l * n + i];
    else
        for (int l = 0; l < t; l++) s1 += (double)x[(long long)i * t + l];
    s1 /= t;
    double s2 = 0.0;
    for (int l = 0; l < t; l++) {
        float xv = layout == 0 ? x[(long long)l * n + i] : x[(long long)i * t + l];
        double d = (double)xv - s1;
        double p = d * d;
        s2 += p;
    }
    mu[i] = s1;
    sd[i] = sqrt(s2);
}

// 64 voxels x 64 time points per block, transposed through LDS so that both
// the input and the voxel-major outputs are coalesced.
__global__ void zscore_write(const float *__restrict__ x, int layout, int n, int t,
                             const double *__restrict__ mu, const double *__restrict__ sd,
                             double *__restrict__ z64, float *__restrict__ p32,
                             uint16_t *__restrict__ phi, uint16_t *__restrict__ plo, int ldp)
{
#pragma clang fp contract(off)
    __shared__ float tile[64][65];
    const int i0 = blockIdx.x * 64, l0 = blockIdx.y * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        int li, ii;
        if (layout == 0) { li = e >> 6; ii = e & 63; }
        else { ii = e >> 6; li = e & 63; }
        int i = i0 + ii, l = l0 + li;
        float v = 0.0f;
        if (i < n && l < t) v = layout == 0 ? x[(long long)l * n + i] : x[(long long)i * t + l];
        tile[li][ii] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        int ii = e >> 6, li = e & 63;
        int i = i0 + ii, l = l0 + li;
        if (i >= n || l >= t) continue;
        double s = sd[i];
        double z = s > 0.0 ? ((double)tile[li][ii] - mu[i]) / s : 0.0;
        z64[(long long)i * t + l] = z;
        float zf = (float)z;
        long long q = (long long)i * ldp + l;
        if (p32) p32[q] = zf;
        if (phi) {
            uint16_t h = f2bf(zf);
            phi[q] = h;
            plo[q] = f2bf(zf - bf2f(h));
        }
    }
}

// ------------------------------------------------------------ Gram tiles

// Upper-triangle tile index -> (bi, bj), bi <= bj, row-major over bi.
__device__ __forceinline__ void tile_coords(long long tid, int nb, int &bi, int &bj)
{
    // start(bi) = bi*nb - bi*(bi-1)/2; solve start(bi) <= tid < start(bi+1)
    double b = 2.0 * nb + 1.0;
    int r = (int)floor((b - sqrt(b * b - 8.0 * (double)tid)) * 0.5);
    if (r < 0) r = 0;
    if (r > nb - 1) r = nb - 1;
    while (r > 0 && (long long)r * nb - (long long)r * (r - 1) / 2 > tid) r--;
    while (r < nb - 1 && (long long)(r + 1) * nb - (long long)(r + 1) * r / 2 <= tid) r++;
    bi = r;
    bj = r + (int)(tid - ((long long)r * nb - (long long)r * (r - 1) / 2));
}

template <int PREC>
__device__ __forceinline__ void load_chunk(const GramArgs &a, char *smem, int bi, int bj, int kc)
{
    const int tid = threadIdx.x;
    if (PREC == PREC_BF16X3) {
        uint16_t *sAh = (uint16_t *)smem;
        uint16_t *sAl = sAh + TILE * LDB;
        uint16_t *sBh = sAl + TILE * LDB;
        uint16_t *sBl = sBh + TILE * LDB;
        const uint16_t *gh = (const uint16_t *)a.p0;
        const uint16_t *gl = (const uint16_t *)a.p1;
        constexpr int PIECES = TILE * KC / 8;  // 16-byte pieces per panel chunk
#pragma unroll
        for (int it = 0; it < PIECES / THREADS; it++) {
            int id = it * THREADS + tid;
            int r = id / (KC / 8), c = (id % (KC / 8)) * 8;
            long long ga = (long long)(bi * TILE + r) * a.ldp + kc * KC + c;
            long long gb = (long long)(bj * TILE + r) * a.ldp + kc * KC + c;
            u32x4 vah = *(const u32x4 *)(gh + ga);
            u32x4 val = *(const u32x4 *)(gl + ga);
            u32x4 vbh = *(const u32x4 *)(gh + gb);
            u32x4 vbl = *(const u32x4 *)(gl + gb);
            *(u32x4 *)(sAh + r * LDB + c) = vah;
            *(u32x4 *)(sAl + r * LDB + c) = val;
            *(u32x4 *)(sBh + r * LDB + c) = vbh;
            *(u32x4 *)(sBl + r * LDB + c) = vbl;
        }
    } else {
        float *sA = (float *)smem;
        float *sB = sA + TILE * LDF;
        const float *g = (const float *)a.p0;
        constexpr int PIECES = TILE * KC / 4;
#pragma unroll
        for (int it = 0; it < PIECES / THREADS; it++) {
            int id = it * THREADS + tid;
            int r = id / (KC / 4), c = (id % (KC / 4)) * 4;
            long long ga = (long long)(bi * TILE + r) * a.ldp + kc * KC + c;
            long long gb = (long long)(bj * TILE + r) * a.ldp + kc * KC + c;
            u32x4 va = *(const u32x4 *)(g + ga);
            u32x4 vb = *(const u32x4 *)(g + gb);
            *(u32x4 *)(sA + r * LDF + c) = va;
            *(u32x4 *)(sB + r * LDF + c) = vb;
        }
    }
}

template <int PREC>
__device__ __forceinline__ void mfma_chunk(const char *smem, f32x16 (&acc)[2][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w >> 1, wn = w & 1, lr = lane & 31;
    if (PREC == PREC_BF16X3) {
        const uint16_t *sAh = (const uint16_t *)smem;
        const uint16_t *sAl = sAh + TILE * LDB;
        const uint16_t *sBh = sAl + TILE * LDB;
        const uint16_t *sBl = sBh + TILE * LDB;
        const int lk = (lane >> 5) * 8;
#pragma unroll
        for (int ks = 0; ks < KC / 16; ks++) {
            bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                int ra = (wm * 64 + b * 32 + lr) * LDB + ks * 16 + lk;
                int rb = (wn * 64 + b * 32 + lr) * LDB + ks * 16 + lk;
                ah[b] = *(const bf16x8 *)(sAh + ra);
                al[b] = *(const bf16x8 *)(sAl + ra);
                bh[b] = *(const bf16x8 *)(sBh + rb);
                bl[b] = *(const bf16x8 *)(sBl + rb);
            }
#pragma unroll
            for (int bm = 0; bm < 2; bm++)
#pragma unroll
                for (int bn = 0; bn < 2; bn++) {
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[bm], bh[bn], acc[bm][bn], 0, 0, 0);
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bm], bl[bn], acc[bm][bn], 0, 0, 0);
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bm], bh[bn], acc[bm][bn], 0, 0, 0);
                }
        }
    } else {
        const float *sA = (const float *)smem;
        const float *sB = sA + TILE * LDF;
        const int lk = lane >> 5;
#pragma unroll
        for (int ks = 0; ks < KC / 2; ks++) {
            float av[2], bv[2];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                av[b] = sA[(wm * 64 + b * 32 + lr) * LDF + ks * 2 + lk];
                bv[b] = sB[(wn * 64 + b * 32 + lr) * LDF + ks * 2 + lk];
            }
#pragma unroll
            for (int bm = 0; bm < 2; bm++)
#pragma unroll
                for (int bn = 0; bn < 2; bn++)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[bm], bv[bn], acc[bm][bn], 0, 0, 0);
        }
    }
}

__device__ __forceinline__ unsigned lane_rank(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Persistent over a contiguous range of upper-triangle tiles (neighbouring
// tiles share the row panel bi, which stays in the XCD's L2).
template <int PREC, int MODE>
__global__ __launch_bounds__(THREADS, 2) void gram_kernel(GramArgs a)
{
    extern __shared__ __align__(16) char smem[];
    constexpr int STAGE_BYTES = PREC == PREC_BF16X3 ? 4 * TILE * LDB * 2 : 2 * TILE * LDF * 4;
    uint32_t *s_hist = (uint32_t *)(smem + STAGE_BYTES);
    __shared__ unsigned long long s_seg;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    if (tid == 0) s_seg = 0;
    if (MODE == MODE_HIST)
        for (int b = tid; b < NBIN; b += THREADS) s_hist[b] = 0;

    const long long t0 = (long long)blockIdx.x * a.per;
    const long long t1 = min(a.ntiles, t0 + a.per);
    int bi = 0, bj = 0;
    if (t0 < t1) tile_coords(t0, a.nb, bi, bj);
    int2 *myseg = a.seg + (long long)blockIdx.x * a.seg_cap;

    for (long long tt = t0; tt < t1; tt++) {
        f32x16 acc[2][2];
#pragma unroll
        for (int bm = 0; bm < 2; bm++)
#pragma unroll
            for (int bn = 0; bn < 2; bn++)
#pragma unroll
                for (int q = 0; q < 16; q++) acc[bm][bn][q] = 0.0f;

        for (int kc = 0; kc < a.kchunks; kc++) {
            __syncthreads();
            load_chunk<PREC>(a, smem, bi, bj, kc);
            __syncthreads();
            mfma_chunk<PREC>(smem, acc);
        }

        // epilogue: 32x32 block layout, lane holds column lane%32 and rows
        // (q/4)*8 + (lane/32)*4 + q%4 of each block.
#pragma unroll
        for (int bm = 0; bm < 2; bm++)
#pragma unroll
            for (int bn = 0; bn < 2; bn++)
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const float v = acc[bm][bn][q];
                    const int i = bi * TILE + wm * 64 + bm * 32 + (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
                    const int j = bj * TILE + wn * 64 + bn * 32 + (lane & 31);
                    const bool valid = (j < a.n) && (i < j);
                    if (MODE == MODE_HIST) {
                        if (valid) {
                            int b = (int)((v + 1.0f) * (NBIN * 0.5f));
                            b = b < 0 ? 0 : (b >= NBIN ? NBIN - 1 : b);
                            atomicAdd(&s_hist[b], 1u);
                        }
                    } else {
                        const bool e = valid && (v > a.upper);
                        const bool c = valid && !e && (v >= a.lower);
                        const unsigned long long me = __ballot(e);
                        if (me) {
                            unsigned long long base = 0;
                            if (lane == 0) base = atomicAdd(&s_seg, (unsigned long long)__popcll(me));
                            base = __shfl(base, 0);
                            if (e) {
                                unsigned long long p = base + lane_rank(me);
                                if ((long long)p < a.seg_cap) myseg[p] = make_int2(i, j);
                            }
                        }
                        const unsigned long long mc = __ballot(c);
                        if (mc) {
                            unsigned long long base = 0;
                            if (lane == 0) base = atomicAdd(a.band_cnt, (unsigned long long)__popcll(mc));
                            base = __shfl(base, 0);
                            if (c) {
                                unsigned long long p = base + lane_rank(mc);
                                if ((long long)p < a.band_cap) {
                                    a.band[p] = make_int2(i, j);
                                    a.band_r[p] = v;
                                }
                            }
                        }
                    }
                }
        // next tile in row-major upper-triangle order
        if (++bj == a.nb) {
            bi++;
            bj = bi;
        }
    }
    __syncthreads();
    if (MODE == MODE_HIST) {
        uint32_t *out = a.hist + (long long)blockIdx.x * NBIN;
        for (int b = tid; b < NBIN; b += THREADS) out[b] = s_hist[b];
    } else if (tid == 0) {
        a.seg_cnt[blockIdx.x] = s_seg;
    }
}

// Sum the per-work-group histograms.
__global__ void hist_reduce(const uint32_t *__restrict__ h, int nwg, unsigned long long *__restrict__ out)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= NBIN) return;
    unsigned long long s = 0;
    for (int g = 0; g < nwg; g++) s += h[(long long)g * NBIN + b];
    out[b] = s;
}

// ---------------------------------------------------------- FP64 recheck

// r64 of one pair exactly as the FP64 oracle forms it: sequential sum over t of
// rounded products, clamped to [-1, 1].
__device__ __forceinline__ double r64_pair(const double *__restrict__ z, int t, int i, int j)
{
#pragma clang fp contract(off)
    const double *a = z + (long long)i * t;
    const double *b = z + (long long)j * t;
    double s = 0.0;
    for (int l = 0; l < t; l++) {
        double p = a[l] * b[l];
        s += p;
    }
    return s > 1.0 ? 1.0 : (s < -1.0 ? -1.0 : s);
}

__global__ void recheck_kernel(const int2 *__restrict__ pairs, const unsigned long long *__restrict__ cnt,
                               long long cap, const double *__restrict__ z, int t,
                               double *__restrict__ r_out)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long c = (long long)min((unsigned long long)cap, *cnt);
    if (k >= c) return;
    int2 p = pairs[k];
    r_out[k] = r64_pair(z, t, p.x, p.y);
}

// r-mode: keep rechecked pairs with r64 > thr; count decisions that differed
// from the fast test r~ > thr.
__global__ void accept_kernel(const int2 *__restrict__ pairs, const float *__restrict__ rfast,
                              const double *__restrict__ r64, const unsigned long long *__restrict__ cnt,
                              long long cap, double thr, int2 *__restrict__ extra,
                              unsigned long long *__restrict__ extra_cnt,
                              unsigned long long *__restrict__ flips)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long c = (long long)min((unsigned long long)cap, *cnt);
    if (k >= c) return;
    bool keep = r64[k] > thr;
    bool fast = (double)rfast[k] > thr;
    if (keep != fast) atomicAdd(flips, 1ull);
    if (keep) {
        unsigned long long p = atomicAdd(extra_cnt, 1ull);
        extra[p] = pairs[k];
    }
}

// ------------------------------------------------------------- CSR build

struct EdgeSrc {
    const int2 *seg;
    long long seg_cap;
    const unsigned long long *seg_cnt;
    int nseg;
    const int2 *extra;
    const unsigned long long *extra_cnt;
};

__device__ __forceinline__ long long src_count(const EdgeSrc &s, int g)
{
    return g < s.nseg ? (long long)s.seg_cnt[g] : (long long)*s.extra_cnt;
}
__device__ __forceinline__ int2 src_edge(const EdgeSrc &s, int g, long long k)
{
    return g < s.nseg ? s.seg[(long long)g * s.seg_cap + k] : s.extra[k];
}

// One block per source segment (the last block takes the accepted rechecks).
__global__ void degree_count(EdgeSrc s, int *__restrict__ deg)
{
    const int g = blockIdx.x;
    const long long c = src_count(s, g);
    for (long long k = threadIdx.x; k < c; k += blockDim.x) {
        int2 e = src_edge(s, g, k);
        atomicAdd(&deg[e.x], 1);
        atomicAdd(&deg[e.y], 1);
    }
}

// Exclusive scan of deg[0..n) into row[0..n], single block of 1024 threads.
__global__ __launch_bounds__(1024) void scan_rows(const int *__restrict__ deg, int n,
                                                   int *__restrict__ row, int *__restrict__ cur)
{
    __shared__ long long part[1024];
    const int tid = threadIdx.x;
    const long long per = (n + 1023) / 1024;
    const long long a = tid * per, b = min((long long)n, a + per);
    long long s = 0;
    for (long long i = a; i < b; i++) s += deg[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    long long run = part[tid] - s;
    for (long long i = a; i < b; i++) {
        row[i] = (int)run;
        cur[i] = 0;
        run += deg[i];
    }
    if (tid == 1023) row[n] = (int)part[1023];
}

__global__ void scatter_edges(EdgeSrc s, const int *__restrict__ row, int *__restrict__ cur,
                              int *__restrict__ ctmp, float *__restrict__ vtmp,
                              const double *__restrict__ z, int t)
{
    const int g = blockIdx.x;
    const long long c = src_count(s, g);
    for (long long k = threadIdx.x; k < c; k += blockDim.x) {
        int2 e = src_edge(s, g, k);
        int p = atomicAdd(&cur[e.x], 1);
        int q = atomicAdd(&cur[e.y], 1);
        ctmp[row[e.x] + p] = e.y;
        ctmp[row[e.y] + q] = e.x;
        if (vtmp) {
            float r = (float)r64_pair(z, t, e.x, e.y);
            vtmp[row[e.x] + p] = r;
            vtmp[row[e.y] + q] = r;
        }
    }
}

// Sort each row's columns ascending: mark columns in an LDS bitmap over
// [0, n), prefix-count the words, then place each entry at its rank.
__global__ __launch_bounds__(256) void sort_rows(const int *__restrict__ row, const int *__restrict__ ctmp,
                                                 const float *__restrict__ vtmp, int n,
                                                 int *__restrict__ col, float *__restrict__ val)
{
    extern __shared__ uint32_t sm[];
    __shared__ int s_part[256];
    const int nw = (n + 31) >> 5;
    uint32_t *bits = sm;
    uint32_t *pre = sm + nw;
    const int r = blockIdx.x, tid = threadIdx.x;
    const int a = row[r], b = row[r + 1];
    const int d = b - a;
    if (d == 0) return;
    if (d == 1) {
        if (tid == 0) {
            col[a] = ctmp[a];
            if (val) val[a] = vtmp[a];
        }
        return;
    }
    for (int q = tid; q < nw; q += 256) bits[q] = 0u;
    __syncthreads();
    for (int e = a + tid; e < b; e += 256) {
        int c = ctmp[e];
        atomicOr(&bits[c >> 5], 1u << (c & 31));
    }
    __syncthreads();
    const int per = (nw + 255) / 256;
    const int w0 = tid * per, w1 = min(nw, w0 + per);
    int s = 0;
    for (int q = w0; q < w1; q++) s += __popc(bits[q]);
    s_part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int v = tid >= off ? s_part[tid - off] : 0;
        __syncthreads();
        s_part[tid] += v;
        __syncthreads();
    }
    int run = s_part[tid] - s;
    for (int q = w0; q < w1; q++) {
        pre[q] = run;
        run += __popc(bits[q]);
    }
    __syncthreads();
    for (int e = a + tid; e < b; e += 256) {
        int c = ctmp[e];
        int pos = pre[c >> 5] + __popc(bits[c >> 5] & ((1u << (c & 31)) - 1u));
        col[a + pos] = c;
        if (val) val[a + pos] = vtmp[e];
    }
}

// ------------------------------------------------------------------ host

// Rigorous |r~ - r64| bound for unit-norm FP64 rows of length t (u = 2^-24):
//  z32 rounding of z64: (2u + u^2) sum|z z| <= 2u + u^2;
//  FP32: every product and add rounded once in any order: gamma_{t+1};
//  BF16x3: z32 = hi + lo + e, |lo| <= 2^-8 |z|, |e| <= 2^-16 |z|; the dropped
//  lo.lo and e terms cost <= 3 * 2^-16, and 3t exact bf16 products summed in
//  FP32 inside MFMA (at most one rounding per term plus per instruction):
//  gamma_{3.2t + 16}.  The FP64 side (t products, t adds, 2^-53) is added too.
//  A factor 1.5 covers the rest.
double error_band(int prec, int t)
{
    const double u = std::ldexp(1.0, -24);
    double z32 = 2.0 * u + u * u;
    double acc;
    if (prec == PREC_F32) {
        double m = (double)(t + 1) * u;
        acc = m / (1.0 - m);
    } else {
        double m = (3.2 * t + 16.0) * u;
        acc = 3.0 * std::ldexp(1.0, -16) + m / (1.0 - m);
    }
    double f64 = 2.0 * t * std::ldexp(1.0, -53);
    return 1.5 * (z32 + acc + f64);
}

float round_up_f(double x)
{
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}
float round_down_f(double x)
{
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}

struct HostCand {
    double r;
    int i, j;
};

}  // namespace

extern "C" pg_status pg_corr_build(pg_ctx *ctx, const float *bold, int32_t n, int32_t t,
                                   const pg_corr_opts *opts, pg_corr_info *info)
{
    if (!ctx) return PG_EINVAL;
    if (!bold || !opts || n < 0 || t < 2) return pg_fail(ctx, PG_EINVAL, "bad arguments (n=%d t=%d)", n, t);
    if (opts->prec != PG_PREC_FP32 && opts->prec != PG_PREC_BF16X3)
        return pg_fail(ctx, PG_EINVAL, "unknown precision %d", opts->prec);
    if (opts->layout != PG_BOLD_TIME_MAJOR && opts->layout != PG_BOLD_VOXEL_MAJOR)
        return pg_fail(ctx, PG_EINVAL, "unknown layout %d", opts->layout);
    const bool smode = opts->thr_kind == PG_THR_SPARSITY;
    if (!smode && opts->thr_kind != PG_THR_R) return pg_fail(ctx, PG_EINVAL, "unknown threshold kind");
    if (smode && !(opts->thr > 0.0f && opts->thr <= 100.0f))
        return pg_fail(ctx, PG_ETHRESH, "sparsity %g%% outside (0, 100]", (double)opts->thr);
    if (!smode && !(opts->thr >= -1.0f && opts->thr <= 1.0f))
        return pg_fail(ctx, PG_ETHRESH, "r threshold %g outside [-1, 1]", (double)opts->thr);
    const int nw = (n + 31) / 32;
    if ((size_t)nw * 8 + 64 > ctx->lds_per_block)
        return pg_fail(ctx, PG_ENOTSUP, "n=%d voxels exceeds the row-sort LDS bitmap", n);
    PG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int prec = opts->prec == PG_PREC_BF16X3 ? PREC_BF16X3 : PREC_F32;
    const bool weighted = opts->weighted != 0;

    pg_corr_info inf;
    std::memset(&inf, 0, sizeof inf);
    inf.n = n;
    inf.t = t;
    const long long m1 = (long long)n * (n - 1) / 2;

    // threshold bookkeeping, as the reference names its outputs
    // (U CorMat.cpp:306-313,333; W CorMat.cpp:310)
    long long kwant = 0;
    float thr_eff = opts->thr;
    if (smode) {
        float prod = (float)(2 * m1) * opts->thr;
        long long clen = (long long)((double)prod / 100.0);
        clen += clen % 2;
        kwant = clen / 2;
        inf.s_thresh = opts->thr;
    } else {
        thr_eff = weighted ? (float)((double)opts->thr - 1e-6) : opts->thr;
        inf.r_thresh = opts->thr;
        inf.t_star = (double)thr_eff;
    }

    // ---- buffers
    const int nb = (int)pg_cdiv(std::max(n, 1), TILE);
    const int npad = nb * TILE;
    const int ldp = (int)pg_cdiv(t, KC) * KC;
    const long long nt = (long long)t;
    double *mu, *sd, *z64;
    PG_TRY(pg_bufT(ctx, "corr.mu", n + 1, &mu));
    PG_TRY(pg_bufT(ctx, "corr.sd", n + 1, &sd));
    PG_TRY(pg_bufT(ctx, "corr.z64", (size_t)n * nt + 1, &z64));
    void *p0 = nullptr, *p1 = nullptr;
    if (prec == PREC_F32) {
        PG_TRY(pg_buf(ctx, "corr.p32", (size_t)npad * ldp * 4, &p0));
        PG_HIP(ctx, hipMemsetAsync(p0, 0, (size_t)npad * ldp * 4, st));
    } else {
        PG_TRY(pg_buf(ctx, "corr.phi", (size_t)npad * ldp * 2, &p0));
        PG_TRY(pg_buf(ctx, "corr.plo", (size_t)npad * ldp * 2, &p1));
        PG_HIP(ctx, hipMemsetAsync(p0, 0, (size_t)npad * ldp * 2, st));
        PG_HIP(ctx, hipMemsetAsync(p1, 0, (size_t)npad * ldp * 2, st));
    }
    const float *dbold = nullptr;
    PG_TRY(pg_in(ctx, bold, (size_t)n * nt, "corr.bold", &dbold));

    PG_HIP(ctx, hipEventRecord(ctx->ev[0], st));
    if (n > 0) {
        zscore_stats<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(dbold, opts->layout, n, t, mu, sd);
        dim3 g((unsigned)pg_cdiv(n, 64), (unsigned)pg_cdiv(t, 64));
        zscore_write<<<g, 256, 0, st>>>(dbold, opts->layout, n, t, mu, sd, z64,
                                         prec == PREC_F32 ? (float *)p0 : nullptr,
                                         prec == PREC_BF16X3 ? (uint16_t *)p0 : nullptr,
                                         prec == PREC_BF16X3 ? (uint16_t *)p1 : nullptr, ldp);
        PG_KERNEL_CHECK(ctx);
    }
    PG_HIP(ctx, hipEventRecord(ctx->ev[1], st));

    // ---- Gram launch geometry
    const long long ntiles = (long long)nb * (nb + 1) / 2;
    long long nwg = std::min<long long>((long long)ctx->num_cu * 2, std::max<long long>(ntiles, 1));
    const long long per = pg_cdiv(std::max<long long>(ntiles, 1), nwg);
    nwg = pg_cdiv(std::max<long long>(ntiles, 1), per);
    const double band = error_band(prec, t);
    inf.band = band;

    GramArgs ga;
    std::memset(&ga, 0, sizeof ga);
    ga.p0 = p0;
    ga.p1 = p1;
    ga.ldp = ldp;
    ga.kchunks = ldp / KC;
    ga.n = n;
    ga.nb = nb;
    ga.ntiles = ntiles;
    ga.per = per;
    const size_t stage = prec == PREC_BF16X3 ? 4 * TILE * LDB * 2 : 2 * TILE * LDF * 4;

    bool ran_hist = false, ran_emit = false;
    auto launch_gram = [&](int mode) -> pg_status {
        size_t shm = stage + (mode == MODE_HIST ? NBIN * 4 : 0);
        const void *fn;
        if (prec == PREC_BF16X3)
            fn = mode == MODE_HIST ? (const void *)gram_kernel<PREC_BF16X3, MODE_HIST>
                                   : (const void *)gram_kernel<PREC_BF16X3, MODE_EMIT>;
        else
            fn = mode == MODE_HIST ? (const void *)gram_kernel<PREC_F32, MODE_HIST>
                                   : (const void *)gram_kernel<PREC_F32, MODE_EMIT>;
        PG_HIP(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        const int e0 = mode == MODE_HIST ? 6 : 8;
        PG_HIP(ctx, hipEventRecord(ctx->ev[e0], st));
        if (prec == PREC_BF16X3) {
            if (mode == MODE_HIST) gram_kernel<PREC_BF16X3, MODE_HIST><<<(unsigned)nwg, THREADS, shm, st>>>(ga);
            else gram_kernel<PREC_BF16X3, MODE_EMIT><<<(unsigned)nwg, THREADS, shm, st>>>(ga);
        } else {
            if (mode == MODE_HIST) gram_kernel<PREC_F32, MODE_HIST><<<(unsigned)nwg, THREADS, shm, st>>>(ga);
            else gram_kernel<PREC_F32, MODE_EMIT><<<(unsigned)nwg, THREADS, shm, st>>>(ga);
        }
        PG_KERNEL_CHECK(ctx);
        PG_HIP(ctx, hipEventRecord(ctx->ev[e0 + 1], st));
        if (mode == MODE_HIST) ran_hist = true;
        else ran_emit = true;
        inf.gram_passes++;
        return PG_OK;
    };

    // ---- sparsity mode: histogram pass to bracket the k-th largest r~
    double lo_edge = 0.0, hi_edge = 0.0;
    float ms_hist = 0.0f;
    if (smode && kwant > 0 && n > 1) {
        uint32_t *hist;
        unsigned long long *htot;
        PG_TRY(pg_bufT(ctx, "corr.hist", (size_t)nwg * NBIN, &hist));
        PG_TRY(pg_bufT(ctx, "corr.htot", NBIN, &htot));
        ga.hist = hist;
        PG_TRY(launch_gram(MODE_HIST));
        hist_reduce<<<NBIN / 256, 256, 0, st>>>(hist, (int)nwg, htot);
        PG_KERNEL_CHECK(ctx);
        std::vector<unsigned long long> h(NBIN);
        PG_HIP(ctx, hipMemcpyAsync(h.data(), htot, NBIN * 8, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        long long above = 0;
        int bstar = -1;
        for (int b = NBIN - 1; b >= 0; b--) {
            if (above + (long long)h[b] >= kwant) { bstar = b; break; }
            above += (long long)h[b];
        }
        if (bstar < 0) return pg_fail(ctx, PG_ETHRESH, "histogram holds fewer than k=%lld pairs", kwant);
        const double w = 2.0 / NBIN;
        lo_edge = bstar * w - 1.0 - 1e-6;
        hi_edge = (bstar + 1) * w - 1.0 + 1e-6;
        ga.upper = round_up_f(hi_edge + 2.0 * band);
        ga.lower = round_down_f(lo_edge - 2.0 * band);
    } else if (smode) {
        ga.upper = 2.0f;  // no edges requested
        ga.lower = 2.0f;
    } else {
        ga.upper = round_up_f((double)thr_eff + band);
        ga.lower = round_down_f((double)thr_eff - band);
    }
    PG_HIP(ctx, hipEventRecord(ctx->ev[2], st));

    // ---- emission pass (re-run with larger buffers on overflow)
    long long seg_cap;
    if (smode) seg_cap = pg_cdiv(kwant, nwg) * 3 / 2 + 4096;
    else seg_cap = pg_cdiv(std::max<long long>((long long)(m1 * 0.03), 1 << 20), nwg) + 4096;
    long long band_cap = 1 << 22;
    int2 *seg = nullptr, *bandl = nullptr;
    float *band_r = nullptr;
    unsigned long long *seg_cnt = nullptr, *counters = nullptr;
    PG_TRY(pg_bufT(ctx, "corr.segcnt", (size_t)nwg, &seg_cnt));
    PG_TRY(pg_bufT(ctx, "corr.counters", 8, &counters));
    std::vector<unsigned long long> hseg(nwg);
    unsigned long long hcnt[8];
    for (int attempt = 0; attempt < 3; attempt++) {
        PG_TRY(pg_bufT(ctx, "corr.seg", (size_t)(nwg * seg_cap), &seg));
        PG_TRY(pg_bufT(ctx, "corr.band", (size_t)band_cap, &bandl));
        PG_TRY(pg_bufT(ctx, "corr.bandr", (size_t)band_cap, &band_r));
        PG_HIP(ctx, hipMemsetAsync(counters, 0, 8 * sizeof(unsigned long long), st));
        PG_HIP(ctx, hipMemsetAsync(seg_cnt, 0, nwg * sizeof(unsigned long long), st));
        ga.seg = seg;
        ga.seg_cap = seg_cap;
        ga.seg_cnt = seg_cnt;
        ga.band = bandl;
        ga.band_r = band_r;
        ga.band_cap = band_cap;
        ga.band_cnt = counters + 0;
        if (n > 1 && !(smode && kwant == 0)) PG_TRY(launch_gram(MODE_EMIT));
        PG_HIP(ctx, hipMemcpyAsync(hseg.data(), seg_cnt, nwg * 8, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipMemcpyAsync(hcnt, counters, sizeof hcnt, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        long long mx = 0;
        for (auto c : hseg) mx = std::max<long long>(mx, (long long)c);
        bool ok = mx <= seg_cap && (long long)hcnt[0] <= band_cap;
        if (ok) break;
        if (attempt == 2) return pg_fail(ctx, PG_ENOMEM, "edge buffers overflowed repeatedly");
        seg_cap = std::max(seg_cap, mx + mx / 8 + 4096);
        band_cap = std::max(band_cap, (long long)hcnt[0] + (long long)hcnt[0] / 8 + 4096);
    }
    PG_HIP(ctx, hipEventRecord(ctx->ev[3], st));
    long long definite = 0;
    for (auto c : hseg) definite += (long long)c;
    const long long nband = (long long)hcnt[0];
    inf.band_pairs = nband;

    // ---- FP64 recheck
    double *band64;
    int2 *extra;
    PG_TRY(pg_bufT(ctx, "corr.band64", (size_t)std::max<long long>(nband, 1), &band64));
    PG_TRY(pg_bufT(ctx, "corr.extra", (size_t)std::max<long long>(nband, 1), &extra));
    unsigned long long *extra_cnt = counters + 1, *flips = counters + 2;
    if (nband > 0)
        recheck_kernel<<<(unsigned)pg_cdiv(nband, 256), 256, 0, st>>>(bandl, counters + 0, band_cap, z64, t, band64);
    if (!smode) {
        if (nband > 0)
            accept_kernel<<<(unsigned)pg_cdiv(nband, 256), 256, 0, st>>>(
                bandl, band_r, band64, counters + 0, band_cap, (double)thr_eff, extra, extra_cnt, flips);
        PG_KERNEL_CHECK(ctx);
        PG_HIP(ctx, hipMemcpyAsync(hcnt, counters, sizeof hcnt, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        inf.band_flips = (long long)hcnt[2];
        long long nnz = 2 * (definite + (long long)hcnt[1]);
        inf.nnz = nnz;
        inf.s_thresh = (float)(m1 > 0 ? 100.0 * (double)nnz / (double)m1 / 2.0 : 0.0);
    } else {
        // exact k-th largest among the candidates: order (r64 desc, i, j) is
        // the reference's rule (all r > t*, then ties in row-major order)
        long long need = kwant - definite;
        long long nsel = 0;
        if (kwant > 0) {
            if (need < 1 || need > nband)
                return pg_fail(ctx, PG_ETHRESH,
                               "selection window invalid (k=%lld definite=%lld candidates=%lld)",
                               kwant, definite, nband);
            PG_KERNEL_CHECK(ctx);
            std::vector<int2> hp(nband);
            std::vector<double> hr(nband);
            std::vector<float> hf(nband);
            PG_HIP(ctx, hipMemcpyAsync(hp.data(), bandl, nband * sizeof(int2), hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipMemcpyAsync(hr.data(), band64, nband * 8, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipMemcpyAsync(hf.data(), band_r, nband * 4, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            std::vector<HostCand> c(nband);
            for (long long q = 0; q < nband; q++) c[q] = {hr[q], hp[q].x, hp[q].y};
            auto cmp = [](const HostCand &a, const HostCand &b) {
                if (a.r != b.r) return a.r > b.r;
                if (a.i != b.i) return a.i < b.i;
                return a.j < b.j;
            };
            std::nth_element(c.begin(), c.begin() + (need - 1), c.end(), cmp);
            const HostCand pivot = c[need - 1];
            std::sort(c.begin(), c.begin() + need, cmp);
            inf.t_star = pivot.r;
            std::vector<int2> sel(need);
            for (long long q = 0; q < need; q++) sel[q] = make_int2(c[q].i, c[q].j);
            PG_HIP(ctx, hipMemcpyAsync(extra, sel.data(), need * sizeof(int2), hipMemcpyHostToDevice, st));
            unsigned long long nn = (unsigned long long)need;
            PG_HIP(ctx, hipMemcpyAsync(extra_cnt, &nn, 8, hipMemcpyHostToDevice, st));
            // fast-path flips relative to t*: candidates whose r~ ordering
            // against t* disagrees with r64's
            long long fl = 0;
            for (long long q = 0; q < nband; q++) fl += ((double)hf[q] > pivot.r) != (hr[q] > pivot.r);
            inf.band_flips = fl;
            nsel = need;
            PG_HIP(ctx, hipStreamSynchronize(st));
        }
        inf.nnz = 2 * (definite + nsel);
        inf.r_thresh = (float)inf.t_star;
        if (kwant == 0) inf.r_thresh = 1.0f;
    }
    PG_HIP(ctx, hipEventRecord(ctx->ev[4], st));
    if (inf.nnz > INT32_MAX) return pg_fail(ctx, PG_ENOTSUP, "nnz %lld exceeds int32 CSR", (long long)inf.nnz);

    // ---- CSR build
    int *deg, *row, *cur, *ctmp, *col;
    float *vtmp = nullptr, *val = nullptr;
    const size_t nnzs = (size_t)std::max<long long>(inf.nnz, 1);
    PG_TRY(pg_bufT(ctx, "corr.deg", (size_t)n + 1, &deg));
    PG_TRY(pg_bufT(ctx, "corr.row", (size_t)n + 2, &row));
    PG_TRY(pg_bufT(ctx, "corr.cur", (size_t)n + 1, &cur));
    PG_TRY(pg_bufT(ctx, "corr.ctmp", nnzs, &ctmp));
    PG_TRY(pg_bufT(ctx, "corr.col", nnzs, &col));
    if (weighted) {
        PG_TRY(pg_bufT(ctx, "corr.vtmp", nnzs, &vtmp));
        PG_TRY(pg_bufT(ctx, "corr.val", nnzs, &val));
    }
    PG_HIP(ctx, hipMemsetAsync(deg, 0, ((size_t)n + 1) * 4, st));
    EdgeSrc es;
    es.seg = seg;
    es.seg_cap = seg_cap;
    es.seg_cnt = seg_cnt;
    es.nseg = (int)nwg;
    es.extra = extra;
    es.extra_cnt = extra_cnt;
    if (n > 0) {
        degree_count<<<(unsigned)(nwg + 1), 256, 0, st>>>(es, deg);
        scan_rows<<<1, 1024, 0, st>>>(deg, n, row, cur);
        scatter_edges<<<(unsigned)(nwg + 1), 256, 0, st>>>(es, row, cur, ctmp, vtmp, z64, t);
        size_t shm = (size_t)nw * 8;
        sort_rows<<<(unsigned)n, 256, shm, st>>>(row, ctmp, vtmp, n, col, val);
        PG_KERNEL_CHECK(ctx);
    } else {
        PG_HIP(ctx, hipMemsetAsync(row, 0, 4, st));
    }
    PG_HIP(ctx, hipEventRecord(ctx->ev[5], st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    int rn = 0;
    PG_HIP(ctx, hipMemcpy(&rn, row + n, 4, hipMemcpyDeviceToHost));
    if ((long long)rn != inf.nnz)
        return pg_fail(ctx, PG_EFORMAT, "CSR build mismatch: row[n]=%d nnz=%lld", rn, (long long)inf.nnz);

    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    inf.ms_zscore = ms;
    (void)hipEventElapsedTime(&ms_hist, ctx->ev[1], ctx->ev[2]);
    (void)hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
    inf.ms_gram = ms + ms_hist;
    (void)hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]);
    inf.ms_select = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]);
    inf.ms_csr = ms;
    if (ran_hist && hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]) == hipSuccess) inf.ms_gram_hist = ms;
    if (ran_emit && hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]) == hipSuccess) inf.ms_gram_emit = ms;

    ctx->cn = n;
    ctx->ct = t;
    ctx->cnnz = inf.nnz;
    ctx->cweighted = weighted;
    ctx->c_row = row;
    ctx->c_col = col;
    ctx->c_val = val;
    ctx->c_z64 = z64;
    if (info) *info = inf;
    return PG_OK;
}

extern "C" pg_status pg_corr_csr(pg_ctx *ctx, int32_t *row, int32_t *col, float *val)
{
    if (!ctx) return PG_EINVAL;
    if (ctx->cnnz < 0) return pg_fail(ctx, PG_ESTATE, "no CSR built yet");
    auto kind = [](void *p) { return pg_is_device_ptr(p) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost; };
    if (row) PG_HIP(ctx, hipMemcpyAsync(row, ctx->c_row, ((size_t)ctx->cn + 1) * 4, kind(row), ctx->stream));
    if (col && ctx->cnnz > 0)
        PG_HIP(ctx, hipMemcpyAsync(col, ctx->c_col, (size_t)ctx->cnnz * 4, kind(col), ctx->stream));
    if (val && ctx->cweighted && ctx->cnnz > 0)
        PG_HIP(ctx, hipMemcpyAsync(val, ctx->c_val, (size_t)ctx->cnnz * 4, kind(val), ctx->stream));
    PG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PG_OK;
}

extern "C" pg_status pg_corr_csr_view(pg_ctx *ctx, const int32_t **row, const int32_t **col,
                                      const float **val)
{
    if (!ctx) return PG_EINVAL;
    if (ctx->cnnz < 0) return pg_fail(ctx, PG_ESTATE, "no CSR built yet");
    if (row) *row = ctx->c_row;
    if (col) *col = ctx->c_col;
    if (val) *val = ctx->cweighted ? ctx->c_val : nullptr;
    return PG_OK;
}

extern "C" pg_status pg_corr_z64_view(pg_ctx *ctx, const double **z64)
{
    if (!ctx || !z64) return PG_EINVAL;
    if (ctx->cnnz < 0) return pg_fail(ctx, PG_ESTATE, "no build yet");
    *z64 = ctx->c_z64;
    return PG_OK;
}
