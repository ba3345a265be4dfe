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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "pg_internal.hpp"
#include "pg_rowsort.hpp"

namespace {

#ifndef PG_GRAM_TM
#define PG_GRAM_TM 128
#endif
#ifndef PG_GRAM_KC
#define PG_GRAM_KC 64
#endif
constexpr int TM = PG_GRAM_TM;  // Gram tile rows (128 or 256)
constexpr int TN = 128;         // Gram tile columns
constexpr int RT = TM / TN;     // tile aspect ratio
constexpr int WR = TM / 2;      // rows per wave (2 x 2 waves)
constexpr int WC = TN / 2;      // columns per wave
constexpr int BMW = WR / 32, BNW = WC / 32;  // 32 x 32 MFMA blocks per wave
constexpr int KC = PG_GRAM_KC; // K (time) chunk staged in LDS
constexpr int THREADS = 256;   // 4 wave64s in a 2 x 2 arrangement
constexpr int NBIN = 8192;     // r histogram bins over [-1, 1] (sparsity mode)
constexpr int LDB = KC + 8;    // bf16 LDS row stride (elements), 16-byte aligned rows
constexpr int LDF = KC + 1;    // fp32 LDS row stride (elements): conflict-free column reads

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
    int nr, nc;            // row tiles (npad / TM) and column tiles (npad / TN)
    int rt;                // TM / TN of the kernel that walks the tiles
    int npad;              // padded voxel count (panel rows)
    const int2 *st_list;   // supertiles (SR, SC), host-built
    long long nst;         // supertile count
    int super;             // supertile edge in tiles
    long long sstep;       // > 0: visit every sstep-th supertile only (histogram sample)
    float upper;           // r~ > upper: edge without recheck
    float lower;           // lower <= r~ <= upper: FP64 recheck
    const float *lower_dev;  // if set, lower is read from device memory
    int2 *seg;             // [gridDim.x][seg_cap] private edge segments
    long long seg_cap;
    unsigned long long *seg_cnt;   // [gridDim.x]
    int2 *band;            // recheck list
    float *band_r;         // r~ of each recheck entry
    long long band_cap;
    unsigned long long *band_cnt;  // [1]
    uint32_t *hist;        // [gridDim.x][NBIN]
    float *seg_r;          // optional: r~ of each segment entry
};

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Geometry of the LDS-DMA staged BF16x3 Gram (gram2_kernel) and its panels.
namespace g2 {
constexpr int TM = 256, TN = 128, RT = TM / TN;
constexpr int WR = 128, WC = 64, BMW = WR / 32, BNW = WC / 32;
constexpr int KC = 16;
constexpr int NSTAGE = 3;
constexpr int A_BYTES = TM * KC * 2;              // one array of the row slice
constexpr int B_BYTES = TN * KC * 2;              // one array of the column slice
constexpr int STAGE = 2 * A_BYTES + 2 * B_BYTES;  // 24 KB
constexpr int DMA_WAVE = STAGE / 1024 / 4;        // 1 KB wave loads per wave per stage (6)
static_assert(DMA_WAVE * 4 * 1024 == STAGE, "stage must split into 1 KB wave loads");
// hits of one wave's block staged in LDS before the global writes: a quarter
// of the ring slot that is free during the epilogue, pairs then r~
constexpr int STAGE_HITS = STAGE / 4 / 12;  // 512
constexpr int RING_RECS = STAGE / 4 / 8;     // 768 8-byte records per wave in the free ring slot
constexpr int DEDIC_RECS = 255;              // records per wave in the dedicated area
}  // namespace g2

__host__ __device__ __forceinline__ int g2_swz(int i) { return ((i >> 2) ^ (i >> 3)) & 1; }

// element offset of (voxel i, time l) in the chunk-major panel
__host__ __device__ __forceinline__ long long g2_index(int i, int l, int npad)
{
    const int c = l >> 4, h = (l >> 3) & 1;
    return ((long long)c * npad + i) * 16 + ((h ^ g2_swz(i)) << 3) + (l & 7);
}


// The dynamic LDS limit of a kernel raised to at least `bytes` on the
// context's device; calls that would not raise it are skipped (the attribute
// is per device and process, and only ever raised here).
inline hipError_t set_shared_limit(pg_ctx *ctx, const void *fn, size_t bytes)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, size_t> set;
    std::lock_guard<std::mutex> lock(mu);
    size_t &cur = set[{ctx->device, fn}];
    if (cur >= bytes) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) cur = bytes;
    return e;
}

// Stage-timing event idx of pg_corr_build on stream st, recorded when the
// context's timing level asks for it (events 8 and 9 bracket the emitting
// Gram kernel, kept at level 1; the others, level 2).  Each recorded event
// costs the device a few microseconds of idle time.
inline bool stage_timed(const pg_ctx *ctx, int idx) { return ctx->timing >= (idx >= 8 ? 1 : 2); }
inline hipError_t stage_event(pg_ctx *ctx, int idx, hipStream_t st)
{
    return stage_timed(ctx, idx) ? hipEventRecord(ctx->ev[idx], st) : hipSuccess;
}

// Zero fill of one or two buffers of 32-bit words (grid over the larger).
__global__ void fill_u32(uint32_t *__restrict__ p, long long count, uint32_t *__restrict__ q = nullptr,
                         long long qcount = 0)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) p[k] = 0u;
    if (k < qcount) q[k] = 0u;
}

// Clears on the stream as one fill kernel: hipMemsetAsync of a small buffer
// goes through the runtime's blit kernels, several microseconds of dispatch
// each.  Sizes in bytes, multiples of 4.
inline pg_status clear_dev(pg_ctx *ctx, hipStream_t st, void *p, size_t bytes, void *q = nullptr, size_t qbytes = 0)
{
    const long long c = (long long)(bytes / 4), d = q ? (long long)(qbytes / 4) : 0;
    const long long m = std::max(c, d);
    if (m > 0) fill_u32<<<(unsigned)pg_cdiv(m, 256), 256, 0, st>>>((uint32_t *)p, c, (uint32_t *)q, d);
    PG_KERNEL_CHECK(ctx);
    return PG_OK;
}

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
                             uint16_t *__restrict__ phi, uint16_t *__restrict__ plo, int ldp, int npad)
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
        if (p32) p32[(long long)i * ldp + l] = zf;
        if (phi) {
            const long long q = g2_index(i, l, npad);
            uint16_t h = f2bf(zf);
            phi[q] = h;
            plo[q] = f2bf(zf - bf2f(h));
        }
    }
}

// Fused form for T up to ZS_TMAX: a block stages the whole T x ZV tile of its
// voxels in LDS (coalesced), ZV threads form the sequential FP64 sums from
// LDS, then all threads write z64 and the panels.  With panels the grid
// covers the npad padded rows (npad a multiple of ZV) and the blocks write
// the padding zeros too.
constexpr int ZS_TMAX = 360;

constexpr int ZV = 32;  // voxels per block of zscore_fused (LDS [t][ZV+1] floats)

__global__ __launch_bounds__(256) void zscore_fused(const float *__restrict__ x, int layout, int n, int t,
                                                   double *__restrict__ z64, float *__restrict__ p32,
                                                   uint16_t *__restrict__ phi, uint16_t *__restrict__ plo, int ldp,
                                                   int npad)
{
#pragma clang fp contract(off)
    extern __shared__ float zt[];  // [t][ZV + 1]
    __shared__ double smu[ZV], ssd[ZV];
    constexpr int LD = ZV + 1;
    const int i0 = blockIdx.x * ZV, tid = threadIdx.x;
    // every load of the thread issued before the first LDS store (a loop
    // that stores each value as it arrives waits for each load in turn)
    if (layout == 0) {  // time-major: ZV consecutive voxels of one time point per 32 lanes
        constexpr int LS = 256 / ZV, LN = (ZS_TMAX + LS - 1) / LS;
        const int ii = tid & (ZV - 1), i = i0 + ii, l0 = tid / ZV;
        float v[LN];
#pragma unroll
        for (int k = 0; k < LN; k++) {
            const int l = l0 + k * LS;
            v[k] = i < n && l < t ? x[(long long)l * n + i] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < LN; k++)
            if (l0 + k * LS < t) zt[(l0 + k * LS) * LD + ii] = v[k];
    } else {  // voxel-major: one voxel's series per wave and pass
        constexpr int LN = (ZS_TMAX + 63) / 64;
        const int lane = tid & 63;
#pragma unroll
        for (int ii = tid >> 6; ii < ZV; ii += 4) {
            const int i = i0 + ii;
            float v[LN];
#pragma unroll
            for (int k = 0; k < LN; k++) {
                const int l = lane + 64 * k;
                v[k] = i < n && l < t ? x[(long long)i * t + l] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < LN; k++)
                if (lane + 64 * k < t) zt[(lane + 64 * k) * LD + ii] = v[k];
        }
    }
    __syncthreads();
    if (tid < ZV) {
        double s1 = 0.0;
        for (int l = 0; l < t; l++) s1 += (double)zt[l * LD + tid];
        s1 /= t;
        double s2 = 0.0;
        for (int l = 0; l < t; l++) {
            double d = (double)zt[l * LD + tid] - s1;
            double p = d * d;
            s2 += p;
        }
        smu[tid] = s1;
        ssd[tid] = sqrt(s2);
    }
    __syncthreads();
    // with panels, every element of rows [0, npad) x [0, ldp) is written
    // (zeros outside [0, n) x [0, t)), so the panels need no clearing first.
    // Each wave takes rows ii = wave, wave + 4, ...; each lane the time
    // points 2 m, 2 m + 1 for m = lane, lane + 64, ... (ldp is even, and in
    // the bf16 panels the two share a 32-bit word).
    const bool panels = p32 || phi;
    const int w = panels ? ldp : t, rows = panels ? npad : n;
    const int lane = tid & 63;
    for (int ii = tid >> 6; ii < ZV; ii += 4) {
        const int i = i0 + ii;
        if (i >= rows) break;
        const double sdv = ssd[ii], mu = smu[ii];
        for (int l = 2 * lane; l < w; l += 128) {
            float zf[2] = {0.0f, 0.0f};
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (i < n && l + h < t) {
                    const double z = sdv > 0.0 ? ((double)zt[(l + h) * LD + ii] - mu) / sdv : 0.0;
                    z64[(long long)i * t + l + h] = z;
                    zf[h] = (float)z;
                }
            }
            if (p32) {
                p32[(long long)i * ldp + l] = zf[0];
                if (l + 1 < w) p32[(long long)i * ldp + l + 1] = zf[1];
            }
            if (phi) {
                const long long q = g2_index(i, l, npad);  // l even: q even, q + 1 holds l + 1
                const uint16_t h0 = f2bf(zf[0]), h1 = f2bf(zf[1]);
                const uint16_t g0 = f2bf(zf[0] - bf2f(h0)), g1 = f2bf(zf[1] - bf2f(h1));
                *(uint32_t *)(phi + q) = (uint32_t)h0 | ((uint32_t)h1 << 16);
                *(uint32_t *)(plo + q) = (uint32_t)g0 | ((uint32_t)g1 << 16);
            }
        }
    }
}

// ------------------------------------------------------------ Gram tiles
//
// Tiles are TM x TN (rows x columns) of the N x N Gram; a tile (R, C) covers
// rows [R*TM, R*TM+TM) and columns [C*TN, C*TN+TN) and is needed when it holds
// some i < j, i.e. C >= RT*R.  Four waves split a tile 2 x 2; each wave owns a
// (TM/2) x (TN/2) block of 32 x 32 MFMA tiles.
//
// Schedule.  Work-groups are dispatched round-robin over the 8 XCDs, so
// work-group g runs on XCD g % 8.  Tiles are grouped in 8 x 8 supertiles
// (listed by the host); the work-groups of one XCD walk the same supertile
// together, so the panels it touches (8 row panels + 8 column panels) stay in
// that XCD's L2 instead of each work-group streaming its own from HBM/MALL.
constexpr int SUPER = 8;  // default supertile edge (GramArgs::super; tunable 4, 8, 16)
constexpr int NXCD = 8;
#ifndef PG_GRAM_WG_PER_CU
#define PG_GRAM_WG_PER_CU 2
#endif

struct TileWalk {
    const int2 *st;     // supertile list (SR, SC)
    long long nst;      // its length
    long long sstep;    // visit list entries 0, sstep, 2 sstep, ... (sampling)
    int nr, nc;         // row and column tiles
    int rt;             // tile aspect (rows / columns)
    int sup, sup2;      // supertile edge and area in tiles
    int L, xcd;
    long long k, kend;  // position in this XCD's tile sequence

    // This XCD's sequence is the concatenation of its supertiles
    // (xcd + 8 m) * sstep, SUPER*SUPER slots each; work-group slot l takes
    // positions l, l + L, ... so the XCD's work-groups move together.
    __device__ __forceinline__ void init(const GramArgs &a)
    {
        st = a.st_list;
        nst = a.nst;
        sstep = a.sstep > 0 ? a.sstep : 1;
        nr = a.nr;
        nc = a.nc;
        rt = a.rt;
        sup = a.super;
        sup2 = sup * sup;
        const int g = blockIdx.x, G = gridDim.x;
        xcd = g % NXCD;
        L = G / NXCD;
        k = g / NXCD;
        const long long mx = nst > 0 ? (nst - 1) / sstep : -1;
        const long long mine = mx >= xcd ? (mx - xcd) / NXCD + 1 : 0;
        kend = mine * sup2;
    }
    __device__ __forceinline__ bool next(int &R, int &C)
    {
        while (k < kend) {
            const long long kk = k;
            k += L;
            const long long s = (xcd + (kk / sup2) * NXCD) * sstep;
            const int q = (int)(kk % sup2);
            // read through the constant address space: a scalar load, which
            // does not make the wave wait for the ring's copies in flight as
            // a vector load's vmcnt wait would (the list is read-only here)
            const __attribute__((address_space(4))) int *stc = (const __attribute__((address_space(4))) int *)st;
            const int2 sc = make_int2(stc[2 * s], stc[2 * s + 1]);
            const int r = sc.x * sup + q / sup, c = sc.y * sup + q % sup;
            if (r < nr && c < nc && c >= rt * r) {
                R = r;
                C = c;
                return true;
            }
        }
        return false;
    }
};

// Global -> register staging of one K chunk of the row (TM) and column (TN) panels.
template <int PREC>
struct Stage {
    static constexpr int VEC = PREC == PREC_BF16X3 ? 8 : 4;   // elements per 16 B
    static constexpr int IA = TM * KC / VEC / THREADS;        // 16-B pieces per thread, row panel
    static constexpr int IB = TN * KC / VEC / THREADS;        // column panel
    static constexpr int NARR = PREC == PREC_BF16X3 ? 2 : 1;  // hi/lo or f32
    u32x4 a[NARR][IA], b[NARR][IB];

    __device__ __forceinline__ static void piece(int it, int &r, int &c)
    {
        const int id = it * THREADS + (int)threadIdx.x;
        r = id / (KC / VEC);
        c = (id % (KC / VEC)) * VEC;
    }
    __device__ __forceinline__ void load(const GramArgs &g, int R, int C, int kc)
    {
#pragma unroll
        for (int it = 0; it < IA; it++) {
            int r, c;
            piece(it, r, c);
            const long long ga = (long long)(R * TM + r) * g.ldp + kc * KC + c;
            if (PREC == PREC_BF16X3) {
                a[0][it] = *(const u32x4 *)((const uint16_t *)g.p0 + ga);
                a[NARR - 1][it] = *(const u32x4 *)((const uint16_t *)g.p1 + ga);
            } else {
                a[0][it] = *(const u32x4 *)((const float *)g.p0 + ga);
            }
        }
#pragma unroll
        for (int it = 0; it < IB; it++) {
            int r, c;
            piece(it, r, c);
            const long long gb = (long long)(C * TN + r) * g.ldp + kc * KC + c;
            if (PREC == PREC_BF16X3) {
                b[0][it] = *(const u32x4 *)((const uint16_t *)g.p0 + gb);
                b[NARR - 1][it] = *(const u32x4 *)((const uint16_t *)g.p1 + gb);
            } else {
                b[0][it] = *(const u32x4 *)((const float *)g.p0 + gb);
            }
        }
    }
    __device__ __forceinline__ void store(char *smem) const
    {
        if (PREC == PREC_BF16X3) {
            uint16_t *sAh = (uint16_t *)smem;
            uint16_t *sAl = sAh + TM * LDB;
            uint16_t *sBh = sAl + TM * LDB;
            uint16_t *sBl = sBh + TN * LDB;
#pragma unroll
            for (int it = 0; it < IA; it++) {
                int r, c;
                piece(it, r, c);
                *(u32x4 *)(sAh + r * LDB + c) = a[0][it];
                *(u32x4 *)(sAl + r * LDB + c) = a[NARR - 1][it];
            }
#pragma unroll
            for (int it = 0; it < IB; it++) {
                int r, c;
                piece(it, r, c);
                *(u32x4 *)(sBh + r * LDB + c) = b[0][it];
                *(u32x4 *)(sBl + r * LDB + c) = b[NARR - 1][it];
            }
        } else {
            // odd row stride (conflict-free column reads): 4-byte stores
            float *sA = (float *)smem;
            float *sB = sA + TM * LDF;
#pragma unroll
            for (int it = 0; it < IA; it++) {
                int r, c;
                piece(it, r, c);
#pragma unroll
                for (int e = 0; e < 4; e++) sA[r * LDF + c + e] = __uint_as_float(a[0][it][e]);
            }
#pragma unroll
            for (int it = 0; it < IB; it++) {
                int r, c;
                piece(it, r, c);
#pragma unroll
                for (int e = 0; e < 4; e++) sB[r * LDF + c + e] = __uint_as_float(b[0][it][e]);
            }
        }
    }
};

template <int PREC>
__device__ __forceinline__ void mfma_chunk(const char *smem, f32x16 (&acc)[BMW][BNW])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w >> 1, wn = w & 1, lr = lane & 31;
    if (PREC == PREC_BF16X3) {
        const uint16_t *sAh = (const uint16_t *)smem;
        const uint16_t *sAl = sAh + TM * LDB;
        const uint16_t *sBh = sAl + TM * LDB;
        const uint16_t *sBl = sBh + TN * LDB;
        const int lk = (lane >> 5) * 8;
#pragma unroll
        for (int ks = 0; ks < KC / 16; ks++) {
            bf16x8 ah[BMW], al[BMW], bh[BNW], bl[BNW];
#pragma unroll
            for (int b = 0; b < BMW; b++) {
                const int ra = (wm * WR + b * 32 + lr) * LDB + ks * 16 + lk;
                ah[b] = *(const bf16x8 *)(sAh + ra);
                al[b] = *(const bf16x8 *)(sAl + ra);
            }
#pragma unroll
            for (int b = 0; b < BNW; b++) {
                const int rb = (wn * WC + b * 32 + lr) * LDB + ks * 16 + lk;
                bh[b] = *(const bf16x8 *)(sBh + rb);
                bl[b] = *(const bf16x8 *)(sBl + rb);
            }
            // small cross terms first; independent accumulators interleaved
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[bm], bh[bn], acc[bm][bn], 0, 0, 0);
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bm], bl[bn], acc[bm][bn], 0, 0, 0);
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bm], bh[bn], acc[bm][bn], 0, 0, 0);
        }
    } else {
        const float *sA = (const float *)smem;
        const float *sB = sA + TM * LDF;
        const int lk = lane >> 5;
#pragma unroll
        for (int ks = 0; ks < KC / 2; ks++) {
            float av[BMW], bv[BNW];
#pragma unroll
            for (int b = 0; b < BMW; b++) av[b] = sA[(wm * WR + b * 32 + lr) * LDF + ks * 2 + lk];
#pragma unroll
            for (int b = 0; b < BNW; b++) bv[b] = sB[(wn * WC + b * 32 + lr) * LDF + ks * 2 + lk];
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[bm], bv[bn], acc[bm][bn], 0, 0, 0);
        }
    }
}

__device__ __forceinline__ unsigned lane_rank(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

template <int PREC>
constexpr int stage_bytes()
{
    return PREC == PREC_BF16X3 ? (2 * TM + 2 * TN) * LDB * 2 : (TM + TN) * LDF * 4;
}

// Persistent: each work-group walks its tiles (TileWalk), prefetching the next
// K chunk into registers while the MFMAs consume the current one from LDS.
template <int PREC, int MODE>
__global__ __launch_bounds__(THREADS, PG_GRAM_WG_PER_CU) void gram_kernel(GramArgs a)
{
    extern __shared__ __align__(16) char smem[];
    uint32_t *s_hist = (uint32_t *)(smem + stage_bytes<PREC>());
    __shared__ unsigned long long s_seg;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    if (tid == 0) s_seg = 0;
    if (MODE == MODE_HIST)
        for (int b = tid; b < NBIN; b += THREADS) s_hist[b] = 0;
    const float lower0 = a.lower_dev ? *a.lower_dev : a.lower;

    TileWalk walk;
    walk.init(a);
    int2 *myseg = a.seg + (long long)blockIdx.x * a.seg_cap;
    int R = 0, C = 0;
    bool have = walk.next(R, C);
    Stage<PREC> pre;
    if (have) pre.load(a, R, C, 0);

    while (have) {
        f32x16 acc[BMW][BNW];
#pragma unroll
        for (int bm = 0; bm < BMW; bm++)
#pragma unroll
            for (int bn = 0; bn < BNW; bn++)
#pragma unroll
                for (int q = 0; q < 16; q++) acc[bm][bn][q] = 0.0f;
        int nR = 0, nC = 0;
        bool nhave = false;
        for (int kc = 0; kc < a.kchunks; kc++) {
            __syncthreads();
            pre.store(smem);
            __syncthreads();
            if (kc + 1 < a.kchunks) pre.load(a, R, C, kc + 1);
            mfma_chunk<PREC>(smem, acc);
        }
        // next tile's first chunk: issued here (not during the last chunk)
        // so the staging registers are not live across the epilogue
        nhave = walk.next(nR, nC);
        if (nhave) pre.load(a, nR, nC, 0);

        // epilogue: 32x32 block layout, lane holds column lane%32 and rows
        // (q/4)*8 + (lane/32)*4 + q%4 of each block.
        const int i0 = R * TM + wm * WR, j0 = C * TN + wn * WC;
        const bool interior = C * TN >= R * TM + TM && (C + 1) * TN <= a.n;
        if (MODE == MODE_HIST) {
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
#pragma unroll
                    for (int q = 0; q < 16; q++) {
                        const float v = acc[bm][bn][q];
                        const int i = i0 + bm * 32 + (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
                        const int j = j0 + bn * 32 + (lane & 31);
                        if (interior || ((j < a.n) && (i < j))) {
                            int b = (int)((v + 1.0f) * (NBIN * 0.5f));
                            b = b < 0 ? 0 : (b >= NBIN ? NBIN - 1 : b);
                            atomicAdd(&s_hist[b], 1u);
                        }
                    }
        } else {
            // pass 1: count edges and rechecks with one compare per value and
            // scalar popcounts of the ballots; pass 2 writes them after one
            // reservation per wave and tile.
            unsigned ne = 0, nc = 0;
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
#pragma unroll
                    for (int q = 0; q < 16; q++) {
                        const float v = acc[bm][bn][q];
                        const int i = i0 + bm * 32 + (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
                        const int j = j0 + bn * 32 + (lane & 31);
                        bool hit = v >= lower0;
                        if (!interior) hit = hit && (j < a.n) && (i < j);
                        const unsigned long long mh = __ballot(hit);
                        if (mh) {
                            const unsigned long long me = __ballot(hit && v > a.upper);
                            ne += __popcll(me);
                            nc += __popcll(mh & ~me);
                        }
                    }
            if (ne + nc) {
                unsigned long long be = 0, bc = 0;
                if (lane == 0) {
                    if (ne) be = atomicAdd(&s_seg, (unsigned long long)ne);
                    if (nc) bc = atomicAdd(a.band_cnt, (unsigned long long)nc);
                }
                be = __shfl(be, 0);
                bc = __shfl(bc, 0);
                // opaque copies: pass 2 recomputes its ballots instead of the
                // compiler holding all of pass 1's live (register pressure)
                float lo = lower0, up = a.upper;
                asm volatile("" : "+s"(lo), "+s"(up));
#pragma unroll
                for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                    for (int bn = 0; bn < BNW; bn++)
#pragma unroll
                        for (int q = 0; q < 16; q++) {
                            const float v = acc[bm][bn][q];
                            const int i = i0 + bm * 32 + (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
                            const int j = j0 + bn * 32 + (lane & 31);
                            bool hit = v >= lo;
                            if (!interior) hit = hit && (j < a.n) && (i < j);
                            const unsigned long long mh = __ballot(hit);
                            if (mh == 0ull) continue;
                            const bool e = hit && v > up;
                            const unsigned long long me = __ballot(e);
                            const unsigned long long mc = mh & ~me;
                            if (e) {
                                const unsigned long long p = be + lane_rank(me);
                                if ((long long)p < a.seg_cap) {
                                    myseg[p] = make_int2(i, j);
                                    if (a.seg_r) a.seg_r[(long long)blockIdx.x * a.seg_cap + p] = v;
                                }
                            } else if (hit) {
                                const unsigned long long p = bc + lane_rank(mc);
                                if ((long long)p < a.band_cap) {
                                    a.band[p] = make_int2(i, j);
                                    a.band_r[p] = v;
                                }
                            }
                            be += __popcll(me);
                            bc += __popcll(mc);
                        }
            }
        }
        R = nR;
        C = nC;
        have = nhave;
    }
    __syncthreads();
    if (MODE == MODE_HIST) {
        uint32_t *out = a.hist + (long long)blockIdx.x * NBIN;
        for (int b = tid; b < NBIN; b += THREADS) out[b] = s_hist[b];
    } else if (tid == 0) {
        a.seg_cnt[blockIdx.x] = s_seg;
    }
}

// --------------------------------------------- BF16x3 Gram, LDS-DMA staged
//
// Panel layout (bf16, hi and lo arrays): chunk-major.  Time is cut in chunks
// of 16; chunk c of voxel i is 32 bytes at ((c * npad + i) * 16) elements,
// its two 16-byte halves (k 0-7, k 8-15) stored in the order h ^ swz(i).  A
// 256-row or 128-row panel slice of one chunk is then one contiguous block
// that the work-group copies global -> LDS with 16-byte-per-lane direct loads
// (no VGPR staging), and the swizzle makes the MFMA operand reads
// (ds_read_b128: lane l reads row l % 32, half l / 32) bank-conflict free.
//
// Tile 256 x 128; four waves in 2 x 2, each a 128 x 64 block of 32 x 32 MFMA
// tiles (8 accumulators).  A three-slot ring of 16-deep K stages: chunk c + 2
// is in flight while chunk c is multiplied; the ring runs on across tile
// boundaries.
#define G2_LDS(p) ((__attribute__((address_space(3))) void *)(p))

// This wave's share of the copies of chunk kc of tile (R, C) into LDS stage
// slot `dst`.  A stage is [A-hi 8 KB][A-lo 8 KB][B-hi 4 KB][B-lo 4 KB], copied
// in 1 KB wave loads; wave w takes A pieces 2w, 2w+1 of both arrays and B
// piece w of both arrays.
__device__ __forceinline__ void g2_dma(const GramArgs &a, char *dst, int R, int C, int kc)
{
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const char *hi = (const char *)a.p0, *lo = (const char *)a.p1;
    const long long arow = ((long long)kc * a.npad + (long long)R * g2::TM) * 32 + 2 * w * 1024;
    const long long brow = ((long long)kc * a.npad + (long long)C * g2::TN) * 32 + w * 1024;
    const int la = lane * 16;
    char *da = dst + 2 * w * 1024, *db = dst + 2 * g2::A_BYTES + w * 1024;
    __builtin_amdgcn_global_load_lds((const void *)(hi + arow + la), G2_LDS(da), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((const void *)(hi + arow + 1024 + la), G2_LDS(da + 1024), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((const void *)(lo + arow + la), G2_LDS(da + g2::A_BYTES), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((const void *)(lo + arow + 1024 + la), G2_LDS(da + g2::A_BYTES + 1024), 16,
                                     0, 0);
    __builtin_amdgcn_global_load_lds((const void *)(hi + brow + la), G2_LDS(db), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((const void *)(lo + brow + la), G2_LDS(db + g2::B_BYTES), 16, 0, 0);
}

// s_waitcnt vmcnt(n) (gfx9 encoding, other counters left alone)
#define G2_WAIT_VM(n) __builtin_amdgcn_s_waitcnt((0x3F70) | ((n) & 15) | ((((n) >> 4) & 3) << 14))


// Pass 2 of the epilogue: each lane writes its hits in order from its slots
// be (edges) and bc (rechecks); CHECK guards every write against capacity.
template <bool BOUNDED, bool SPLIT, bool CHECK>
__device__ __forceinline__ void g2_write(const GramArgs &a, const f32x16 (&acc)[g2::BMW][g2::BNW], float lo, float up,
                                         int ib, int jb, int nn, int2 *myseg, float *segr, unsigned long long be,
                                         unsigned long long bc)
{
#pragma unroll
    for (int bm = 0; bm < g2::BMW; bm++)
#pragma unroll
        for (int bn = 0; bn < g2::BNW; bn++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const float v = acc[bm][bn][q];
                const int i = ib + bm * 32 + (q >> 2) * 8 + (q & 3), j = jb + bn * 32;
                bool hit = v >= lo;
                if (!BOUNDED) hit = hit & (j < nn) & (i < j);
                if (!hit) continue;
                if (SPLIT && !(v > up)) {
                    if (!CHECK || (long long)bc < a.band_cap) {
                        a.band[bc] = make_int2(i, j);
                        a.band_r[bc] = v;
                    }
                    bc++;
                } else {
                    if (!CHECK || (long long)be < a.seg_cap) {
                        myseg[be] = make_int2(i, j);
                        if (!SPLIT) segr[be] = v;
                    }
                    be++;
                }
            }
}

// LDS byte address of a pointer into the kernel's LDS
__device__ __forceinline__ unsigned g2_lds_addr(const void *p)
{
    return (unsigned)(size_t)(const __attribute__((address_space(3))) char *)p;
}

// The staging area's LDS accesses are written in inline assembly: the
// compiler cannot tell the staging slot from the slots the ring's copies
// are still filling, and would wait for every copy (vmcnt(0)) before each
// access; no copy targets the staging slot while it is in use.
__device__ __forceinline__ void g2_lds_put(unsigned a8, unsigned a4, int i, int j, float v)
{
    const unsigned long long ij = ((unsigned long long)(unsigned)j << 32) | (unsigned)i;
    asm volatile("ds_write_b64 %0, %1\n\tds_write_b32 %2, %3" ::"v"(a8), "v"(ij), "v"(a4), "v"(v) : "memory");
}
// the work-group's segment cursor, reserved the same way (the compiler would
// otherwise wait for every copy in flight before the atomic)
__device__ __forceinline__ unsigned long long g2_lds_add(unsigned long long *p, unsigned long long v)
{
    unsigned long long old;
    asm volatile("ds_add_rtn_u64 %0, %1, %2\n\ts_waitcnt lgkmcnt(0)" : "=v"(old) : "v"(g2_lds_addr(p)), "v"(v) : "memory");
    return old;
}
// histogram bins likewise (no result, so no wait at all)
__device__ __forceinline__ void g2_lds_inc(uint32_t *p)
{
    asm volatile("ds_add_u32 %0, %1" ::"v"(g2_lds_addr(p)), "v"(1u) : "memory");
}
__device__ __forceinline__ void g2_lds_get(unsigned a8, unsigned a4, unsigned long long &ij, float &v)
{
    asm volatile("ds_read_b64 %0, %2\n\tds_read_b32 %1, %3\n\ts_waitcnt lgkmcnt(0)"
                 : "=v"(ij), "=v"(v)
                 : "v"(a8), "v"(a4)
                 : "memory");
}

// Pass 2 into LDS: the lane's hits in order, edges from index pe and rechecks
// from index pc (placed after every edge of the wave), of which those in
// [base, base + STAGE_HITS) are written at index - base.  spair / sval: LDS
// byte addresses of the staged pairs and values.
template <bool BOUNDED, bool SPLIT>
__device__ __forceinline__ void g2_stage(const f32x16 (&acc)[g2::BMW][g2::BNW], float lo, float up, int ib, int jb,
                                         int nn, unsigned spair, unsigned sval, unsigned pe, unsigned pc, unsigned base)
{
#pragma unroll
    for (int bm = 0; bm < g2::BMW; bm++)
#pragma unroll
        for (int bn = 0; bn < g2::BNW; bn++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const float v = acc[bm][bn][q];
                const int i = ib + bm * 32 + (q >> 2) * 8 + (q & 3), j = jb + bn * 32;
                bool hit = v >= lo;
                if (!BOUNDED) hit = hit & (j < nn) & (i < j);
                if (!hit) continue;
                const unsigned p = ((SPLIT && !(v > up)) ? pc++ : pe++) - base;
                if (p < (unsigned)g2::STAGE_HITS) g2_lds_put(spair + p * 8u, sval + p * 4u, i, j, v);
            }
}

// A wave's staged hits waiting for their global stores.  The stores are
// issued at the next tile's first step, after its wait for the ring's copy
// and before its barrier: a wait for a copy also waits for every store issued
// before it, so stores issued in the epilogue would hold the next tile's
// first wait for their full latency, while these have a whole step of MFMAs
// to complete before the next wait.
// Without a band (SPLIT false) the hits are staged as 8-byte records by
// g2_emit_blocks: r~ bits in the low word, the pair relative to the wave's
// block origin (i0, j0) as (i - i0) << 6 | (j - j0) in the high word.
struct G2Pending {
    unsigned te, tc;       // edges, then rechecks, staged
    unsigned be0, bc0;     // their positions in the segment and the recheck list
    unsigned spair, sval;  // LDS byte addresses of the staged pairs and r~ (records: spair)
    int i0, j0;            // records: the block origin
    bool on;
};

// The stores of a wave's staged hits (the staging slot's next copy is issued
// only after the work-group barrier that follows).
template <bool SPLIT>
__device__ __forceinline__ void g2_flush(const GramArgs &a, const G2Pending &pd, int lane, int2 *myseg)
{
    float *segr = SPLIT ? nullptr : a.seg_r + (long long)blockIdx.x * a.seg_cap;
    const unsigned long long cap = (unsigned long long)a.seg_cap, bcap = (unsigned long long)a.band_cap;
    if (!SPLIT) {
        for (unsigned k = lane; k < pd.te; k += 64) {
            unsigned long long rec;
            asm volatile("ds_read_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(rec) : "v"(pd.spair + k * 8u) : "memory");
            const unsigned idx = (unsigned)(rec >> 32);
            const unsigned long long p = (unsigned long long)pd.be0 + k;
            if (p < cap) {
                myseg[p] = make_int2(pd.i0 + (int)(idx >> 6), pd.j0 + (int)(idx & 63u));
                segr[p] = __uint_as_float((unsigned)rec);
            }
        }
        return;
    }
    for (unsigned k = lane; k < pd.te + pd.tc; k += 64) {
        unsigned long long ij;
        float v;
        g2_lds_get(pd.spair + k * 8u, pd.sval + k * 4u, ij, v);
        const int2 e = make_int2((int)(unsigned)ij, (int)(unsigned)(ij >> 32));
        if (k < pd.te) {
            const unsigned long long p = (unsigned long long)pd.be0 + k;
            if (p < cap) {
                myseg[p] = e;
                if (!SPLIT) segr[p] = v;
            }
        } else {
            const unsigned long long p = (unsigned long long)pd.bc0 + (k - pd.te);
            if (p < bcap) {
                a.band[p] = e;
                a.band_r[p] = v;
            }
        }
    }
}

// Epilogue of one wave's 128 x 64 block: pass 1 counts per lane (packed: low
// 16 bits edges v > upper, high 16 bits rechecks lower <= v <= upper; each at
// most 8192 per wave); one reservation per wave; then the wave's hits (edges,
// then rechecks) are staged in LDS in output order and left to g2_flush, or,
// beyond STAGE_HITS of them, each lane writes its own (g2_write).  BOUNDED:
// every pair is inside the matrix and above the diagonal.
template <bool BOUNDED, bool SPLIT>
__device__ __forceinline__ void g2_emit(const GramArgs &a, float lower, const f32x16 (&acc)[g2::BMW][g2::BNW], int i0,
                                        int j0, int lane, int2 *myseg, unsigned long long *s_seg, char *stage,
                                        G2Pending &pd)
{
    const int ib0 = i0 + (lane >> 5) * 4, jb0 = j0 + (lane & 31);
    unsigned cnt = 0;
#pragma unroll
    for (int bm = 0; bm < g2::BMW; bm++)
#pragma unroll
        for (int bn = 0; bn < g2::BNW; bn++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const float v = acc[bm][bn][q];
                bool hit = v >= lower;
                if (!BOUNDED) {
                    const int i = ib0 + bm * 32 + (q >> 2) * 8 + (q & 3), j = jb0 + bn * 32;
                    hit = hit & (j < a.n) & (i < j);
                }
                if (SPLIT) cnt += hit ? (v > a.upper ? 1u : 0x10000u) : 0u;
                else cnt += hit ? 1u : 0u;
            }
    unsigned x = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    const unsigned tot = __builtin_amdgcn_readfirstlane(__shfl(x, 63));
    if (tot == 0) return;
    unsigned long long be = 0, bc = 0;
    if (lane == 0) {
        if (tot & 0xFFFFu) be = g2_lds_add(s_seg, (unsigned long long)(tot & 0xFFFFu));
        if (tot >> 16) bc = atomicAdd(a.band_cnt, (unsigned long long)(tot >> 16));
    }
    const unsigned ex = x - cnt;
    const unsigned be0 = (unsigned)__shfl(be, 0), bc0 = (unsigned)__shfl(bc, 0);
    // opaque copies: the staging pass recomputes its predicates instead of
    // the compiler keeping the count pass's 128 of them live
    float lo = lower, up = a.upper;
    int ib = ib0, jb = jb0, nn = a.n;
    asm volatile("" : "+v"(lo), "+v"(up), "+v"(ib), "+v"(jb), "+v"(nn));
    const unsigned te = tot & 0xFFFFu, tt = te + (tot >> 16);
    const unsigned seg_cap = (unsigned)a.seg_cap, band_cap = (unsigned)a.band_cap;
    // r~ of every segment entry is kept exactly when there is no band (the
    // sparsity single pass)
    float *segr = SPLIT ? nullptr : a.seg_r + (long long)blockIdx.x * a.seg_cap;
    if (!SPLIT || tt > (unsigned)g2::STAGE_HITS) {
        // more hits than the staging area holds (or the band-free form, whose
        // staging is g2_emit_blocks'): each lane writes its own
        const bool fits = (unsigned long long)be0 + te <= (unsigned long long)a.seg_cap &&
                          (unsigned long long)bc0 + (tt - te) <= (unsigned long long)a.band_cap;
        const unsigned long long bl = be0 + (ex & 0xFFFFu), cl = bc0 + (ex >> 16);
        if (fits) g2_write<BOUNDED, SPLIT, false>(a, acc, lo, up, ib, jb, nn, myseg, segr, bl, cl);
        else g2_write<BOUNDED, SPLIT, true>(a, acc, lo, up, ib, jb, nn, myseg, segr, bl, cl);
        return;
    }
    const unsigned spair = g2_lds_addr(stage), sval = spair + g2::STAGE_HITS * 8;
    g2_stage<BOUNDED, SPLIT>(acc, lo, up, ib, jb, nn, spair, sval, ex & 0xFFFFu, te + (ex >> 16), 0u);
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    pd.te = te;
    pd.tc = tt - te;
    pd.be0 = be0;
    pd.bc0 = bc0;
    pd.spair = spair;
    pd.sval = sval;
    pd.on = true;
}

// One accumulator's 16 elements as a lane mask, element q at bit 15 - q:
// per element a compare and an add-with-carry (m = 2 m + hit).
__device__ __forceinline__ unsigned g2_mask16(const f32x16 &a, float lower)
{
    unsigned m = 0;
#pragma unroll
    for (int q = 0; q < 16; q++)
        asm volatile("v_cmp_le_f32 vcc, %1, %2\n\tv_addc_co_u32 %0, vcc, %0, %0, vcc"
                     : "+v"(m)
                     : "s"(lower), "v"(a[q])
                     : "vcc");
    return m;
}

// Element q of an accumulator, q per lane: a select tree on q's bits, in
// assembly (the compiler turns the same tree written in C++ into a scratch
// array indexed by q).
__device__ __forceinline__ float g2_pick16(const f32x16 &a, unsigned q)
{
    float t0, t1, t2, t3, t4, t5, t6, t7, r;
    unsigned b;
    asm volatile(
        "v_and_b32 %[b], 1, %[q]\n\t"
        "v_cmp_ne_u32 vcc, 0, %[b]\n\t"
        "v_cndmask_b32 %[t0], %[a0], %[a1], vcc\n\t"
        "v_cndmask_b32 %[t1], %[a2], %[a3], vcc\n\t"
        "v_cndmask_b32 %[t2], %[a4], %[a5], vcc\n\t"
        "v_cndmask_b32 %[t3], %[a6], %[a7], vcc\n\t"
        "v_cndmask_b32 %[t4], %[a8], %[a9], vcc\n\t"
        "v_cndmask_b32 %[t5], %[a10], %[a11], vcc\n\t"
        "v_cndmask_b32 %[t6], %[a12], %[a13], vcc\n\t"
        "v_cndmask_b32 %[t7], %[a14], %[a15], vcc\n\t"
        "v_and_b32 %[b], 2, %[q]\n\t"
        "v_cmp_ne_u32 vcc, 0, %[b]\n\t"
        "v_cndmask_b32 %[t0], %[t0], %[t1], vcc\n\t"
        "v_cndmask_b32 %[t1], %[t2], %[t3], vcc\n\t"
        "v_cndmask_b32 %[t2], %[t4], %[t5], vcc\n\t"
        "v_cndmask_b32 %[t3], %[t6], %[t7], vcc\n\t"
        "v_and_b32 %[b], 4, %[q]\n\t"
        "v_cmp_ne_u32 vcc, 0, %[b]\n\t"
        "v_cndmask_b32 %[t0], %[t0], %[t1], vcc\n\t"
        "v_cndmask_b32 %[t1], %[t2], %[t3], vcc\n\t"
        "v_and_b32 %[b], 8, %[q]\n\t"
        "v_cmp_ne_u32 vcc, 0, %[b]\n\t"
        "v_cndmask_b32 %[r], %[t0], %[t1], vcc"
        : [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3), [t4] "=&v"(t4), [t5] "=&v"(t5),
          [t6] "=&v"(t6), [t7] "=&v"(t7), [r] "=&v"(r), [b] "=&v"(b)
        : [q] "v"(q), [a0] "v"(a[0]), [a1] "v"(a[1]), [a2] "v"(a[2]), [a3] "v"(a[3]), [a4] "v"(a[4]),
          [a5] "v"(a[5]), [a6] "v"(a[6]), [a7] "v"(a[7]), [a8] "v"(a[8]), [a9] "v"(a[9]), [a10] "v"(a[10]),
          [a11] "v"(a[11]), [a12] "v"(a[12]), [a13] "v"(a[13]), [a14] "v"(a[14]), [a15] "v"(a[15])
        : "vcc");
    return r;
}

// The band-free epilogue by accumulator (16 elements at a time): each lane's
// hit mask, then rounds in which every lane with hits left writes its next
// one at the wave's count plus the round's mbcnt (as many LDS record writes
// per accumulator as the most hits any lane has in it, about two, instead of
// one per element with a hit).  8-byte records (G2Pending) in the staging
// area; more than STAGE_RECS hits: the staged ones are dropped and g2_emit
// writes them all directly.  The staging order (accumulator-major) differs
// from g2_emit's, which the segment does not care about.
template <bool BOUNDED, unsigned STAGE_RECS>
__device__ __forceinline__ void g2_emit_blocks(const GramArgs &a, float lower, const f32x16 (&acc)[g2::BMW][g2::BNW],
                                               int i0, int j0, int lane, int2 *myseg, unsigned long long *s_seg,
                                               char *stage, G2Pending &pd)
{
    const int ib0 = i0 + (lane >> 5) * 4, jb0 = j0 + (lane & 31);
    const unsigned lidx = ((unsigned)(lane >> 5) << 8) | (unsigned)(lane & 31);
    const unsigned srec = g2_lds_addr(stage);
    unsigned nst = 0;  // hits so far (wave-uniform)
#pragma unroll
    for (int bm = 0; bm < g2::BMW; bm++)
#pragma unroll
        for (int bn = 0; bn < g2::BNW; bn++) {
            unsigned m;
            if (BOUNDED) {
                m = g2_mask16(acc[bm][bn], lower);
            } else {
                m = 0;
                const int j = jb0 + bn * 32;
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const int i = ib0 + bm * 32 + (q >> 2) * 8 + (q & 3);
                    const bool hit = (acc[bm][bn][q] >= lower) & (j < a.n) & (i < j);
                    m = 2 * m + (hit ? 1u : 0u);
                }
            }
            // round r: every lane with more than r hits writes its next one;
            // positions from the round's ballot (mbcnt), the count by popcount
            const unsigned eb = ((unsigned)(bm * 32) << 6) | (unsigned)(bn * 32);
            while (true) {
                const unsigned long long act = __builtin_amdgcn_ballot_w64(m != 0u);
                if (act == 0ull) break;
                if (m) {
                    const unsigned p = __builtin_amdgcn_mbcnt_hi((unsigned)(act >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((unsigned)act, nst));
                    const unsigned k = 31u - (unsigned)__builtin_clz(m);  // highest set bit: element 15 - k
                    m &= ~(1u << k);
                    const unsigned q = 15u - k;
                    const float v = g2_pick16(acc[bm][bn], q);
                    const unsigned idx = lidx | eb | ((q & 12u) << 7) | ((q & 3u) << 6);
                    const unsigned long long rec = ((unsigned long long)idx << 32) | (unsigned long long)__float_as_uint(v);
                    if (p < STAGE_RECS) asm volatile("ds_write_b64 %0, %1" ::"v"(srec + p * 8u), "v"(rec) : "memory");
                }
                nst += (unsigned)__popcll(act);
            }
        }
    if (nst > STAGE_RECS) {
        float lo2 = lower;
        int i02 = i0, j02 = j0;
        asm volatile("" : "+v"(lo2), "+v"(i02), "+v"(j02));
        g2_emit<BOUNDED, false>(a, lo2, acc, i02, j02, lane, myseg, s_seg, stage, pd);
        return;
    }
    if (nst == 0) return;
    unsigned long long be = 0;
    if (lane == 0) be = g2_lds_add(s_seg, (unsigned long long)nst);
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    pd.te = nst;
    pd.tc = 0;
    pd.be0 = (unsigned)__shfl(be, 0);
    pd.bc0 = 0;
    pd.spair = srec;
    pd.sval = 0;
    pd.i0 = i0;
    pd.j0 = j0;
    pd.on = true;
}

// Epilogue of one wave's block of tile (R, C): lane holds column lane%32 and
// rows (q/4)*8 + (lane/32)*4 + q%4 of each 32 x 32 block.  Histogram mode bins
// every pair i < j < n; emit mode writes the hits (g2_emit).
template <int MODE, bool DEDIC>
__device__ __forceinline__ void g2_tile_epilogue(const GramArgs &a, const f32x16 (&acc)[g2::BMW][g2::BNW], int R,
                                                 int C, int wm, int wn, int lane, float lower, bool split,
                                                 int2 *myseg, unsigned long long *s_seg, uint32_t *s_hist,
                                                 char *stage, G2Pending &pd)
{
    const int i0 = R * g2::TM + wm * g2::WR, j0 = C * g2::TN + wn * g2::WC;
    const bool interior = C * g2::TN >= R * g2::TM + g2::TM && (C + 1) * g2::TN <= a.n;
    if (MODE == MODE_HIST) {
#pragma unroll
        for (int bm = 0; bm < g2::BMW; bm++)
#pragma unroll
            for (int bn = 0; bn < g2::BNW; bn++)
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const float v = acc[bm][bn][q];
                    const int i = i0 + bm * 32 + (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
                    const int j = j0 + bn * 32 + (lane & 31);
                    if (interior || ((j < a.n) && (i < j))) {
                        int b = (int)((v + 1.0f) * (NBIN * 0.5f));
                        b = b < 0 ? 0 : (b >= NBIN ? NBIN - 1 : b);
                        g2_lds_inc(&s_hist[b]);
                    }
                }
    } else if (split) {
        if (interior) g2_emit<true, true>(a, lower, acc, i0, j0, lane, myseg, s_seg, stage, pd);
        else g2_emit<false, true>(a, lower, acc, i0, j0, lane, myseg, s_seg, stage, pd);
    } else {
        constexpr unsigned cap = DEDIC ? g2::DEDIC_RECS : g2::RING_RECS;
        if (interior) g2_emit_blocks<true, cap>(a, lower, acc, i0, j0, lane, myseg, s_seg, stage, pd);
        else g2_emit_blocks<false, cap>(a, lower, acc, i0, j0, lane, myseg, s_seg, stage, pd);
    }
}

// SPLIT (emit mode): hits above a.upper are edges, the others go to the
// recheck list; without it every hit is an edge stored with its r~.
template <int MODE, bool SPLIT = false, bool DEDIC = false>
__global__ __launch_bounds__(THREADS, 2) void gram2_kernel(GramArgs a)
{
    constexpr int TM = g2::TM, TN = g2::TN, WR = g2::WR, WC = g2::WC, BMW = g2::BMW, BNW = g2::BNW;
    constexpr int NSTAGE = g2::NSTAGE, STAGE = g2::STAGE, A_BYTES = g2::A_BYTES, B_BYTES = g2::B_BYTES;
    constexpr int DMA_WAVE = g2::DMA_WAVE;
    extern __shared__ __align__(16) char smem[];
    uint32_t *s_hist = (uint32_t *)(smem + NSTAGE * STAGE);
    __shared__ unsigned long long s_seg;
    // DEDIC (band-free emit only): hits staged in an area of their own, so
    // the epilogue needs no barrier for a ring slot to come free
    constexpr bool dedic = DEDIC && !SPLIT;
    __shared__ unsigned long long s_rec[4][g2::DEDIC_RECS];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    if (tid == 0) s_seg = 0;
    if (MODE == MODE_HIST)
        for (int b = tid; b < NBIN; b += THREADS) s_hist[b] = 0;
    const float lower = a.lower_dev ? *a.lower_dev : a.lower;
    // upper below -1: every hit is an edge (no recheck band)
    constexpr bool split = SPLIT;

    // per-lane operand offsets inside a stage (row l % 32, half l / 32)
    const int lr = lane & 31, lh = lane >> 5;
    const int opoff = lr * 32 + ((lh ^ g2_swz(lr)) << 4);
    const int aoff = (wm * WR) * 32 + opoff;
    const int boff = 2 * A_BYTES + (wn * WC) * 32 + opoff;

    TileWalk walk;
    walk.init(a);
    int2 *myseg = a.seg + (long long)blockIdx.x * a.seg_cap;
    const int K = a.kchunks;  // >= 2
    int R = 0, C = 0, R1 = 0, C1 = 0;
    bool have = walk.next(R, C);
    bool have1 = have && walk.next(R1, C1);
    if (have) {
        g2_dma(a, smem, R, C, 0);
        g2_dma(a, smem + STAGE, R, C, 1);
    }
    int slot = 0;
    G2Pending pend;
    pend.on = false;

    while (have) {
        f32x16 acc[BMW][BNW];
#pragma unroll
        for (int bm = 0; bm < BMW; bm++)
#pragma unroll
            for (int bn = 0; bn < BNW; bn++)
#pragma unroll
                for (int q = 0; q < 16; q++) acc[bm][bn][q] = 0.0f;

        for (int kc = 0; kc < K; kc++) {
            // chunk kc has landed once at most the operations issued after
            // it are outstanding (vector memory operations complete in
            // order): the next chunk's copies, and for chunks 0 and 1 the
            // previous epilogue's stores
            if (kc + 1 < K || have1) G2_WAIT_VM(DMA_WAVE);
            else G2_WAIT_VM(0);
            // the last tile's staged hits leave before the barrier after
            // which their slot is refilled
            if (MODE == MODE_EMIT && kc == 0 && pend.on) {
                g2_flush<SPLIT>(a, pend, lane, myseg);
                pend.on = false;
            }
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");  // no LDS access moves across the barrier
            // the slot two ahead was last read one step ago by every wave
            {
                int s2 = slot + 2;
                s2 -= s2 >= NSTAGE ? NSTAGE : 0;
                if (kc + 2 < K) g2_dma(a, smem + s2 * STAGE, R, C, kc + 2);
                else if (have1) g2_dma(a, smem + s2 * STAGE, R1, C1, kc + 2 - K);
            }
            const char *sp = smem + slot * STAGE;
            bf16x8 ah[BMW], al[BMW], bh[BNW], bl[BNW];
#pragma unroll
            for (int b = 0; b < BMW; b++) {
                ah[b] = *(const bf16x8 *)(sp + aoff + b * 32 * 32);
                al[b] = *(const bf16x8 *)(sp + A_BYTES + aoff + b * 32 * 32);
            }
#pragma unroll
            for (int b = 0; b < BNW; b++) {
                bh[b] = *(const bf16x8 *)(sp + boff + b * 32 * 32);
                bl[b] = *(const bf16x8 *)(sp + B_BYTES + boff + b * 32 * 32);
            }
            // same product order as the register-staged kernel: lo.hi, hi.lo, hi.hi
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[bm], bh[bn], acc[bm][bn], 0, 0, 0);
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bm], bl[bn], acc[bm][bn], 0, 0, 0);
#pragma unroll
            for (int bm = 0; bm < BMW; bm++)
#pragma unroll
                for (int bn = 0; bn < BNW; bn++)
                    acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bm], bh[bn], acc[bm][bn], 0, 0, 0);
            slot = slot + 1 == NSTAGE ? 0 : slot + 1;
        }

        // the slot of the tile's last chunk is free once every wave has
        // multiplied it (the next tile's chunks 0 and 1 are in the other two)
        char *stage = nullptr;
        if (MODE == MODE_EMIT && dedic) {
            stage = (char *)s_rec[w];
        } else if (MODE == MODE_EMIT) {
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            int fs = slot + NSTAGE - 1;
            fs -= fs >= NSTAGE ? NSTAGE : 0;
            stage = smem + fs * STAGE + w * (STAGE / 4);
        }
        g2_tile_epilogue<MODE, dedic>(a, acc, R, C, wm, wn, lane, lower, split, myseg, &s_seg, s_hist, stage, pend);
        R = R1;
        C = C1;
        have = have1;
        if (have) have1 = walk.next(R1, C1);
    }
    if (MODE == MODE_EMIT && pend.on) g2_flush<SPLIT>(a, pend, lane, myseg);
    G2_WAIT_VM(0);
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // the histogram's inline-asm LDS adds
    __syncthreads();
    if (MODE == MODE_HIST) {
        uint32_t *out = a.hist + (long long)blockIdx.x * NBIN;
        for (int b = tid; b < NBIN; b += THREADS) out[b] = s_hist[b];
    } else if (tid == 0) {
        a.seg_cnt[blockIdx.x] = s_seg;
    }
}

// Sum the per-work-group histograms: 32 bins per block, 8 row groups each.
__global__ __launch_bounds__(256) void hist_reduce(const uint32_t *__restrict__ h, int nwg,
                                                   unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long part[8][32];
    const int bl = threadIdx.x & 31, rg = threadIdx.x >> 5;
    const int b = blockIdx.x * 32 + bl;
    unsigned long long s = 0;
    int g = rg;
    for (; g + 8 * 15 < nwg; g += 8 * 16) {  // 16 rows in flight per thread
        unsigned v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = h[(long long)(g + 8 * u) * NBIN + b];
#pragma unroll
        for (int u = 0; u < 16; u++) s += v[u];
    }
    for (; g < nwg; g += 8) s += h[(long long)g * NBIN + b];
    part[rg][bl] = s;
    __syncthreads();
    if (rg == 0) {
        for (int q = 1; q < 8; q++) s += part[q][bl];
        out[b] = s;
    }
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
    if ((t & 1) == 0) {  // rows 16-byte aligned: paired loads, same order of sums
        const double2 *a2 = (const double2 *)a, *b2 = (const double2 *)b;
#pragma unroll 8
        for (int l = 0; l < t / 2; l++) {
            const double2 x = a2[l], y = b2[l];
            const double p0 = x.x * y.x;
            s += p0;
            const double p1 = x.y * y.y;
            s += p1;
        }
    } else {
        for (int l = 0; l < t; l++) {
            double p = a[l] * b[l];
            s += p;
        }
    }
    return s > 1.0 ? 1.0 : (s < -1.0 ? -1.0 : s);
}

// Device-side decisions of the single-pass sparsity selection, so the pass
// runs without host round trips.
struct SelCtl {
    float lower;   // emit bound L
    float L2, U2;  // FP64 candidate window on r~
    int ok;        // cleared by any step whose assumption fails
};

// Selection histogram: SEL_NB bins of the candidates' FP64 r over the window
// widened by the band (r64 lies in [L2 - band, U2 + band]); bin 0 holds the
// largest r.  Built by the recheck, read by select_kernel.
constexpr int SEL_NB = 4096, SEL_CAP = 2048;
struct SelHist {
    unsigned *h;  // SEL_NB counts, zeroed before the recheck; nullptr: none
    const SelCtl *ctl;
    double band;
};
__device__ __forceinline__ void sel_range(const SelHist &s, double &top, double &scale)
{
    const double lo = (double)s.ctl->L2 - s.band, hi = (double)s.ctl->U2 + s.band;
    top = hi;
    scale = hi > lo ? (double)SEL_NB / (hi - lo) : 1.0;
}
// The cut a selection found: the selected candidates are those whose
// (key, (i, j)) is at most (k1, k2); ok = 0 when the selection failed.
struct SelCut {
    unsigned long long k1, k2;
    double ts;
    long long c;
    int ok;
};

// clamp(floor((top - r) * scale)): non-decreasing as r decreases, i.e. in the
// selection key (r descending), so the bins cut the key order into ranges
__device__ __forceinline__ int sel_bin(double r, double top, double scale)
{
    const double x = (top - r) * scale;
    return x < 1.0 ? 0 : (x >= (double)(SEL_NB - 1) ? SEL_NB - 1 : (int)x);
}

// FP64 r of every listed pair; with rfast, also the largest |r~ - r64| over
// them into *errmax (bit pattern of a non-negative double, whose unsigned
// order is its numeric order), one atomic per wave.
__global__ void recheck_kernel(const int2 *__restrict__ pairs, const unsigned long long *__restrict__ cnt,
                               long long cap, const double *__restrict__ z, int t,
                               double *__restrict__ r_out, const float *__restrict__ rfast,
                               unsigned long long *__restrict__ errmax, SelHist sh)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long c = (long long)min((unsigned long long)cap, *cnt);
    double err = 0.0;
    if (k < c) {
        int2 p = pairs[k];
        const double r = r64_pair(z, t, p.x, p.y);
        r_out[k] = r;
        if (rfast) err = fabs((double)rfast[k] - r);
        if (sh.h) {
            double top, scale;
            sel_range(sh, top, scale);
            atomicAdd(&sh.h[sel_bin(r, top, scale)], 1u);
        }
    }
    if (!rfast) return;
    for (int o = 32; o > 0; o >>= 1) err = fmax(err, __shfl_xor(err, o));
    if ((threadIdx.x & 63) == 0 && err > 0.0)
        atomicMax(errmax, (unsigned long long)__double_as_longlong(err));
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

// The same FP64 r, RCK_P pairs per wave: the wave walks its pairs' rows
// RCK_T time points at a time; in load m a lane reads element
// lane % RCK_T of pair m * (64 / RCK_T) + lane / RCK_T, so every load covers
// whole RCK_T * 8-byte runs of a few rows (a thread-per-pair loop touches 64
// rows per load) and a pair's t points take t / RCK_T dependent rounds.  The
// rounded products go through LDS and lane p < RCK_P adds pair p's in time
// order: the same products and the same sequential sum as r64_pair.  Row
// offsets i * t in 32 bits (n * t < 2^32, checked by the caller).  The loop
// bounds are uniform over the block.
constexpr int RCK_P = 64, RCK_T = 1024 / RCK_P, RCK_M = 16;
__global__ __launch_bounds__(256) void recheck_rows(const int2 *__restrict__ pairs,
                                                    const unsigned long long *__restrict__ cnt, long long cap,
                                                    const double *__restrict__ z, int t, double *__restrict__ r_out,
                                                    const float *__restrict__ rfast,
                                                    unsigned long long *__restrict__ errmax, SelHist sh)
{
#pragma clang fp contract(off)
    __shared__ double prod[4][RCK_P][RCK_T + 1];
    constexpr int PPL = 64 / RCK_T;  // pairs per load
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lt = lane % RCK_T, lp = lane / RCK_T;
    const long long c = (long long)min((unsigned long long)cap, *cnt);
    double err = 0.0;
    for (long long b0 = (long long)blockIdx.x * 4 * RCK_P; b0 < c; b0 += (long long)gridDim.x * 4 * RCK_P) {
        const long long q = b0 + w * RCK_P + lane;
        const int2 e = lane < RCK_P && q < c ? pairs[q] : make_int2(0, 0);
        // row offsets of the pairs this lane loads for
        unsigned oa[RCK_M], ob[RCK_M];
#pragma unroll
        for (int m = 0; m < RCK_M; m++) {
            const int src = m * PPL + lp;
            oa[m] = (unsigned)__shfl(e.x, src) * (unsigned)t;
            ob[m] = (unsigned)__shfl(e.y, src) * (unsigned)t;
        }
        double sum = 0.0;
        for (int l0 = 0; l0 < t; l0 += RCK_T) {
            const int l = l0 + lt;
            const bool in = l < t;
            double x[RCK_M], y[RCK_M];
#pragma unroll
            for (int m = 0; m < RCK_M; m++) {
                x[m] = in ? z[oa[m] + (unsigned)l] : 0.0;
                y[m] = in ? z[ob[m] + (unsigned)l] : 0.0;
            }
#pragma unroll
            for (int m = 0; m < RCK_M; m++) prod[w][m * PPL + lp][lt] = x[m] * y[m];
            __syncthreads();
            if (lane < RCK_P) {
                const int kt = min(RCK_T, t - l0);
                for (int k = 0; k < kt; k++) {
                    const double pk = prod[w][lane][k];
                    sum += pk;
                }
            }
            __syncthreads();
        }
        if (lane < RCK_P && q < c) {
            const double r = sum > 1.0 ? 1.0 : (sum < -1.0 ? -1.0 : sum);
            r_out[q] = r;
            if (rfast) err = fmax(err, fabs((double)rfast[q] - r));
            if (sh.h) {
                double top, scale;
                sel_range(sh, top, scale);
                atomicAdd(&sh.h[sel_bin(r, top, scale)], 1u);
            }
        }
    }
    if (!rfast) return;
    for (int o = 32; o > 0; o >>= 1) err = fmax(err, __shfl_xor(err, o));
    if (lane == 0 && err > 0.0) atomicMax(errmax, (unsigned long long)__double_as_longlong(err));
}

// FP64 r of the first min(cap, *cnt) listed pairs (recheck_rows when the row
// offsets fit 32 bits; grid of at most 8 work-groups per CU, bound by cap).
// sh.h: also the selection histogram of the r (zeroed by the caller).
inline void launch_recheck(const pg_ctx *ctx, hipStream_t st, const int2 *pairs, const unsigned long long *cnt,
                           long long cap, const double *z, int n, int t, double *r_out, const float *rfast,
                           unsigned long long *errmax, SelHist sh = SelHist{nullptr, nullptr, 0.0})
{
    if (cap <= 0) return;
    if ((unsigned long long)n * (unsigned long long)t < (1ull << 32)) {
        const long long g = std::min<long long>(pg_cdiv(cap, 4 * RCK_P), 8LL * std::max(ctx->num_cu, 1));
        recheck_rows<<<(unsigned)g, 256, 0, st>>>(pairs, cnt, cap, z, t, r_out, rfast, errmax, sh);
    } else {
        recheck_kernel<<<(unsigned)pg_cdiv(cap, 256), 256, 0, st>>>(pairs, cnt, cap, z, t, r_out, rfast, errmax,
                                                                     sh);
    }
}

// Sampled histogram pass: every stride-th supertile of the nst listed (at
// most 1 in PG_SAMPLE_MAX, at least 16 sampled; 1 = no sampled pass).
#ifndef PG_SAMPLE_MAX
#define PG_SAMPLE_MAX 100
#endif
inline long long sample_stride(long long nst)
{
    return nst >= 8 ? std::max<long long>(2, std::min<long long>(PG_SAMPLE_MAX, nst / 16)) : 1;
}

// Whether a band-free emit with hit fraction d stages in the dedicated area:
// a wave's block holds 8192 pairs, so about 8192 d hits; at most 160 keeps
// six standard deviations under the area's 255 records (a block beyond it
// writes its hits directly, correct but slower).
inline bool dedic_fits(double d) { return d * (g2::WR * g2::WC) <= 160.0; }

// ------------------------------------------------------------- CSR build

struct EdgeSrc {
    const int2 *seg;
    long long seg_cap;
    const unsigned long long *seg_cnt;
    int nseg;
    const int2 *extra;
    const unsigned long long *extra_cnt;
    const float *seg_r;   // optional filter: keep segment entries with r~ > keep_above
    float keep_above;
    const float *keep_above_dev;  // if set, the bound is read from device memory
    long long extra_cap;
};

__device__ __forceinline__ float src_bound(const EdgeSrc &s)
{
    return s.keep_above_dev ? *s.keep_above_dev : s.keep_above;
}

// entries of source g, clamped to what its buffer holds
__device__ __forceinline__ long long src_count(const EdgeSrc &s, int g)
{
    return g < s.nseg ? min((long long)s.seg_cnt[g], s.seg_cap) : min((long long)*s.extra_cnt, s.extra_cap);
}
__device__ __forceinline__ int2 src_edge(const EdgeSrc &s, int g, long long k)
{
    return g < s.nseg ? s.seg[(long long)g * s.seg_cap + k] : s.extra[k];
}

// Two-level counting sort into CSR.  Both directions of every kept edge go to
// the bucket of BROWS rows holding their source row: per-segment bucket counts
// in LDS, a 2-D exclusive scan (over the segments inside each bucket) that fixes
// every segment's output range, a scatter of packed 32-bit entries
// (row-in-bucket << 25 | column), then one work-group per bucket derives the
// row offsets and sorts each row by LDS-bitmap rank.  No global atomics.
#ifndef PG_CSR_BSH
#define PG_CSR_BSH 4
#endif
constexpr int BSH = PG_CSR_BSH;  // log2 of the rows per bucket
constexpr int BROWS = 1 << BSH;
constexpr unsigned COLMASK = (1u << 25) - 1u;

#ifndef PG_SEG_UNROLL
#define PG_SEG_UNROLL 8
#endif
#ifndef PG_SEG_PIPE
#define PG_SEG_PIPE 1
#endif
constexpr int SEG_THREADS = 1024;          // one work-group per segment: 16 waves in flight
constexpr int SEG_UNROLL = PG_SEG_UNROLL;  // segment entries per thread per batch

// One thread's batch of source g: entries k0 + u * SEG_THREADS, their r~ when
// the source is filtered (else NaN, which every keep test passes).
struct SegBatch {
    int2 e[SEG_UNROLL];
    float r[SEG_UNROLL];
    bool in[SEG_UNROLL];
    __device__ __forceinline__ void fetch(const EdgeSrc &s, int g, long long k0, long long c)
    {
        const bool filt = s.seg_r && g < s.nseg;
#pragma unroll
        for (int u = 0; u < SEG_UNROLL; u++) {
            const long long k = k0 + (long long)u * SEG_THREADS;
            in[u] = k < c;
            e[u] = in[u] ? src_edge(s, g, k) : make_int2(0, 0);
            r[u] = in[u] && filt ? s.seg_r[(long long)g * s.seg_cap + k] : __int_as_float(0x7fc00000);
        }
    }
    __device__ __forceinline__ bool keep(int u, float ka) const { return in[u] && !(r[u] <= ka); }
};

// f(batch) over this thread's batches of source g; pipelined: the next
// batch's loads are issued before f works on the current one.
template <class F>
__device__ __forceinline__ void seg_batches(const EdgeSrc &s, int g, long long c, F &&f)
{
    const long long step = (long long)SEG_THREADS * SEG_UNROLL;
#if PG_SEG_PIPE
    SegBatch a;
    a.fetch(s, g, threadIdx.x, c);
    for (long long k0 = threadIdx.x; k0 < c; k0 += step) {
        SegBatch b;
        b.fetch(s, g, k0 + step, c);
        f(a);
        a = b;
    }
#else
    for (long long k0 = threadIdx.x; k0 < c; k0 += step) {
        SegBatch a;
        a.fetch(s, g, k0, c);
        f(a);
    }
#endif
}

// Work-group x counts source g0 + x; counts of source g land in row g of the
// segment-major table cnt[nsrc][nbk].
__global__ __launch_bounds__(SEG_THREADS) void bucket_count(EdgeSrc s, int nbk, int g0, int nsrc,
                                                            unsigned *__restrict__ cnt)
{
    extern __shared__ unsigned sc[];
    for (int b = threadIdx.x; b < nbk; b += SEG_THREADS) sc[b] = 0u;
    __syncthreads();
    const int g = g0 + blockIdx.x;
    const long long c = src_count(s, g);
    const float ka = src_bound(s);
    seg_batches(s, g, c, [&](const SegBatch &a) {
#pragma unroll
        for (int u = 0; u < SEG_UNROLL; u++)
            if (a.keep(u, ka)) {
                atomicAdd(&sc[a.e[u].x >> BSH], 1u);
                atomicAdd(&sc[a.e[u].y >> BSH], 1u);
            }
    });
    __syncthreads();
    for (int b = threadIdx.x; b < nbk; b += SEG_THREADS) cnt[(long long)g * nbk + b] = sc[b];
}

// The window split and the segments' bucket counts in one pass over the
// segments (one work-group each, as bucket_count): r~ > U2 (more exactly,
// not r~ <= U2: the CSR filter's test) are definite edges, counted into the
// segment's row of cnt (both directions) and its definite total dpart[g];
// L2 <= r~ <= U2 go to the candidate list, staged in LDS and reserved once
// per work-group (beyond WC_STAGE, one by one).  Without a window
// (ctl->ok == 0) nothing is kept.  Does the work of window_kernel + bucket_count over
// the segments; work-group 0 clears the selected list's row (row nseg), which
// select_write counts as it writes the list.
constexpr int WC_STAGE = 1024;
__global__ __launch_bounds__(SEG_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void window_count(EdgeSrc s, const SelCtl *__restrict__ ctl, int nbk,
                                                            int nsrc, unsigned *__restrict__ cnt,
                                                            int2 *__restrict__ cand, float *__restrict__ cand_r,
                                                            long long cand_cap,
                                                            unsigned long long *__restrict__ cand_cnt,
                                                            unsigned long long *__restrict__ dpart)
{
    extern __shared__ unsigned sc[];  // [nbk]
    __shared__ int2 s_ce[WC_STAGE];
    __shared__ float s_cr[WC_STAGE];
    __shared__ unsigned s_nc;
    __shared__ unsigned long long s_nd, s_base;
    const int tid = threadIdx.x, g = blockIdx.x;
    for (int b = tid; b < nbk; b += SEG_THREADS) sc[b] = 0u;
    if (g == 0)  // the selected list's row, counted by select_write
        for (int b = tid; b < nbk; b += SEG_THREADS) cnt[(long long)gridDim.x * nbk + b] = 0u;
    if (tid == 0) {
        s_nc = 0u;
        s_nd = 0ull;
    }
    __syncthreads();
    const long long c = ctl->ok ? src_count(s, g) : 0;
    const float lo = ctl->L2, hi = ctl->U2;
    unsigned nd = 0;
    seg_batches(s, g, c, [&](const SegBatch &a) {
#pragma unroll
        for (int u = 0; u < SEG_UNROLL; u++) {
            if (!a.in[u]) continue;
            const float r = a.r[u];
            const int2 e = a.e[u];
            if (!(r <= hi)) {
                atomicAdd(&sc[e.x >> BSH], 1u);
                atomicAdd(&sc[e.y >> BSH], 1u);
                nd++;
            } else if (r >= lo) {
                const unsigned p = atomicAdd(&s_nc, 1u);
                if (p < (unsigned)WC_STAGE) {
                    s_ce[p] = e;
                    s_cr[p] = r;
                } else {
                    const unsigned long long q = atomicAdd(cand_cnt, 1ull);
                    if ((long long)q < cand_cap) {
                        cand[q] = e;
                        cand_r[q] = r;
                    }
                }
            }
        }
    });
    unsigned long long ndw = nd;
    for (int o = 32; o > 0; o >>= 1) ndw += __shfl_xor(ndw, o);
    if ((tid & 63) == 0 && ndw) atomicAdd(&s_nd, ndw);
    __syncthreads();
    for (int b = tid; b < nbk; b += SEG_THREADS) cnt[(long long)g * nbk + b] = sc[b];
    const unsigned nl = min(s_nc, (unsigned)WC_STAGE);
    if (tid == 0) {
        dpart[g] = s_nd;
        s_base = nl ? atomicAdd(cand_cnt, (unsigned long long)nl) : 0ull;
    }
    __syncthreads();
    for (unsigned k = tid; k < nl; k += SEG_THREADS) {
        const unsigned long long q = s_base + k;
        if ((long long)q < cand_cap) {
            cand[q] = s_ce[k];
            cand_r[q] = s_cr[k];
        }
    }
}

// Per bucket b: tot[b] and the within-bucket exclusive scan over segments,
// off[g][b] = sum_{g' < g} cnt[g'][b] (counts and offsets segment-major, so
// the count kernels write and the scatter reads whole rows).  One block per
// 32 consecutive buckets: lane b % 32 of each wave sums its bucket over one
// of SCAN_SG runs of segments (32 buckets' counts of a segment are one
// 128-byte read), the partial sums are scanned per bucket, then written back.
constexpr int SCAN_BK = 32, SCAN_SG = 32, SCAN_PER = 24;
__global__ __launch_bounds__(SCAN_BK * SCAN_SG) void bucket_scan_local(const unsigned *__restrict__ cnt, int nsrc,
                                                                         int nbk, long long *__restrict__ off,
                                                                         long long *__restrict__ tot)
{
    __shared__ long long part[SCAN_SG][SCAN_BK];
    const int tid = threadIdx.x, bl = tid & (SCAN_BK - 1), sg = tid / SCAN_BK;
    const int b = blockIdx.x * SCAN_BK + bl;
    const int per = (nsrc + SCAN_SG - 1) / SCAN_SG;
    const int g0 = sg * per, g1 = min(nsrc, g0 + per);
    // runs of at most SCAN_PER segments are held in registers between the two
    // passes (all their reads in flight at once); longer runs read twice
    const bool held = per <= SCAN_PER;
    unsigned v[SCAN_PER];
    long long sum = 0;
    if (held) {
#pragma unroll
        for (int u = 0; u < SCAN_PER; u++)
            v[u] = b < nbk && g0 + u < g1 ? cnt[(long long)(g0 + u) * nbk + b] : 0u;
#pragma unroll
        for (int u = 0; u < SCAN_PER; u++) sum += v[u];
    } else if (b < nbk) {
#pragma unroll 8
        for (int g = g0; g < g1; g++) sum += cnt[(long long)g * nbk + b];
    }
    part[sg][bl] = sum;
    __syncthreads();
    long long run = 0;
    for (int q = 0; q < sg; q++) run += part[q][bl];
    if (b < nbk) {
        if (held) {
#pragma unroll
            for (int u = 0; u < SCAN_PER; u++)
                if (g0 + u < g1) {
                    off[(long long)(g0 + u) * nbk + b] = run;
                    run += v[u];
                }
        } else {
            for (int g = g0; g < g1; g++) {
                off[(long long)g * nbk + b] = run;
                run += cnt[(long long)g * nbk + b];
            }
        }
        if (sg == SCAN_SG - 1) tot[b] = run;
    }
}

// bstart = exclusive scan of the bucket totals (64-bit: a selection that did
// not hold can count more than INT32_MAX entries); one wave scan per wave,
// then each wave adds the totals of the waves before it.
__global__ __launch_bounds__(1024) void bucket_scan_global(const long long *__restrict__ tot, int nbk,
                                                           long long *__restrict__ bstart)
{
    __shared__ long long wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int per = (nbk + 1023) / 1024;
    const int b0 = tid * per, b1 = min(nbk, b0 + per);
    long long s = 0;
    for (int b = b0; b < b1; b++) s += tot[b];
    long long incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    long long run = incl - s;
    for (int q = 0; q < w; q++) run += wsum[q];
    for (int b = b0; b < b1; b++) {
        bstart[b] = run;
        run += tot[b];
    }
    if (tid == 1023) bstart[nbk] = run;
}

// Entries beyond `cap` (a selection that did not hold; the caller checks
// row[n] and rebuilds) are dropped rather than written.
// W: the FP64 r of every entry goes along (weighted CSR).
template <bool W>
__global__ __launch_bounds__(SEG_THREADS) void bucket_scatter(EdgeSrc s, int nbk, const long long *__restrict__ off,
                                                              const long long *__restrict__ bstart,
                                                              unsigned *__restrict__ ent, float *__restrict__ entv,
                                                              const double *__restrict__ z, int t, long long cap)
{
    // per-bucket: this segment's output base (bstart + off; below cap, which
    // is at most INT32_MAX, whenever anything is written) and a cursor
    extern __shared__ unsigned sbase[];
    unsigned *sc = sbase + nbk;
    const int g = blockIdx.x, nsrc = gridDim.x;
    for (int b = threadIdx.x; b < nbk; b += SEG_THREADS) {
        sbase[b] = (unsigned)(bstart[b] + off[(long long)g * nbk + b]);
        sc[b] = 0u;
    }
    __syncthreads();
    const long long c = src_count(s, g);
    const float ka = src_bound(s);
    if (bstart[nbk] > cap) return;
    seg_batches(s, g, c, [&](const SegBatch &a) {
#pragma unroll
        for (int u = 0; u < SEG_UNROLL; u++) {
            if (!a.keep(u, ka)) continue;
            const int2 e = a.e[u];
            const int bf = e.x >> BSH, br = e.y >> BSH;
            const unsigned pf = sbase[bf] + atomicAdd(&sc[bf], 1u);
            const unsigned pr = sbase[br] + atomicAdd(&sc[br], 1u);
            ent[pf] = ((unsigned)(e.x & (BROWS - 1)) << 25) | (unsigned)e.y;
            ent[pr] = ((unsigned)(e.y & (BROWS - 1)) << 25) | (unsigned)e.x;
            if (W) {
                const float r = (float)r64_pair(z, t, e.x, e.y);
                entv[pf] = r;
                entv[pr] = r;
            }
        }
    });
}

// Bitonic sort of one row held lane-major in a wave (element e = lane*P + r),
// M = 64*P keys, ascending; padding keys sort last.  Key type: 32-bit column
// or 64-bit (column << 32 | value bits) to carry the weight along.
constexpr int ROWS_THREADS = 1024;
constexpr int ROWS_UNROLL = 8;  // bucket entries per thread in flight
constexpr int SORT_MAX = 64 * 16;  // longest row sorted in registers

// LDS of one bitmap wave of bucket_rows: a bitmap over the columns [0, n) in
// groups of 128 (16 bytes) and one 32-bit prefix count per group.
__host__ __device__ inline long long rows_groups(int n) { return ((long long)n + 127) / 128; }

// Place the d = z1 - a entries of one row (columns in ctmp, unsorted, no
// repeats) in column order by rank: set one bit per column, prefix-count the
// groups (each lane a contiguous run of `per` groups, `per` odd so the
// 16-byte reads of a wave spread over the banks), then write every entry at
// its rank; the bits are cleared again by the entries.  Rows of at most
// 64 * HOLD entries are held in registers across the three passes.
__device__ __forceinline__ void bitmap_prefix(const uint4 *bits4, unsigned *gpre, long long ng, int per, int lane)
{
    const long long g0 = (long long)lane * per;
    unsigned s = 0;
    for (int q = 0; q < per; q++) {
        const long long g = g0 + q;
        if (g < ng) {
            const uint4 v = bits4[g];
            gpre[g] = s;
            s += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
        }
    }
    unsigned incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const unsigned base = incl - s;
    if (base)
        for (int q = 0; q < per; q++) {
            const long long g = g0 + q;
            if (g < ng) atomicAdd(&gpre[g], base);
        }
}

// rank of column c among the set bits
__device__ __forceinline__ unsigned bitmap_rank(const uint4 *bits4, const unsigned *gpre, int c)
{
    const int g = c >> 7;
    const uint4 v = bits4[g];
    const int w = (c >> 5) & 3;
    unsigned pos = gpre[g];
    pos += w > 0 ? __popc(v.x) : 0u;
    pos += w > 1 ? __popc(v.y) : 0u;
    pos += w > 2 ? __popc(v.z) : 0u;
    const unsigned word = w == 0 ? v.x : (w == 1 ? v.y : (w == 2 ? v.z : v.w));
    return pos + __popc(word & ((1u << (c & 31)) - 1u));
}

template <int HOLD>
__device__ __forceinline__ void rank_row_bitmap(const int *__restrict__ ctmp, const float *__restrict__ vtmp,
                                                long long a, long long z1, int lane, uint4 *bits4, unsigned *gpre,
                                                long long ng, int per, int *__restrict__ col,
                                                float *__restrict__ val)
{
    unsigned *bits = (unsigned *)bits4;
    if (z1 - a <= 64 * HOLD) {
        int hc[HOLD];
#pragma unroll
        for (int r = 0; r < HOLD; r++) {
            const long long e = a + r * 64 + lane;
            hc[r] = e < z1 ? ctmp[e] : -1;
        }
#pragma unroll
        for (int r = 0; r < HOLD; r++)
            if (hc[r] >= 0) atomicOr(&bits[hc[r] >> 5], 1u << (hc[r] & 31));
        __builtin_amdgcn_wave_barrier();
        bitmap_prefix(bits4, gpre, ng, per, lane);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < HOLD; r++)
            if (hc[r] >= 0) {
                const unsigned pos = bitmap_rank(bits4, gpre, hc[r]);
                col[a + pos] = hc[r];
                if (val) val[a + pos] = vtmp[a + r * 64 + lane];
            }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < HOLD; r++)
            if (hc[r] >= 0) bits[hc[r] >> 5] = 0u;
    } else {
        for (long long e = a + lane; e < z1; e += 64) {
            const int c = ctmp[e];
            atomicOr(&bits[c >> 5], 1u << (c & 31));
        }
        __builtin_amdgcn_wave_barrier();
        bitmap_prefix(bits4, gpre, ng, per, lane);
        __builtin_amdgcn_wave_barrier();
        for (long long e = a + lane; e < z1; e += 64) {
            const int c = ctmp[e];
            const unsigned pos = bitmap_rank(bits4, gpre, c);
            col[a + pos] = c;
            if (val) val[a + pos] = vtmp[e];
        }
        __builtin_amdgcn_wave_barrier();
        for (long long e = a + lane; e < z1; e += 64) bits[ctmp[e] >> 5] = 0u;
    }
    __builtin_amdgcn_wave_barrier();
}

// One work-group per bucket: row counts -> row offsets, scatter into the
// rows, then the rows are sorted: rows of more than `dbit` entries by bitmap
// rank on the first nbm waves (each with its own bitmap), the others by a
// register bitonic sort on every wave.  W: weights travel with the columns.
template <bool W>
__global__ __launch_bounds__(ROWS_THREADS, W ? 4 : 8) void bucket_rows(const long long *__restrict__ bstart,
                                                               const unsigned *__restrict__ ent,
                                                               const float *__restrict__ entv, int n, int nbk,
                                                               int nbm, int dbit, int *__restrict__ row,
                                                               int *__restrict__ ctmp, float *__restrict__ vtmp,
                                                               int *__restrict__ col, float *__restrict__ val,
                                                               long long cap)
{
    if (!W) {
        entv = nullptr;
        vtmp = nullptr;
        val = nullptr;
    }
    extern __shared__ __align__(16) uint4 sm4[];
    __shared__ unsigned rc[BROWS], rb[BROWS + 1];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int NWAVE = ROWS_THREADS / 64;
    if (bstart[nbk] > cap) {  // nothing was scattered; row[n] reports the total
        if (b == nbk - 1 && tid == 0) row[n] = (int)min(bstart[nbk], (long long)INT_MAX);
        return;
    }
    const long long ng = rows_groups(n);
    const int per = (int)(((ng + 63) / 64) | 1);
    uint4 *bits4 = sm4 + (long long)wv * (ng + (ng + 3) / 4);  // rows_wave_u4
    unsigned *gpre = (unsigned *)(bits4 + ng);
    if (wv < nbm)
        for (long long q = lane; q < ng; q += 64) bits4[q] = make_uint4(0u, 0u, 0u, 0u);
    const long long e0 = bstart[b], e1 = bstart[b + 1];
    if (tid < BROWS) rc[tid] = 0u;
    __syncthreads();
    for (long long k0 = e0 + tid; k0 < e1; k0 += (long long)ROWS_THREADS * ROWS_UNROLL) {
        unsigned x[ROWS_UNROLL];
#pragma unroll
        for (int u = 0; u < ROWS_UNROLL; u++) {
            const long long k = k0 + (long long)u * ROWS_THREADS;
            x[u] = k < e1 ? ent[k] : ~0u;
        }
#pragma unroll
        for (int u = 0; u < ROWS_UNROLL; u++)
            if (x[u] != ~0u) atomicAdd(&rc[x[u] >> 25], 1u);
    }
    __syncthreads();
    if (tid < 64) {
        // exclusive scan of the BROWS row counts by one wave
        constexpr int PL = (BROWS + 63) / 64;
        unsigned own = 0;
#pragma unroll
        for (int q = 0; q < PL; q++) own += tid * PL + q < BROWS ? rc[tid * PL + q] : 0u;
        unsigned incl = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned y = __shfl_up(incl, o);
            if (tid >= o) incl += y;
        }
        unsigned run = incl - own;
#pragma unroll
        for (int q = 0; q < PL; q++)
            if (tid * PL + q < BROWS) {
                rb[tid * PL + q] = run;
                run += rc[tid * PL + q];
            }
        if (tid == 63) rb[BROWS] = incl;
    }
    __syncthreads();
    const int r0 = b * BROWS;
    if (tid < BROWS && r0 + tid < n) row[r0 + tid] = (int)(e0 + rb[tid]);
    if (b == nbk - 1 && tid == 0) row[n] = (int)e1;
    if (tid < BROWS) rc[tid] = 0u;
    __syncthreads();
    for (long long k0 = e0 + tid; k0 < e1; k0 += (long long)ROWS_THREADS * ROWS_UNROLL) {
        unsigned x[ROWS_UNROLL];
        float v[ROWS_UNROLL];
#pragma unroll
        for (int u = 0; u < ROWS_UNROLL; u++) {
            const long long k = k0 + (long long)u * ROWS_THREADS;
            x[u] = k < e1 ? ent[k] : ~0u;
            v[u] = (vtmp && k < e1) ? entv[k] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < ROWS_UNROLL; u++) {
            if (x[u] == ~0u) continue;
            const unsigned lr = x[u] >> 25;
            const long long p = e0 + rb[lr] + atomicAdd(&rc[lr], 1u);
            ctmp[p] = (int)(x[u] & COLMASK);
            if (vtmp) vtmp[p] = v[u];
        }
    }
    __syncthreads();
    // short rows on every wave: wave wv takes rows wv, wv + NWAVE, ...
    for (int lr = wv; lr < BROWS && r0 + lr < n; lr += NWAVE) {
        const long long a = e0 + rb[lr];
        const int d = (int)(rb[lr + 1] - rb[lr]);
        if (d == 0 || d > dbit) continue;
        if (d <= 64) pgdev::sort_row_regs<1>(ctmp, vtmp, a, d, lane, col, val);
        else if (d <= 128) pgdev::sort_row_regs<2>(ctmp, vtmp, a, d, lane, col, val);
        else if (d <= 256) pgdev::sort_row_regs<4>(ctmp, vtmp, a, d, lane, col, val);
        else if (d <= 512) pgdev::sort_row_regs<8>(ctmp, vtmp, a, d, lane, col, val);
        else pgdev::sort_row_regs<16>(ctmp, vtmp, a, d, lane, col, val);
    }
    // long rows on the bitmap waves wv < nbm
    if (wv >= nbm) return;
    for (int lr = wv; lr < BROWS && r0 + lr < n; lr += nbm) {
        const long long a = e0 + rb[lr], z1 = e0 + rb[lr + 1];
        if (z1 - a <= dbit) continue;
        rank_row_bitmap<W ? 8 : 16>(ctmp, vtmp, a, z1, lane, bits4, gpre, ng, per, col, val);
    }
}

// Unweighted rows of a bucket from one LDS bitmap per row over the columns
// [0, n): every entry sets its bit and a row's degree is the popcount of its
// bitmap, so the rows come out sorted with one read of the entries and one
// write of the columns.  Work-groups loop over buckets; the bitmap is zero
// between buckets.  Wave r counts row r: each lane owns the run of 128-column
// groups [lane * ng / 64, (lane + 1) * ng / 64) of the row.  Then, with RANK
// (the bitmaps and one 32-bit base per group fit LDS), every group's base
// (its first column's place in its row) goes to LDS with the row's count,
// and every entry's place is its row's start (a scan of the row counts in
// lanes 0..15) + its group's base + the set bits below it in the group, so
// the bucket's work is spread evenly over the work-group; without RANK each
// lane reads its run's columns back in order, one per step.  The first
// BITS_PF entries per thread of a work-group's next bucket are loaded before
// the current one is written out, so their HBM latency is spent under it;
// the loads are clamped to the bucket's last entry (a repeated entry sets a
// bit that is already set) and only the uses are masked.
constexpr int BITS_THREADS = 64 * BROWS;
constexpr int BITS_UNROLL = 8;
constexpr int BITS_PF = 12;
// longest run of groups per lane: the BROWS bitmaps fit 160 KB of LDS only
// while ng <= 64 * BITS_RUN
constexpr int BITS_RUN = 10;
static_assert((size_t)BROWS * 16 * (64 * BITS_RUN + 1) + 64 > 160 * 1024, "a lane's run can exceed BITS_RUN groups");
// with RANK the bitmaps and the group bases (20 bytes per group and row) fit
// only while ng <= 64 * BITS_RUN_RANK
constexpr int BITS_RUN_RANK = 8;
static_assert((size_t)BROWS * 20 * (64 * BITS_RUN_RANK + 1) + 64 > 160 * 1024,
              "a lane's run can exceed BITS_RUN_RANK groups");

// Inclusive sum over the 64 lanes of a wave (DPP: shifts within each row of
// 16 lanes, then the row totals broadcast into the rows above); every lane
// of the wave must be active.
__device__ __forceinline__ unsigned wave_incl_sum(unsigned x)
{
    x += (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1, 3
    x += (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2, 3
    return x;
}

__device__ __forceinline__ void bits_set(unsigned *bm, unsigned ng, unsigned x)
{
    const unsigned c = x & COLMASK;
    atomicOr(&bm[(x >> 25) * ng * 4u + (c >> 5)], 1u << (c & 31));
}

// Entry x's place in its row: its group's base within the row plus the set
// bits below its column in the group.
__device__ __forceinline__ unsigned bits_rank(const uint4 *bm4, const unsigned *gbase, unsigned ng, unsigned x)
{
    const unsigned c = x & COLMASK;
    const unsigned gi = (x >> 25) * ng + (c >> 7);
    const uint4 v = bm4[gi];
    const unsigned w = (c >> 5) & 3u, m = (1u << (c & 31)) - 1u;
    unsigned pos = gbase[gi];
    pos += __popc(v.x & (w > 0u ? ~0u : m));
    pos += __popc(v.y & (w > 1u ? ~0u : (w == 1u ? m : 0u)));
    pos += __popc(v.z & (w > 2u ? ~0u : (w == 2u ? m : 0u)));
    pos += __popc(v.w & (w == 3u ? m : 0u));
    return pos;
}

template <bool RANK>
__global__ __launch_bounds__(BITS_THREADS) void bucket_rows_bits(const long long *__restrict__ bstart,
                                                                  const unsigned *__restrict__ ent, int n, int nbk,
                                                                  int *__restrict__ row, int *__restrict__ col,
                                                                  long long cap)
{
    extern __shared__ __align__(16) uint4 bm4[];  // [BROWS][ng], then (RANK) the group bases [BROWS][ng]
    __shared__ unsigned rdeg[BROWS];
    const int tid = threadIdx.x, lane = tid & 63, r = tid >> 6;
    if (bstart[nbk] > cap) {  // nothing was scattered; row[n] reports the total
        if (blockIdx.x == 0 && tid == 0) row[n] = (int)min(bstart[nbk], (long long)INT_MAX);
        return;
    }
    const int ng = (n + 127) >> 7;
    unsigned *bm = (unsigned *)bm4;
    unsigned *gbase = (unsigned *)(bm4 + (long long)BROWS * ng);
    uint4 *rb4 = bm4 + (long long)r * ng;
    for (int q = lane; q < ng; q += 64) rb4[q] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    const int g0 = (lane * ng) >> 6, g1 = ((lane + 1) * ng) >> 6;
    // this bucket's first entries (x) and the next bucket's (xn); bucket
    // sizes are below cap <= INT32_MAX, so offsets within one are 32-bit
    unsigned x[BITS_PF] = {}, xn[BITS_PF] = {};
    long long e0 = 0, e1 = 0;
    if (blockIdx.x < nbk) {
        e0 = bstart[blockIdx.x];
        e1 = bstart[blockIdx.x + 1];
        if (e1 > e0) {
            const unsigned *xb = ent + e0;
            const unsigned last = (unsigned)(e1 - e0) - 1u;
#pragma unroll
            for (int u = 0; u < BITS_PF; u++) x[u] = xb[min((unsigned)(tid + u * BITS_THREADS), last)];
        }
    }
    for (int b = blockIdx.x; b < nbk; b += gridDim.x) {
        const unsigned len = (unsigned)(e1 - e0);
#pragma unroll
        for (int u = 0; u < BITS_PF; u++)
            if ((unsigned)(tid + u * BITS_THREADS) < len) bits_set(bm, (unsigned)ng, x[u]);
        for (unsigned k0 = (unsigned)(tid + BITS_PF * BITS_THREADS); k0 < len; k0 += BITS_THREADS * BITS_UNROLL) {
            unsigned y[BITS_UNROLL];
#pragma unroll
            for (int u = 0; u < BITS_UNROLL; u++) y[u] = ent[e0 + min(k0 + (unsigned)(u * BITS_THREADS), len - 1u)];
#pragma unroll
            for (int u = 0; u < BITS_UNROLL; u++)
                if (k0 + (unsigned)(u * BITS_THREADS) < len) bits_set(bm, (unsigned)ng, y[u]);
        }
        const long long c0 = e0, c1 = e1;
        if (b + (int)gridDim.x < nbk) {
            e0 = bstart[b + gridDim.x];
            e1 = bstart[b + gridDim.x + 1];
            if (e1 > e0) {
                const unsigned *xb = ent + e0;
                const unsigned last = (unsigned)(e1 - e0) - 1u;
#pragma unroll
                for (int u = 0; u < BITS_PF; u++) xn[u] = xb[min((unsigned)(tid + u * BITS_THREADS), last)];
            }
        }
        __syncthreads();
        // this wave's row: per-lane popcounts (the run's reads all issued
        // before the first is used; reads past the run are clamped into the
        // row and not counted), one wave scan
        constexpr int RUN = RANK ? BITS_RUN_RANK : BITS_RUN;
        unsigned s = 0, cq[RUN];
#pragma unroll
        for (int q = 0; q < RUN; q++) {
            const uint4 v = rb4[min(g0 + q, ng - 1)];
            cq[q] = g0 + q < g1 ? __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w) : 0u;
            s += cq[q];
        }
        const unsigned incl = wave_incl_sum(s);
        if (lane == 63) rdeg[r] = incl;
        if (RANK) {  // group bases within the row, before the row starts are known
            unsigned base = incl - s;
            unsigned *gb = gbase + (long long)r * ng;
#pragma unroll
            for (int q = 0; q < RUN; q++)
                if (g0 + q < g1) {
                    gb[g0 + q] = base;
                    base += cq[q];
                }
        }
        __syncthreads();
        // row q's start within the bucket in lane q of every wave
        const unsigned rd = lane < BROWS ? rdeg[lane] : 0u;
        const unsigned rstart = wave_incl_sum(rd) - rd;
        const unsigned before = (unsigned)__builtin_amdgcn_readlane((int)rstart, r);
        const int gr = b * BROWS + r;
        if (lane == 0 && gr < n) row[gr] = (int)(c0 + before);
        if (b == nbk - 1 && tid == 0) row[n] = (int)c1;
        if (RANK) {
            int *colb = col + c0;
            const unsigned words = (unsigned)(BROWS * ng * 4);
            if (len <= min((unsigned)(BITS_PF * BITS_THREADS), words)) {
                // every rank first; then the bitmaps' LDS takes the bucket's
                // columns at their places, and the bucket goes out in whole
                // 256-byte wave stores (stores straight to the ranks leave
                // most L2 write requests carrying one or two columns); the
                // bitmaps are zero again when the bucket is done
                unsigned pos[BITS_PF];
#pragma unroll
                for (int u = 0; u < BITS_PF; u++) {
                    const unsigned rs = (unsigned)__shfl((int)rstart, (int)(x[u] >> 25));
                    pos[u] = (unsigned)(tid + u * BITS_THREADS) < len ? rs + bits_rank(bm4, gbase, (unsigned)ng, x[u])
                                                                      : 0u;
                }
                __syncthreads();  // every rank read before the bitmaps are overwritten
#pragma unroll
                for (int u = 0; u < BITS_PF; u++)
                    if ((unsigned)(tid + u * BITS_THREADS) < len) bm[pos[u]] = x[u] & COLMASK;
                const unsigned l4 = (len + 3u) >> 2;
                if ((unsigned)tid < 4u * l4 - len) bm[len + tid] = 0u;
                for (unsigned q = l4 + tid; q < (unsigned)(BROWS * ng); q += BITS_THREADS)
                    bm4[q] = make_uint4(0u, 0u, 0u, 0u);
                __syncthreads();  // the staged columns complete
                unsigned v[BITS_PF];
                const unsigned last = len ? len - 1u : 0u;
#pragma unroll
                for (int u = 0; u < BITS_PF; u++) v[u] = bm[min((unsigned)(tid + u * BITS_THREADS), last)];
#pragma unroll
                for (int u = 0; u < BITS_PF; u++) {
                    const unsigned k = (unsigned)(tid + u * BITS_THREADS);
                    if (k < len) {
                        colb[k] = (int)v[u];
                        bm[k] = 0u;
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < BITS_PF; u++) {
                    const unsigned rs = (unsigned)__shfl((int)rstart, (int)(x[u] >> 25));
                    if ((unsigned)(tid + u * BITS_THREADS) < len)
                        colb[rs + bits_rank(bm4, gbase, (unsigned)ng, x[u])] = (int)(x[u] & COLMASK);
                }
                for (unsigned k0 = (unsigned)(tid + BITS_PF * BITS_THREADS); k0 < len;
                     k0 += BITS_THREADS * BITS_UNROLL) {
                    unsigned y[BITS_UNROLL];
#pragma unroll
                    for (int u = 0; u < BITS_UNROLL; u++)
                        y[u] = ent[c0 + min(k0 + (unsigned)(u * BITS_THREADS), len - 1u)];
#pragma unroll
                    for (int u = 0; u < BITS_UNROLL; u++) {
                        const unsigned rs = (unsigned)__shfl((int)rstart, (int)(y[u] >> 25));
                        if (k0 + (unsigned)(u * BITS_THREADS) < len)
                            colb[rs + bits_rank(bm4, gbase, (unsigned)ng, y[u])] = (int)(y[u] & COLMASK);
                    }
                }
                __syncthreads();  // every rank read before the clears
                if (s)
                    for (int q = g0; q < g1; q++) rb4[q] = make_uint4(0u, 0u, 0u, 0u);
            }
        } else {
            // read the lane's s columns back in order (one per step, moving
            // to the next non-empty group when the current one is spent; the
            // step takes the lowest bit of the low or high half without a
            // branch)
            int *out = col + (c0 + before + (incl - s));
            int g = g0;
            unsigned long long lo = 0, hi = 0;
            uint4 nx = rb4[min(g, ng - 1)];  // the next group, read one ahead
            for (unsigned it = 0; it < s; it++) {
                while (!(lo | hi)) {
                    lo = ((unsigned long long)nx.y << 32) | nx.x;
                    hi = ((unsigned long long)nx.w << 32) | nx.z;
                    g++;
                    nx = rb4[min(g, ng - 1)];
                }
                const bool inlo = lo != 0ull;
                const unsigned long long w = inlo ? lo : hi;
                const int c = __builtin_ctzll(w) + (inlo ? 0 : 64);
                const unsigned long long w2 = w & (w - 1ull);
                hi = inlo ? hi : w2;
                lo = inlo ? w2 : 0ull;
                *out++ = (g - 1) * 128 + c;
            }
            if (s)
                for (int q = g0; q < g1; q++) rb4[q] = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < BITS_PF; u++) x[u] = xn[u];
    }
}

// LDS of bucket_rows_bits<RANK> (0 when it does not fit): the bitmaps of a
// bucket, 16 bytes per 128 columns and row, and with RANK 4 bytes more per
// group for its base
size_t rows_bits_lds(int n, size_t lds_per_block, bool rank)
{
    const size_t b = (size_t)BROWS * (size_t)((n + 127) >> 7) * (rank ? 20 : 16);
    return b + 64 <= std::min<size_t>(lds_per_block, 160 * 1024) ? b : 0;
}

// The row kernel's variant and LDS: RANK where its bases fit, else the
// read-back where the bitmaps fit, else lds = 0.
struct BitsPlan {
    bool rank = false;
    size_t lds = 0;
};
BitsPlan rows_bits_plan(int n, size_t lds_per_block)
{
    BitsPlan p;
    p.lds = rows_bits_lds(n, lds_per_block, true);
    p.rank = p.lds != 0;
    if (!p.rank) p.lds = rows_bits_lds(n, lds_per_block, false);
    return p;
}

const void *rows_bits_fn(bool rank)
{
    return rank ? (const void *)bucket_rows_bits<true> : (const void *)bucket_rows_bits<false>;
}

// its grid: as many work-groups as the LDS of every CU holds (at most four
// per CU), no more than there are buckets
unsigned rows_bits_grid(const pg_ctx *ctx, int n, int nbk)
{
    const size_t l = std::max<size_t>(rows_bits_plan(n, ctx->lds_per_block).lds, 1);
    const long long per_cu = std::max<long long>(1, std::min<long long>(4, (long long)(160 * 1024 / (l + 64))));
    return (unsigned)std::max<long long>(1, std::min<long long>(nbk, (long long)ctx->num_cu * per_cu));
}

void launch_rows_bits(const pg_ctx *ctx, const BitsPlan &bp, hipStream_t st, const long long *bstart,
                      const unsigned *ent, int n, int nbk, int *row, int *col, long long cap)
{
    const unsigned grid = rows_bits_grid(ctx, n, nbk);
    if (bp.rank)
        bucket_rows_bits<true><<<grid, BITS_THREADS, bp.lds, st>>>(bstart, ent, n, nbk, row, col, cap);
    else
        bucket_rows_bits<false><<<grid, BITS_THREADS, bp.lds, st>>>(bstart, ent, n, nbk, row, col, cap);
}

// Bitmap waves and the smallest row length they take: as many waves as fit
// half the LDS (two work-groups per CU; at least one wave when a bitmap fits
// at all, else nbm = 0 and the build is refused), and rows longer than dbit,
// where dbit is the longest row for which the register bitonic sort costs
// less than the bitmap's fixed per-row scan (per groups per lane).
struct RowsPlan {
    int nbm, dbit;
    size_t lds;
};
__host__ inline long long rows_wave_u4(int n) { return rows_groups(n) + (rows_groups(n) + 3) / 4; }
RowsPlan rows_plan(int n, size_t lds_per_block)
{
    RowsPlan p;
    const long long wb = rows_wave_u4(n) * 16;
    const long long lds = (long long)std::min<size_t>(lds_per_block, 160 * 1024);
    const long long budget = lds / 2 - 1024;
    p.nbm = (int)std::max<long long>(0, std::min<long long>(ROWS_THREADS / 64, budget / wb));
    if (p.nbm == 0 && wb + 1024 <= lds) p.nbm = 1;
    const long long per = ((rows_groups(n) + 63) / 64) | 1;
    // VALU work per lane: bitonic ~ 4 P log2(64P)(log2(64P)+1)/2, bitmap
    // ~ 10 per + 20 P; the first P where the bitmap wins sets dbit = 32 P
    p.dbit = SORT_MAX;
    for (int P = 2; P <= 16; P <<= 1) {
        const int lg = 6 + (P == 2 ? 1 : P == 4 ? 2 : P == 8 ? 3 : 4);
        const long long bitonic = 4LL * P * lg * (lg + 1) / 2, bitmap = 10 * per + 20LL * P;
        if (bitmap < bitonic) {
            p.dbit = 32 * P;
            break;
        }
    }
    p.lds = (size_t)p.nbm * (size_t)wb;
    return p;
}

// float nearest to x on the far side (outward rounding of window bounds)
__host__ __device__ inline float round_up_f(double x)
{
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}
__host__ __device__ inline float round_down_f(double x)
{
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}

// ------------------------------------------------------ sparsity selection

constexpr int NBIN2 = 16384;   // fine histogram over the emitted set


// Largest bin b of h[0..nb) with sum_{b' >= b} h[b'] >= target (-1 if none),
// by one block: per-thread bin runs, a suffix scan over threads, one scan of
// the crossing run.
__device__ int top_bin_reaching(const unsigned long long *h, int nb, double target, unsigned long long *total)
{
    __shared__ unsigned long long wsum[16];
    __shared__ int s_bin;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, per = nb / 1024;
    unsigned long long own = 0;
#pragma unroll 16
    for (int q = 0; q < per; q++) own += h[tid * per + q];
    // prefix sums over the threads: a wave scan, then the waves' totals
    unsigned long long incl = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    if (tid == 0) s_bin = -1;
    __syncthreads();
    unsigned long long before = 0, all = 0;
    for (int q = 0; q < 16; q++) {
        const unsigned long long v = wsum[q];
        all += v;
        before += q < w ? v : 0ull;
    }
    const unsigned long long above = all - before - incl;  // threads after this one
    if ((double)(above + own) >= target && (double)above < target) {
        unsigned long long cum = above;
        for (int q = per - 1; q >= 0; q--) {
            cum += h[tid * per + q];
            if ((double)cum >= target) {
                s_bin = tid * per + q;
                break;
            }
        }
    }
    if (target <= 0.0 && tid == 1023) s_bin = nb - 1;
    __syncthreads();
    if (total) *total = all;
    const int r = s_bin;
    __syncthreads();  // shared state is reused by the next call
    return r;
}

// L from the sampled histogram: the bin where the top fraction f of the
// sampled pairs is reached, minus one bin edge (bins of width 2/NBIN on [-1, 1]).
__global__ __launch_bounds__(1024) void sample_bound(const unsigned long long *__restrict__ htot, double f,
                                                     SelCtl *__restrict__ ctl)
{
    unsigned long long tot = 0;
    // the target needs the total: a first call (target 0) returns it
    top_bin_reaching(htot, NBIN, 0.0, &tot);
    const double target = f * (double)tot;
    const int bl = top_bin_reaching(htot, NBIN, target, nullptr);
    if (threadIdx.x == 0) {
        const bool ok = bl > 0 && tot > 0;
        ctl->lower = ok ? round_down_f(bl * (2.0 / NBIN) - 1.0 - 1e-6) : 2.0f;  // 2: emit nothing
        ctl->ok = ok ? 1 : 0;
    }
}

constexpr double FINE_TOP = 1.0 + 1e-6;  // upper end of the fine histogram

// Candidate window from the fine histogram of the emitted r~: the bin holding
// the k-th largest, widened by twice the error band on each side.
// clr_a[0..na), clr_b[0..nb): buffers the next steps need zeroed (the
// candidate split's counters and the selection histogram), cleared here
// instead of by a fill launch of their own.
__global__ __launch_bounds__(1024) void window_bounds(const unsigned long long *__restrict__ h2, long long kwant,
                                                      double band, SelCtl *__restrict__ ctl,
                                                      unsigned long long *__restrict__ clr_a = nullptr, int na = 0,
                                                      unsigned *__restrict__ clr_b = nullptr, int nb = 0)
{
    for (int k = threadIdx.x; k < na; k += 1024) clr_a[k] = 0ull;
    for (int k = threadIdx.x; k < nb; k += 1024) clr_b[k] = 0u;
    unsigned long long tot = 0;
    const int b2 = top_bin_reaching(h2, NBIN2, (double)kwant, &tot);
    if (threadIdx.x == 0) {
        const float lower = ctl->lower;
        const float scale = (float)(NBIN2 / (FINE_TOP - (double)lower));
        const double w2 = 1.0 / (double)scale;
        const double lo2 = (double)lower + b2 * w2 - 1e-6;
        const double hi2 = (double)lower + (b2 + 1) * w2 + 1e-6;
        const float L2 = round_down_f(lo2 - 2.0 * band);
        const float U2 = round_up_f(hi2 + 2.0 * band);
        ctl->L2 = L2;
        ctl->U2 = U2;
        if (!(b2 >= 0 && (long long)tot >= kwant && (double)L2 >= (double)lower)) ctl->ok = 0;
    }
}

// Histogram of r~ over the kept segment entries, bins over [lo, lo + NBIN2/scale).
// Segments start at multiples of 4 floats (seg_cap % 4 == 0): float4 loads.
// tot: the NBIN2 totals that hist_sum accumulates next, zeroed here (each
// work-group a slice) instead of by a fill launch of their own.
__global__ __launch_bounds__(1024) void emitted_hist(const float *__restrict__ seg_r, long long seg_cap,
                                                     const unsigned long long *__restrict__ seg_cnt, int nseg,
                                                     const SelCtl *__restrict__ ctl, uint32_t *__restrict__ out,
                                                     unsigned long long *__restrict__ tot)
{
    __shared__ uint32_t h[NBIN2];
    {
        const int per = (NBIN2 + (int)gridDim.x - 1) / (int)gridDim.x;
        const int b0 = blockIdx.x * per, b1 = min(NBIN2, b0 + per);
        for (int b = b0 + threadIdx.x; b < b1; b += 1024) tot[b] = 0ull;
    }
    const float lo = ctl->lower;
    const float scale = (float)(NBIN2 / (FINE_TOP - (double)lo));
    for (int b = threadIdx.x; b < NBIN2; b += 1024) h[b] = 0;
    __syncthreads();
    auto bin = [&](float r) {
        int b = (int)((r - lo) * scale);
        b = b < 0 ? 0 : (b >= NBIN2 ? NBIN2 - 1 : b);
        atomicAdd(&h[b], 1u);
    };
    for (int g = blockIdx.x; g < nseg; g += gridDim.x) {
        const long long c = min((long long)seg_cnt[g], seg_cap);
        const float *r = seg_r + (long long)g * seg_cap;
        const long long c4 = c >> 2;
        for (long long k = threadIdx.x; k < c4; k += 1024) {
            const float4 v = ((const float4 *)r)[k];
            bin(v.x);
            bin(v.y);
            bin(v.z);
            bin(v.w);
        }
        for (long long k = (c4 << 2) + threadIdx.x; k < c; k += 1024) bin(r[k]);
    }
    __syncthreads();
    uint32_t *o = out + (long long)blockIdx.x * NBIN2;
    for (int b = threadIdx.x; b < NBIN2; b += 1024) o[b] = h[b];
}

// out[b] += sum of rows [y*rpb, (y+1)*rpb) of column b (out zeroed before)
__global__ void hist_sum(const uint32_t *__restrict__ h, int nrow, int nbin, unsigned long long *__restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbin) return;
    const int rpb = (nrow + gridDim.y - 1) / gridDim.y;
    const int g0 = blockIdx.y * rpb, g1 = min(nrow, g0 + rpb);
    unsigned long long s = 0;
    int g = g0;
    for (; g + 15 < g1; g += 16) {  // 16 rows in flight per thread
        unsigned v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = h[(long long)(g + u) * nbin + b];
#pragma unroll
        for (int u = 0; u < 16; u++) s += v[u];
    }
    for (; g < g1; g++) s += h[(long long)g * nbin + b];
    if (s) atomicAdd(&out[b], s);
}

// Split the emitted set: r~ > hi counts as a certain edge (kept by the CSR
// filter), lo <= r~ <= hi goes to the FP64 candidate list.  Block (x, g)
// scans entries [x*WIN_CHUNK, (x+1)*WIN_CHUNK) of segment g: per-thread
// counts, one block scan, one global reservation per block.
constexpr int WIN_CHUNK = 8192;
__global__ __launch_bounds__(256) void window_kernel(const int2 *__restrict__ seg, const float *__restrict__ seg_r,
                                                     long long seg_cap, const unsigned long long *__restrict__ seg_cnt,
                                                     int nseg, const SelCtl *__restrict__ ctl,
                                                     int2 *__restrict__ cand, float *__restrict__ cand_r,
                                                     long long cand_cap, unsigned long long *__restrict__ cand_cnt,
                                                     unsigned long long *__restrict__ dpart)
{
    constexpr int PER = WIN_CHUNK / 256 / 4;  // float4 loads per thread
    __shared__ unsigned s_scan[256];
    __shared__ unsigned long long s_base, s_def;
    const int g = blockIdx.y, tid = threadIdx.x;
    unsigned long long *my_def = dpart + (long long)blockIdx.y * gridDim.x + blockIdx.x;
    const long long c = min((long long)seg_cnt[g], seg_cap);
    const long long k0 = (long long)blockIdx.x * WIN_CHUNK;
    if (!ctl->ok || k0 >= c) {
        if (tid == 0) *my_def = 0;
        return;
    }
    const float lo = ctl->L2, hi = ctl->U2;
    if (tid == 0) s_def = 0;
    const long long base = (long long)g * seg_cap;
    float4 v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const long long k = k0 + ((long long)u * 256 + tid) * 4;
        if (k + 3 < c) {
            v[u] = *(const float4 *)(seg_r + base + k);
        } else {
            v[u].x = k < c ? seg_r[base + k] : -2.0f;
            v[u].y = k + 1 < c ? seg_r[base + k + 1] : -2.0f;
            v[u].z = k + 2 < c ? seg_r[base + k + 2] : -2.0f;
            v[u].w = -2.0f;
        }
    }
    unsigned nd = 0, nc = 0;
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const float e[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            nd += e[q] > hi;
            nc += e[q] >= lo && e[q] <= hi;
        }
    }
    // block exclusive scan of nc
    s_scan[tid] = nc;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const unsigned y = tid >= o ? s_scan[tid - o] : 0u;
        __syncthreads();
        s_scan[tid] += y;
        __syncthreads();
    }
    const unsigned tot = s_scan[255];
    unsigned long long ndb = nd;
    for (int o = 32; o > 0; o >>= 1) ndb += __shfl_xor(ndb, o);
    if ((tid & 63) == 0 && ndb) atomicAdd(&s_def, ndb);
    __syncthreads();
    if (tid == 0) *my_def = s_def;  // summed by the selection, no global hot spot
    if (tot == 0) return;
    if (tid == 0) s_base = atomicAdd(cand_cnt, (unsigned long long)tot);
    __syncthreads();
    unsigned long long p = s_base + s_scan[tid] - nc;
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const float e[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (e[q] >= lo && e[q] <= hi) {
                const long long k = k0 + ((long long)u * 256 + tid) * 4 + q;
                if ((long long)p < cand_cap) {
                    cand[p] = seg[base + k];
                    cand_r[p] = e[q];
                }
                p++;
            }
        }
    }
}

// Keys ordered like the reference's selection: r64 descending, then row-major.
__device__ __forceinline__ unsigned long long key_r_desc(double x)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    b = (b >> 63) ? ~b : (b | 0x8000000000000000ull);  // ascending order
    return ~b;                                        // descending
}

// m-th smallest (1-based) key among the entries passing `want`, by 8-bit
// radix select in one work-group.  Returns the key; *rank_left gets m minus
// the number of strictly smaller keys.
template <class KeyF, class WantF>
__device__ unsigned long long radix_select(long long c, long long m, KeyF key, WantF want,
                                           unsigned *hist, unsigned long long *sh, long long *rank_left)
{
    // one block; every pass reads RB keys per thread before using them, so
    // the (L2-resident) loads overlap instead of serialising on latency
    constexpr int RB = 8;
    const long long stride = blockDim.x;
    auto batch = [&](long long q0, unsigned long long (&kk)[RB], bool (&ww)[RB]) {
#pragma unroll
        for (int u = 0; u < RB; u++) {
            const long long q = q0 + u * stride;
            ww[u] = q < c && want(q);
        }
#pragma unroll
        for (int u = 0; u < RB; u++) kk[u] = ww[u] ? key(q0 + u * stride) : 0ull;
    };
    // bytes shared by every wanted key are fixed up front: candidates sit in
    // a narrow r window, so their top bytes agree and a digit histogram over
    // them would serialise on one LDS bin
    unsigned long long kmin = ~0ull, kmax = 0ull;
    for (long long q0 = threadIdx.x; q0 < c; q0 += RB * stride) {
        unsigned long long kk[RB];
        bool ww[RB];
        batch(q0, kk, ww);
#pragma unroll
        for (int u = 0; u < RB; u++)
            if (ww[u]) {
                kmin = kk[u] < kmin ? kk[u] : kmin;
                kmax = kk[u] > kmax ? kk[u] : kmax;
            }
    }
    if (threadIdx.x == 0) {
        sh[0] = ~0ull;
        sh[1] = 0ull;
    }
    __syncthreads();
    atomicMin(&sh[0], kmin);
    atomicMax(&sh[1], kmax);
    __syncthreads();
    kmin = sh[0];
    kmax = sh[1];
    __syncthreads();
    int top = 56;
    while (top >= 0 && ((kmin ^ kmax) >> top) == 0) top -= 8;
    unsigned long long mask = top >= 0 ? ~0ull << (top + 8) : ~0ull;
    if (top == 56) mask = 0;
    unsigned long long prefix = kmin & mask;
    long long mr = m;
    for (int shift = top; shift >= 0; shift -= 8) {
        for (int d = threadIdx.x; d < 256; d += blockDim.x) hist[d] = 0;
        __syncthreads();
        for (long long q0 = threadIdx.x; q0 < c; q0 += RB * stride) {
            unsigned long long kk[RB];
            bool ww[RB];
            batch(q0, kk, ww);
#pragma unroll
            for (int u = 0; u < RB; u++)
                if (ww[u] && (kk[u] & mask) == prefix) atomicAdd(&hist[(kk[u] >> shift) & 255u], 1u);
        }
        __syncthreads();
        // digit holding the mr-th key: the first wave scans the 256 bins,
        // four per lane, with one wave prefix sum
        if (threadIdx.x < 64) {
            const int l = threadIdx.x;
            const long long h0 = hist[4 * l], h1 = hist[4 * l + 1], h2 = hist[4 * l + 2], h3 = hist[4 * l + 3];
            const long long own = h0 + h1 + h2 + h3;
            long long incl = own;
            for (int o = 1; o < 64; o <<= 1) {
                const long long y = __shfl_up(incl, o);
                if (l >= o) incl += y;
            }
            const long long before = incl - own;
            if (before < mr && incl >= mr) {
                long long cum = before;
                int d = 4 * l;
                const long long hh[4] = {h0, h1, h2, h3};
                for (int q = 0; q < 4; q++) {
                    if (cum + hh[q] >= mr) { d = 4 * l + q; break; }
                    cum += hh[q];
                }
                sh[0] = (unsigned long long)d;
                sh[1] = (unsigned long long)(mr - cum);
            }
            if (l == 63 && incl < mr) {  // fewer keys than mr: last digit, as a serial scan gives
                sh[0] = 255ull;
                sh[1] = (unsigned long long)(mr - incl);
            }
        }
        __syncthreads();
        const unsigned long long dsel = sh[0];
        const bool single = hist[dsel] == 1u;  // one key left under the new prefix
        prefix |= dsel << shift;
        mask |= 255ull << shift;
        mr = (long long)sh[1];
        __syncthreads();
        if (single && shift > 0) {
            // its remaining bytes in one scan instead of a pass per byte
            for (long long q0 = threadIdx.x; q0 < c; q0 += RB * stride) {
                unsigned long long kk[RB];
                bool ww[RB];
                batch(q0, kk, ww);
#pragma unroll
                for (int u = 0; u < RB; u++)
                    if (ww[u] && (kk[u] & mask) == prefix) sh[0] = kk[u];
            }
            __syncthreads();
            prefix = sh[0];
            __syncthreads();
            break;
        }
    }
    *rank_left = mr;
    return prefix;
}

// Exact selection (U CorMat.cpp:693-773 on FP64 r): need = k - definite
// candidates, all with r > t* plus the earliest ties in row-major order.
// The selected set is every candidate whose (key, (i, j)) is at most the
// need-th smallest such pair (key = r64 descending), found in one work-group:
// a pass for the key range, a 4096-bin histogram of the keys' offsets above
// the smallest, and the bin holding the need-th key; its members (a handful
// for a window of near-uniform r) are gathered into LDS and ranked by
// comparison.  A bin of more than SEL_CAP members takes the byte-wise radix
// select instead.
__global__ __launch_bounds__(1024) void select_kernel(const int2 *__restrict__ cand, const double *__restrict__ r64,
                                                      const float *__restrict__ rfast,
                                                      const unsigned long long *__restrict__ cand_cnt, long long cand_cap,
                                                      long long k, const unsigned long long *__restrict__ dsrc, int dn,
                                                      int2 *__restrict__ extra, unsigned long long *__restrict__ extra_cnt,
                                                      double *__restrict__ tstar, unsigned long long *__restrict__ flips,
                                                      int *__restrict__ err, SelHist shh,
                                                      unsigned long long *__restrict__ def_out, SelCut *__restrict__ cut)
{
    __shared__ unsigned hist[SEL_NB];
    __shared__ unsigned long long mkey[SEL_CAP];
    __shared__ unsigned mq[SEL_CAP];
    __shared__ unsigned long long sh[2], wtot[16];
    __shared__ unsigned long long s_def, s_k1, s_k2;
    __shared__ double s_ts;
    __shared__ unsigned s_m;
    __shared__ int s_bin;
    __shared__ long long s_rank;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int RB = 16;
    if (tid == 0) {
        s_def = 0;
        s_m = 0;
        s_bin = 0;
        s_rank = 1;
        s_k1 = 0;
        s_k2 = 0;
        s_ts = 0.0;
    }
    __syncthreads();
    unsigned long long dpart = 0;
    for (int g = tid; g < dn; g += blockDim.x) dpart += dsrc[g];
    atomicAdd(&s_def, dpart);
    __syncthreads();
    if (def_out && tid == 0) *def_out = s_def;  // the definite count, for the host's check
    if (cut && tid == 0) cut->ok = 0;
    const long long c = (long long)min((unsigned long long)cand_cap, *cand_cnt);
    const long long need = k - (long long)s_def;
    if (need < 1 || need > c || (long long)*cand_cnt > cand_cap) {
        if (tid == 0) *err = 1;
        return;
    }
    auto ij_of = [&](long long q) {
        const int2 e = cand[q];
        return ((unsigned long long)(unsigned)e.x << 32) | (unsigned)e.y;
    };
    // the bins: from the recheck's histogram of the r (sel_bin), else a pass
    // for the key range and a histogram of (key - kmin) >> shift
    unsigned long long kmin = ~0ull, kmax = 0ull;
    int shift = 0;
    double htop = 0.0, hscale = 1.0;
    if (shh.h) {
        sel_range(shh, htop, hscale);
        for (int b = tid; b < SEL_NB; b += 1024) hist[b] = shh.h[b];
        __syncthreads();
    } else {
        for (long long q0 = tid; q0 < c; q0 += RB * 1024) {
            unsigned long long kk[RB];
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const long long q = q0 + u * 1024;
                kk[u] = q < c ? key_r_desc(r64[q]) : 0ull;
            }
#pragma unroll
            for (int u = 0; u < RB; u++)
                if (q0 + u * 1024 < c) {
                    kmin = kk[u] < kmin ? kk[u] : kmin;
                    kmax = kk[u] > kmax ? kk[u] : kmax;
                }
        }
        for (int b = tid; b < SEL_NB; b += 1024) hist[b] = 0u;
        if (tid == 0) {
            sh[0] = ~0ull;
            sh[1] = 0ull;
        }
        __syncthreads();
        atomicMin(&sh[0], kmin);
        atomicMax(&sh[1], kmax);
        __syncthreads();
        kmin = sh[0];
        kmax = sh[1];
        while (((kmax - kmin) >> shift) >= (unsigned long long)SEL_NB) shift++;
        // histogram of (key - kmin) >> shift
        for (long long q0 = tid; q0 < c; q0 += RB * 1024) {
            unsigned long long kk[RB];
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const long long q = q0 + u * 1024;
                kk[u] = q < c ? key_r_desc(r64[q]) : 0ull;
            }
#pragma unroll
            for (int u = 0; u < RB; u++)
                if (q0 + u * 1024 < c) atomicAdd(&hist[(kk[u] - kmin) >> shift], 1u);
        }
        __syncthreads();
    }
    // bin holding the need-th smallest key: thread t owns bins 4t .. 4t+3
    {
        const unsigned h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2], h3 = hist[4 * tid + 3];
        const unsigned long long own = (unsigned long long)h0 + h1 + h2 + h3;
        unsigned long long incl = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == 63) wtot[wv] = incl;
        __syncthreads();
        unsigned long long wbase = 0;
        for (int w = 0; w < wv; w++) wbase += wtot[w];
        const unsigned long long before = wbase + incl - own;
        if (before < (unsigned long long)need && (unsigned long long)need <= before + own) {
            const unsigned hh[4] = {h0, h1, h2, h3};
            unsigned long long cum = before;
            int d = 0;
            for (int q = 0; q < 4; q++) {
                if (cum + hh[q] >= (unsigned long long)need) {
                    d = q;
                    break;
                }
                cum += hh[q];
            }
            s_bin = 4 * tid + d;
            s_rank = need - (long long)cum;  // 1-based rank inside the bin
        }
    }
    __syncthreads();
    const int bsel = s_bin;
    const long long rin = s_rank;
    const unsigned msel = hist[bsel];
    if (msel <= (unsigned)SEL_CAP) {
        // members of the bin into LDS
        for (long long q0 = tid; q0 < c; q0 += RB * 1024) {
            unsigned long long kk[RB];
            double rr[RB];
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const long long q = q0 + u * 1024;
                rr[u] = q < c ? r64[q] : 0.0;
                kk[u] = key_r_desc(rr[u]);
            }
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const long long q = q0 + u * 1024;
                const int bq = shh.h ? sel_bin(rr[u], htop, hscale) : (int)((kk[u] - kmin) >> shift);
                if (q < c && bq == bsel) {
                    const unsigned p = atomicAdd(&s_m, 1u);
                    mkey[p] = kk[u];
                    mq[p] = (unsigned)q;
                }
            }
        }
        __syncthreads();
        // the member of rank rin by (key, (i, j))
        for (unsigned x = tid; x < msel; x += 1024) {
            const unsigned long long kx = mkey[x];
            long long ijx = -1;
            unsigned rank = 0;
            for (unsigned y = 0; y < msel; y++) {
                const unsigned long long ky = mkey[y];
                if (ky < kx) {
                    rank++;
                } else if (ky == kx && y != x) {
                    if (ijx < 0) ijx = (long long)ij_of(mq[x]);
                    rank += ij_of(mq[y]) < (unsigned long long)ijx;
                }
            }
            if ((long long)rank == rin - 1) {
                s_k1 = kx;
                s_k2 = ij_of(mq[x]);
                s_ts = r64[mq[x]];
                tstar[0] = s_ts;
            }
        }
    } else {
        // crowded bin: byte-wise radix select over all candidates
        long long ties_left = 0;
        const unsigned long long k1 = radix_select(
            c, need, [&](long long q) { return key_r_desc(r64[q]); }, [](long long) { return true; }, hist, sh,
            &ties_left);
        const unsigned long long k2 = radix_select(
            c, ties_left, [&](long long q) { return ij_of(q); },
            [&](long long q) { return key_r_desc(r64[q]) == k1; }, hist, sh, &ties_left);
        if (tid == 0) {
            s_k1 = k1;
            s_k2 = k2;
        }
        for (long long q = tid; q < c; q += 1024)
            if (key_r_desc(r64[q]) == k1) {
                s_ts = r64[q];
                tstar[0] = r64[q];
            }
    }
    __syncthreads();
    const unsigned long long k1 = s_k1, k2 = s_k2;
    const double ts = s_ts;
    if (cut) {  // the selected candidates are written by select_write
        if (tid == 0) {
            cut->k1 = k1;
            cut->k2 = k2;
            cut->ts = ts;
            cut->c = c;
            cut->ok = 1;
        }
        return;
    }
    // the selected candidates: this thread's count (and decision flips),
    // one block scan and one reservation, then the writes
    auto selected = [&](long long q, const int2 &e, double rq) {
        const unsigned long long kq = key_r_desc(rq);
        const unsigned long long k2q = ((unsigned long long)(unsigned)e.x << 32) | (unsigned)e.y;
        return kq < k1 || (kq == k1 && k2q <= k2);
    };
    // one pass: per batch, each wave reserves its selected entries' slots
    // with one LDS atomic and writes them (the list's order is free: the
    // CSR build sorts every row), then one global reservation of the total
    if (tid == 0) s_m = 0;
    __syncthreads();
    const unsigned long long base0 = *extra_cnt;
    unsigned long long fl = 0;
    // trip count uniform over the block (the scan below spans whole waves)
    for (long long qb = 0; qb < c; qb += RB * 1024) {
        const long long q0 = qb + tid;
        double rr[RB];
        float rf[RB];
        int2 ee[RB];
#pragma unroll
        for (int u = 0; u < RB; u++) {
            const long long q = q0 + u * 1024;
            rr[u] = q < c ? r64[q] : 0.0;
            rf[u] = q < c ? rfast[q] : 0.0f;
            ee[u] = q < c ? cand[q] : make_int2(0, 0);
        }
        unsigned sel = 0, ns = 0;
#pragma unroll
        for (int u = 0; u < RB; u++)
            if (q0 + u * 1024 < c) {
                const unsigned b = selected(q0 + u * 1024, ee[u], rr[u]) ? 1u : 0u;
                sel |= b << u;
                ns += b;
                fl += ((double)rf[u] > ts) != (rr[u] > ts);
            }
        unsigned incl = ns;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const unsigned wall = __shfl(incl, 63);
        unsigned wb = 0;
        if (lane == 63 && wall) wb = atomicAdd(&s_m, wall);
        wb = __shfl(wb, 63);
        unsigned long long p = base0 + wb + incl - ns;
#pragma unroll
        for (int u = 0; u < RB; u++)
            if ((sel >> u) & 1u) extra[p++] = ee[u];
    }
    __syncthreads();
    if (tid == 0) atomicAdd(extra_cnt, (unsigned long long)s_m);
    atomicAdd(flips, fl);
}

// The selected candidates of a select_kernel run that left its cut in *cut:
// every candidate whose (key, (i, j)) is at most (k1, k2), appended to extra
// with one reservation per wave (the list's order is free: the CSR build
// sorts every row); decision flips against t* into *flips.  Grid-stride over
// the candidates, trip count uniform over each block.
__global__ __launch_bounds__(256) void select_write(const int2 *__restrict__ cand, const double *__restrict__ r64,
                                                   const float *__restrict__ rfast, const SelCut *__restrict__ cut,
                                                   int2 *__restrict__ extra,
                                                   unsigned long long *__restrict__ extra_cnt,
                                                   unsigned long long *__restrict__ flips,
                                                   unsigned *__restrict__ bcnt_row)
{
    if (!cut->ok) return;
    const long long c = cut->c;
    const unsigned long long k1 = cut->k1, k2 = cut->k2;
    const double ts = cut->ts;
    const int lane = threadIdx.x & 63;
    unsigned long long fl = 0;
    for (long long qb = (long long)blockIdx.x * 256; qb < c; qb += (long long)gridDim.x * 256) {
        const long long q = qb + threadIdx.x;
        bool sel = false;
        int2 e = make_int2(0, 0);
        if (q < c) {
            e = cand[q];
            const double rq = r64[q];
            const unsigned long long kq = key_r_desc(rq);
            const unsigned long long k2q = ((unsigned long long)(unsigned)e.x << 32) | (unsigned)e.y;
            sel = kq < k1 || (kq == k1 && k2q <= k2);
            fl += ((double)rfast[q] > ts) != (rq > ts);
        }
        const unsigned long long m = __builtin_amdgcn_ballot_w64(sel);
        if (m == 0ull) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(extra_cnt, (unsigned long long)__popcll(m));
        base = __shfl(base, 0);
        if (sel) {
            extra[base + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = e;
            if (bcnt_row) {  // the list's CSR bucket counts, both directions
                atomicAdd(&bcnt_row[e.x >> BSH], 1u);
                atomicAdd(&bcnt_row[e.y >> BSH], 1u);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) fl += __shfl_xor(fl, o);
    if (lane == 0 && fl) atomicAdd(flips, fl);
}

// out = sum of part[0..np) (one block)
__global__ __launch_bounds__(1024) void definite_sum(const unsigned long long *__restrict__ part, int np,
                                                     unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long r[1024];
    unsigned long long s = 0;
    for (int i = threadIdx.x; i < np; i += 1024) s += part[i];
    r[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) r[threadIdx.x] += r[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = r[0];
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
double host_err(unsigned long long bits)
{
    double e;
    std::memcpy(&e, &bits, 8);
    return e;
}

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



}  // namespace

// Supertiles (SR, SC) of edge `sup` tiles holding some needed tile
// (C >= RT*R), row-major.
void supertiles(int nr, int nc, int rt, int sup, std::vector<int2> &stl)
{
    stl.clear();
    const int nsr = (int)pg_cdiv(nr, sup), nsc = (int)pg_cdiv(nc, sup);
    for (int sr = 0; sr < nsr; sr++)
        for (int sc = rt * sr; sc < nsc; sc++)
            if (sc * sup + sup - 1 >= rt * sr * sup) stl.push_back(make_int2(sr, sc));
}

// Supertile edge of the Gram walk: PAGANI_GRAM_SUPER, else a value forced by
// the autotuner's search, else the tuned value for (arch, n, t), else 8.
int gram_super(pg_ctx *ctx, int n, int t)
{
    int v = 0;
    if (const char *e = std::getenv("PAGANI_GRAM_SUPER")) v = std::atoi(e);
    if (v <= 0) v = ctx->force_gram_super;
    if (v <= 0) v = pg_tune_get(ctx, "gram_super", n, t, 0);
    return (v == 4 || v == 8 || v == 16) ? v : SUPER;
}

// Pairs i < j < n inside the tiles of every stride-th supertile of the list:
// what a histogram pass with that sampling stride counts.
long long sampled_pairs(const std::vector<int2> &stl, long long stride, int nr, int nc, int rt, int tm, int tn,
                        int n, int sup)
{
    long long tot = 0;
    for (long long q = 0; q < (long long)stl.size(); q += stride) {
        const int2 sc = stl[(size_t)q];
        for (int a = 0; a < sup; a++)
            for (int b = 0; b < sup; b++) {
                const int r = sc.x * sup + a, c = sc.y * sup + b;
                if (r >= nr || c >= nc || c < rt * r) continue;
                const long long c0 = (long long)c * tn, c1 = std::min<long long>(n, c0 + tn);
                const long long i1 = std::min<long long>(n, (long long)r * tm + tm);
                if (c0 >= i1) {  // every row of the tile is below every column
                    if (c1 > c0 && i1 > (long long)r * tm) tot += (i1 - (long long)r * tm) * (c1 - c0);
                    continue;
                }
                for (long long i = (long long)r * tm; i < i1; i++) {
                    const long long lo = std::max(i + 1, c0);
                    if (c1 > lo) tot += c1 - lo;
                }
            }
    }
    return tot;
}

namespace {

pg_status check_opts(pg_ctx *ctx, const float *bold, int32_t n, int32_t t, const pg_corr_opts *opts)
{
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
    if (rows_plan(n, ctx->lds_per_block).nbm == 0)
        return pg_fail(ctx, PG_ENOTSUP, "n=%d voxels exceeds the row-sort LDS bitmap", n);
    return PG_OK;
}

// k of the sparsity rule with the reference's float arithmetic
// (U CorMat.cpp:311-313): Clength = (long long)(2 M1 s / 100) rounded up to
// even, k = Clength / 2.
long long sparsity_k(long long m1, float s)
{
    float prod = (float)(2 * m1) * s;
    long long clen = (long long)((double)prod / 100.0);
    clen += clen % 2;
    return clen / 2;
}

// r threshold as the edge test applies it (W CorMat.cpp:310: the weighted
// tool passes thr - 1e-6 into a float)
float thr_effective(float thr, bool weighted) { return weighted ? (float)((double)thr - 1e-6) : thr; }

// One construction's geometry, device buffers and launch helpers, shared by
// pg_corr_build and the pg_corr_emit / pg_corr_select pair.
// What the host checks after the single sparsity pass, gathered on the
// device so that one small copy into pinned memory brings it back.
struct SpReport {
    unsigned long long seg_max;      // the fullest segment's count
    unsigned long long counters[16]; // the pass's counters
    SelCtl ctl;
    double tstar;
    int rn;                          // row[n] of the CSR built
    int pad;
};
static_assert(sizeof(SpReport) <= PG_REPORT_BYTES, "report exceeds the pinned buffer");

__global__ __launch_bounds__(256) void sp_report(const unsigned long long *__restrict__ seg_cnt, int nseg,
                                                 const unsigned long long *__restrict__ counters,
                                                 const SelCtl *__restrict__ ctl, const double *__restrict__ tstar,
                                                 const int *__restrict__ rown, SpReport *__restrict__ out)
{
    __shared__ unsigned long long part[256];
    unsigned long long m = 0;
    for (int g = threadIdx.x; g < nseg; g += 256) m = seg_cnt[g] > m ? seg_cnt[g] : m;
    part[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && part[threadIdx.x + o] > part[threadIdx.x]) part[threadIdx.x] = part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 16) out->counters[threadIdx.x] = counters[threadIdx.x];
    if (threadIdx.x == 0) {
        out->seg_max = part[0];
        out->ctl = *ctl;
        out->tstar = *tstar;
        out->rn = rown ? *rown : -1;
    }
}

struct CorrRun {
    pg_ctx *ctx = nullptr;
    hipStream_t st = nullptr;
    int n = 0, t = 0, prec = PREC_BF16X3;
    bool weighted = false, g2path = true;
    int tm = 0, tn = 0, rt = 0, nr = 0, npad = 0, nc = 0, ldp = 0, nbk = 0, nsrc = 0, sup = SUPER;
    long long m1 = 0, nst = 0, nwg = 0, nwg_hist = 0;
    double band = 0.0;
    double *z64 = nullptr;
    void *p0 = nullptr, *p1 = nullptr;
    std::vector<int2> stl;
    GramArgs ga;
    size_t stage = 0;
    unsigned long long *counters = nullptr, *seg_cnt = nullptr, *htot = nullptr;
    int2 *seg = nullptr, *bandl = nullptr;
    float *seg_r = nullptr, *band_r = nullptr;
    long long seg_cap = 0, band_cap = 0;
    int *row = nullptr, *col = nullptr;
    float *val = nullptr;
    bool ran_hist = false, ran_emit = false;
    // band-free emits: stage hits in the dedicated LDS area (no epilogue
    // barrier), chosen when a wave's 128 x 64 block is expected to hold well
    // under its DEDIC_RECS records (else the free ring slot's 768)
    bool dedic = false;
    long long gram_passes = 0;
    std::vector<unsigned long long> hseg;
    unsigned long long hcnt[16] = {};

    // z-score and MFMA panels of the series (events ev[0], ev[1] around it),
    // supertile list and Gram arguments.
    // FP32: register-staged kernel, TM x TN tiles, voxel-major panel [npad][ldp].
    // BF16x3: gram2_kernel, 256 x 128 tiles, chunk-major panels of 16-deep
    // time chunks (at least two chunks).
    pg_status setup(pg_ctx *c, const float *bold, int n_, int t_, const pg_corr_opts *opts)
    {
        ctx = c;
        st = c->stream;
        n = n_;
        t = t_;
        prec = opts->prec == PG_PREC_BF16X3 ? PREC_BF16X3 : PREC_F32;
        weighted = opts->weighted != 0;
        m1 = (long long)n * (n - 1) / 2;
        g2path = prec == PREC_BF16X3;
        tm = g2path ? g2::TM : TM;
        tn = g2path ? g2::TN : TN;
        rt = tm / tn;
        nr = (int)pg_cdiv(std::max(n, 1), tm);  // row tiles
        npad = nr * tm;
        nc = npad / tn;  // column tiles
        ldp = g2path ? (int)std::max<long long>(2 * g2::KC, pg_cdiv(t, g2::KC) * g2::KC) : (int)pg_cdiv(t, KC) * KC;
        const long long nt = (long long)t;
        double *mu, *sd;
        PG_TRY(pg_bufT(ctx, "corr.mu", n + 1, &mu));
        PG_TRY(pg_bufT(ctx, "corr.sd", n + 1, &sd));
        PG_TRY(pg_bufT(ctx, "corr.z64", (size_t)n * nt + 1, &z64));
        // the fused z-score writes the panels' padding; the two-kernel form
        // (long series) does not, so the panels are cleared first
        const bool fused = n > 0 && t <= ZS_TMAX;
        if (prec == PREC_F32) {
            PG_TRY(pg_buf(ctx, "corr.p32", (size_t)npad * ldp * 4, &p0));
            if (!fused) PG_HIP(ctx, hipMemsetAsync(p0, 0, (size_t)npad * ldp * 4, st));
        } else {
            PG_TRY(pg_buf(ctx, "corr.phi", (size_t)npad * ldp * 2, &p0));
            PG_TRY(pg_buf(ctx, "corr.plo", (size_t)npad * ldp * 2, &p1));
            if (!fused) {
                PG_HIP(ctx, hipMemsetAsync(p0, 0, (size_t)npad * ldp * 2, st));
                PG_HIP(ctx, hipMemsetAsync(p1, 0, (size_t)npad * ldp * 2, st));
            }
        }
        const float *dbold = nullptr;
        PG_TRY(pg_in(ctx, bold, (size_t)n * nt, "corr.bold", &dbold));

        PG_HIP(ctx, stage_event(ctx, 0, st));
        if (fused) {
            const size_t shz = (size_t)t * (ZV + 1) * 4;
            PG_HIP(ctx, set_shared_limit(ctx, (const void *)zscore_fused, (int)shz));
            zscore_fused<<<(unsigned)pg_cdiv(npad, ZV), 256, shz, st>>>(
                dbold, opts->layout, n, t, z64, prec == PREC_F32 ? (float *)p0 : nullptr,
                prec == PREC_BF16X3 ? (uint16_t *)p0 : nullptr, prec == PREC_BF16X3 ? (uint16_t *)p1 : nullptr, ldp,
                npad);
            PG_KERNEL_CHECK(ctx);
        } else if (n > 0) {
            zscore_stats<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(dbold, opts->layout, n, t, mu, sd);
            dim3 g((unsigned)pg_cdiv(n, 64), (unsigned)pg_cdiv(t, 64));
            zscore_write<<<g, 256, 0, st>>>(dbold, opts->layout, n, t, mu, sd, z64,
                                             prec == PREC_F32 ? (float *)p0 : nullptr,
                                             prec == PREC_BF16X3 ? (uint16_t *)p0 : nullptr,
                                             prec == PREC_BF16X3 ? (uint16_t *)p1 : nullptr, ldp, npad);
            PG_KERNEL_CHECK(ctx);
        }
        PG_HIP(ctx, stage_event(ctx, 1, st));

        // Gram launch geometry: PG_GRAM_WG_PER_CU work-groups per CU, a
        // multiple of the XCD count (see TileWalk); supertiles (SR, SC)
        // holding some needed tile (C >= RT*R), row-major
        sup = gram_super(ctx, n, t);
        supertiles(nr, nc, rt, sup, stl);
        int2 *dstl;
        PG_TRY(pg_bufT(ctx, "corr.stl", stl.size() + 1, &dstl));
        // the list depends on the geometry only: copied again only when it or
        // its buffer changed
        const long long key[5] = {nr, nc, rt, sup, (long long)stl.size()};
        if (ctx->stl_dev != (void *)dstl || std::memcmp(ctx->stl_key, key, sizeof key) != 0) {
            PG_HIP(ctx, hipMemcpyAsync(dstl, stl.data(), stl.size() * sizeof(int2), hipMemcpyHostToDevice, st));
            std::memcpy(ctx->stl_key, key, sizeof key);
            ctx->stl_dev = (void *)dstl;
        }
        nst = (long long)stl.size();
        nwg = (long long)ctx->num_cu * PG_GRAM_WG_PER_CU;
        nwg = std::max<long long>(NXCD, nwg / NXCD * NXCD);
        band = error_band(prec, t);

        std::memset(&ga, 0, sizeof ga);
        ga.p0 = p0;
        ga.p1 = p1;
        ga.ldp = ldp;
        ga.kchunks = g2path ? ldp / g2::KC : ldp / KC;
        ga.n = n;
        ga.nr = nr;
        ga.nc = nc;
        ga.rt = rt;
        ga.npad = npad;
        ga.st_list = dstl;
        ga.nst = nst;
        ga.super = sup;
        ga.sstep = 0;
        ga.lower_dev = nullptr;
        stage = g2path ? (size_t)g2::NSTAGE * g2::STAGE : (size_t)stage_bytes<PREC_F32>();
        // the histogram form of gram2_kernel holds one work-group per CU (LDS)
        nwg_hist = g2path ? std::max<long long>(NXCD, ctx->num_cu / NXCD * NXCD) : nwg;

        // counters: [0] band/candidate count, [1] accepted/selected count,
        // [2] flips, [3] definite (sparsity single pass), [4] error flag,
        // [5] largest |r~ - r64| of the rechecked pairs (double bits)
        PG_TRY(pg_bufT(ctx, "corr.counters", 16, &counters));
        PG_TRY(pg_bufT(ctx, "corr.segcnt", (size_t)nwg, &seg_cnt));
        hseg.assign((size_t)nwg, 0ull);
        nbk = (int)pg_cdiv(std::max(n, 1), BROWS);
        nsrc = (int)nwg + 1;  // segments + the accepted/selected list
        return PG_OK;
    }

    pg_status launch_gram(int mode)
    {
        size_t shm = stage + (mode == MODE_HIST ? NBIN * 4 : 0);
        const void *fn;
        if (g2path)
            fn = mode == MODE_HIST        ? (const void *)gram2_kernel<MODE_HIST>
                 : ga.upper >= -1.0f      ? (const void *)gram2_kernel<MODE_EMIT, true>
                 : dedic                  ? (const void *)gram2_kernel<MODE_EMIT, false, true>
                                          : (const void *)gram2_kernel<MODE_EMIT, false>;
        else
            fn = mode == MODE_HIST ? (const void *)gram_kernel<PREC_F32, MODE_HIST>
                                   : (const void *)gram_kernel<PREC_F32, MODE_EMIT>;
        PG_HIP(ctx, set_shared_limit(ctx, fn, shm));
        // the kernel's own start and stop events (ev[e0], ev[e0 + 1]) ride on
        // its dispatch: no separate event packets around it
        const int e0 = mode == MODE_HIST ? 6 : 8;
        hipEvent_t e_start = stage_timed(ctx, e0) ? ctx->ev[e0] : nullptr;
        hipEvent_t e_stop = stage_timed(ctx, e0) ? ctx->ev[e0 + 1] : nullptr;
        const unsigned grid = (unsigned)(g2path && mode == MODE_HIST ? nwg_hist : nwg);
        void *args[] = {&ga};
        PG_HIP(ctx, hipExtLaunchKernel(fn, dim3(grid), dim3(THREADS), args, shm, st, e_start, e_stop, 0));
        if (mode == MODE_HIST) ran_hist = true;
        else ran_emit = true;
        if (ga.sstep <= 1) gram_passes++;
        return PG_OK;
    }

    // One emitting Gram pass without a host check (the caller checks
    // seg_cnt); store_r keeps r~ of every segment entry.
    // clear: zero the counters first (the recheck-list count, read by
    // emit_pass); every work-group writes its own segment count.
    pg_status emit_launch(long long seg_cap0, long long band_cap0, bool store_r, bool clear = true)
    {
        seg_cap = (seg_cap0 + 63) / 64 * 64;  // segments 256-byte aligned (vector loads)
        band_cap = band_cap0;
        PG_TRY(pg_bufT(ctx, "corr.seg", (size_t)(nwg * seg_cap), &seg));
        if (store_r) PG_TRY(pg_bufT(ctx, "corr.segr", (size_t)(nwg * seg_cap), &seg_r));
        else seg_r = nullptr;
        PG_TRY(pg_bufT(ctx, "corr.band", (size_t)band_cap, &bandl));
        PG_TRY(pg_bufT(ctx, "corr.bandr", (size_t)band_cap, &band_r));
        if (clear) PG_TRY(clear_dev(ctx, st, counters, 16 * sizeof(unsigned long long)));
        ga.seg = seg;
        ga.seg_r = seg_r;
        ga.seg_cap = seg_cap;
        ga.seg_cnt = seg_cnt;
        ga.band = bandl;
        ga.band_r = band_r;
        ga.band_cap = band_cap;
        ga.band_cnt = counters + 0;
        ga.sstep = 0;
        if (n > 1) PG_TRY(launch_gram(MODE_EMIT));
        else PG_TRY(clear_dev(ctx, st, seg_cnt, nwg * sizeof(unsigned long long)));
        return PG_OK;
    }

    // The same with a host check of the segment and band counts (hseg, hcnt),
    // re-run with larger buffers on overflow.
    pg_status emit_pass(long long seg_cap0, long long band_cap0, bool store_r)
    {
        long long sc = seg_cap0, bcap = band_cap0;
        for (int attempt = 0; attempt < 3; attempt++) {
            PG_TRY(emit_launch(sc, bcap, store_r));
            PG_HIP(ctx, hipMemcpyAsync(hseg.data(), seg_cnt, nwg * 8, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipMemcpyAsync(hcnt, counters, sizeof hcnt, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            const long long mx = max_segment();
            if (mx <= seg_cap && (long long)hcnt[0] <= band_cap) return PG_OK;
            if (attempt == 2) break;
            sc = std::max(seg_cap, mx + mx / 8 + 4096);
            bcap = std::max(band_cap, (long long)hcnt[0] + (long long)hcnt[0] / 8 + 4096);
        }
        return pg_fail(ctx, PG_ENOMEM, "edge buffers overflowed repeatedly");
    }

    long long max_segment() const
    {
        long long mx = 0;
        for (auto c : hseg) mx = std::max<long long>(mx, (long long)c);
        return mx;
    }

    // Histogram Gram pass over all tiles (stride 0) or a strided sample of
    // tiles, reduced into htot on the device.
    pg_status hist_launch(long long stride)
    {
        uint32_t *hist;
        PG_TRY(pg_bufT(ctx, "corr.hist", (size_t)nwg_hist * NBIN, &hist));
        PG_TRY(pg_bufT(ctx, "corr.htot", NBIN, &htot));
        ga.hist = hist;
        ga.sstep = stride;
        PG_TRY(launch_gram(MODE_HIST));
        ga.sstep = 0;
        hist_reduce<<<NBIN / 32, 256, 0, st>>>(hist, (int)nwg_hist, htot);
        PG_KERNEL_CHECK(ctx);
        return PG_OK;
    }

    pg_status hist_pass(long long stride, std::vector<unsigned long long> &h)
    {
        PG_TRY(hist_launch(stride));
        h.resize(NBIN);
        PG_HIP(ctx, hipMemcpyAsync(h.data(), htot, NBIN * 8, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        return PG_OK;
    }

    // CSR build (bucketed counting sort), launched without a host check.
    // seg_counted: every source's bucket counts are in place (the segments'
    // from window_count, the selected list's from select_write).
    pg_status launch_csr(const EdgeSrc &src, long long nnz, bool seg_counted = false)
    {
        if (nnz > INT32_MAX) return pg_fail(ctx, PG_ENOTSUP, "nnz %lld exceeds int32 CSR", (long long)nnz);
        const size_t nnzs = (size_t)std::max<long long>(nnz, 1);
        unsigned *bcnt, *ent;
        long long *boff, *bstart, *btot;
        int *ctmp = nullptr;
        float *entv = nullptr, *vtmp = nullptr;
        PG_TRY(pg_bufT(ctx, "corr.bcnt", (size_t)nsrc * nbk, &bcnt));
        PG_TRY(pg_bufT(ctx, "corr.boff", (size_t)nsrc * nbk, &boff));
        PG_TRY(pg_bufT(ctx, "corr.bstart", (size_t)nbk + 1, &bstart));
        PG_TRY(pg_bufT(ctx, "corr.btot", (size_t)nbk + 1, &btot));
        PG_TRY(pg_bufT(ctx, "corr.ent", nnzs, &ent));
        PG_TRY(pg_bufT(ctx, "corr.row", (size_t)n + 2, &row));
        PG_TRY(pg_bufT(ctx, "corr.col", nnzs, &col));
        const BitsPlan bp = weighted ? BitsPlan() : rows_bits_plan(n, ctx->lds_per_block);
        const size_t lbits = bp.lds;
        if (!lbits) PG_TRY(pg_bufT(ctx, "corr.ctmp", nnzs, &ctmp));
        val = nullptr;
        if (weighted) {
            PG_TRY(pg_bufT(ctx, "corr.entv", nnzs, &entv));
            PG_TRY(pg_bufT(ctx, "corr.vtmp", nnzs, &vtmp));
            PG_TRY(pg_bufT(ctx, "corr.val", nnzs, &val));
        }
        if (n > 0) {
            const size_t shb = (size_t)nbk * 4;
            if (!seg_counted) bucket_count<<<(unsigned)nsrc, SEG_THREADS, shb, st>>>(src, nbk, 0, nsrc, bcnt);
            bucket_scan_local<<<(unsigned)pg_cdiv(nbk, SCAN_BK), SCAN_BK * SCAN_SG, 0, st>>>(bcnt, nsrc, nbk, boff,
                                                                                          btot);
            bucket_scan_global<<<1, 1024, 0, st>>>(btot, nbk, bstart);
            const size_t shs = (size_t)nbk * (4 + 4);  // bases + cursors
            const void *scatter_fn = weighted ? (const void *)bucket_scatter<true> : (const void *)bucket_scatter<false>;
            if (shs > 64 * 1024)
                PG_HIP(ctx, set_shared_limit(ctx, scatter_fn, shs));
            if (weighted)
                bucket_scatter<true><<<(unsigned)nsrc, SEG_THREADS, shs, st>>>(src, nbk, boff, bstart, ent, entv, z64,
                                                                               t, (long long)nnzs);
            else
                bucket_scatter<false><<<(unsigned)nsrc, SEG_THREADS, shs, st>>>(src, nbk, boff, bstart, ent, nullptr,
                                                                                z64, t, (long long)nnzs);
            if (lbits) {
                PG_HIP(ctx, set_shared_limit(ctx, rows_bits_fn(bp.rank), (int)lbits));
                launch_rows_bits(ctx, bp, st, bstart, ent, n, nbk, row, col, (long long)nnzs);
                PG_KERNEL_CHECK(ctx);
                return PG_OK;
            }
            const RowsPlan rp = rows_plan(n, ctx->lds_per_block);
            if (rp.nbm == 0) return pg_fail(ctx, PG_ENOTSUP, "n = %d: the row bitmap does not fit LDS", n);
            const void *rows_fn = weighted ? (const void *)bucket_rows<true> : (const void *)bucket_rows<false>;
            PG_HIP(ctx, set_shared_limit(ctx, rows_fn, rp.lds));
            if (weighted)
                bucket_rows<true><<<(unsigned)nbk, ROWS_THREADS, rp.lds, st>>>(
                    bstart, ent, entv, n, nbk, rp.nbm, rp.dbit, row, ctmp, vtmp, col, val, (long long)nnzs);
            else
                bucket_rows<false><<<(unsigned)nbk, ROWS_THREADS, rp.lds, st>>>(
                    bstart, ent, nullptr, n, nbk, rp.nbm, rp.dbit, row, ctmp, nullptr, col, nullptr, (long long)nnzs);
            PG_KERNEL_CHECK(ctx);
        } else {
            PG_HIP(ctx, hipMemsetAsync(row, 0, 4, st));
        }
        return PG_OK;
    }

    // launch, wait, and check row[n]
    pg_status finish_csr(const EdgeSrc &src, long long nnz)
    {
        PG_HIP(ctx, stage_event(ctx, 4, st));
        PG_TRY(launch_csr(src, nnz));
        PG_HIP(ctx, stage_event(ctx, 5, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        int rn = 0;
        PG_HIP(ctx, hipMemcpy(&rn, row + n, 4, hipMemcpyDeviceToHost));
        if ((long long)rn != nnz)
            return pg_fail(ctx, PG_EFORMAT, "CSR build mismatch: row[n]=%d nnz=%lld", rn, (long long)nnz);
        return PG_OK;
    }

    EdgeSrc edges(int2 *extra, unsigned long long *extra_cnt, long long extra_cap) const
    {
        EdgeSrc es;
        std::memset(&es, 0, sizeof es);
        es.seg = seg;
        es.seg_cap = seg_cap;
        es.seg_cnt = seg_cnt;
        es.nseg = (int)nwg;
        es.extra = extra;
        es.extra_cnt = extra_cnt;
        es.extra_cap = extra_cap;
        return es;
    }

    // Stage times from the events, the result kept in the context.
    void finish(pg_corr_info &inf)
    {
        // stages not timed at the context's timing level report 0
        float ms = 0.0f, ms_hist = 0.0f;
        if (ctx->timing >= 2) {
            (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
            inf.ms_zscore = ms;
            (void)hipEventElapsedTime(&ms_hist, ctx->ev[1], ctx->ev[2]);
            (void)hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
            inf.ms_gram = ms + ms_hist;
            (void)hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]);
            inf.ms_select = ms;
            (void)hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]);
            inf.ms_csr = ms;
        }
        if (ctx->timing >= 2 && ran_hist && hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]) == hipSuccess)
            inf.ms_gram_hist = ms;
        if (ctx->timing >= 1 && ran_emit && hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]) == hipSuccess)
            inf.ms_gram_emit = ms;
        inf.gram_passes = gram_passes;
        ctx->cn = n;
        ctx->ct = t;
        ctx->cnnz = inf.nnz;
        ctx->cweighted = weighted;
        ctx->c_row = row;
        ctx->c_col = col;
        ctx->c_val = val;
        ctx->c_z64 = z64;
    }
};

// Fine histogram of the kept r~ (bins over [ctl->lower, 1 + 1e-6]) into *h2t.
pg_status fine_hist(CorrRun &R, const SelCtl *ctl, unsigned long long **h2t_out)
{
    pg_ctx *ctx = R.ctx;
    hipStream_t st = R.st;
    const int hb = (int)std::min<long long>(256, R.nwg);
    uint32_t *h2;
    unsigned long long *h2t;
    PG_TRY(pg_bufT(ctx, "corr.h2", (size_t)hb * NBIN2, &h2));
    PG_TRY(pg_bufT(ctx, "corr.h2t", NBIN2, &h2t));
    emitted_hist<<<hb, 1024, 0, st>>>(R.seg_r, R.seg_cap, R.seg_cnt, (int)R.nwg, ctl, h2, h2t);
    hist_sum<<<dim3(NBIN2 / 256, 16), 256, 0, st>>>(h2, hb, NBIN2, h2t);
    PG_KERNEL_CHECK(ctx);
    *h2t_out = h2t;
    return PG_OK;
}

struct Window {
    int2 *cand = nullptr, *extra = nullptr;
    float *cand_r = nullptr;
    double *cand64 = nullptr;
    unsigned long long *dpart = nullptr;
    int nd = 0;  // entries of dpart
    SelHist sh{nullptr, nullptr, 0.0};  // the candidates' selection histogram
};

// Split the kept segment entries by ctl's window [L2, U2]: r~ > U2 are
// counted as definite (per-block counts in dpart), the rest of the window
// goes to the candidate list (count in counters[0], at most ccap written),
// whose FP64 r is then computed (largest |r~ - r64| into counters[5]).
// count_segments: the split also fills the segments' rows of the CSR
// build's bucket counts (window_count, the definite edges being exactly the
// segment entries the build keeps: r~ above ctl's U2); launch_csr then counts
// only the selected list.
// cleared: the counters and the selection histogram were zeroed already
// (window_bounds).
pg_status window_candidates(CorrRun &R, const SelCtl *ctl, long long ccap, Window *win, bool count_segments = false,
                            bool cleared = false)
{
    pg_ctx *ctx = R.ctx;
    hipStream_t st = R.st;
    PG_TRY(pg_bufT(ctx, "corr.cand", (size_t)ccap, &win->cand));
    PG_TRY(pg_bufT(ctx, "corr.candr", (size_t)ccap, &win->cand_r));
    PG_TRY(pg_bufT(ctx, "corr.cand64", (size_t)ccap, &win->cand64));
    PG_TRY(pg_bufT(ctx, "corr.extra", (size_t)ccap, &win->extra));
    PG_TRY(pg_bufT(ctx, "corr.selhist", (size_t)SEL_NB, &win->sh.h));
    win->sh.ctl = ctl;
    win->sh.band = R.band;
    if (!cleared)
        PG_TRY(clear_dev(ctx, st, R.counters, 16 * sizeof(unsigned long long), win->sh.h, SEL_NB * sizeof(unsigned)));
    if (count_segments) {
        win->nd = (int)R.nwg;
        PG_TRY(pg_bufT(ctx, "corr.dpart", (size_t)win->nd, &win->dpart));
        unsigned *bcnt;
        PG_TRY(pg_bufT(ctx, "corr.bcnt", (size_t)R.nsrc * R.nbk, &bcnt));
        EdgeSrc es = R.edges(nullptr, nullptr, 0);
        es.seg_r = R.seg_r;
        PG_HIP(ctx, set_shared_limit(ctx, (const void *)window_count, (int)((size_t)R.nbk * 4)));
        window_count<<<(unsigned)R.nwg, SEG_THREADS, (size_t)R.nbk * 4, st>>>(
            es, ctl, R.nbk, R.nsrc, bcnt, win->cand, win->cand_r, ccap, R.counters + 0, win->dpart);
    } else {
        const unsigned wx = (unsigned)std::max<long long>(1, pg_cdiv(R.seg_cap, WIN_CHUNK));
        win->nd = (int)(wx * R.nwg);
        PG_TRY(pg_bufT(ctx, "corr.dpart", (size_t)win->nd, &win->dpart));
        window_kernel<<<dim3(wx, (unsigned)R.nwg), 256, 0, st>>>(R.seg, R.seg_r, R.seg_cap, R.seg_cnt, (int)R.nwg,
                                                                  ctl, win->cand, win->cand_r, ccap, R.counters + 0,
                                                                  win->dpart);
    }
    launch_recheck(ctx, st, win->cand, R.counters + 0, ccap, R.z64, R.n, R.t, win->cand64, win->cand_r,
                   R.counters + 5, win->sh);
    PG_KERNEL_CHECK(ctx);
    return PG_OK;
}

// Sparsity: the exact selection of the k - definite best candidates
// (select_kernel; t* into tstar_d, errors into counters[4]) and the definite
// count into counters[3].
pg_status window_select(CorrRun &R, const SelCtl *ctl, long long ccap, long long kwant, double *tstar_d,
                        Window *win, bool count_segments = false, bool cleared = false)
{
    PG_TRY(window_candidates(R, ctl, ccap, win, count_segments, cleared));
    hipStream_t st = R.st;
    unsigned long long *c = R.counters;
    SelCut *cut;
    PG_TRY(pg_bufT(R.ctx, "corr.selcut", 1, &cut));
    select_kernel<<<1, 1024, 0, st>>>(win->cand, win->cand64, win->cand_r, c + 0, ccap, kwant, win->dpart, win->nd,
                                      win->extra, c + 1, tstar_d, c + 2, (int *)(c + 4), win->sh, c + 3, cut);
    unsigned *bcnt_row = nullptr;  // with the segments' counts, the list's row of them
    if (count_segments) {
        unsigned *bcnt;
        PG_TRY(pg_bufT(R.ctx, "corr.bcnt", (size_t)R.nsrc * R.nbk, &bcnt));
        bcnt_row = bcnt + (size_t)R.nwg * R.nbk;
    }
    select_write<<<(unsigned)std::min<long long>(pg_cdiv(ccap, 256), 4LL * std::max(R.ctx->num_cu, 1)), 256, 0, st>>>(
        win->cand, win->cand64, win->cand_r, cut, win->extra, c + 1, c + 2, bcnt_row);
    PG_KERNEL_CHECK(R.ctx);
    return PG_OK;
}

// The band is error_band()'s proven bound times 1.5; a rechecked pair whose
// fast r~ is further from its FP64 r than the bound itself means the rounding
// model behind the band does not hold on this device, and pairs outside the
// band may then be misclassified.
pg_status check_band(pg_ctx *ctx, const pg_corr_info &inf)
{
    if (inf.band_max_err > inf.band / 1.5)
        return pg_fail(ctx, PG_EDEVICE,
                       "fast-path error %.3g exceeds the proven bound %.3g (band %.3g): edge set not guaranteed",
                       inf.band_max_err, inf.band / 1.5, inf.band);
    return PG_OK;
}

}  // namespace

extern "C" pg_status pg_corr_build(pg_ctx *ctx, const float *bold, int32_t n, int32_t t,
                                   const pg_corr_opts *opts, pg_corr_info *info)
{
    if (!ctx) return PG_EINVAL;
    PG_TRY(check_opts(ctx, bold, n, t, opts));
    PG_HIP(ctx, hipSetDevice(ctx->device));
    pg_emit_invalidate(ctx);
    if (n > 1 && pg_autotune_on_first_use(ctx) && pg_tune_get(ctx, "gram_super", n, t, 0) == 0) {
        pg_tune_result tr;
        PG_TRY(pg_autotune_corr(ctx, bold, n, t, opts, &tr));
    }
    const bool smode = opts->thr_kind == PG_THR_SPARSITY;
    CorrRun R;
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
        kwant = sparsity_k(m1, opts->thr);
        inf.s_thresh = opts->thr;
    } else {
        thr_eff = thr_effective(opts->thr, opts->weighted != 0);
        inf.r_thresh = opts->thr;
        inf.t_star = (double)thr_eff;
    }
    PG_TRY(R.setup(ctx, bold, n, t, opts));
    hipStream_t st = R.st;
    const double band = R.band;
    inf.band = band;
    double *tstar_d = nullptr;
    PG_TRY(pg_bufT(ctx, "corr.tstar", 2, &tstar_d));
    unsigned long long *counters = R.counters;
    unsigned long long *hcnt = R.hcnt;
    const long long nwg = R.nwg, nst = R.nst;

    int2 *extra = nullptr;
    unsigned long long *extra_cnt = counters + 1, *flips = counters + 2;

    if (!smode) {
        // ---------------------------------------------------- r threshold
        R.ga.upper = round_up_f((double)thr_eff + band);
        R.ga.lower = round_down_f((double)thr_eff - band);
        PG_HIP(ctx, stage_event(ctx, 2, st));
        const long long seg0 = pg_cdiv(std::max<long long>((long long)(m1 * 0.03), 1 << 20), nwg) + 4096;
        PG_TRY(R.emit_pass(seg0, 1 << 22, false));
        PG_HIP(ctx, stage_event(ctx, 3, st));
        long long definite = 0;
        for (auto c : R.hseg) definite += (long long)c;
        const long long nband = (long long)hcnt[0];
        inf.band_pairs = nband;
        double *band64;
        PG_TRY(pg_bufT(ctx, "corr.band64", (size_t)std::max<long long>(nband, 1), &band64));
        PG_TRY(pg_bufT(ctx, "corr.extra", (size_t)std::max<long long>(nband, 1), &extra));
        if (nband > 0) {
            launch_recheck(ctx, st, R.bandl, counters + 0, std::min(nband, R.band_cap), R.z64, n, t, band64, R.band_r,
                           counters + 5);
            accept_kernel<<<(unsigned)pg_cdiv(nband, 256), 256, 0, st>>>(
                R.bandl, R.band_r, band64, counters + 0, R.band_cap, (double)thr_eff, extra, extra_cnt, flips);
        }
        PG_KERNEL_CHECK(ctx);
        PG_HIP(ctx, hipMemcpyAsync(hcnt, counters, sizeof R.hcnt, hipMemcpyDeviceToHost, st));
        PG_HIP(ctx, hipStreamSynchronize(st));
        inf.band_flips = (long long)hcnt[2];
        inf.band_max_err = host_err(hcnt[5]);
        inf.nnz = 2 * (definite + (long long)hcnt[1]);
        inf.s_thresh = (float)(m1 > 0 ? 100.0 * (double)inf.nnz / (double)m1 / 2.0 : 0.0);
        PG_TRY(R.finish_csr(R.edges(extra, extra_cnt, std::max<long long>(nband, 1)), inf.nnz));
    } else if (kwant == 0 || n < 2) {
        // ---------------------------------------------------- nothing asked
        PG_HIP(ctx, stage_event(ctx, 2, st));
        PG_TRY(clear_dev(ctx, st, counters, 16 * sizeof(unsigned long long), R.seg_cnt,
                         nwg * sizeof(unsigned long long)));
        PG_HIP(ctx, stage_event(ctx, 3, st));
        PG_TRY(pg_bufT(ctx, "corr.extra", 1, &extra));
        PG_TRY(pg_bufT(ctx, "corr.seg", 1, &R.seg));
        R.seg_cap = 0;
        inf.nnz = 0;
        inf.r_thresh = 1.0f;
        inf.t_star = 1.0;
        PG_TRY(R.finish_csr(R.edges(extra, extra_cnt, 1), 0));
    } else {
        // ---------------------------------------------------- sparsity
        const double w = 2.0 / NBIN;
        bool done = false;
        // about 1% of the supertiles, when there are enough to sample
        const long long stride = sample_stride(nst);
        if (stride > 1) {
            // Single pass, no host round trip until the end:
            // (1) sampled histogram -> a conservative lower bound L on the
            //     k-th largest r~ (device: sample_bound);
            // (2) one emitting pass keeps every pair with r~ >= L and its r~;
            // (3) fine histogram of the kept r~ -> the bin of the k-th,
            //     widened by 2 bands each side (device: window_bounds);
            // (4) candidates in the window, FP64 recheck, exact selection;
            // (5) CSR build; then one check of every step's assumption.
            const long long msamp = sampled_pairs(R.stl, stride, R.nr, R.nc, R.rt, R.tm, R.tn, n, R.sup);
            const double p = (double)kwant / (double)m1;
            const double meff = std::max(1.0, (double)msamp / 64.0);  // pairs inside a tile are correlated
            const double f = std::min(1.0, 1.15 * p + 6.0 * std::sqrt(p * (1.0 - p) / meff) + 1.0 / meff);
            SelCtl *ctl;
            PG_TRY(pg_bufT(ctx, "corr.selctl", 1, &ctl));
            PG_TRY(R.hist_launch(stride));
            sample_bound<<<1, 1024, 0, st>>>(R.htot, f, ctl);
            PG_HIP(ctx, stage_event(ctx, 2, st));
            R.ga.lower_dev = &ctl->lower;
            R.ga.upper = -4.0f;  // every hit is a segment entry
            const double expect = std::min(1.0, f * 1.2 + 2.0 * w) * (double)m1;
            R.dedic = dedic_fits(expect / std::max<double>((double)m1, 1.0));
            // the counters are zeroed by window_bounds below, before their first use
            PG_TRY(R.emit_launch(pg_cdiv((long long)expect, nwg) * 5 / 4 + 4096, 1 << 10, true, false));
            R.dedic = false;
            R.ga.lower_dev = nullptr;
            unsigned long long *h2t;
            PG_TRY(fine_hist(R, ctl, &h2t));
            unsigned *selhist;
            PG_TRY(pg_bufT(ctx, "corr.selhist", (size_t)SEL_NB, &selhist));
            window_bounds<<<1, 1024, 0, st>>>(h2t, kwant, band, ctl, counters, 16, selhist, SEL_NB);
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, stage_event(ctx, 3, st));
            const long long ccap = std::min<long long>(1 << 22, std::max<long long>(1 << 16, kwant / 4));
            Window win;
            PG_TRY(window_select(R, ctl, ccap, kwant, tstar_d, &win, true, true));
            extra = win.extra;
            EdgeSrc es = R.edges(extra, extra_cnt, ccap);
            es.seg_r = R.seg_r;
            es.keep_above_dev = &ctl->U2;
            PG_HIP(ctx, stage_event(ctx, 4, st));
            PG_TRY(R.launch_csr(es, 2 * kwant, true));
            PG_HIP(ctx, stage_event(ctx, 5, st));
            // every check's inputs gathered on the device, one small copy back
            SpReport *rep_d;
            PG_TRY(pg_bufT(ctx, "corr.report", 1, &rep_d));
            sp_report<<<1, 256, 0, st>>>(R.seg_cnt, (int)nwg, counters, ctl, tstar_d, n > 0 ? R.row + n : nullptr,
                                         rep_d);
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, hipMemcpyAsync(ctx->h_report, rep_d, sizeof(SpReport), hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            const SpReport &rep = *(const SpReport *)ctx->h_report;
            std::memcpy(hcnt, rep.counters, sizeof R.hcnt);
            const SelCtl hc = rep.ctl;
            const int rn = rep.rn;
            inf.t_star = rep.tstar;
            const long long mx = (long long)rep.seg_max;
            done = hc.ok == 1 && mx <= R.seg_cap && (long long)hcnt[0] <= ccap && ((int *)hcnt)[8] == 0 &&
                   (long long)hcnt[3] + (long long)hcnt[1] == kwant && (long long)rn == 2 * kwant;
            if (done) {
                inf.band_pairs = (long long)hcnt[0];
                inf.band_flips = (long long)hcnt[2];
                inf.band_max_err = host_err(hcnt[5]);
            }
        }
        if (!done) {
            // (fallback) exact bracket from a histogram of every pair, then an
            // emitting pass with a definite/candidate split
            std::vector<unsigned long long> h;
            PG_TRY(R.hist_pass(0, h));
            long long above = 0;
            int bstar = -1;
            for (int b = NBIN - 1; b >= 0; b--) {
                if (above + (long long)h[b] >= kwant) {
                    bstar = b;
                    break;
                }
                above += (long long)h[b];
            }
            if (bstar < 0) return pg_fail(ctx, PG_ETHRESH, "histogram holds fewer than k=%lld pairs", kwant);
            const double lo_edge = bstar * w - 1.0 - 1e-6;
            const double hi_edge = (bstar + 1) * w - 1.0 + 1e-6;
            R.ga.upper = round_up_f(hi_edge + 2.0 * band);
            R.ga.lower = round_down_f(lo_edge - 2.0 * band);
            PG_HIP(ctx, stage_event(ctx, 2, st));
            PG_TRY(R.emit_pass(pg_cdiv(kwant, nwg) * 3 / 2 + 4096, 1 << 22, false));
            PG_HIP(ctx, stage_event(ctx, 3, st));
            const long long nband = (long long)hcnt[0];
            double *band64;
            PG_TRY(pg_bufT(ctx, "corr.band64", (size_t)std::max<long long>(nband, 1), &band64));
            PG_TRY(pg_bufT(ctx, "corr.extra", (size_t)std::max<long long>(nband, 1), &extra));
            if (nband > 0)
                launch_recheck(ctx, st, R.bandl, counters + 0, std::min(nband, R.band_cap), R.z64, n, t, band64,
                               R.band_r, counters + 5);
            select_kernel<<<1, 1024, 0, st>>>(R.bandl, band64, R.band_r, counters + 0, R.band_cap, kwant, R.seg_cnt,
                                              (int)nwg, extra, extra_cnt, tstar_d, flips, (int *)(counters + 4),
                                              SelHist{nullptr, nullptr, 0.0}, nullptr, nullptr);
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, hipMemcpyAsync(hcnt, counters, sizeof R.hcnt, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipMemcpyAsync(&inf.t_star, tstar_d, 8, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            if (((int *)hcnt)[8] != 0)
                return pg_fail(ctx, PG_ETHRESH, "selection window invalid (k=%lld candidates=%lld)", kwant,
                               (long long)hcnt[0]);
            inf.band_pairs = nband;
            inf.band_flips = (long long)hcnt[2];
            inf.band_max_err = host_err(hcnt[5]);
            PG_TRY(R.finish_csr(R.edges(extra, extra_cnt, std::max<long long>(nband, 1)), 2 * kwant));
        }
        inf.nnz = 2 * kwant;
        inf.r_thresh = (float)inf.t_star;
    }

    PG_TRY(check_band(ctx, inf));
    R.finish(inf);
    if (info) *info = inf;
    return PG_OK;
}

// ------------------------------------------------ one Gram pass, many thresholds

// State of the last pg_corr_emit: the kept pairs (segments with r~), the
// emission bound and, in sparsity mode, the fine histogram of the kept r~.
struct PgEmit {
    bool valid = false;
    CorrRun R;
    int kind = PG_THR_R;
    float lower = 2.0f;           // every pair with r~ >= lower is kept
    long long kwant_max = 0;      // sparsity: the k the emission serves
    SelCtl *ctl = nullptr;        // device control block (lower, window)
    unsigned long long *h2t = nullptr;
    pg_corr_info base;            // the emission's own figures
};

void pg_emit_invalidate(pg_ctx *ctx)
{
    if (ctx && ctx->emit) ctx->emit->valid = false;
}

void pg_emit_free(PgEmit *e) { delete e; }

extern "C" pg_status pg_corr_emit(pg_ctx *ctx, const float *bold, int32_t n, int32_t t, const pg_corr_opts *opts,
                                  pg_corr_info *info)
{
    if (!ctx) return PG_EINVAL;
    PG_TRY(check_opts(ctx, bold, n, t, opts));
    PG_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->emit) ctx->emit = new PgEmit();
    PgEmit &E = *ctx->emit;
    E.valid = false;
    ctx->cnnz = -1;  // no CSR until pg_corr_select
    E.R = CorrRun();
    CorrRun &R = E.R;
    const bool smode = opts->thr_kind == PG_THR_SPARSITY;
    E.kind = smode ? PG_THR_SPARSITY : PG_THR_R;
    pg_corr_info inf;
    std::memset(&inf, 0, sizeof inf);
    inf.n = n;
    inf.t = t;
    PG_TRY(R.setup(ctx, bold, n, t, opts));
    hipStream_t st = R.st;
    const double band = R.band;
    const long long m1 = R.m1, nwg = R.nwg;
    inf.band = band;
    PG_TRY(pg_bufT(ctx, "corr.selctl", 1, &E.ctl));
    SelCtl hc;
    std::memset(&hc, 0, sizeof hc);
    if (!smode) {
        // every pair the loosest threshold could keep, with its r~
        const float thr_eff = thr_effective(opts->thr, opts->weighted != 0);
        E.lower = round_down_f((double)thr_eff - band);
        // segment sizes from a sampled histogram (about 1% of the supertiles)
        // when there are enough supertiles to sample, else a fixed guess
        long long seg0 = pg_cdiv(std::max<long long>((long long)(m1 * 0.03), 1 << 20), nwg) + 4096;
        double dens = 1.0;  // expected hit fraction (unknown without the sampled pass)
        const long long stride = sample_stride(R.nst);
        if (stride > 1 && n > 1) {
            std::vector<unsigned long long> h;
            PG_TRY(R.hist_pass(stride, h));
            const long long msamp = sampled_pairs(R.stl, stride, R.nr, R.nc, R.rt, R.tm, R.tn, n, R.sup);
            int b0 = (int)(((double)E.lower + 1.0) * (NBIN * 0.5));
            b0 = std::max(0, std::min(NBIN - 1, b0));
            unsigned long long above = 0;
            for (int b = b0; b < NBIN; b++) above += h[b];
            const double frac = msamp > 0 ? (double)above / (double)msamp : 1.0;
            const double expect = std::min(1.0, frac * 1.25 + 0.002) * (double)m1;
            seg0 = pg_cdiv((long long)expect, nwg) * 5 / 4 + 4096;
            dens = expect / std::max<double>((double)m1, 1.0);
        }
        R.ga.lower = E.lower;
        R.ga.upper = -4.0f;
        PG_HIP(ctx, stage_event(ctx, 2, st));
        R.dedic = dedic_fits(dens);
        PG_TRY(R.emit_pass(seg0, 1 << 10, true));
        R.dedic = false;
        PG_HIP(ctx, stage_event(ctx, 3, st));
        hc.lower = E.lower;
        hc.ok = 1;
        PG_HIP(ctx, hipMemcpyAsync(E.ctl, &hc, sizeof hc, hipMemcpyHostToDevice, st));
        inf.r_thresh = opts->thr;
        inf.t_star = (double)thr_eff;
    } else {
        E.kwant_max = sparsity_k(m1, opts->thr);
        inf.s_thresh = opts->thr;
        const double w = 2.0 / NBIN;
        bool ok = false;
        const long long stride = sample_stride(R.nst);
        if (E.kwant_max > 0 && n > 1 && stride > 1) {
            // the single pass's sampled bound for the largest k
            const long long msamp = sampled_pairs(R.stl, stride, R.nr, R.nc, R.rt, R.tm, R.tn, n, R.sup);
            const double p = (double)E.kwant_max / (double)m1;
            const double meff = std::max(1.0, (double)msamp / 64.0);
            const double f = std::min(1.0, 1.15 * p + 6.0 * std::sqrt(p * (1.0 - p) / meff) + 1.0 / meff);
            PG_TRY(R.hist_launch(stride));
            sample_bound<<<1, 1024, 0, st>>>(R.htot, f, E.ctl);
            PG_HIP(ctx, stage_event(ctx, 2, st));
            R.ga.lower_dev = &E.ctl->lower;
            R.ga.upper = -4.0f;
            const double expect = std::min(1.0, f * 1.2 + 2.0 * w) * (double)m1;
            long long cap = pg_cdiv((long long)expect, nwg) * 5 / 4 + 4096;
            R.dedic = dedic_fits(expect / std::max<double>((double)m1, 1.0));
            for (int attempt = 0; attempt < 2; attempt++) {
                PG_TRY(R.emit_launch(cap, 1 << 10, true));
                PG_HIP(ctx, hipMemcpyAsync(R.hseg.data(), R.seg_cnt, nwg * 8, hipMemcpyDeviceToHost, st));
                PG_HIP(ctx, hipStreamSynchronize(st));
                const long long mx = R.max_segment();
                if (mx <= R.seg_cap) break;
                cap = mx + mx / 8 + 4096;
            }
            R.ga.lower_dev = nullptr;
            R.dedic = false;
            PG_TRY(fine_hist(R, E.ctl, &E.h2t));
            window_bounds<<<1, 1024, 0, st>>>(E.h2t, E.kwant_max, band, E.ctl);
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, hipMemcpyAsync(&hc, E.ctl, sizeof hc, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            ok = hc.ok == 1 && R.max_segment() <= R.seg_cap;
            E.lower = hc.lower;
        }
        if (E.kwant_max > 0 && n > 1 && !ok) {
            // the bound from a histogram of every pair: the bin of the k-th
            // largest r~, widened by two bands and one bin
            std::vector<unsigned long long> h;
            PG_TRY(R.hist_pass(0, h));
            long long above = 0;
            int bstar = -1;
            for (int b = NBIN - 1; b >= 0; b--) {
                if (above + (long long)h[b] >= E.kwant_max) {
                    bstar = b;
                    break;
                }
                above += (long long)h[b];
            }
            if (bstar < 0) return pg_fail(ctx, PG_ETHRESH, "histogram holds fewer than k=%lld pairs", E.kwant_max);
            E.lower = round_down_f(bstar * w - 1.0 - 1e-6 - 2.0 * band - w);
            hc.lower = E.lower;
            hc.ok = 1;
            PG_HIP(ctx, hipMemcpyAsync(E.ctl, &hc, sizeof hc, hipMemcpyHostToDevice, st));
            PG_HIP(ctx, stage_event(ctx, 2, st));
            R.ga.lower = E.lower;
            R.ga.upper = -4.0f;
            PG_TRY(R.emit_pass(pg_cdiv(above + (long long)h[bstar], nwg) * 3 / 2 + 4096, 1 << 10, true));
            PG_TRY(fine_hist(R, E.ctl, &E.h2t));
            window_bounds<<<1, 1024, 0, st>>>(E.h2t, E.kwant_max, band, E.ctl);
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, hipMemcpyAsync(&hc, E.ctl, sizeof hc, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            if (hc.ok != 1) return pg_fail(ctx, PG_ETHRESH, "selection window below the emitted pairs");
        }
        if (E.kwant_max == 0 || n < 2) {
            // nothing to keep: empty segments, no Gram pass
            PG_HIP(ctx, stage_event(ctx, 2, st));
            R.seg_cap = 64;
            PG_TRY(pg_bufT(ctx, "corr.seg", (size_t)(nwg * R.seg_cap), &R.seg));
            PG_TRY(pg_bufT(ctx, "corr.segr", (size_t)(nwg * R.seg_cap), &R.seg_r));
            PG_TRY(clear_dev(ctx, st, R.seg_cnt, nwg * sizeof(unsigned long long)));
            hc.lower = 2.0f;
            hc.ok = 1;
            PG_HIP(ctx, hipMemcpyAsync(E.ctl, &hc, sizeof hc, hipMemcpyHostToDevice, st));
        }
        PG_HIP(ctx, stage_event(ctx, 3, st));
    }
    PG_HIP(ctx, hipStreamSynchronize(st));
    long long kept = 0;
    PG_HIP(ctx, hipMemcpy(R.hseg.data(), R.seg_cnt, nwg * 8, hipMemcpyDeviceToHost));
    for (auto c : R.hseg) kept += (long long)c;
    inf.band_pairs = kept;  // pairs kept for the selections
    float ms = 0.0f;
    if (ctx->timing >= 2) {
        (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
        inf.ms_zscore = ms;
        float ms_hist = 0.0f;
        (void)hipEventElapsedTime(&ms_hist, ctx->ev[1], ctx->ev[2]);
        (void)hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
        inf.ms_gram = ms + ms_hist;
    }
    if (ctx->timing >= 2 && R.ran_hist && hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]) == hipSuccess)
        inf.ms_gram_hist = ms;
    if (ctx->timing >= 1 && R.ran_emit && hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]) == hipSuccess)
        inf.ms_gram_emit = ms;
    inf.gram_passes = R.gram_passes;
    inf.nnz = -1;
    E.base = inf;
    E.valid = true;
    if (info) *info = inf;
    return PG_OK;
}

extern "C" pg_status pg_corr_select(pg_ctx *ctx, float thr, pg_corr_info *info)
{
    if (!ctx) return PG_EINVAL;
    if (!ctx->emit || !ctx->emit->valid) return pg_fail(ctx, PG_ESTATE, "pg_corr_emit first");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    PgEmit &E = *ctx->emit;
    CorrRun &R = E.R;
    hipStream_t st = R.st;
    const double band = R.band;
    const long long m1 = R.m1;
    const int n = R.n;
    pg_corr_info inf = E.base;
    inf.band_pairs = inf.band_flips = 0;
    inf.band_max_err = 0.0;
    inf.ms_select = inf.ms_csr = 0.0;
    unsigned long long *c = R.counters;
    double *tstar_d = nullptr;
    PG_TRY(pg_bufT(ctx, "corr.tstar", 2, &tstar_d));
    SelCtl hc;
    std::memset(&hc, 0, sizeof hc);
    hc.lower = E.lower;
    hc.ok = 1;
    long long ccap = 1 << 16;
    PG_HIP(ctx, stage_event(ctx, 3, st));
    if (E.kind == PG_THR_R) {
        if (!(thr >= -1.0f && thr <= 1.0f)) return pg_fail(ctx, PG_ETHRESH, "r threshold %g outside [-1, 1]", (double)thr);
        const float thr_eff = thr_effective(thr, R.weighted);
        hc.L2 = round_down_f((double)thr_eff - band);
        hc.U2 = round_up_f((double)thr_eff + band);
        if (hc.L2 < E.lower)
            return pg_fail(ctx, PG_ETHRESH, "r threshold %g is looser than the emitted one", (double)thr);
        long long nnz = 0;
        Window win;
        for (int attempt = 0; attempt < 2; attempt++) {
            PG_HIP(ctx, hipMemcpyAsync(E.ctl, &hc, sizeof hc, hipMemcpyHostToDevice, st));
            PG_TRY(window_candidates(R, E.ctl, ccap, &win));
            accept_kernel<<<(unsigned)pg_cdiv(ccap, 256), 256, 0, st>>>(win.cand, win.cand_r, win.cand64, c + 0, ccap,
                                                                        (double)thr_eff, win.extra, c + 1, c + 2);
            definite_sum<<<1, 1024, 0, st>>>(win.dpart, win.nd, c + 3);
            PG_KERNEL_CHECK(ctx);
            PG_HIP(ctx, hipMemcpyAsync(R.hcnt, c, sizeof R.hcnt, hipMemcpyDeviceToHost, st));
            PG_HIP(ctx, hipStreamSynchronize(st));
            if ((long long)R.hcnt[0] <= ccap) break;
            ccap = (long long)R.hcnt[0] + (long long)R.hcnt[0] / 8 + 4096;
        }
        if ((long long)R.hcnt[0] > ccap) return pg_fail(ctx, PG_ENOMEM, "candidate list overflowed");
        inf.band_pairs = (long long)R.hcnt[0];
        inf.band_flips = (long long)R.hcnt[2];
        inf.band_max_err = host_err(R.hcnt[5]);
        nnz = 2 * ((long long)R.hcnt[3] + (long long)R.hcnt[1]);
        inf.nnz = nnz;
        inf.r_thresh = thr;
        inf.t_star = (double)thr_eff;
        inf.s_thresh = (float)(m1 > 0 ? 100.0 * (double)nnz / (double)m1 / 2.0 : 0.0);
        EdgeSrc es = R.edges(win.extra, c + 1, ccap);
        es.seg_r = R.seg_r;
        es.keep_above_dev = &E.ctl->U2;
        PG_TRY(R.finish_csr(es, nnz));
    } else {
        if (!(thr > 0.0f && thr <= 100.0f)) return pg_fail(ctx, PG_ETHRESH, "sparsity %g%% outside (0, 100]", (double)thr);
        const long long kwant = sparsity_k(m1, thr);
        if (kwant > E.kwant_max)
            return pg_fail(ctx, PG_ETHRESH, "sparsity %g%% keeps more pairs than the emitted one", (double)thr);
        inf.s_thresh = thr;
        if (kwant == 0 || n < 2) {
            Window win;
            PG_TRY(pg_bufT(ctx, "corr.extra", 1, &win.extra));
            PG_TRY(clear_dev(ctx, st, c, 16 * sizeof(unsigned long long)));
            EdgeSrc es = R.edges(win.extra, c + 1, 1);
            es.seg_r = R.seg_r;
            es.keep_above = 2.0f;  // no r~ passes: no segment entry is kept
            PG_TRY(R.finish_csr(es, 0));
            inf.nnz = 0;
            inf.r_thresh = 1.0f;
            inf.t_star = 1.0;
        } else {
            ccap = std::min<long long>(1 << 22, std::max<long long>(1 << 16, kwant / 4));
            bool done = false;
            for (int attempt = 0; attempt < 2 && !done; attempt++) {
                PG_HIP(ctx, hipMemcpyAsync(E.ctl, &hc, sizeof hc, hipMemcpyHostToDevice, st));
                window_bounds<<<1, 1024, 0, st>>>(E.h2t, kwant, band, E.ctl);
                Window win;
                PG_TRY(window_select(R, E.ctl, ccap, kwant, tstar_d, &win));
                EdgeSrc es = R.edges(win.extra, c + 1, ccap);
                es.seg_r = R.seg_r;
                es.keep_above_dev = &E.ctl->U2;
                PG_HIP(ctx, stage_event(ctx, 4, st));
                PG_TRY(R.launch_csr(es, 2 * kwant));
                PG_HIP(ctx, stage_event(ctx, 5, st));
                SelCtl hw;
                int rn = -1;
                PG_HIP(ctx, hipMemcpyAsync(R.hcnt, c, sizeof R.hcnt, hipMemcpyDeviceToHost, st));
                PG_HIP(ctx, hipMemcpyAsync(&hw, E.ctl, sizeof hw, hipMemcpyDeviceToHost, st));
                PG_HIP(ctx, hipMemcpyAsync(&inf.t_star, tstar_d, 8, hipMemcpyDeviceToHost, st));
                PG_HIP(ctx, hipMemcpyAsync(&rn, R.row + n, 4, hipMemcpyDeviceToHost, st));
                PG_HIP(ctx, hipStreamSynchronize(st));
                if (hw.ok != 1) return pg_fail(ctx, PG_ETHRESH, "selection window below the emitted pairs");
                if ((long long)R.hcnt[0] > ccap) {
                    ccap = (long long)R.hcnt[0] + (long long)R.hcnt[0] / 8 + 4096;
                    continue;
                }
                if (((int *)R.hcnt)[8] != 0 || (long long)R.hcnt[3] + (long long)R.hcnt[1] != kwant ||
                    (long long)rn != 2 * kwant)
                    return pg_fail(ctx, PG_ETHRESH, "selection of k=%lld failed (definite %lld, selected %lld)", kwant,
                                   (long long)R.hcnt[3], (long long)R.hcnt[1]);
                done = true;
            }
            if (!done) return pg_fail(ctx, PG_ENOMEM, "candidate list overflowed");
            inf.band_pairs = (long long)R.hcnt[0];
            inf.band_flips = (long long)R.hcnt[2];
            inf.band_max_err = host_err(R.hcnt[5]);
            inf.nnz = 2 * kwant;
            inf.r_thresh = (float)inf.t_star;
        }
    }
    PG_TRY(check_band(ctx, inf));
    float ms = 0.0f;
    if (ctx->timing >= 2 && hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]) == hipSuccess) inf.ms_select = ms;
    if (ctx->timing >= 2 && hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]) == hipSuccess) inf.ms_csr = ms;
    ctx->cn = n;
    ctx->ct = R.t;
    ctx->cnnz = inf.nnz;
    ctx->cweighted = R.weighted;
    ctx->c_row = R.row;
    ctx->c_col = R.col;
    ctx->c_val = R.val;
    ctx->c_z64 = R.z64;
    if (info) *info = inf;
    return PG_OK;
}

// ------------------------------------------- captured cohort: construction

namespace {

// Every assumption of the single sparsity pass that pg_corr_build checks on
// the host after it, checked on the device; a failed one sets a bit of
// *status and the subject is rebuilt outside the graph.
__global__ __launch_bounds__(256) void cohort_corr_check(const SelCtl *__restrict__ ctl,
                                                         const unsigned long long *__restrict__ counters,
                                                         const unsigned long long *__restrict__ seg_cnt, int nseg,
                                                         long long seg_cap, long long ccap, long long kwant,
                                                         const int *__restrict__ row, int n,
                                                         unsigned *__restrict__ status)
{
    bool over = false;
    for (int g = threadIdx.x; g < nseg; g += 256) over |= (long long)seg_cnt[g] > seg_cap;
    if (over) atomicOr(status, PG_COHORT_SEGMENTS);
    if (threadIdx.x == 0) {
        unsigned bad = 0;
        if (ctl->ok != 1) bad |= PG_COHORT_WINDOW;
        if ((long long)counters[0] > ccap) bad |= PG_COHORT_CANDIDATES;
        if (((const int *)(counters + 4))[0] != 0 || (long long)counters[3] + (long long)counters[1] != kwant)
            bad |= PG_COHORT_SELECTION;
        if ((long long)row[n] != 2 * kwant) bad |= PG_COHORT_CSR;
        if (bad) atomicOr(status, bad);
    }
}

}  // namespace

// Plan of the construction part of a captured cohort: per-subject geometry
// and launch sizes (all fixed on the host, as capture requires) and one
// workspace sized for the largest subject, reused by every subject in turn.
struct PgCohortCorr {
    struct Subj {
        int n = 0, t = 0, ldp = 0, nr = 0, npad = 0, nc = 0, nbk = 0, nbm = 1, dbit = SORT_MAX, wx = 1;
        size_t rows_lds = 0, bits_lds = 0;
        bool bits_rank = false;
        long long m1 = 0, kwant = 0, nst = 0, stride = 1, stl_off = 0, seg_cap = 0, ccap = 0;
        double f = 0.0, band = 0.0;
        bool graph = false;  // built in the graph (else outside it)
        bool dedic = false;  // Gram hits staged in the dedicated LDS area
    };
    pg_ctx *ctx = nullptr;
    std::vector<Subj> subj;
    long long nwg = 0, nwg_hist = 0;
    int sup = SUPER;
    double *z64 = nullptr, *cand64 = nullptr, *tstar = nullptr;
    uint16_t *phi = nullptr, *plo = nullptr;
    int2 *stl = nullptr, *seg = nullptr, *cand = nullptr, *extra = nullptr, *band = nullptr;
    uint32_t *hist = nullptr, *h2 = nullptr;
    unsigned long long *htot = nullptr, *seg_cnt = nullptr, *counters = nullptr, *h2t = nullptr, *dpart = nullptr;
    float *seg_r = nullptr, *cand_r = nullptr, *band_r = nullptr;
    SelCtl *ctl = nullptr;
    unsigned *selhist = nullptr;  // the candidates' selection histogram
    SelCut *selcut = nullptr;     // and the cut found on it
    unsigned *bcnt = nullptr, *ent = nullptr;
    long long *boff = nullptr, *bstart = nullptr, *btot = nullptr;
    int *ctmp = nullptr;
};

pg_status pg_cohort_corr_plan(pg_ctx *ctx, int nsub, const int32_t *n, const int32_t *t, const float *spars,
                              PgCohortCorr **out)
{
    PgCohortCorr *P = new PgCohortCorr();
    std::unique_ptr<PgCohortCorr> guard(P);
    P->ctx = ctx;
    P->subj.resize((size_t)nsub);
    P->nwg = std::max<long long>(NXCD, (long long)ctx->num_cu * PG_GRAM_WG_PER_CU / NXCD * NXCD);
    P->nwg_hist = std::max<long long>(NXCD, ctx->num_cu / NXCD * NXCD);
    int max_n0 = 1, max_t0 = 2;
    for (int s = 0; s < nsub; s++) {
        max_n0 = std::max(max_n0, n[s]);
        max_t0 = std::max(max_t0, t[s]);
    }
    P->sup = gram_super(ctx, max_n0, max_t0);
    const long long nwg = P->nwg;
    size_t max_zt = 1, max_panel = 1, max_rows_lds = 16, max_bits_lds = 16;
    long long max_ctmp = 1;
    long long max_seg = 64, max_ccap = 1, max_dpart = 1, max_nnz = 1, max_bsrc = 1, max_nbk = 1;
    int max_t = 2, max_n = 1;
    std::vector<int2> stl_all;
    for (int s = 0; s < nsub; s++) {
        auto &S = P->subj[(size_t)s];
        S.n = n[s];
        S.t = t[s];
        S.m1 = (long long)S.n * (S.n - 1) / 2;
        S.kwant = sparsity_k(S.m1, spars[s]);
        S.ldp = (int)std::max<long long>(2 * g2::KC, pg_cdiv(S.t, g2::KC) * g2::KC);
        S.nr = (int)pg_cdiv(std::max(S.n, 1), g2::TM);
        S.npad = S.nr * g2::TM;
        S.nc = S.npad / g2::TN;
        std::vector<int2> stl;
        supertiles(S.nr, S.nc, g2::RT, P->sup, stl);
        S.nst = (long long)stl.size();
        S.stl_off = (long long)stl_all.size();
        stl_all.insert(stl_all.end(), stl.begin(), stl.end());
        S.stride = sample_stride(S.nst);
        // the single pass needs a sample of the supertiles and T within the
        // LDS-staged z-score; anything else is built outside the graph
        S.graph = S.kwant > 0 && 2 * S.kwant <= INT32_MAX && S.n > 1 && S.stride > 1 && S.t <= ZS_TMAX;
        if (!S.graph) continue;
        const long long msamp = sampled_pairs(stl, S.stride, S.nr, S.nc, g2::RT, g2::TM, g2::TN, S.n, P->sup);
        const double p = (double)S.kwant / (double)S.m1;
        const double meff = std::max(1.0, (double)msamp / 64.0);
        S.f = std::min(1.0, 1.15 * p + 6.0 * std::sqrt(p * (1.0 - p) / meff) + 1.0 / meff);
        const double w = 2.0 / NBIN;
        const double expect = std::min(1.0, S.f * 1.2 + 2.0 * w) * (double)S.m1;
        S.seg_cap = (pg_cdiv((long long)expect, nwg) * 5 / 4 + 4096 + 63) / 64 * 64;
        S.dedic = dedic_fits(expect / std::max<double>((double)S.m1, 1.0));
        // PAGANI_COHORT_SEG_SCALE scales the segment headroom (a value far
        // below 1 makes segments overflow: the subject is then flagged and
        // rebuilt outside the graph; writes past a segment are dropped)
        if (const char *e = std::getenv("PAGANI_COHORT_SEG_SCALE"))
            S.seg_cap = std::max<long long>(64, (long long)((double)S.seg_cap * std::atof(e)) / 64 * 64);
        S.ccap = std::min<long long>(1 << 22, std::max<long long>(1 << 16, S.kwant / 4));
        S.wx = (int)std::max<long long>(1, pg_cdiv(S.seg_cap, WIN_CHUNK));
        S.band = error_band(PREC_BF16X3, S.t);
        S.nbk = (int)pg_cdiv(S.n, BROWS);
        const BitsPlan bp = rows_bits_plan(S.n, ctx->lds_per_block);
        S.bits_lds = bp.lds;
        S.bits_rank = bp.rank;
        if (S.bits_lds) {
            max_bits_lds = std::max(max_bits_lds, S.bits_lds);
        } else {
            const RowsPlan rp = rows_plan(S.n, ctx->lds_per_block);
            if (rp.nbm == 0)
                return pg_fail(ctx, PG_ENOTSUP, "subject %d: n = %d: the row bitmap does not fit LDS", s, S.n);
            S.nbm = rp.nbm;
            S.dbit = rp.dbit;
            S.rows_lds = rp.lds;
            max_rows_lds = std::max(max_rows_lds, rp.lds);
            max_ctmp = std::max(max_ctmp, 2 * S.kwant);
        }
        max_zt = std::max(max_zt, (size_t)S.n * S.t);
        max_panel = std::max(max_panel, (size_t)S.npad * S.ldp);
        max_seg = std::max(max_seg, S.seg_cap);
        max_ccap = std::max(max_ccap, S.ccap);
        max_dpart = std::max(max_dpart, (long long)S.wx * nwg);
        max_nnz = std::max(max_nnz, 2 * S.kwant);
        max_nbk = std::max<long long>(max_nbk, S.nbk);
        max_bsrc = std::max<long long>(max_bsrc, (nwg + 1) * S.nbk);
        max_t = std::max(max_t, S.t);
        max_n = std::max(max_n, S.n);
    }
    PG_TRY(pg_bufT(ctx, "coh.z64", max_zt, &P->z64));
    PG_TRY(pg_bufT(ctx, "coh.phi", max_panel, &P->phi));
    PG_TRY(pg_bufT(ctx, "coh.plo", max_panel, &P->plo));
    PG_TRY(pg_bufT(ctx, "coh.stl", stl_all.size() + 1, &P->stl));
    PG_HIP(ctx, hipMemcpy(P->stl, stl_all.data(), stl_all.size() * sizeof(int2), hipMemcpyHostToDevice));
    PG_TRY(pg_bufT(ctx, "coh.hist", (size_t)P->nwg_hist * NBIN, &P->hist));
    PG_TRY(pg_bufT(ctx, "coh.htot", NBIN, &P->htot));
    PG_TRY(pg_bufT(ctx, "coh.ctl", 1, &P->ctl));
    PG_TRY(pg_bufT(ctx, "coh.seg", (size_t)(nwg * max_seg), &P->seg));
    PG_TRY(pg_bufT(ctx, "coh.segr", (size_t)(nwg * max_seg), &P->seg_r));
    PG_TRY(pg_bufT(ctx, "coh.segcnt", (size_t)nwg, &P->seg_cnt));
    PG_TRY(pg_bufT(ctx, "coh.counters", 16, &P->counters));
    PG_TRY(pg_bufT(ctx, "coh.band", 1, &P->band));
    PG_TRY(pg_bufT(ctx, "coh.bandr", 1, &P->band_r));
    const int hb = (int)std::min<long long>(256, nwg);
    PG_TRY(pg_bufT(ctx, "coh.h2", (size_t)hb * NBIN2, &P->h2));
    PG_TRY(pg_bufT(ctx, "coh.h2t", NBIN2, &P->h2t));
    PG_TRY(pg_bufT(ctx, "coh.cand", (size_t)max_ccap, &P->cand));
    PG_TRY(pg_bufT(ctx, "coh.candr", (size_t)max_ccap, &P->cand_r));
    PG_TRY(pg_bufT(ctx, "coh.cand64", (size_t)max_ccap, &P->cand64));
    PG_TRY(pg_bufT(ctx, "coh.extra", (size_t)max_ccap, &P->extra));
    PG_TRY(pg_bufT(ctx, "coh.dpart", (size_t)max_dpart, &P->dpart));
    PG_TRY(pg_bufT(ctx, "coh.tstar", 2, &P->tstar));
    PG_TRY(pg_bufT(ctx, "coh.selhist", (size_t)SEL_NB, &P->selhist));
    PG_TRY(pg_bufT(ctx, "coh.selcut", 1, &P->selcut));
    PG_TRY(pg_bufT(ctx, "coh.bcnt", (size_t)max_bsrc, &P->bcnt));
    PG_TRY(pg_bufT(ctx, "coh.boff", (size_t)max_bsrc, &P->boff));
    PG_TRY(pg_bufT(ctx, "coh.bstart", (size_t)max_nbk + 1, &P->bstart));
    PG_TRY(pg_bufT(ctx, "coh.btot", (size_t)max_nbk + 1, &P->btot));
    PG_TRY(pg_bufT(ctx, "coh.ent", (size_t)max_nnz, &P->ent));
    PG_TRY(pg_bufT(ctx, "coh.ctmp", (size_t)max_ctmp, &P->ctmp));
    // dynamic LDS limits for the largest subject, set before any capture
    const size_t shz = (size_t)max_t * (ZV + 1) * 4;
    PG_HIP(ctx, set_shared_limit(ctx, (const void *)zscore_fused, (int)shz));
    const size_t stage = (size_t)g2::NSTAGE * g2::STAGE;
    PG_HIP(ctx, set_shared_limit(ctx, (const void *)gram2_kernel<MODE_HIST>, (int)(stage + NBIN * 4)));
    PG_HIP(ctx, set_shared_limit(ctx, (const void *)gram2_kernel<MODE_EMIT, false>, stage));
    PG_HIP(ctx, set_shared_limit(ctx, (const void *)gram2_kernel<MODE_EMIT, false, true>, stage));
    PG_HIP(ctx, set_shared_limit(ctx, (const void *)window_count, (int)((size_t)max_nbk * 4)));
    const size_t shs = (size_t)max_nbk * 8;
    if (shs > 64 * 1024)
        PG_HIP(ctx, set_shared_limit(ctx, (const void *)bucket_scatter<false>, (int)shs));
    PG_HIP(ctx, set_shared_limit(ctx, (const void *)bucket_rows<false>, (int)max_rows_lds));
    PG_HIP(ctx, set_shared_limit(ctx, rows_bits_fn(false), (int)max_bits_lds));
    PG_HIP(ctx, set_shared_limit(ctx, rows_bits_fn(true), (int)max_bits_lds));
    *out = guard.release();
    return PG_OK;
}

bool pg_cohort_corr_in_graph(const PgCohortCorr *P, int s) { return P->subj[(size_t)s].graph; }

// The subject's launches on the context's stream: no allocation, no host
// synchronisation, no memset (fills are kernels), so the sequence can be
// captured.  row/col: the subject's CSR outputs (device, n + 1 and 2k).
pg_status pg_cohort_corr_enqueue(PgCohortCorr *P, int s, const float *dbold, int32_t *row, int32_t *col,
                                 unsigned *status)
{
    pg_ctx *ctx = P->ctx;
    hipStream_t st = ctx->stream;
    const auto &S = P->subj[(size_t)s];
    const long long nwg = P->nwg;
    if (!S.graph) return PG_OK;
    // the z-score writes the panels whole, padding included
    const size_t shz = (size_t)S.t * (ZV + 1) * 4;
    zscore_fused<<<(unsigned)pg_cdiv(S.npad, ZV), 256, shz, st>>>(dbold, PG_BOLD_TIME_MAJOR, S.n, S.t, P->z64, nullptr,
                                                               P->phi, P->plo, S.ldp, S.npad);
    GramArgs ga;
    std::memset(&ga, 0, sizeof ga);
    ga.p0 = P->phi;
    ga.p1 = P->plo;
    ga.ldp = S.ldp;
    ga.kchunks = S.ldp / g2::KC;
    ga.n = S.n;
    ga.nr = S.nr;
    ga.nc = S.nc;
    ga.rt = g2::RT;
    ga.npad = S.npad;
    ga.st_list = P->stl + S.stl_off;
    ga.nst = S.nst;
    ga.super = P->sup;
    ga.sstep = S.stride;
    ga.hist = P->hist;
    const size_t stage = (size_t)g2::NSTAGE * g2::STAGE;
    gram2_kernel<MODE_HIST><<<(unsigned)P->nwg_hist, THREADS, stage + NBIN * 4, st>>>(ga);
    hist_reduce<<<NBIN / 32, 256, 0, st>>>(P->hist, (int)P->nwg_hist, P->htot);
    sample_bound<<<1, 1024, 0, st>>>(P->htot, S.f, P->ctl);
    // (the counters are zeroed by window_bounds before their first use; the
    // band-free emit writes every segment count itself)
    ga.sstep = 0;
    ga.lower_dev = &P->ctl->lower;
    ga.upper = -4.0f;  // every hit is a segment entry, with its r~
    ga.seg = P->seg;
    ga.seg_r = P->seg_r;
    ga.seg_cap = S.seg_cap;
    ga.seg_cnt = P->seg_cnt;
    ga.band = P->band;
    ga.band_r = P->band_r;
    ga.band_cap = 0;
    ga.band_cnt = P->counters + 0;
    if (S.dedic) gram2_kernel<MODE_EMIT, false, true><<<(unsigned)nwg, THREADS, stage, st>>>(ga);
    else gram2_kernel<MODE_EMIT, false><<<(unsigned)nwg, THREADS, stage, st>>>(ga);
    const int hb = (int)std::min<long long>(256, nwg);
    emitted_hist<<<hb, 1024, 0, st>>>(P->seg_r, S.seg_cap, P->seg_cnt, (int)nwg, P->ctl, P->h2, P->h2t);
    hist_sum<<<dim3(NBIN2 / 256, 16), 256, 0, st>>>(P->h2, hb, NBIN2, P->h2t);
    window_bounds<<<1, 1024, 0, st>>>(P->h2t, S.kwant, S.band, P->ctl, P->counters, 16, P->selhist, SEL_NB);
    const SelHist shh{P->selhist, P->ctl, S.band};
    unsigned long long *c = P->counters;
    // the segments as edge sources (r~ above the window kept), then the
    // window split with their bucket counts (window_count)
    EdgeSrc es;
    std::memset(&es, 0, sizeof es);
    es.seg = P->seg;
    es.seg_cap = S.seg_cap;
    es.seg_cnt = P->seg_cnt;
    es.nseg = (int)nwg;
    es.extra = P->extra;
    es.extra_cnt = c + 1;
    es.extra_cap = S.ccap;
    es.seg_r = P->seg_r;
    es.keep_above_dev = &P->ctl->U2;
    const int nsrcb = (int)nwg + 1;
    window_count<<<(unsigned)nwg, SEG_THREADS, (size_t)S.nbk * 4, st>>>(es, P->ctl, S.nbk, nsrcb, P->bcnt, P->cand,
                                                                        P->cand_r, S.ccap, c + 0, P->dpart);
    launch_recheck(ctx, st, P->cand, c + 0, S.ccap, P->z64, S.n, S.t, P->cand64, P->cand_r, c + 5, shh);
    const int nd = (int)nwg;
    select_kernel<<<1, 1024, 0, st>>>(P->cand, P->cand64, P->cand_r, c + 0, S.ccap, S.kwant, P->dpart, nd, P->extra,
                                      c + 1, P->tstar, c + 2, (int *)(c + 4), shh, c + 3, P->selcut);
    select_write<<<(unsigned)std::min<long long>(pg_cdiv(S.ccap, 256), 4LL * std::max(ctx->num_cu, 1)), 256, 0, st>>>(
        P->cand, P->cand64, P->cand_r, P->selcut, P->extra, c + 1, c + 2, P->bcnt + (size_t)nwg * S.nbk);
    // CSR straight into the subject's outputs (every source's counts in place)
    const long long nnz = 2 * S.kwant;
    bucket_scan_local<<<(unsigned)pg_cdiv(S.nbk, SCAN_BK), SCAN_BK * SCAN_SG, 0, st>>>(P->bcnt, nsrcb, S.nbk,
                                                                                     P->boff, P->btot);
    bucket_scan_global<<<1, 1024, 0, st>>>(P->btot, S.nbk, P->bstart);
    bucket_scatter<false><<<(unsigned)nsrcb, SEG_THREADS, (size_t)S.nbk * 8, st>>>(es, S.nbk, P->boff, P->bstart,
                                                                                   P->ent, nullptr, P->z64, S.t, nnz);
    if (S.bits_lds) {
        BitsPlan bp;
        bp.rank = S.bits_rank;
        bp.lds = S.bits_lds;
        launch_rows_bits(ctx, bp, st, P->bstart, P->ent, S.n, S.nbk, row, col, nnz);
    } else {
        bucket_rows<false><<<(unsigned)S.nbk, ROWS_THREADS, S.rows_lds, st>>>(
            P->bstart, P->ent, nullptr, S.n, S.nbk, S.nbm, S.dbit, row, P->ctmp, nullptr, col, nullptr, nnz);
    }
    cohort_corr_check<<<1, 256, 0, st>>>(P->ctl, c, P->seg_cnt, (int)nwg, S.seg_cap, S.ccap, S.kwant, row, S.n,
                                         status);
    PG_KERNEL_CHECK(ctx);
    return PG_OK;
}

void pg_cohort_corr_free(PgCohortCorr *P) { delete P; }

long long pg_sparsity_k(long long n, float s) { return sparsity_k(n * (n - 1) / 2, s); }

pg_status pg_zscore64(pg_ctx *ctx, const float *dbold, int layout, int n, int t, double *z64)
{
    if (n <= 0 || t <= 0) return PG_OK;
    hipStream_t st = ctx->stream;
    if (t <= ZS_TMAX) {
        const size_t shz = (size_t)t * (ZV + 1) * 4;
        PG_HIP(ctx, set_shared_limit(ctx, (const void *)zscore_fused, (int)shz));
        zscore_fused<<<(unsigned)pg_cdiv(n, ZV), 256, shz, st>>>(dbold, layout, n, t, z64, nullptr, nullptr,
                                                                  nullptr, 0, 0);
    } else {
        double *mu, *sd;
        PG_TRY(pg_bufT(ctx, "zs.mu", (size_t)n + 1, &mu));
        PG_TRY(pg_bufT(ctx, "zs.sd", (size_t)n + 1, &sd));
        zscore_stats<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(dbold, layout, n, t, mu, sd);
        dim3 g((unsigned)pg_cdiv(n, 64), (unsigned)pg_cdiv(t, 64));
        zscore_write<<<g, 256, 0, st>>>(dbold, layout, n, t, mu, sd, z64, nullptr, nullptr, nullptr, 0, 0);
    }
    PG_KERNEL_CHECK(ctx);
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
