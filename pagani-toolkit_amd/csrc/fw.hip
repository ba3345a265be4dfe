// fw.hip -- blocked Floyd-Warshall APSP on MI355X (gfx950) VALUs.
//
// Replaces cuAPSP (W/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW.cpp:19-168) and the phase
// kernels of CUBFW_ph1.cu / CUBFW_ph2.cu / CUBFW_ph3.cu.  The whole padded
// matrix stays resident in HBM (the reference streams every tile over PCIe each
// round).  Per k-block round:
//   phase 1: the diagonal tile, sequential k over register micro-tiles (one
//            work-group);
//   phase 2: row and column panel tiles, the same;
//   phase 3: every other tile, min over k of (ik + kj), 4x4 register
//            micro-tiles over LDS panels.
// Min-plus is not a dense contraction, so this runs on VALU (v_add_f32 +
// v_min3_f32), not MFMA.  Each candidate path length is one FP32 add of the
// same operands as the blocked CPU algorithm (W/Lp/CUBFW_Lp/CUBFW_Lp/BFW.cpp:
// 40-108) with the same tile edge, and min is exact, so the result is
// bit-identical to it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "pg_internal.hpp"

namespace {

__device__ __forceinline__ float relax(float cur, float cand) { return cand < cur ? cand : cur; }

// Phase 1/2 tile update with sequential k.  role 0: diagonal tile (ik = kj =
// self); role 1: row panel (ik = diag, kj = self); role 2: column panel
// (ik = self, kj = diag).  The tile lives in registers, an M x M micro-tile per
// thread (M = B/16, 256 threads); at step k the threads holding row k and
// column k of the tile publish them to a double-buffered LDS row / column
// (one barrier per step), everyone relaxes its micro-tile against them.  Row
// and column k do not change during step k (d_kk = 0), as in the sequential
// reference loop, so every element sees the same operands there.
template <int B>
__global__ __launch_bounds__(256) void fw_seq_tiles(float *__restrict__ d, int np, int kb, int phase)
{
    constexpr int M = B / 16;
    __shared__ float diag[B][B + 1];
    __shared__ float rowk[2][B], colk[2][B];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    int tr, tc, role;
    const int nbk = np / B;
    if (phase == 1) {
        tr = kb;
        tc = kb;
        role = 0;
    } else {
        int t = blockIdx.x;  // 0 .. 2*(nbk-1)-1
        int h = nbk - 1;
        if (t < h) { tr = kb; tc = t < kb ? t : t + 1; role = 1; }
        else { t -= h; tr = t < kb ? t : t + 1; tc = kb; role = 2; }
    }
    const long long r0 = (long long)tr * B, c0 = (long long)tc * B, k0 = (long long)kb * B;
    if (role != 0)
        for (int e = tid; e < B * B; e += 256) {
            const int i = e / B, j = e % B;
            diag[i][j] = d[(k0 + i) * np + k0 + j];
        }
    float acc[M][M];
#pragma unroll
    for (int a = 0; a < M; a++)
#pragma unroll
        for (int b = 0; b < M; b++) acc[a][b] = d[(r0 + ty * M + a) * np + c0 + tx * M + b];
    __syncthreads();
    for (int q = 0; q < 16; q++) {  // k = q M + kk
#pragma unroll
        for (int kk = 0; kk < M; kk++) {
            const int k = q * M + kk, buf = k & 1;
            if (role != 2 && ty == q)
#pragma unroll
                for (int b = 0; b < M; b++) rowk[buf][tx * M + b] = acc[kk][b];
            if (role != 1 && tx == q)
#pragma unroll
                for (int a = 0; a < M; a++) colk[buf][ty * M + a] = acc[a][kk];
            __syncthreads();
            float ik[M], kj[M];
#pragma unroll
            for (int a = 0; a < M; a++) ik[a] = role == 1 ? diag[ty * M + a][k] : colk[buf][ty * M + a];
#pragma unroll
            for (int b = 0; b < M; b++) kj[b] = role == 2 ? diag[k][tx * M + b] : rowk[buf][tx * M + b];
#pragma unroll
            for (int a = 0; a < M; a++)
#pragma unroll
                for (int b = 0; b < M; b++) acc[a][b] = relax(acc[a][b], ik[a] + kj[b]);
        }
    }
#pragma unroll
    for (int a = 0; a < M; a++)
#pragma unroll
        for (int b = 0; b < M; b++) d[(r0 + ty * M + a) * np + c0 + tx * M + b] = acc[a][b];
}

// Phase 3: tile (r, c) = min(tile, min_k ik[i][k] + kj[k][j]) for r, c != kb.
// 4 x 4 register micro-tile per thread, (B/4)^2 threads: 256 for B = 64, 1024
// for B = 128 (whose 134 KB of LDS panels allow one work-group per CU).
template <int B>
constexpr int fw3_threads() { return (B / 4) * (B / 4); }

template <int B>
__global__ __launch_bounds__(fw3_threads<B>()) void fw_phase3(float *__restrict__ d, int np, int kb)
{
    constexpr int M = 4;
    constexpr int TS = B / M;           // threads per tile edge
    constexpr int NT = fw3_threads<B>();
    __shared__ float sa[B][B + 1];  // ik panel, [i][k]
    __shared__ float sb[B][B + 4];  // kj panel, [k][j]
    const int nbk = np / B;
    int r = blockIdx.y, c = blockIdx.x;
    r = r < kb ? r : r + 1;
    c = c < kb ? c : c + 1;
    if (r >= nbk || c >= nbk) return;
    const int tid = threadIdx.x, tx = tid % TS, ty = tid / TS;
    const long long r0 = (long long)r * B, c0 = (long long)c * B, k0 = (long long)kb * B;
    for (int e = tid; e < B * B; e += NT) {
        int i = e / B, j = e % B;
        sa[i][j] = d[(r0 + i) * np + k0 + j];
        sb[i][j] = d[(k0 + i) * np + c0 + j];
    }
    float acc[M][M];
#pragma unroll
    for (int a = 0; a < M; a++)
#pragma unroll
        for (int b = 0; b < M; b++) acc[a][b] = d[(r0 + ty * M + a) * np + c0 + tx * M + b];
    __syncthreads();
    for (int k = 0; k < B; k += 2) {
        float x0[M], x1[M], y0[M], y1[M];
#pragma unroll
        for (int a = 0; a < M; a++) {
            x0[a] = sa[ty * M + a][k];
            x1[a] = sa[ty * M + a][k + 1];
        }
#pragma unroll
        for (int b = 0; b < M; b++) {
            y0[b] = sb[k][tx * M + b];
            y1[b] = sb[k + 1][tx * M + b];
        }
#pragma unroll
        for (int a = 0; a < M; a++)
#pragma unroll
            for (int b = 0; b < M; b++) acc[a][b] = fminf(acc[a][b], fminf(x0[a] + y0[b], x1[a] + y1[b]));
    }
#pragma unroll
    for (int a = 0; a < M; a++)
#pragma unroll
        for (int b = 0; b < M; b++) d[(r0 + ty * M + a) * np + c0 + tx * M + b] = acc[a][b];
}

__global__ void fw_pad_in(const float *__restrict__ src, int n, float *__restrict__ dst, int np)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)np * np) return;
    long long i = k / np, j = k % np;
    dst[k] = (i < n && j < n) ? src[i * n + j] : (i == j ? 0.0f : INFINITY);
}

__global__ void fw_pad_out(const float *__restrict__ src, int np, float *__restrict__ dst, int n)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)n * n) return;
    long long i = k / n, j = k % n;
    dst[k] = src[i * np + j];
}

__global__ void fw_fill_inf(float *__restrict__ d, long long total)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < total) d[k] = INFINITY;
}

// W FormBlock.cpp:52-74 in row-major: edge distances, then 0 on the diagonal.
__global__ void fw_from_csr(const int *__restrict__ row, const int *__restrict__ col,
                            const float *__restrict__ val, int n, int kind, float *__restrict__ d)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = row[i]; e < row[i + 1]; e++) {
        float w = val ? val[e] : 1.0f;
        float x = kind == 0 ? 1.0f / w : (kind == 1 ? 1.0f - w : 1.0f);
        d[(long long)i * n + col[e]] = x;
    }
    d[(long long)i * n + i] = 0.0f;
}

// One block per row u: e_u and the FP64 row sum for Lp (CUBFW_Lp.cpp:70-113).
__global__ __launch_bounds__(256) void fw_eff_rows(const float *__restrict__ d, int n, float *__restrict__ eff,
                                                  double *__restrict__ rowsum, int *__restrict__ iso)
{
    __shared__ float rf[256];
    __shared__ double rd[256];
    const int u = blockIdx.x, tid = threadIdx.x;
    float sf = 0.0f;
    double sd = 0.0;
    for (int v = tid; v < n; v += 256) {
        const float x = d[(long long)u * n + v];
        if (v == u || !(x >= 0.1f)) continue;  // the reference skips d < 0.1
        const float q = 1.0f / x;
        sf += q;
        sd += (double)q;
    }
    rf[tid] = sf;
    rd[tid] = sd;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            rf[tid] += rf[tid + o];
            rd[tid] += rd[tid + o];
        }
        __syncthreads();
    }
    if (tid == 0) {
        float e = n > 1 ? rf[0] / (float)(n - 1) : 0.0f;
        if (e < 1.0f / (float)n) {
            e = 0.0f;
            atomicAdd(iso, 1);
        }
        if (eff) eff[u] = e;
        rowsum[u] = rd[0];
    }
}

__global__ __launch_bounds__(256) void fw_lp_sum(const double *__restrict__ rowsum, int n, double *__restrict__ out)
{
    __shared__ double r[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += rowsum[i];
    r[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) r[threadIdx.x] += r[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = r[0];
}

template <int B>
pg_status fw_run(pg_ctx *ctx, float *dm, int np)
{
    hipStream_t st = ctx->stream;
    const int nbk = np / B;
    for (int kb = 0; kb < nbk; kb++) {
        fw_seq_tiles<B><<<1, 256, 0, st>>>(dm, np, kb, 1);
        if (nbk > 1) {
            fw_seq_tiles<B><<<2 * (nbk - 1), 256, 0, st>>>(dm, np, kb, 2);
            dim3 g(nbk - 1, nbk - 1);
            fw_phase3<B><<<g, fw3_threads<B>(), 0, st>>>(dm, np, kb);
        }
    }
    PG_KERNEL_CHECK(ctx);
    return PG_OK;
}

}  // namespace

extern "C" pg_status pg_fw(pg_ctx *ctx, float *d, int32_t n, int32_t block)
{
    if (!ctx) return PG_EINVAL;
    if (!d || n < 0) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    if (block == 0 && n > 0 && pg_autotune_on_first_use(ctx) && pg_tune_get(ctx, "fw_block", n, 0, 0) == 0) {
        pg_tune_result tr;
        PG_TRY(pg_autotune_fw(ctx, d, n, &tr));
    }
    if (block == 0) {
        // a value forced by the autotuner's search, else the tuned one for
        // (arch, n), else 64 (measured faster than 128 at N = 8192)
        block = ctx->force_fw_block;
        if (block <= 0) block = pg_tune_get(ctx, "fw_block", n, 0, 0);
        if (block != 64 && block != 128) block = 64;
    }
    if (block != 64 && block != 128) return pg_fail(ctx, PG_EINVAL, "tile edge %d: use 64 or 128", block);
    if (n == 0) return PG_OK;
    PG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int np = (int)pg_cdiv(n, block) * block;
    const long long nn = (long long)n * n, npp = (long long)np * np;
    const bool dev = pg_is_device_ptr(d);
    float *dm = nullptr;
    float *din = d;
    if (!dev) {
        PG_TRY(pg_bufT(ctx, "fw.in", (size_t)nn, &din));
        PG_HIP(ctx, hipMemcpyAsync(din, d, nn * 4, hipMemcpyHostToDevice, st));
    }
    if (np == n) {
        dm = din;
    } else {
        PG_TRY(pg_bufT(ctx, "fw.pad", (size_t)npp, &dm));
        fw_pad_in<<<(unsigned)pg_cdiv(npp, 256), 256, 0, st>>>(din, n, dm, np);
    }
    if (block == 64) PG_TRY(fw_run<64>(ctx, dm, np));
    else PG_TRY(fw_run<128>(ctx, dm, np));
    if (np != n) fw_pad_out<<<(unsigned)pg_cdiv(nn, 256), 256, 0, st>>>(dm, np, din, n);
    PG_KERNEL_CHECK(ctx);
    if (!dev) PG_HIP(ctx, hipMemcpyAsync(d, din, nn * 4, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    return PG_OK;
}

extern "C" pg_status pg_fw_init(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val,
                                int32_t n, int64_t nnz, int32_t kind, float *d)
{
    if (!ctx) return PG_EINVAL;
    if (!d || kind < 0 || kind > 2 || (kind != 2 && !val)) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_TRY(pg_check_csr_args(ctx, row, col, n, nnz));
    PG_HIP(ctx, hipSetDevice(ctx->device));
    const int *drow, *dcol;
    const float *dval = nullptr;
    PG_TRY(pg_in(ctx, row, (size_t)n + 1, "in.row", &drow));
    PG_TRY(pg_in(ctx, col, (size_t)nnz, "in.col", &dcol));
    if (val) PG_TRY(pg_in(ctx, val, (size_t)nnz, "in.val", &dval));
    PG_TRY(pg_validate_csr(ctx, drow, dcol, n, nnz, false));
    const long long nn = (long long)n * n;
    float *dd;
    PG_TRY(pg_out(ctx, d, (size_t)nn, "out.fw", &dd));
    hipStream_t st = ctx->stream;
    if (nn > 0) {
        fw_fill_inf<<<(unsigned)pg_cdiv(nn, 256), 256, 0, st>>>(dd, nn);
        fw_from_csr<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(drow, dcol, dval, n, kind, dd);
    }
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, d, dd, (size_t)nn));
    PG_HIP(ctx, hipStreamSynchronize(st));
    return PG_OK;
}

extern "C" pg_status pg_fw_efficiency(pg_ctx *ctx, const float *d, int32_t n, float *eff, double *lp,
                                      int32_t *isolated)
{
    if (!ctx) return PG_EINVAL;
    if (!d || n < 1) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const float *dd;
    PG_TRY(pg_in(ctx, d, (size_t)n * n, "fw.in", &dd));
    float *de;
    double *rs, *tot;
    int *iso;
    PG_TRY(pg_out(ctx, eff, (size_t)n, "out.eff", &de));
    PG_TRY(pg_bufT(ctx, "fw.rowsum", (size_t)n, &rs));
    PG_TRY(pg_bufT(ctx, "fw.lpsum", 1, &tot));
    PG_TRY(pg_bufT(ctx, "fw.iso", 1, &iso));
    PG_HIP(ctx, hipMemsetAsync(iso, 0, 4, st));
    fw_eff_rows<<<(unsigned)n, 256, 0, st>>>(dd, n, de, rs, iso);
    fw_lp_sum<<<1, 256, 0, st>>>(rs, n, tot);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, eff, de, (size_t)n));
    double s = 0.0;
    int ni = 0;
    PG_HIP(ctx, hipMemcpyAsync(&s, tot, 8, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipMemcpyAsync(&ni, iso, 4, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (lp) {
        double x = s;
        x /= n;
        x /= (n - 1);
        *lp = 1.0 / x;
    }
    if (isolated) *isolated = ni;
    return PG_OK;
}
