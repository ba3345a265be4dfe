// rewire.hip -- degree-preserving random rewiring (Maslov-Sneppen double edge
// swaps) of an undirected CSR network on MI355X (gfx950).
//
// Replaces Maslov / Rewire (U BFS_MulCPU/Maslov.cpp:8-110) and
// Maslov_weighted_1 (W Maslov/Maslov/Maslov.cpp:50-196), the random reference
// networks behind Cp's gamma (U Cp_main.cpp:150-195), the Lp tools' lambda (W
// CUBFW_Lp_main.cpp:210-249, U MUBFS_LP_main.cpp:129-150) and the Maslov tool
// (W Maslov_main.cpp:11-81).  The reference algorithm: the M = nnz/2 edges go
// into an edge list E in row-major order of the upper triangle; 2M attempts
// each draw two edges i = (V1, V2), j = (V3, V4), give up when they share a
// vertex, else a fair coin picks 1-2,3-4 -> 1-3,2-4 (E[i].C = V3, E[j].R = V2)
// or -> 1-4,3-2 (E[i].C = V4, E[j].C = V2); the attempt is given up when either
// new edge already exists.  A weight stays with its edge slot, so each edge
// keeps its weight in both directions of the output CSR; rows keep their
// lengths, so the row offsets are the input's.
//
// MI355X form.  Attempts run in rounds of up to M/2 at once.  Round r pairs the
// edges through a keyed bijection of [0, M) (a 4-round Feistel network on the
// next even power of two, cycle-walked into range): attempt t takes edges
// perm(2t) and perm(2t+1), so no edge is in two attempts of one round.  The
// round-start edges sit in an open-addressing hash (linear probing, 64-bit
// keys) with owner 0.  Kernel rw_propose finds or inserts each attempt's two
// new-edge keys and, when neither is an existing edge, takes both with an
// atomicMin of its attempt id; kernel rw_commit applies the swaps whose
// attempt owns both keys.  The winners are a function of the seed alone (no
// dependence on thread timing), every applied swap is valid for the graph at
// the end of the round, and degrees, symmetry and the absence of self-loops
// and multi-edges hold after every round.  Deviations from the sequential
// order, rare at brain-network densities: an edge removed in a round cannot be
// re-created in the same round, and an attempt that loses one of its keys
// blocks the other for the rest of the round.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "pg_internal.hpp"
#include "pg_rowsort.hpp"
#include "pg_scan.hpp"

namespace {

constexpr unsigned long long EMPTY = ~0ull;
constexpr unsigned NO_OWNER = 0xffffffffu;
constexpr int SHORT_ROW = 64 * 16;  // rows up to this length are sorted in registers

// splitmix64 output function (Steele, Lea, Flood 2014)
__host__ __device__ __forceinline__ unsigned long long mix64(unsigned long long x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

__device__ __forceinline__ unsigned long long edge_key(int a, int b)
{
    const unsigned lo = (unsigned)min(a, b), hi = (unsigned)max(a, b);
    return ((unsigned long long)lo << 32) | hi;
}

struct RwRound {
    unsigned long long pkey;  // permutation key
    unsigned long long ckey;  // coin key
    unsigned half_mask;       // Feistel half: h bits
    int half_bits;
    int m;           // edges
    long long p;     // attempts this round
    unsigned long long mask;  // hash capacity - 1
};

// Keyed bijection of [0, m): 4 Feistel rounds on 2h bits, cycle-walked.
__device__ __forceinline__ unsigned rw_perm(unsigned x, const RwRound &r)
{
    do {
        unsigned L = x >> r.half_bits, R = x & r.half_mask;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const unsigned f = (unsigned)mix64(r.pkey ^ ((unsigned long long)q << 56) ^ R) & r.half_mask;
            const unsigned t = L ^ f;
            L = R;
            R = t;
        }
        x = (L << r.half_bits) | R;
    } while (x >= (unsigned)r.m);
    return x;
}

// Slot holding key k, inserting it into the first empty slot of its probe run
// when absent.  The table is never more than half full.
__device__ __forceinline__ unsigned long long hs_find_or_insert(unsigned long long *tab, unsigned long long mask,
                                                                unsigned long long k)
{
    unsigned long long s = mix64(k) & mask;
    for (;;) {
        const unsigned long long cur = __hip_atomic_load(&tab[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == k) return s;
        if (cur == EMPTY) {
            const unsigned long long prev = atomicCAS(&tab[s], EMPTY, k);
            if (prev == EMPTY || prev == k) return s;
        }
        s = (s + 1) & mask;
    }
}

// Per row: entries above the diagonal (columns sorted ascending), a self-loop
// flag (bit 0) and a long-row flag (bit 1).
__global__ void rw_upper_count(const int *__restrict__ row, const int *__restrict__ col, int n,
                               int *__restrict__ cnt, int *__restrict__ flags)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const int a = row[u], b = row[u + 1];
    int lo = a, hi = b;  // first column > u
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    cnt[u] = b - lo;
    int f = 0;
    if (lo > a && col[lo - 1] == u) f |= 1;
    if (b - a > SHORT_ROW) f |= 2;
    if (f) atomicOr(flags, f);
}

// Exclusive scan of n ints into out[0..n], by one work-group.
__global__ __launch_bounds__(1024) void rw_scan(const int *__restrict__ in, int n, long long *__restrict__ out)
{
    pgdev::block_scan_1024(in, n, out, 1);
}

// One wave per row: the row's upper entries into E at base[u], in column order.
__global__ __launch_bounds__(256) void rw_extract(const int *__restrict__ row, const int *__restrict__ col,
                                                  const float *__restrict__ val, const int *__restrict__ cnt,
                                                  const long long *__restrict__ base, int n, int *__restrict__ eu,
                                                  int *__restrict__ ev, float *__restrict__ ew)
{
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (u >= n) return;
    const int b = row[u + 1], c = cnt[u], a = b - c;
    const long long o = base[u];
    for (int k = lane; k < c; k += 64) {
        eu[o + k] = u;
        ev[o + k] = col[a + k];
        if (ew) ew[o + k] = val[a + k];
    }
}

// Every upper entry (u, v) must have its mirror (v, u), with the same weight.
__global__ void rw_check_mirror(const int *__restrict__ row, const int *__restrict__ col,
                                const float *__restrict__ val, const int *__restrict__ eu,
                                const int *__restrict__ ev, const float *__restrict__ ew, int m,
                                int *__restrict__ flags)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int u = eu[e], v = ev[e];
    int lo = row[v], hi = row[v + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    bool ok = lo < row[v + 1] && col[lo] == u;
    if (ok && ew) ok = __float_as_uint(val[lo]) == __float_as_uint(ew[e]);
    if (!ok) atomicOr(flags, 4);
}

__global__ void rw_clear(unsigned long long *__restrict__ tab, unsigned *__restrict__ own, unsigned long long cap)
{
    const unsigned long long s = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= cap) return;
    tab[s] = EMPTY;
    own[s] = NO_OWNER;
}

// The current edges, owner 0.
__global__ void rw_insert_edges(const int *__restrict__ eu, const int *__restrict__ ev, int m,
                                unsigned long long *__restrict__ tab, unsigned *__restrict__ own,
                                unsigned long long mask)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const unsigned long long s = hs_find_or_insert(tab, mask, edge_key(eu[e], ev[e]));
    own[s] = 0u;
}

struct Attempt {
    int i, j, v1, v2, v3, v4;
    bool form13;  // 1-3,2-4 (else 1-4,3-2)
    bool disjoint;
};

__device__ __forceinline__ Attempt rw_attempt(long long t, const RwRound &r, const int *__restrict__ eu,
                                              const int *__restrict__ ev)
{
    Attempt a;
    a.i = (int)rw_perm((unsigned)(2 * t), r);
    a.j = (int)rw_perm((unsigned)(2 * t + 1), r);
    a.v1 = eu[a.i];
    a.v2 = ev[a.i];
    a.v3 = eu[a.j];
    a.v4 = ev[a.j];
    a.disjoint = !(a.v1 == a.v3 || a.v1 == a.v4 || a.v2 == a.v3 || a.v2 == a.v4);
    a.form13 = (mix64(r.ckey ^ (unsigned long long)t) & 1ull) != 0;
    return a;
}

// Attempt t (id t + 1) proposes its two new edges: both slots are found or
// inserted; unless one is an existing edge, both are claimed by atomicMin.
__global__ void rw_propose(RwRound r, const int *__restrict__ eu, const int *__restrict__ ev,
                           unsigned long long *__restrict__ tab, unsigned *__restrict__ own,
                           uint2 *__restrict__ slots)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= r.p) return;
    const Attempt a = rw_attempt(t, r, eu, ev);
    uint2 sl = make_uint2(NO_OWNER, NO_OWNER);
    if (a.disjoint) {
        const int x = a.form13 ? a.v3 : a.v4, y = a.form13 ? a.v4 : a.v3;
        const unsigned long long sa = hs_find_or_insert(tab, r.mask, edge_key(a.v1, x));
        const unsigned long long sb = hs_find_or_insert(tab, r.mask, edge_key(a.v2, y));
        // owner 0 marks a round-start edge; it was written by an earlier launch
        if (own[sa] != 0u && own[sb] != 0u) {
            const unsigned id = (unsigned)(t + 1);
            atomicMin(&own[sa], id);
            atomicMin(&own[sb], id);
            sl = make_uint2((unsigned)sa, (unsigned)sb);
        }
    }
    slots[t] = sl;
}

// Apply the swaps whose attempt owns both of its keys; count them.
__global__ __launch_bounds__(256) void rw_commit(RwRound r, int *__restrict__ eu, int *__restrict__ ev,
                                                 const unsigned *__restrict__ own, const uint2 *__restrict__ slots,
                                                 unsigned long long *__restrict__ swaps)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool done = false;
    if (t < r.p) {
        const uint2 sl = slots[t];
        const unsigned id = (unsigned)(t + 1);
        if (sl.x != NO_OWNER && own[sl.x] == id && own[sl.y] == id) {
            const Attempt a = rw_attempt(t, r, eu, ev);
            if (a.form13) {
                ev[a.i] = a.v3;
                eu[a.j] = a.v2;
            } else {
                ev[a.i] = a.v4;
                ev[a.j] = a.v2;
            }
            done = true;
        }
    }
    const unsigned long long b = __ballot(done);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(swaps, (unsigned long long)__popcll(b));
}

__global__ void rw_cursor_init(const int *__restrict__ row, int n, int *__restrict__ cur)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) cur[u] = row[u];
}

// Both directions of every edge into their rows (unordered inside a row).
// Swaps preserve degrees, so each row receives exactly its length; a write
// that would leave its row is dropped and flagged (bit 3), and the call fails.
__global__ void rw_fill(const int *__restrict__ row, const int *__restrict__ eu, const int *__restrict__ ev,
                        const float *__restrict__ ew, int m, int *__restrict__ cur, int *__restrict__ ctmp,
                        float *__restrict__ vtmp, int *__restrict__ flags)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int u = eu[e], v = ev[e];
    const int pu = atomicAdd(&cur[u], 1), pv = atomicAdd(&cur[v], 1);
    const float w = ew ? ew[e] : 0.0f;
    if (pu < row[u + 1]) {
        ctmp[pu] = v;
        if (ew) vtmp[pu] = w;
    }
    if (pv < row[v + 1]) {
        ctmp[pv] = u;
        if (ew) vtmp[pv] = w;
    }
    if (pu >= row[u + 1] || pv >= row[v + 1]) atomicOr(flags, 8);
}

// One wave per row: rows up to SHORT_ROW entries sorted by column in registers.
__global__ __launch_bounds__(256) void rw_sort_short(const int *__restrict__ row, int n, const int *__restrict__ ctmp,
                                                     const float *__restrict__ vtmp, int *__restrict__ col,
                                                     float *__restrict__ val)
{
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (u >= n) return;
    const long long a = row[u];
    const int d = row[u + 1] - row[u];
    if (d == 0 || d > SHORT_ROW) return;
    if (d <= 64) pgdev::sort_row_regs<1>(ctmp, vtmp, a, d, lane, col, val);
    else if (d <= 128) pgdev::sort_row_regs<2>(ctmp, vtmp, a, d, lane, col, val);
    else if (d <= 256) pgdev::sort_row_regs<4>(ctmp, vtmp, a, d, lane, col, val);
    else if (d <= 512) pgdev::sort_row_regs<8>(ctmp, vtmp, a, d, lane, col, val);
    else pgdev::sort_row_regs<16>(ctmp, vtmp, a, d, lane, col, val);
}

// One work-group per row longer than SHORT_ROW: rank by an LDS bitmap over
// [0, n) and prefix popcounts (the columns of a row are distinct).
__global__ __launch_bounds__(256) void rw_sort_long(const int *__restrict__ row, int n, const int *__restrict__ ctmp,
                                                    const float *__restrict__ vtmp, int *__restrict__ col,
                                                    float *__restrict__ val)
{
    extern __shared__ uint32_t sm[];
    __shared__ int part[256];
    const int u = blockIdx.x, tid = threadIdx.x;
    const int a = row[u], b = row[u + 1];
    if (b - a <= SHORT_ROW) return;
    const int nw = (n + 31) >> 5;
    uint32_t *bits = sm, *pre = sm + nw;
    for (int q = tid; q < nw; q += 256) bits[q] = 0u;
    __syncthreads();
    for (int e = a + tid; e < b; e += 256) {
        const int c = ctmp[e];
        atomicOr(&bits[c >> 5], 1u << (c & 31));
    }
    __syncthreads();
    const int per = (nw + 255) / 256;
    const int w0 = min(nw, tid * per), w1 = min(nw, w0 + per);
    int s = 0;
    for (int q = w0; q < w1; q++) s += __popc(bits[q]);
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int y = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += y;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (int q = w0; q < w1; q++) {
        pre[q] = run;
        run += __popc(bits[q]);
    }
    __syncthreads();
    for (int e = a + tid; e < b; e += 256) {
        const int c = ctmp[e];
        const int pos = pre[c >> 5] + __popc(bits[c >> 5] & ((1u << (c & 31)) - 1u));
        col[a + pos] = c;
        if (val) val[a + pos] = vtmp[e];
    }
}

}  // namespace

extern "C" pg_status pg_rewire(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val, int32_t n,
                               int64_t nnz, uint64_t seed, int64_t attempts, int32_t *out_col, float *out_val,
                               int64_t *swaps)
{
    if (!ctx) return PG_EINVAL;
    if (!out_col || attempts < 0) return pg_fail(ctx, PG_EINVAL, "bad arguments");
    PG_TRY(pg_check_csr_args(ctx, row, col, n, nnz));
    PG_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (swaps) *swaps = 0;
    if (n == 0 || nnz == 0) return PG_OK;
    const int *drow, *dcol;
    const float *dval = nullptr;
    PG_TRY(pg_in(ctx, row, (size_t)n + 1, "in.row", &drow));
    PG_TRY(pg_in(ctx, col, (size_t)nnz, "in.col", &dcol));
    if (val) PG_TRY(pg_in(ctx, val, (size_t)nnz, "in.val", &dval));
    PG_TRY(pg_validate_csr(ctx, drow, dcol, n, nnz, true));
    const bool weighted = val && out_val;
    if (!weighted) dval = nullptr;

    // the edge list
    int *cnt, *flags;
    long long *base;
    PG_TRY(pg_bufT(ctx, "rw.cnt", (size_t)n, &cnt));
    PG_TRY(pg_bufT(ctx, "rw.flags", 1, &flags));
    PG_TRY(pg_bufT(ctx, "rw.base", (size_t)n + 1, &base));
    PG_HIP(ctx, hipMemsetAsync(flags, 0, 4, st));
    rw_upper_count<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(drow, dcol, n, cnt, flags);
    rw_scan<<<1, 1024, 0, st>>>(cnt, n, base);
    PG_KERNEL_CHECK(ctx);
    long long m64 = 0;
    int fl = 0;
    PG_HIP(ctx, hipMemcpyAsync(&m64, base + n, 8, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipMemcpyAsync(&fl, flags, 4, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (fl & 1) return pg_fail(ctx, PG_EFORMAT, "self-loop in the network");
    if (2 * m64 != nnz) return pg_fail(ctx, PG_EFORMAT, "network is not symmetric (%lld entries above the diagonal of %lld)",
                                       m64, (long long)nnz);
    // hash slots are 32-bit with 0xffffffff reserved: capacity 4M <= 2^31
    if (m64 > (1ll << 29)) return pg_fail(ctx, PG_ENOTSUP, "%lld edges: at most 2^29 are supported", m64);
    const int m = (int)m64;
    const bool long_rows = (fl & 2) != 0;
    const size_t long_lds = 2 * (size_t)((n + 31) / 32) * 4;  // bitmap + prefix words
    constexpr size_t LDS_MAX = 159 * 1024;                     // 160 KB less the static scan array
    if (long_rows && long_lds > LDS_MAX)
        return pg_fail(ctx, PG_ENOTSUP, "rows above %d entries need n <= %d", SHORT_ROW, (int)(LDS_MAX * 4));
    if (long_rows)
        PG_HIP(ctx, hipFuncSetAttribute((const void *)rw_sort_long, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)long_lds));
    int *eu, *ev;
    float *ew = nullptr;
    PG_TRY(pg_bufT(ctx, "rw.eu", (size_t)m, &eu));
    PG_TRY(pg_bufT(ctx, "rw.ev", (size_t)m, &ev));
    if (weighted) PG_TRY(pg_bufT(ctx, "rw.ew", (size_t)m, &ew));
    rw_extract<<<(unsigned)pg_cdiv(n, 4), 256, 0, st>>>(drow, dcol, dval, cnt, base, n, eu, ev, ew);
    rw_check_mirror<<<(unsigned)pg_cdiv(m, 256), 256, 0, st>>>(drow, dcol, dval, eu, ev, ew, m, flags);
    PG_KERNEL_CHECK(ctx);
    PG_HIP(ctx, hipMemcpyAsync(&fl, flags, 4, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (fl & 4) return pg_fail(ctx, PG_EFORMAT, "network is not symmetric (an entry or its weight has no mirror)");

    // equal rounds of at most m/2 disjoint attempts (the reference makes 2M
    // attempts); each round costs a hash rebuild, so no short tail round
    const long long total = attempts == 0 ? 2LL * m : attempts;
    const long long max_round = m / 2;
    const long long rounds = max_round > 0 ? (total + max_round - 1) / max_round : 0;
    const long long per_round = rounds > 0 ? (total + rounds - 1) / rounds : 0;
    unsigned long long *dswaps;
    PG_TRY(pg_bufT(ctx, "rw.swaps", 1, &dswaps));
    PG_HIP(ctx, hipMemsetAsync(dswaps, 0, 8, st));
    if (per_round > 0 && total > 0) {
        unsigned long long cap = 1024;
        while (cap < 4ull * (unsigned long long)m) cap <<= 1;
        unsigned long long *tab;
        unsigned *own;
        uint2 *slots;
        PG_TRY(pg_bufT(ctx, "rw.tab", (size_t)cap, &tab));
        PG_TRY(pg_bufT(ctx, "rw.own", (size_t)cap, &own));
        PG_TRY(pg_bufT(ctx, "rw.slots", (size_t)per_round, &slots));
        int bits = 0;
        while ((1ll << bits) < m) bits++;
        bits += bits & 1;
        RwRound r;
        r.half_bits = bits / 2;
        r.half_mask = (1u << r.half_bits) - 1u;
        r.m = m;
        r.mask = cap - 1;
        unsigned long long key = mix64(seed ^ 0x5851f42d4c957f2dull);
        for (long long done = 0; done < total; done += per_round) {
            r.p = std::min(per_round, total - done);
            key = mix64(key + 0x9e3779b97f4a7c15ull);
            r.pkey = key;
            r.ckey = mix64(key ^ 0xd1b54a32d192ed03ull);
            rw_clear<<<(unsigned)pg_cdiv((long long)cap, 256), 256, 0, st>>>(tab, own, cap);
            rw_insert_edges<<<(unsigned)pg_cdiv(m, 256), 256, 0, st>>>(eu, ev, m, tab, own, r.mask);
            rw_propose<<<(unsigned)pg_cdiv(r.p, 256), 256, 0, st>>>(r, eu, ev, tab, own, slots);
            rw_commit<<<(unsigned)pg_cdiv(r.p, 256), 256, 0, st>>>(r, eu, ev, own, slots, dswaps);
        }
        PG_KERNEL_CHECK(ctx);
    }

    // the rewired CSR: same row offsets, columns (and weights) sorted per row
    int *dout, *cur, *ctmp;
    float *dvout = nullptr, *vtmp = nullptr;
    PG_TRY(pg_out(ctx, out_col, (size_t)nnz, "out.rwcol", &dout));
    if (weighted) PG_TRY(pg_out(ctx, out_val, (size_t)nnz, "out.rwval", &dvout));
    PG_TRY(pg_bufT(ctx, "rw.cur", (size_t)n, &cur));
    PG_TRY(pg_bufT(ctx, "rw.ctmp", (size_t)nnz, &ctmp));
    if (weighted) PG_TRY(pg_bufT(ctx, "rw.vtmp", (size_t)nnz, &vtmp));
    rw_cursor_init<<<(unsigned)pg_cdiv(n, 256), 256, 0, st>>>(drow, n, cur);
    rw_fill<<<(unsigned)pg_cdiv(m, 256), 256, 0, st>>>(drow, eu, ev, ew, m, cur, ctmp, vtmp, flags);
    rw_sort_short<<<(unsigned)pg_cdiv(n, 4), 256, 0, st>>>(drow, n, ctmp, vtmp, dout, dvout);
    if (long_rows) rw_sort_long<<<(unsigned)n, 256, long_lds, st>>>(drow, n, ctmp, vtmp, dout, dvout);
    PG_KERNEL_CHECK(ctx);
    PG_TRY(pg_out_finish(ctx, out_col, dout, (size_t)nnz));
    if (weighted) PG_TRY(pg_out_finish(ctx, out_val, dvout, (size_t)nnz));
    unsigned long long ns = 0;
    PG_HIP(ctx, hipMemcpyAsync(&ns, dswaps, 8, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipMemcpyAsync(&fl, flags, 4, hipMemcpyDeviceToHost, st));
    PG_HIP(ctx, hipStreamSynchronize(st));
    if (fl & 8) return pg_fail(ctx, PG_EDEVICE, "internal error: the rewiring changed a vertex degree");
    if (swaps) *swaps = (int64_t)ns;
    return PG_OK;
}
