// pg_rowsort.hpp -- in-register sorts of one CSR row by a 64-lane wave,
// shared by the CSR builders (corr.hip, rewire.hip).  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

namespace pgdev {

// x from lane (lane ^ lj) of the wave.  Distances inside a 16-lane row are DPP
// moves on the VALU (quad_perm for 1 and 2, row_ror:8 for 8, row_shl:4 /
// row_shr:4 picked by lane bit 2 for 4); 16 and 32 go through ds_bpermute.
__device__ __forceinline__ unsigned xor_lanes32(unsigned x, int lj, int lane)
{
    switch (lj) {
    case 1: return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    case 2: return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    case 8: return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, false);  // row_ror:8
    case 4: {
        const unsigned up = (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x104, 0xF, 0xF, false);  // row_shl:4
        const unsigned dn = (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
        return (lane & 4) ? dn : up;
    }
    default: return (unsigned)__shfl_xor((int)x, lj);
    }
}

template <class K>
__device__ __forceinline__ K xor_lanes(K x, int lj, int lane)
{
    if constexpr (sizeof(K) == 8) {
        const unsigned lo = xor_lanes32((unsigned)x, lj, lane);
        const unsigned hi = xor_lanes32((unsigned)((unsigned long long)x >> 32), lj, lane);
        return (K)(((unsigned long long)hi << 32) | lo);
    } else {
        return (K)xor_lanes32((unsigned)x, lj, lane);
    }
}

// Bitonic sort of 64*P keys held P per lane (lane l holds keys l*P .. l*P+P-1),
// ascending in that order: register compare-exchanges for partner distances
// below P, cross-lane shuffles above.
template <int P, class K>
__device__ __forceinline__ void wave_bitonic(K (&x)[P], int lane)
{
    constexpr int M = 64 * P;
#pragma unroll
    for (int k = 2; k <= M; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < P) {
#pragma unroll
                for (int r = 0; r < P; r++) {
                    if (r & j) continue;
                    const int r2 = r | j;
                    const bool asc = k >= P ? (((lane * P) & k) == 0) : ((r & k) == 0);
                    const K a = x[r], b = x[r2];
                    const K lo = a < b ? a : b, hi = a < b ? b : a;
                    x[r] = asc ? lo : hi;
                    x[r2] = asc ? hi : lo;
                }
            } else {
                const int lj = j / P;
                const bool lower = (lane & lj) == 0;
                const bool asc = k == M ? true : ((lane & (k / P)) == 0);
                const bool keep_min = lower == asc;
#pragma unroll
                for (int r = 0; r < P; r++) {
                    const K y = xor_lanes(x[r], lj, lane);
                    x[r] = keep_min ? (x[r] < y ? x[r] : y) : (x[r] < y ? y : x[r]);
                }
            }
        }
    }
}

// Sort the d <= 64*P columns (and values) of the row starting at a: ctmp /
// vtmp hold the unsorted entries, col / val receive them in column order.
// Keys are loaded lane-interleaved (entry r*64 + lane into register r) and the
// sorted keys written lane-major (lane*P + r), the network's order.
template <int P>
__device__ __forceinline__ void sort_row_regs(const int *__restrict__ ctmp, const float *__restrict__ vtmp,
                                              long long a, int d, int lane, int *__restrict__ col,
                                              float *__restrict__ val)
{
    if (vtmp) {
        unsigned long long x[P];
#pragma unroll
        for (int r = 0; r < P; r++) {
            const int e = r * 64 + lane;  // coalesced: the sort does not care where a key starts
            x[r] = e < d ? ((unsigned long long)(unsigned)ctmp[a + e] << 32) | __float_as_uint(vtmp[a + e]) : ~0ull;
        }
        wave_bitonic<P>(x, lane);
#pragma unroll
        for (int r = 0; r < P; r++) {
            const int e = lane * P + r;
            if (e < d) {
                col[a + e] = (int)(x[r] >> 32);
                val[a + e] = __uint_as_float((unsigned)x[r]);
            }
        }
    } else {
        int x[P];
#pragma unroll
        for (int r = 0; r < P; r++) {
            const int e = r * 64 + lane;
            x[r] = e < d ? ctmp[a + e] : INT_MAX;
        }
        wave_bitonic<P>(x, lane);
#pragma unroll
        for (int r = 0; r < P; r++) {
            const int e = lane * P + r;
            if (e < d) col[a + e] = x[r];
        }
    }
}

}  // namespace pgdev
