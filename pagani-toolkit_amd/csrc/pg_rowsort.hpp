// pg_rowsort.hpp -- in-register sorts of one CSR row by a 64-lane wave,
// shared by the CSR builders (corr.hip, rewire.hip).  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

namespace pgdev {

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
                    const K y = __shfl_xor(x[r], lj);
                    x[r] = keep_min ? (x[r] < y ? x[r] : y) : (x[r] < y ? y : x[r]);
                }
            }
        }
    }
}

// Sort the d <= 64*P columns (and values) of the row starting at a: ctmp /
// vtmp hold the unsorted entries, col / val receive them in column order.
template <int P>
__device__ __forceinline__ void sort_row_regs(const int *__restrict__ ctmp, const float *__restrict__ vtmp,
                                              long long a, int d, int lane, int *__restrict__ col,
                                              float *__restrict__ val)
{
    if (vtmp) {
        unsigned long long x[P];
#pragma unroll
        for (int r = 0; r < P; r++) {
            const int e = lane * P + r;
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
            const int e = lane * P + r;
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
