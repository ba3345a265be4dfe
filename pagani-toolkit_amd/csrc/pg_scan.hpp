// pg_scan.hpp -- exclusive scan of a short int array by one work-group,
// shared by the CSR builders of the dense average, the rewiring and the
// sliced eigenvector-centrality layout.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

namespace pgdev {

// out[q] = scale * (in[0] + ... + in[q - 1]) for q = 0 .. n (out[n] is the
// scaled total).  One work-group of 1024 threads: contiguous per-thread runs,
// a Hillis-Steele scan of the run totals in LDS, then each run written out.
__device__ __forceinline__ void block_scan_1024(const int *__restrict__ in, long long n, long long *__restrict__ out,
                                                long long scale)
{
    __shared__ long long part[1024];
    const int tid = threadIdx.x;
    const long long per = (n + 1023) / 1024;
    const long long a = std::min<long long>(n, tid * per), b = std::min<long long>(n, a + per);
    long long s = 0;
    for (long long q = a; q < b; q++) s += in[q];
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long y = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += y;
        __syncthreads();
    }
    long long run = part[tid] - s;
    for (long long q = a; q < b; q++) {
        out[q] = scale * run;
        run += in[q];
    }
    if (tid == 1023) out[n] = scale * part[1023];
}

}  // namespace pgdev
