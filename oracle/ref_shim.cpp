// ref_shim.cpp -- extern "C" entry points over the reference's own CPU sources,
// compiled from /root/reference by oracle/Makefile into oracle/_ref/libref.so.
// TEST INFRASTRUCTURE: used only by tests/ (to pin the oracle) and by the
// cpu_baseline leg of bench.py.  No reference source is copied here; this file
// only declares the reference functions and calls them.
#include <cstdint>
#include <iostream>
#include <limits>
#include <vector>

// U/BC/BC_CPU/BC_CPU/BC_CPU.cpp:6 (float sigma/delta Brandes, all sources)
void Betweenness_CPU(int* row, int* col, int numVertices, int numEdges, float* BC);
// W/Lp/CUBFW_Lp/CUBFW_Lp/BFW.cpp:40 (blocked Floyd-Warshall on a tile-major matrix)
void BFW_C(float* costmat, long long numVertices, long long block_size);
// W/Lp/CUBFW_Lp/CUBFW_Lp/FormBlock.cpp:5,24,45 (layout helpers, distance = 1/w)
bool FormBlock(float* B, float* A, int N, int Block_size);
bool DeFormBlock(float* A, float* B, int N, int Block_size);
void init_block(float* costmat, const int* row, const int* col, const float* power,
                const int numVertices, const int block_size);

namespace {
// The reference functions print progress on std::cout; silence them.
struct Mute {
    std::streambuf* old;
    Mute() : old(std::cout.rdbuf(nullptr)) {}
    ~Mute() { std::cout.rdbuf(old); std::cout.clear(); }
};
}  // namespace

extern "C" {

void ref_bc_all(const int32_t* row, const int32_t* col, int32_t n, float* bc)
{
    std::vector<int> r(row, row + n + 1), c(col, col + row[n]);
    Mute m;
    Betweenness_CPU(r.data(), c.data(), n, row[n], bc);
}

// d: row-major np x np, np a multiple of b; closed in place by BFW_C.
void ref_fw_dense(float* d, int32_t np, int32_t b)
{
    std::vector<float> blk((size_t)np * np);
    FormBlock(blk.data(), d, np, b);
    {
        Mute m;
        BFW_C(blk.data(), np, b);
    }
    DeFormBlock(d, blk.data(), np, b);
}

// The reference's own initialisation (d = 1/w, inf, 0 on the diagonal blocks'
// diagonals) from a CSR, returned row-major np x np.
void ref_fw_init(const int32_t* row, const int32_t* col, const float* w, int32_t np,
                 int32_t b, float* d)
{
    std::vector<float> blk((size_t)np * np);
    init_block(blk.data(), row, col, w, np, b);
    DeFormBlock(d, blk.data(), np, b);
}

}  // extern "C"
