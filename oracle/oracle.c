/*
 * oracle.c -- FP64 CPU oracle for the PAGANI construction-and-metrics path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so, and only as the checker.  Nothing in
 * the product links it.
 *
 * Reference path prefixes:
 *   U/ = src/BNAPlatform-win64-cuda7.0-20151118/src/
 *   W/ = src/BNAPlatform-weighted-network-win64-cuda7.0-20151118/src/
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_zscore          closed forms (mean 0, unit L2 norm, z of an affine
 *                       series), numpy FP64 of the same expression
 *   orc_r / corr_csr_*  numpy corrcoef FP64 (1e-12), r(x, a x + b) = sign(a),
 *                       brute-force all-pairs threshold with injected exact
 *                       ties (duplicate voxels) for the sparsity tie rule
 *   orc_sparsity_k      the reference's float expression evaluated with numpy
 *                       float32 scalars
 *   orc_degree          R differences; the float expression with numpy float32
 *   orc_strength        numpy float32 sequential sum
 *   orc_triangles/cp    networkx 3.4.2 triangles/clustering fixtures, complete
 *                       graph (Cp = 1), star (Cp = 0), trace(A^3)/6
 *   orc_cp_weighted     networkx clustering(weight=) fixtures (Onnela), brute
 *                       force dense Python loops (Barrat)
 *   orc_bc              networkx betweenness x2 fixtures, closed forms on paths
 *                       and stars, the compiled reference Betweenness_CPU
 *   orc_bfs             networkx shortest_path_length fixtures
 *   orc_bfs_efficiency  networkx global_efficiency (= 1/Lp) and FP64 nodal sums
 *                       from networkx hop counts; the FP32 nodal sums against a
 *                       numpy float32 sequential sum; closed forms on complete
 *                       graphs, stars and paths
 *   orc_fw_blocked      the compiled reference BFW_C (bit-exact), networkx
 *                       floyd_warshall_numpy, unit weights == BFS hops
 *   orc_modularity      disjoint cliques split exactly, Q against networkx
 *                       modularity, leading vector against numpy eigh
 *   orc_ec              numpy eigh leading pair and networkx
 *                       eigenvector_centrality_numpy fixtures, uniform vector on
 *                       regular graphs, 10000 iterations on bipartite graphs
 *   orc_pc              FP64 brute force per vertex, closed forms (one module:
 *                       0; equal weight over K modules: 1 - 1/K, normalised 1)
 *   orc_cormat_*        numpy corrcoef cast to float32 and summed in float32
 *                       (1 ulp: corrcoef rounds differently), numpy arctanh /
 *                       tanh for the Fisher pair; averaging one subject is the
 *                       identity on (float)r
 *   orc_dense_csr_*     brute force with numpy (mask a > thr; argsort for the
 *                       k-th largest and the row-major tie rule)
 *   orc_rewire          invariants (degrees, symmetry, no self-loops or
 *                       multi-edges, each edge's weight in both directions,
 *                       weight multiset); closed cases: graphs where no swap
 *                       is possible (stars, complete graphs) stay unchanged,
 *                       two disjoint edges reach all three perfect matchings;
 *                       clustering of a rewired ring lattice against networkx
 *                       double_edge_swap with the same number of swaps.  The
 *                       reference Maslov.cpp is not built: its index
 *                       expression assumes RAND_MAX = 0x7fff and overflows int
 *                       with glibc's rand().
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* Threads of the OpenMP loops below (timing the CPU baseline at 1 thread). */
void orc_set_threads(int nthreads) { omp_set_num_threads(nthreads > 0 ? nthreads : omp_get_num_procs()); }
int orc_get_threads(void) { return omp_get_max_threads(); }

/* ---------------------------------------------------------------- z-score */

/* U CorMat/CUCorMat/CUCORMAT/CorMat.cu:34-53: per-voxel mean and
 * sqrt(sum (x-mu)^2) in FP64, z = (x - mu) / norm.  A zero-variance voxel gives
 * 0/0 there, which post_block (U CorMat.cpp:834) turns into r = 0; z = 0 here
 * yields the same r = 0 directly. */
void orc_zscore(const float *bold, int64_t n, int t, double *z64, float *z32)
{
    for (int64_t i = 0; i < n; i++) {
        double s1 = 0.0, s2 = 0.0;
        for (int l = 0; l < t; l++) s1 += (double)bold[(int64_t)l * n + i];
        s1 /= t;
        for (int l = 0; l < t; l++) {
            double d = (double)bold[(int64_t)l * n + i] - s1;
            double p = d * d;
            s2 += p;
        }
        s2 = sqrt(s2);
        for (int l = 0; l < t; l++) {
            double z = (s2 > 0.0) ? ((double)bold[(int64_t)l * n + i] - s1) / s2 : 0.0;
            if (z64) z64[i * t + l] = z;
            if (z32) z32[i * t + l] = (float)z;
        }
    }
}

/* r = sum_t z_i[t] z_j[t], sequential, each product rounded before the add
 * (the file is built with -ffp-contract=off).  Clamped to [-1, 1]: the
 * reference drops |r| > 1 from rounding (U CorMat.cpp:834). */
double orc_r(const double *z64, int t, int64_t i, int64_t j)
{
    const double *a = z64 + i * t, *b = z64 + j * t;
    double s = 0.0;
    for (int l = 0; l < t; l++) {
        double p = a[l] * b[l];
        s += p;
    }
    if (s > 1.0) s = 1.0;
    if (s < -1.0) s = -1.0;
    return s;
}

/* U CorMat.cpp:646 tests r > thr; W CorMat.cpp:310 passes thr - 1e-6 (double)
 * into a float parameter, so the weighted threshold is rounded to float. */
float orc_r_threshold_eff(float thr, int weighted)
{
    if (!weighted) return thr;
    return (float)((double)thr - 1e-6);
}

/* U CorMat.cpp:311-313:  Clength = (long long) 2*M1*s/100.0; Clength += Clength%2.
 * 2*M1 (long long) times a float is a float product; /100.0 is double. */
int64_t orc_sparsity_k(int64_t n, float s_percent)
{
    int64_t m1 = (n - 1) * n / 2;
    float prod = (float)(2 * m1) * s_percent;
    int64_t clen = (int64_t)((double)prod / 100.0);
    clen += clen % 2;
    return clen / 2;
}

int64_t orc_corr_count_r(const double *z64, int64_t n, int t, float thr_eff)
{
    int64_t cnt = 0;
    const double th = (double)thr_eff;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : cnt)
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++)
            if (orc_r(z64, t, i, j) > th) cnt++;
    return 2 * cnt;
}

int64_t orc_corr_count_rows(const double *z64, int64_t n, int t, float thr_eff, int64_t i0, int64_t i1)
{
    int64_t cnt = 0;
    const double th = (double)thr_eff;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cnt)
    for (int64_t i = i0; i < i1; i++)
        for (int64_t j = i + 1; j < n; j++)
            if (orc_r(z64, t, i, j) > th) cnt++;
    return cnt;
}

/* FormFullGraph + FormCSRGraph (U CorMat.cpp:633-653, 839-857; weighted V from
 * W CorMat.cpp:853-885): columns ascending, both halves, no self loops. */
void orc_corr_csr_r(const double *z64, int64_t n, int t, float thr_eff,
                    int32_t *row, int32_t *col, float *val)
{
    const double th = (double)thr_eff;
    int64_t pos = 0;
    row[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            double r = (i < j) ? orc_r(z64, t, i, j) : orc_r(z64, t, j, i);
            if (r > th) {
                col[pos] = (int32_t)j;
                if (val) val[pos] = (float)r;
                pos++;
            }
        }
        row[i + 1] = (int32_t)pos;
    }
}

/* Rows rows[0..nrows) of the r-mode CSR above, each formed alone (for
 * networks too large for the whole matrix): deg[q] entries of row rows[q] go
 * to col/val[q * n ...], columns ascending. */
void orc_corr_rows_r(const double *z64, int64_t n, int t, float thr_eff, const int32_t *rows, int64_t nrows,
                     int32_t *deg, int32_t *col, float *val)
{
    const double th = (double)thr_eff;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < nrows; q++) {
        const int64_t i = rows[q];
        int64_t pos = 0;
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            double r = (i < j) ? orc_r(z64, t, i, j) : orc_r(z64, t, j, i);
            if (r > th) {
                col[q * n + pos] = (int32_t)j;
                if (val) val[q * n + pos] = (float)r;
                pos++;
            }
        }
        deg[q] = (int32_t)pos;
    }
}

/* ------------------------------------------------------- sparsity select */

typedef struct { double r; int64_t idx; } rpair;

static void swapd(double *a, int64_t i, int64_t j) { double x = a[i]; a[i] = a[j]; a[j] = x; }

/* k-th largest (1-based) of a[0..m) by quickselect on a scratch copy. */
static double kth_largest(double *a, int64_t m, int64_t k)
{
    int64_t lo = 0, hi = m - 1, want = k - 1;
    uint64_t st = 0x9E3779B97F4A7C15ull;
    while (lo < hi) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        int64_t p = lo + (int64_t)(st % (uint64_t)(hi - lo + 1));
        double pv = a[p];
        /* three-way partition, descending: [> pv][== pv][< pv] */
        int64_t lt = lo, i = lo, gt = hi;
        while (i <= gt) {
            if (a[i] > pv) swapd(a, lt++, i++);
            else if (a[i] < pv) swapd(a, i, gt--);
            else i++;
        }
        if (want < lt) hi = lt - 1;
        else if (want > gt) lo = gt + 1;
        else return pv;
    }
    return a[lo];
}

/* FormFullGraph_s (U CorMat.cpp:693-773): t* = k-th largest r; edges are all
 * r > t*, then ties r == t* in row-major (i<j) order until k edges (the
 * reference pushes ties on a stack and pops the newest when over count). */
double orc_corr_csr_s(const double *z64, int64_t n, int t, int64_t k,
                      int32_t *row, int32_t *col, float *val)
{
    int64_t m1 = (n - 1) * n / 2;
    double *all = (double *)malloc(sizeof(double) * (size_t)(m1 > 0 ? m1 : 1));
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) {
        int64_t base = i * n - i * (i + 1) / 2; /* packed index of (i, i+1) */
        for (int64_t j = i + 1; j < n; j++) all[base + (j - i - 1)] = orc_r(z64, t, i, j);
    }
    double tstar = 1.0;
    unsigned char *edge = NULL;
    if (k > 0 && k <= m1) {
        double *cp = (double *)malloc(sizeof(double) * (size_t)m1);
        memcpy(cp, all, sizeof(double) * (size_t)m1);
        tstar = kth_largest(cp, m1, k);
        free(cp);
        int64_t greater = 0;
        for (int64_t q = 0; q < m1; q++) greater += all[q] > tstar;
        int64_t ties_left = k - greater;
        edge = (unsigned char *)calloc((size_t)m1, 1);
        for (int64_t q = 0; q < m1; q++) {
            if (all[q] > tstar) edge[q] = 1;
            else if (all[q] == tstar && ties_left > 0) { edge[q] = 1; ties_left--; }
        }
    }
    int64_t pos = 0;
    row[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = 0; j < n; j++) {
            if (j == i || !edge) continue;
            int64_t a = i < j ? i : j, b = i < j ? j : i;
            int64_t q = a * n - a * (a + 1) / 2 + (b - a - 1);
            if (edge[q]) {
                col[pos] = (int32_t)j;
                if (val) val[pos] = (float)all[q];
                pos++;
            }
        }
        row[i + 1] = (int32_t)pos;
    }
    free(edge);
    free(all);
    return tstar;
}

/* ----------------------------------------------------------------- degree */

/* U Degree/Degree/Degree.cpp:87: Degree[k] = (float) R[k+1] - R[k], a float
 * subtraction of two rounded row offsets (inexact once offsets pass 2^24). */
void orc_degree(const int32_t *row, int64_t n, int32_t *deg, float *deg_ref)
{
    for (int64_t i = 0; i < n; i++) {
        if (deg) deg[i] = row[i + 1] - row[i];
        if (deg_ref) {
            volatile float a = (float)row[i + 1];
            volatile float b = (float)row[i];
            deg_ref[i] = a - b;
        }
    }
}

/* W Degree/Degree/Degree.cpp:84-88: FP32 running sum in CSR order. */
void orc_strength(const int32_t *row, const float *val, int64_t n, float *s)
{
    for (int64_t i = 0; i < n; i++) {
        float acc = 0.0f;
        for (int32_t e = row[i]; e < row[i + 1]; e++) acc += val[e];
        s[i] = acc;
    }
}

/* ------------------------------------------------------------- clustering */

/* |N(a) ∩ N(b)| restricted to members > floor, by merging sorted rows. */
static int64_t isect_count(const int32_t *col, int32_t a0, int32_t a1, int32_t b0, int32_t b1)
{
    int64_t c = 0;
    while (a0 < a1 && b0 < b1) {
        int32_t x = col[a0], y = col[b0];
        if (x < y) a0++;
        else if (y < x) b0++;
        else { c++; a0++; b0++; }
    }
    return c;
}

/* Triangles through each vertex, T_v = #{(i<j) in N(v) : a_ij}, as counted by
 * the pair loop of U Cp/Cp/Cp.cpp:102-108 (here by sorted intersection). */
void orc_triangles(const int32_t *row, const int32_t *col, int64_t n, int64_t *tri)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t v = 0; v < n; v++) {
        int64_t s = 0;
        for (int32_t e = row[v]; e < row[v + 1]; e++) {
            int32_t u = col[e];
            s += isect_count(col, row[v], row[v + 1], row[u], row[u + 1]);
        }
        tri[v] = s / 2;
    }
}

/* U Cp/Cp/Cp.cpp:95-111: Cp[k] = (double)T/k/(k-1)*2 stored as float, 0 for
 * k < 2; returns the FP64 mean of the stored floats. */
double orc_cp_binary(const int32_t *row, const int32_t *col, int64_t n, float *cp)
{
    int64_t *tri = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    orc_triangles(row, col, n, tri);
    double mean = 0.0;
    for (int64_t v = 0; v < n; v++) {
        int64_t k = row[v + 1] - row[v];
        float c = 0.0f;
        if (k >= 2) c = (float)((double)tri[v] / (double)k / (double)(k - 1) * 2.0);
        cp[v] = c;
        mean += c;
    }
    free(tri);
    return n > 0 ? mean / (double)n : 0.0;
}

/* weight of edge (a,b) by binary search of a's sorted row; 0 if absent. */
static int find_w(const int32_t *row, const int32_t *col, const float *val,
                  int32_t a, int32_t b, double *w)
{
    int32_t lo = row[a], hi = row[a + 1] - 1;
    while (lo <= hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (col[mid] == b) { *w = (double)val[mid]; return 1; }
        if (col[mid] < b) lo = mid + 1; else hi = mid - 1;
    }
    return 0;
}

/* Barrat (W Cp/Cp/Cp_1.cpp:109-127):
 *   C_k = sum_{i<j in N(k), w_ij != 0} (w_ki + w_kj) / (deg_k - 1) / s_k
 * Onnela (W Cp/Cp/Cp_2.cpp:117-125, weights divided by the maximum weight,
 * W Cp/Cp/Cp_2.cpp:27-30 and W Cp_GPU/CUCP/Cp.cu:223-239):
 *   C_k = 2 sum_{i<j in N(k), w_ij != 0} (w^_ij w^_ki w^_kj)^(1/3) / deg_k / (deg_k - 1)
 * Both are 0 for deg_k < 2.  Returns the mean. */
static double cpw_wmax(const int32_t *row, const float *val, int64_t n, int wmax_mode)
{
    double wmax = 0.0;
    int64_t nnz = row[n];
    int64_t lim = (wmax_mode == 1) ? (n < nnz ? n : nnz) : nnz;
    if (lim > 0) wmax = val[0];
    for (int64_t e = 1; e < lim; e++) if (val[e] > wmax) wmax = val[e];
    return wmax;
}

static double cpw_vertex(const int32_t *row, const int32_t *col, const float *val, int kind, double wmax,
                         int64_t k)
{
    int32_t b = row[k], e = row[k + 1];
    int64_t deg = e - b;
    if (deg < 2) return 0.0;
    double sk = 0.0;
    for (int32_t q = b; q < e; q++) sk += (double)val[q];
    double acc = 0.0;
    for (int32_t p = b; p < e; p++) {
        int32_t i = col[p];
        double wki = (double)val[p];
        /* j in N(k) ∩ N(i), j > i */
        int32_t a0 = p + 1, a1 = e, b0 = row[i], b1 = row[i + 1];
        while (a0 < a1 && b0 < b1) {
            int32_t x = col[a0], y = col[b0];
            if (x < y) a0++;
            else if (y < x) b0++;
            else {
                double wij = (double)val[b0], wkj = (double)val[a0];
                if (wij != 0.0) {
                    if (kind == 1) acc += wki + wkj;
                    else acc += pow((wij / wmax) * (wki / wmax) * (wkj / wmax), 1.0 / 3.0);
                }
                a0++; b0++;
            }
        }
    }
    if (kind == 1) return acc / (double)(deg - 1) / sk;
    return 2.0 * acc / (double)deg / (double)(deg - 1);
}

double orc_cp_weighted(const int32_t *row, const int32_t *col, const float *val,
                       int64_t n, int kind, int wmax_mode, double *cp)
{
    const double wmax = cpw_wmax(row, val, n, wmax_mode);
    double mean = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : mean)
    for (int64_t k = 0; k < n; k++) {
        cp[k] = cpw_vertex(row, col, val, kind, wmax, k);
        mean += cp[k];
    }
    (void)find_w;
    return n > 0 ? mean / (double)n : 0.0;
}

/* orc_cp_weighted at the vertices verts[0..nv) only (same normalisation). */
void orc_cp_weighted_at(const int32_t *row, const int32_t *col, const float *val, int64_t n, int kind,
                        int wmax_mode, const int32_t *verts, int64_t nv, double *cp)
{
    const double wmax = cpw_wmax(row, val, n, wmax_mode);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < nv; q++) cp[q] = cpw_vertex(row, col, val, kind, wmax, verts[q]);
}

/* -------------------------------------------------------------------- BC */

/* U BC/BC_CPU/BC_CPU/BC_CPU.cpp:22-75 with uint64 sigma and FP64 delta:
 * BFS from s (queue, CSR order), predecessor lists in discovery order, then
 * the stack in reverse BFS order: delta[v] += sigma[v]/sigma[w]*(1+delta[w]),
 * BC[w] += delta[w] for w != s.  Summed over ordered pairs (no /2). */
void orc_bc(const int32_t *row, const int32_t *col, int64_t n,
            const int32_t *sources, int64_t nsrc, double *bc)
{
    int64_t nnz = row[n];
    if (!sources) nsrc = n;
    for (int64_t v = 0; v < n; v++) bc[v] = 0.0;
#pragma omp parallel
    {
        int32_t *dist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
        uint64_t *sig = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n + 1));
        double *del = (double *)malloc(sizeof(double) * (size_t)(n + 1));
        int32_t *pcnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
        int32_t *pred = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
        int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
        double *part = (double *)calloc((size_t)(n + 1), sizeof(double));
#pragma omp for schedule(static)
        for (int64_t si = 0; si < nsrc; si++) {
            int32_t s = sources ? sources[si] : (int32_t)si;
            for (int64_t v = 0; v < n; v++) { dist[v] = -1; sig[v] = 0; del[v] = 0.0; pcnt[v] = 0; }
            sig[s] = 1; dist[s] = 0;
            int64_t head = 0, tail = 0;
            order[tail++] = s;
            while (head < tail) {
                int32_t v = order[head++];
                for (int32_t e = row[v]; e < row[v + 1]; e++) {
                    int32_t w = col[e];
                    if (dist[w] < 0) { dist[w] = dist[v] + 1; order[tail++] = w; }
                    if (dist[w] == dist[v] + 1) {
                        sig[w] += sig[v];
                        pred[row[w] + pcnt[w]] = v;
                        pcnt[w]++;
                    }
                }
            }
            for (int64_t q = tail - 1; q >= 0; q--) {
                int32_t w = order[q];
                for (int32_t p = 0; p < pcnt[w]; p++) {
                    int32_t v = pred[row[w] + p];
                    del[v] = del[v] + 1.0 * (double)sig[v] / (double)sig[w] * (1.0 + del[w]);
                }
                if (w != s) part[w] += del[w];
            }
        }
#pragma omp critical
        {
            for (int64_t v = 0; v < n; v++) bc[v] += part[v];
        }
        free(dist); free(sig); free(del); free(pcnt); free(pred); free(order); free(part);
    }
}

/* U BFS_MulCPU/BFS_MulCPU/APSP_BFS.cpp:163-190: queue BFS, dist[w] = dist[v]+1. */
void orc_bfs(const int32_t *row, const int32_t *col, int64_t n, int32_t src, int32_t *level)
{
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    for (int64_t v = 0; v < n; v++) level[v] = -1;
    int64_t head = 0, tail = 0;
    level[src] = 0;
    q[tail++] = src;
    while (head < tail) {
        int32_t v = q[head++];
        for (int32_t e = row[v]; e < row[v + 1]; e++) {
            int32_t w = col[e];
            if (level[w] < 0) { level[w] = level[v] + 1; q[tail++] = w; }
        }
    }
    free(q);
}

/* U BFS_MulCPU/BFS_MulCPU/APSP_BFS.cpp:100-133.  The reference fills a dense
 * float hop matrix (0 where unreachable and on the diagonal) and then walks
 * i < j in row-major order: a pair with d < 1 marks the network unconnected,
 * otherwise r = 1.0f / d is added to the FP64 Lp sum and to the FP32 sums of
 * both ends; after row i, e_i /= (n - 1) and a zero e_i counts as isolated.
 * Lp = 1 / (sum / n / (n - 1)) / 2.  Row i's hops come from one BFS from i
 * (d(i -> j), the value the matrix holds at [i][j]), so no n x n buffer. */
double orc_bfs_efficiency(const int32_t *row, const int32_t *col, int64_t n, float *eff,
                          int32_t *isolated, int32_t *unconnected)
{
    int32_t *lev = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    double lp = 0.0;
    int32_t iso = 0, unc = 0;
    for (int64_t v = 0; v < n; v++) eff[v] = 0.0f;
    for (int64_t i = 0; i < n; i++) {
        orc_bfs(row, col, n, (int32_t)i, lev);
        for (int64_t j = i + 1; j < n; j++) {
            const float d = lev[j] < 0 ? 0.0f : (float)lev[j];
            if (d < 1) {
                unc = 1;
            } else {
                const float r = 1.0f / d;
                lp += r;
                eff[i] += r;
                eff[j] += r;
            }
        }
        eff[i] /= (float)(n - 1);
        if (eff[i] == 0) iso++;
    }
    free(lev);
    lp /= (double)n;
    lp /= (double)(n - 1);
    lp = 1 / lp / 2;
    if (isolated) *isolated = iso;
    if (unconnected) *unconnected = unc;
    return lp;
}

/* ----------------------------------------------------------- Floyd-Warshall */

/* W Lp/CUBFW_Lp/CUBFW_Lp/BFW.cpp:26-38: for k, i, j: dst = min(dst, ik + kj). */
static void fw_block(float *d, int64_t ld, int64_t b,
                     int64_t dr, int64_t dc, int64_t ir, int64_t ic, int64_t kr, int64_t kc)
{
    for (int64_t k = 0; k < b; k++)
        for (int64_t i = 0; i < b; i++) {
            float ik = d[(ir + i) * ld + ic + k];
            float *dst = d + (dr + i) * ld + dc;
            const float *kj = d + (kr + k) * ld + kc;
            for (int64_t j = 0; j < b; j++) {
                float c = ik + kj[j];
                if (c < dst[j]) dst[j] = c;
            }
        }
}

/* W BFW.cpp:40-108 (phases 1, 2 row then column, 3) on an n x n row-major
 * matrix padded to a multiple of b with +inf (0 on the padded diagonal), as
 * W FormBlock.cpp:45-75 pads.  Phase 3 tiles are independent. */
void orc_fw_blocked(float *d, int64_t n, int64_t b)
{
    int64_t np = (n + b - 1) / b * b, nb = np / b;
    float *m = (float *)malloc(sizeof(float) * (size_t)(np * np));
    for (int64_t i = 0; i < np; i++)
        for (int64_t j = 0; j < np; j++)
            m[i * np + j] = (i < n && j < n) ? d[i * n + j] : (i == j ? 0.0f : INFINITY);
    for (int64_t kb = 0; kb < nb; kb++) {
        int64_t K = kb * b;
        fw_block(m, np, b, K, K, K, K, K, K);
        for (int64_t c = 0; c < nb; c++) if (c != kb) fw_block(m, np, b, K, c * b, K, K, K, c * b);
        for (int64_t r = 0; r < nb; r++) if (r != kb) fw_block(m, np, b, r * b, K, r * b, K, K, K);
#pragma omp parallel for collapse(2) schedule(static)
        for (int64_t r = 0; r < nb; r++)
            for (int64_t c = 0; c < nb; c++)
                if (r != kb && c != kb) fw_block(m, np, b, r * b, c * b, r * b, K, K, c * b);
    }
    for (int64_t i = 0; i < n; i++) memcpy(d + i * n, m + i * np, sizeof(float) * (size_t)n);
    free(m);
}

/* W FormBlock.cpp:49-74 in row-major: +inf, d_ij from the edge weight, 0 on
 * the diagonal.  kind 0: 1/w (the reference), 1: 1-w, 2: 1 (hops). */
void orc_fw_init(const int32_t *row, const int32_t *col, const float *val, int64_t n,
                 int kind, float *d)
{
    for (int64_t q = 0; q < n * n; q++) d[q] = INFINITY;
    for (int64_t i = 0; i < n; i++)
        for (int32_t e = row[i]; e < row[i + 1]; e++) {
            float w = val ? val[e] : 1.0f;
            float x = kind == 0 ? 1.0f / w : (kind == 1 ? 1.0f - w : 1.0f);
            d[i * n + col[e]] = x;
        }
    for (int64_t i = 0; i < n; i++) d[i * n + i] = 0.0f;
}

/* -------------------------------------------------------------- modularity */

typedef struct {
    int64_t ng;          /* group size */
    int32_t *g;          /* global ids, ascending */
    int32_t *srow;       /* sub-CSR over members, local column ids */
    int32_t *scol;
    double *sval;
    double *kg;          /* k_i of members */
    double *sumbg;       /* k^(g)_i - k_i K_g / 2m */
} mgroup;

/* W Newman_Modularity_GPU/CUmodularity/Modularity_CPU.cpp:224-298 and
 * modularity_GPU.cu:225-325: power iteration on B^(g) + beta I from
 * v_i = global index, until ||v - v0|| or ||v + v0|| <= 1e-6 or max_iter
 * (10000) steps; returns +-||(B + beta I) v0|| with the sign of v0 . v at the
 * largest |v|.  stop = 1 (the Rayleigh rule) instead tracks
 * rho_k = v0 . (B + beta I) v0 / v0 . v0, stops once two consecutive values
 * satisfy |rho_k - rho_(k-1)| <= tol |rho_k| (or at max_iter) and returns the
 * last rho. */
static double lead_vector(const mgroup *G, double m2, double beta, double *v, double *v0, int *iters,
                          int stop, double tol, int max_iter)
{
    int64_t ng = G->ng;
    for (int64_t i = 0; i < ng; i++) v[i] = (double)G->g[i];
    double err1 = 1.0, err2 = 1.0, vnorm = 0.0, rho = 0.0, rho_prev = 0.0;
    int it = 0;
    for (;;) {
        if (it >= max_iter) break;
        if (stop == 0 && !(err1 > 1e-6 && err2 > 1e-6)) break;
        if (stop == 1 && it >= 2 && fabs(rho - rho_prev) <= tol * fabs(rho)) break;
        for (int64_t i = 0; i < ng; i++) v0[i] = v[i];
        double vk = 0.0;
        for (int64_t i = 0; i < ng; i++) vk += v0[i] * G->kg[i];
        /* rows in parallel, each row's sum in CSR order (deterministic) */
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t i = 0; i < ng; i++) {
            double acc = 0.0;
            for (int32_t e = G->srow[i]; e < G->srow[i + 1]; e++) acc += v0[G->scol[e]] * G->sval[e];
            acc -= vk / m2 * G->kg[i] + (G->sumbg[i] - beta) * v0[i];
            v[i] = acc;
        }
        double s = 0.0, sxy = 0.0, sxx = 0.0;
        for (int64_t i = 0; i < ng; i++) {
            s += v[i] * v[i];
            sxy += v0[i] * v[i];
            sxx += v0[i] * v0[i];
        }
        rho_prev = rho;
        rho = sxy / sxx;
        vnorm = sqrt(s);
        for (int64_t i = 0; i < ng; i++) v[i] = v[i] / vnorm;
        double e1 = 0.0, e2 = 0.0;
        for (int64_t i = 0; i < ng; i++) {
            double a = v[i] - v0[i], b = v[i] + v0[i];
            e1 += a * a; e2 += b * b;
        }
        err1 = sqrt(e1); err2 = sqrt(e2);
        it++;
    }
    int64_t mi = 0;
    for (int64_t i = 0; i < ng; i++) if (fabs(v[i]) > fabs(v[mi])) mi = i;
    if (iters) *iters += it;
    if (stop == 1) return rho;
    return (v[mi] * v0[mi] > 0) ? vnorm : -vnorm;
}

static void node_k(const int32_t *row, const float *val, int64_t n, int kmode, double *k, double *m2)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double ki = 0.0;
        if (kmode == 1 || !val) ki = (double)(row[i + 1] - row[i]);
        else for (int32_t e = row[i]; e < row[i + 1]; e++) ki += (double)val[e];
        k[i] = ki;
        s += ki;
    }
    *m2 = s;
}

/* W modularity_GPU.cu:405-525 (Partition_GPU) with Sub_Partition
 * (Modularity_CPU.cpp:164-208 / modularity_GPU.cu:329-397).  Groups are
 * processed in label order Round = 0, 1, 2, ...; a split writes labels
 * Max+1 (the larger side) and Max+2.  The negative-eigenvalue shift uses the
 * reference's size rule: groups above 25000 nodes take the GPU test
 * lambda < -0.01 (modularity_GPU.cu:369, 483), smaller ones lambda < 0
 * (Modularity_CPU.cpp:183). */
double orc_modularity(const int32_t *row, const int32_t *col, const float *val,
                      int64_t n, int kmode, int32_t *labels, double *lambda_out,
                      double *eig_out, int32_t *nrounds_out)
{
    return orc_modularity_ex(row, col, val, n, kmode, 0, 1e-8, 10000, 0, labels, lambda_out, eig_out,
                             nrounds_out);
}

/* max_depth > 0: a group max_depth bisections below the whole network is not
 * split again (it keeps one label, given in round order like any final
 * group). */
double orc_modularity_ex(const int32_t *row, const int32_t *col, const float *val, int64_t n, int kmode,
                         int stop, double tol, int max_iter, int max_depth, int32_t *labels, double *lambda_out,
                         double *eig_out, int32_t *nrounds_out)
{
    double *k = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double m2 = 0.0;
    node_k(row, val, n, kmode, k, &m2);
    int32_t *res = (int32_t *)calloc((size_t)(n + 1), sizeof(int32_t));
    int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n + 2));
    int32_t *loc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t *depth = (int32_t *)calloc((size_t)(2 * n + 4), sizeof(int32_t));
    int64_t nnz = row[n];
    mgroup G;
    G.g = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    G.srow = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 2));
    G.scol = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
    G.sval = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    G.kg = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    G.sumbg = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *v = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *v0 = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    int32_t round = 0, maxr = 0, index = 1, rounds = 0;
    if (m2 <= 0.0) {
        for (int64_t i = 0; i < n; i++) labels[i] = 1;
        if (lambda_out) *lambda_out = 0.0;
        if (nrounds_out) *nrounds_out = 0;
        goto done;
    }
    for (int64_t i = 0; i < n; i++) loc[i] = -1;
    while (round <= maxr) {
        int64_t ng = 0;
        double Kg = 0.0;
        for (int64_t i = 0; i < n; i++) if (res[i] == round) { loc[i] = (int32_t)ng; G.g[ng++] = (int32_t)i; Kg += k[i]; }
        if (ng > 0 && max_depth > 0 && depth[round] >= max_depth) {
            adj[round] = index++;
            for (int64_t a = 0; a < ng; a++) loc[G.g[a]] = -1;
        } else if (ng > 0) {
            rounds++;
            G.ng = ng;
            int64_t pos = 0;
            G.srow[0] = 0;
            for (int64_t a = 0; a < ng; a++) {
                int32_t i = G.g[a];
                double kin = 0.0;
                for (int32_t e = row[i]; e < row[i + 1]; e++) {
                    int32_t j = col[e];
                    if (res[j] != round) continue;
                    double w = (val && kmode != 2) ? (double)val[e] : 1.0;
                    G.scol[pos] = loc[j];
                    G.sval[pos] = w;
                    kin += (kmode == 1 && val) ? 1.0 : w;
                    pos++;
                }
                G.srow[a + 1] = (int32_t)pos;
                G.kg[a] = k[i];
                G.sumbg[a] = kin - k[i] * Kg / m2;
            }
            int its = 0;
            double lam = lead_vector(&G, m2, 0.0, v, v0, &its, stop, tol, max_iter);
            int shift = (ng > 25000) ? (lam < -0.01) : (lam < 0.0);
            if (shift) lam += lead_vector(&G, m2, -lam, v, v0, &its, stop, tol, max_iter);
            if (round == 0) {
                if (lambda_out) *lambda_out = lam;
                if (eig_out) for (int64_t a = 0; a < ng; a++) eig_out[G.g[a]] = v[a];
            }
            int64_t sp = 0, sn = 0;
            for (int64_t a = 0; a < ng; a++) { sp += v[a] > 0; sn += v[a] <= 0; }
            int split = lam > 0.01 && sp > 1 && sn > 1;
            if (split) {
                for (int64_t a = 0; a < ng; a++)
                    res[G.g[a]] = maxr + 1 + (v[a] * ((double)(sp - sn) + 0.5) <= 0);
                depth[maxr + 1] = depth[maxr + 2] = depth[round] + 1;
                maxr += 2;
            } else {
                adj[round] = index++;
            }
            for (int64_t a = 0; a < ng; a++) loc[G.g[a]] = -1;
        }
        round++;
    }
    for (int64_t i = 0; i < n; i++) labels[i] = adj[res[i]];
    if (nrounds_out) *nrounds_out = rounds;
done:;
    double q = orc_modularity_q(row, col, val, n, kmode, labels);
    free(k); free(res); free(adj); free(loc); free(depth);
    free(G.g); free(G.srow); free(G.scol); free(G.sval); free(G.kg); free(G.sumbg);
    free(v); free(v0);
    return q;
}

/* W modularity_GPU.cu:500-509: Q = [sum_ij A_ij d(ci,cj) - sum_ij k_i k_j d / 2m] / 2m,
 * the second sum taken per community (sum of K_c^2). */
double orc_modularity_q(const int32_t *row, const int32_t *col, const float *val,
                        int64_t n, int kmode, const int32_t *labels)
{
    double *k = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double m2 = 0.0;
    node_k(row, val, n, kmode, k, &m2);
    int32_t maxl = 0;
    for (int64_t i = 0; i < n; i++) if (labels[i] > maxl) maxl = labels[i];
    double *kc = (double *)calloc((size_t)maxl + 2, sizeof(double));
    double in = 0.0;
    for (int64_t i = 0; i < n; i++) {
        kc[labels[i]] += k[i];
        for (int32_t e = row[i]; e < row[i + 1]; e++)
            if (labels[col[e]] == labels[i]) in += (kmode == 1 || !val) ? 1.0 : (double)val[e];
    }
    double s2 = 0.0;
    for (int32_t c = 0; c <= maxl; c++) s2 += kc[c] * kc[c];
    free(k); free(kc);
    if (m2 <= 0.0) return 0.0;
    return (in - s2 / m2) / m2;
}

/* ------------------------------------------------- eigenvector centrality */

/* U CUEC/CUEC/CUEC_unweighted.cu:61-196 (W CUEC/weighted-net.cu:55-203 with
 * the stored weights): u = 0..n-1; while err1 > 1e-6 and err2 > 1e-6 and
 * fewer than 10000 iterations: u0 = u, |u| (nrm2), u *= 1/|u| (scal), u = A u
 * (csrmv), err1 = |u - u0|, err2 = |u + u0|.  Sums are sequential here. */
double orc_ec(const int32_t *row, const int32_t *col, const float *val, int64_t n,
              double *vhat, double *avec, int32_t *iters)
{
    double *u = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *u0 = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *h = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    for (int64_t i = 0; i < n; i++) u[i] = (double)i;
    double err1 = 1, err2 = 1, vnorm = 0;
    int32_t it = 0;
    while (err1 > 1e-6 && err2 > 1e-6 && it < 10000) {
        double ss = 0.0;
        for (int64_t i = 0; i < n; i++) {
            u0[i] = u[i];
            ss += u[i] * u[i];
        }
        vnorm = sqrt(ss);
        const double t = 1 / vnorm;
        for (int64_t i = 0; i < n; i++) h[i] = u[i] * t;
        double e1 = 0.0, e2 = 0.0;
        for (int64_t i = 0; i < n; i++) {
            double y = 0.0;
            for (int32_t e = row[i]; e < row[i + 1]; e++) {
                const double w = val ? (double)val[e] : 1.0;
                y += w * h[col[e]];
            }
            u[i] = y;
            const double d1 = y - u0[i], d2 = y + u0[i];
            e1 += d1 * d1;
            e2 += d2 * d2;
        }
        err1 = sqrt(e1);
        err2 = sqrt(e2);
        it++;
    }
    if (vhat) memcpy(vhat, h, sizeof(double) * (size_t)n);
    if (avec) memcpy(avec, u, sizeof(double) * (size_t)n);
    if (iters) *iters = it;
    free(u); free(u0); free(h);
    return vnorm;
}

/* --------------------------------------------- participation coefficient */

/* W PC_CPU/PC_CPU/main.cpp:96-160, in FP32 as there. */
int32_t orc_pc(const int32_t *row, const int32_t *col, const float *val, int64_t n,
               const int32_t *labels, float *pc, float *pc_norm)
{
    int32_t k = 0;
    for (int64_t i = 0; i < n; i++)
        if (labels[i] > k) k = labels[i];
    float *wm = (float *)calloc((size_t)(k > 0 ? k : 1), sizeof(float));
    for (int64_t i = 0; i < n; i++) {
        float deg = 0.0f;
        for (int32_t e = row[i]; e < row[i + 1]; e++) deg += val ? val[e] : 1.0f;
        if (deg == 0) {
            pc[i] = 0.0f;
            continue;
        }
        memset(wm, 0, sizeof(float) * (size_t)k);
        for (int32_t e = row[i]; e < row[i + 1]; e++) wm[labels[col[e]] - 1] += val ? val[e] : 1.0f;
        float sum = 0.0f;
        for (int32_t j = 0; j < k; j++) {
            const float t = wm[j] / deg;
            const float t2 = t * t;
            sum += t2;
        }
        pc[i] = 1.0f - sum;
    }
    const float mx = (float)(k - 1) / (float)k;
    for (int64_t i = 0; i < n; i++) pc_norm[i] = k != 1 ? pc[i] / mx : pc[i];
    free(wm);
    return k;
}

/* ------------------------------------------------------ subject averaging */

/* U CorMat.cpp:35-42 in FP32, the logarithm in FP64. */
static float fisher_f(float r)
{
    if (r == 1) r = (float)((double)r - 1e-6);
    const float q = (1 + r) / (1 - r);
    return (float)(0.5 * log((double)q));
}

/* U CorMat.cpp:44-49 in FP32, the exponential in FP64. */
static float inv_fisher_f(float z)
{
    float r = (float)exp((double)(2 * z));
    r = (r - 1) / (r + 1);
    return r;
}

/* post_block (U CorMat.cpp:806-838): every pair i < j in row-major order adds
 * its value to the packed accumulator.  The reference drops |r| > 1 from FP32
 * rounding; r here is FP64 and clamped, as in orc_r. */
void orc_cormat_add(const double *z64, int64_t n, int t, int fisher, float *acc)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; i++) {
        int64_t idx = i * n - i * (i + 1) / 2;
        for (int64_t j = i + 1; j < n; j++, idx++) {
            const float r = (float)orc_r(z64, t, i, j);
            acc[idx] += fisher ? fisher_f(r) : r;
        }
    }
}

/* U CorMat.cpp:507-517. */
void orc_cormat_finish(float *acc, int64_t m1, int nsub, int fisher)
{
    for (int64_t q = 0; q < m1; q++) {
        acc[q] /= (float)nsub;
        if (fisher) acc[q] = inv_fisher_f(acc[q]);
    }
}

static void dense_csr(const float *a, int64_t n, const unsigned char *keep_upper, int32_t *row, int32_t *col,
                      float *val)
{
    int64_t pos = 0;
    row[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
            const int64_t idx = lo * n - lo * (lo + 1) / 2 + (hi - lo - 1);
            if (keep_upper[idx]) {
                col[pos] = (int32_t)j;
                if (val) val[pos] = a[idx];
                pos++;
            }
        }
        row[i + 1] = (int32_t)pos;
    }
}

/* FormFullGraph (U CorMat.cpp:647-664) + FormCSRGraph. */
int64_t orc_dense_csr_r(const float *a, int64_t n, float thr, int32_t *row, int32_t *col, float *val)
{
    const int64_t m1 = n * (n - 1) / 2;
    unsigned char *keep = (unsigned char *)malloc((size_t)(m1 > 0 ? m1 : 1));
    int64_t cnt = 0;
    for (int64_t q = 0; q < m1; q++) {
        keep[q] = a[q] > thr;
        cnt += keep[q];
    }
    if (row) dense_csr(a, n, keep, row, col, val);
    free(keep);
    return 2 * cnt;
}

static int cmp_desc_f(const void *x, const void *y)
{
    const float a = *(const float *)x, b = *(const float *)y;
    return a < b ? 1 : (a > b ? -1 : 0);
}

/* FormFullGraph_s (U CorMat.cpp:693-773): t = k-th largest; every a > t and
 * the earliest ties a == t (a stack that drops the newest tie whenever the
 * count passes k leaves the earliest ones). */
float orc_dense_csr_s(const float *a, int64_t n, int64_t k, int32_t *row, int32_t *col, float *val)
{
    const int64_t m1 = n * (n - 1) / 2;
    if (k < 1 || k > m1) return NAN;
    float *s = (float *)malloc(sizeof(float) * (size_t)m1);
    memcpy(s, a, sizeof(float) * (size_t)m1);
    qsort(s, (size_t)m1, sizeof(float), cmp_desc_f);
    const float t = s[k - 1];
    free(s);
    unsigned char *keep = (unsigned char *)malloc((size_t)m1);
    int64_t gt = 0;
    for (int64_t q = 0; q < m1; q++) gt += a[q] > t;
    int64_t ties = k - gt;
    for (int64_t q = 0; q < m1; q++) {
        keep[q] = a[q] > t;
        if (a[q] == t && ties > 0) {
            keep[q] = 1;
            ties--;
        }
    }
    if (row) dense_csr(a, n, keep, row, col, val);
    free(keep);
    return t;
}

/* -------------------------------------------------------------- rewiring */

/* splitmix64 (Steele, Lea, Flood 2014), the random stream standing in for the
 * reference's srand(time(0)) / rand(). */
static uint64_t sm64_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* Set of undirected edges: linear probing, backward-shift deletion. */
typedef struct {
    uint64_t *t;
    uint64_t mask;
} edge_set;
#define ES_EMPTY (~0ull)

static uint64_t es_key(int32_t a, int32_t b)
{
    const uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    return ((uint64_t)lo << 32) | hi;
}
static uint64_t es_home(const edge_set *s, uint64_t k)
{
    uint64_t z = k * 0x9e3779b97f4a7c15ull;
    return (z ^ (z >> 29)) & s->mask;
}
static uint64_t es_slot(const edge_set *s, uint64_t k)  /* slot of k, or its first empty */
{
    uint64_t i = es_home(s, k);
    while (s->t[i] != ES_EMPTY && s->t[i] != k) i = (i + 1) & s->mask;
    return i;
}
static int es_has(const edge_set *s, int32_t a, int32_t b)
{
    const uint64_t k = es_key(a, b);
    return s->t[es_slot(s, k)] == k;
}
static void es_add(edge_set *s, int32_t a, int32_t b)
{
    const uint64_t k = es_key(a, b);
    s->t[es_slot(s, k)] = k;
}
static void es_del(edge_set *s, int32_t a, int32_t b)
{
    uint64_t i = es_slot(s, es_key(a, b)), j = i;
    for (;;) {
        j = (j + 1) & s->mask;
        if (s->t[j] == ES_EMPTY) break;
        const uint64_t h = es_home(s, s->t[j]);
        /* t[j] may move back to i unless its home lies cyclically in (i, j] */
        const int stays = i <= j ? (h > i && h <= j) : (h > i || h <= j);
        if (!stays) {
            s->t[i] = s->t[j];
            i = j;
        }
    }
    s->t[i] = ES_EMPTY;
}

typedef struct {
    int32_t c;
    float v;
} col_val;
static int cmp_col_val(const void *a, const void *b)
{
    const int32_t x = ((const col_val *)a)->c, y = ((const col_val *)b)->c;
    return (x > y) - (x < y);
}

/* Maslov / Rewire (U BFS_MulCPU/Maslov.cpp:8-110) and Maslov_weighted_1 (W
 * Maslov/Maslov/Maslov.cpp:50-196): the edges (upper triangle, row-major)
 * into E; per attempt i = rnd % M, j = rnd % M, give up when the two edges
 * share a vertex, else rnd & 1 picks 1-3,2-4 (E[i].C = V3, E[j].R = V2) or
 * 1-4,3-2 (E[i].C = V4, E[j].C = V2), given up when either new edge exists.
 * Weights stay with their E slot.  The output keeps the row offsets; columns
 * (and weights) are sorted per row.  attempts 0 means 2M.  Returns the swap
 * count. */
int64_t orc_rewire(const int32_t *row, const int32_t *col, const float *val, int64_t n, uint64_t seed,
                   int64_t attempts, int32_t *out_col, float *out_val)
{
    const int64_t nnz = row[n];
    int64_t m = 0;
    for (int64_t u = 0; u < n; u++)
        for (int32_t e = row[u]; e < row[u + 1]; e++) m += col[e] > u;
    int32_t *er = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    int32_t *ec = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    float *ev = (float *)malloc(sizeof(float) * (size_t)(m + 1));
    edge_set s;
    uint64_t cap = 1024;
    while (cap < 4 * (uint64_t)m) cap <<= 1;
    s.mask = cap - 1;
    s.t = (uint64_t *)malloc(sizeof(uint64_t) * cap);
    for (uint64_t q = 0; q < cap; q++) s.t[q] = ES_EMPTY;
    int64_t k = 0;
    for (int64_t u = 0; u < n; u++)
        for (int32_t e = row[u]; e < row[u + 1]; e++)
            if (col[e] > u) {
                er[k] = (int32_t)u;
                ec[k] = col[e];
                ev[k] = val ? val[e] : 1.0f;
                es_add(&s, (int32_t)u, col[e]);
                k++;
            }
    const int64_t tries = attempts ? attempts : 2 * m;
    uint64_t st = seed;
    int64_t swaps = 0;
    for (int64_t q = 0; m > 0 && q < tries; q++) {
        const int64_t i = (int64_t)(sm64_next(&st) % (uint64_t)m);
        const int64_t j = (int64_t)(sm64_next(&st) % (uint64_t)m);
        const int32_t v1 = er[i], v2 = ec[i], v3 = er[j], v4 = ec[j];
        if (v1 == v3 || v1 == v4 || v2 == v3 || v2 == v4) continue;
        if (sm64_next(&st) & 1) {
            if (es_has(&s, v1, v3) || es_has(&s, v2, v4)) continue;
            es_del(&s, v1, v2);
            es_del(&s, v3, v4);
            es_add(&s, v1, v3);
            es_add(&s, v2, v4);
            ec[i] = v3;
            er[j] = v2;
        } else {
            if (es_has(&s, v1, v4) || es_has(&s, v2, v3)) continue;
            es_del(&s, v1, v2);
            es_del(&s, v3, v4);
            es_add(&s, v1, v4);
            es_add(&s, v2, v3);
            ec[i] = v4;
            ec[j] = v2;
        }
        swaps++;
    }
    col_val *buf = (col_val *)malloc(sizeof(col_val) * (size_t)(nnz + 1));
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t u = 0; u < n; u++) cur[u] = row[u];
    for (int64_t e = 0; e < m; e++) {
        buf[cur[er[e]]++] = (col_val){ec[e], ev[e]};
        buf[cur[ec[e]]++] = (col_val){er[e], ev[e]};
    }
    for (int64_t u = 0; u < n; u++) qsort(buf + row[u], (size_t)(row[u + 1] - row[u]), sizeof(col_val), cmp_col_val);
    for (int64_t e = 0; e < nnz; e++) {
        out_col[e] = buf[e].c;
        if (out_val) out_val[e] = buf[e].v;
    }
    free(buf);
    free(cur);
    free(s.t);
    free(er);
    free(ec);
    free(ev);
    return swaps;
}
