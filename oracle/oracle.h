/*
 * oracle.h -- FP64 CPU restatement of the PAGANI network-construction-and-metrics
 * path.  TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may call it, as the checker.  The product
 * (libpagani.so, include/pagani.h) never links or calls anything here.
 *
 * Every function cites the reference file:line it restates.  Path prefixes:
 *   U/ = src/BNAPlatform-win64-cuda7.0-20151118/src/
 *   W/ = src/BNAPlatform-weighted-network-win64-cuda7.0-20151118/src/
 *
 * Parity pinning of each function (closed forms, networkx fixtures, the
 * compiled reference in oracle/_ref) is listed in the header of oracle.c.
 */
#ifndef PAGANI_ORACLE_H
#define PAGANI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* z-score of every voxel series.  bold is time-major [T][N] (BOLD[l*N+i]). */
void orc_set_threads(int nthreads);
int orc_get_threads(void);
void orc_zscore(const float *bold, int64_t n, int t, double *z64, float *z32);

/* r(i,j) for i<j from FP64 z rows (sequential sum, no FMA), clamped to [-1,1]. */
double orc_r(const double *z64, int t, int64_t i, int64_t j);

/* Effective FP32 threshold used by the edge test (weighted subtracts 1e-6). */
float orc_r_threshold_eff(float thr, int weighted);

/* Number of undirected edges the sparsity rule asks for (U CorMat.cpp:311-313). */
int64_t orc_sparsity_k(int64_t n, float s_percent);

/* r-mode: count of undirected edges with r > thr_eff. */
int64_t orc_corr_count_r(const double *z64, int64_t n, int t, float thr_eff);

/* r-mode: undirected edge count over rows [i0, i1) only (pairs i < j); a
 * bounded sample of the all-pairs work for CPU timing. */
int64_t orc_corr_count_rows(const double *z64, int64_t n, int t, float thr_eff, int64_t i0, int64_t i1);
void orc_corr_rows_r(const double *z64, int64_t n, int t, float thr_eff, const int32_t *rows, int64_t nrows,
                     int32_t *deg, int32_t *col, float *val);

/* r-mode: CSR.  row[n+1]; col/val sized by orc_corr_count_r()*2. val may be NULL. */
void orc_corr_csr_r(const double *z64, int64_t n, int t, float thr_eff,
                    int32_t *row, int32_t *col, float *val);

/* sparsity-mode: k = orc_sparsity_k(); returns the FP64 k-th largest r (t*) and
 * builds the CSR: all r > t* plus the earliest (row-major) ties. */
double orc_corr_csr_s(const double *z64, int64_t n, int t, int64_t k,
                      int32_t *row, int32_t *col, float *val);

/* Degree: exact int32, plus the float value the reference writes. */
void orc_degree(const int32_t *row, int64_t n, int32_t *deg, float *deg_ref);
/* Strength: sequential FP32 sum over the CSR row (W Degree.cpp:84-88). */
void orc_strength(const int32_t *row, const float *val, int64_t n, float *s);

/* Triangles per vertex (edges among neighbours), exact. */
void orc_triangles(const int32_t *row, const int32_t *col, int64_t n, int64_t *tri);
/* Binary clustering: Cp = (double)T/k/(k-1)*2 stored float (U Cp.cpp:95-110). */
double orc_cp_binary(const int32_t *row, const int32_t *col, int64_t n, float *cp);
/* Weighted clustering, FP64.  kind 1 = Barrat (W Cp_1.cpp), 2 = Onnela (W Cp_2.cpp).
 * wmax_mode 0 = max over all weights, 1 = max over val[0..n-1] (reference CPU). */
double orc_cp_weighted(const int32_t *row, const int32_t *col, const float *val,
                       int64_t n, int kind, int wmax_mode, double *cp);
void orc_cp_weighted_at(const int32_t *row, const int32_t *col, const float *val, int64_t n, int kind,
                        int wmax_mode, const int32_t *verts, int64_t nv, double *cp);

/* Brandes betweenness over the given sources (all if sources == NULL),
 * uint64 path counts, FP64 dependencies, ordered-pair sum (U BC_CPU.cpp:6-82). */
void orc_bc(const int32_t *row, const int32_t *col, int64_t n,
            const int32_t *sources, int64_t nsrc, double *bc);

/* BFS hop levels from one source (-1 unreachable) (U APSP_BFS.cpp:150-192). */
void orc_bfs(const int32_t *row, const int32_t *col, int64_t n, int32_t src, int32_t *level);

/* Binary nodal efficiency and Lp from BFS hop counts (U APSP_BFS.cpp:100-133):
 * the reference's i < j loop order, FP32 nodal sums, FP64 Lp sum; returns Lp.
 * isolated counts e_i == 0; unconnected is 1 when some pair is unreachable. */
double orc_bfs_efficiency(const int32_t *row, const int32_t *col, int64_t n, float *eff,
                          int32_t *isolated, int32_t *unconnected);

/* Blocked Floyd-Warshall, row-major n x n FP32 in place, block size b
 * (W BFW.cpp:26-108 on a padded matrix, W FormBlock.cpp:45-75). */
void orc_fw_blocked(float *d, int64_t n, int64_t b);

/* Distance matrix from CSR: kind 0: d = 1/w (W FormBlock.cpp:64), kind 1: d = 1-w,
 * kind 2: d = 1 (hops); inf where no edge, 0 on the diagonal. */
void orc_fw_init(const int32_t *row, const int32_t *col, const float *val, int64_t n,
                 int kind, float *d);

/* Leading-eigenvector modularity (Newman), recursive bisection.
 * kmode 0: k = strength, 2m = sum of weights (val may be NULL => binary);
 * kmode 1: k = binary degree, 2m = nnz (reference).  Writes labels 1..K and
 * returns Q.  eig_out (optional, n doubles) gets the round-0 leading vector and
 * lambda_out its eigenvalue. */
double orc_modularity(const int32_t *row, const int32_t *col, const float *val,
                      int64_t n, int kmode, int32_t *labels, double *lambda_out,
                      double *eig_out, int32_t *nrounds_out);
double orc_modularity_ex(const int32_t *row, const int32_t *col, const float *val, int64_t n, int kmode,
                         int stop, double tol, int max_iter, int max_depth, int32_t *labels, double *lambda_out,
                         double *eig_out, int32_t *nrounds_out);

/* Modularity Q of a given labelling (same k/m convention as kmode). */
double orc_modularity_q(const int32_t *row, const int32_t *col, const float *val,
                        int64_t n, int kmode, const int32_t *labels);

/* Eigenvector centrality by the reference's power iteration (U CUEC/CUEC/
 * CUEC_unweighted.cu:61-196, W CUEC/weighted-net.cu:55-203): start vector
 * 0..n-1, A weighted by val (1 when NULL).  vhat = normalised vector of the
 * last iteration (the unweighted tool's output), avec = A vhat (the weighted
 * tool's); returns that iteration's norm (the printed eigenvalue). */
double orc_ec(const int32_t *row, const int32_t *col, const float *val, int64_t n,
              double *vhat, double *avec, int32_t *iters);

/* Participation coefficient (W PC_CPU/PC_CPU/main.cpp:96-160): FP32 strength
 * and per-module sums in CSR order, pc = 1 - sum_j (s_ij / s_i)^2 over modules
 * 1..K (K = max label), pc_norm = pc / ((K-1)/K) (pc when K == 1).  val NULL
 * means unit weights.  Returns K. */
int32_t orc_pc(const int32_t *row, const int32_t *col, const float *val, int64_t n,
               const int32_t *labels, float *pc, float *pc_norm);

/* Subject averaging of the dense correlation matrix (U CorMat.cpp:380-540 with
 * post_block :775-838 and fisher_trans / inv_fisher_trans :35-49).  acc is the
 * packed upper triangle, row-major over i < j, M1 = n(n-1)/2 floats.
 * orc_cormat_add adds one subject: (float)clamp(r64) per pair, or its Fisher z
 * (FP32 (1+r)/(1-r), FP64 log, rounded to FP32; r == 1 first lowered by 1e-6).
 * orc_cormat_finish divides by nsub in FP32 and, for Fisher, maps back with
 * (e - 1)/(e + 1), e = exp(2 z) rounded to FP32. */
void orc_cormat_add(const double *z64, int64_t n, int t, int fisher, float *acc);
void orc_cormat_finish(float *acc, int64_t m1, int nsub, int fisher);

/* Thresholds on a packed FP32 matrix (U CorMat.cpp:647-664, 693-773, 839-857;
 * W CorMat.cpp:853-885 for the weights): r-mode keeps a > thr; returns the
 * directed edge count, CSR in row/col/val when row != NULL.  s-mode keeps the k
 * largest (ties at the k-th value in row-major order) and returns that value. */
int64_t orc_dense_csr_r(const float *a, int64_t n, float thr, int32_t *row, int32_t *col, float *val);
float orc_dense_csr_s(const float *a, int64_t n, int64_t k, int32_t *row, int32_t *col, float *val);

/* Degree-preserving rewiring by double edge swaps (U BFS_MulCPU/Maslov.cpp:
 * 8-110, W Maslov/Maslov/Maslov.cpp:50-196), sequential as the reference, with
 * a splitmix64 stream seeded by seed for its rand().  attempts 0 means 2M
 * (M = nnz/2).  out_col (nnz) gets the rewired columns under the same row
 * offsets, sorted per row; out_val (may be NULL) each edge's weight.  Returns
 * the swaps made. */
int64_t orc_rewire(const int32_t *row, const int32_t *col, const float *val, int64_t n, uint64_t seed,
                   int64_t attempts, int32_t *out_col, float *out_val);

#ifdef __cplusplus
}
#endif
#endif
