/*
 * pagani.h -- C ABI of libpagani.so, the MI355X (gfx950) implementation of the
 * PAGANI network-construction-and-metrics path.
 *
 * One context per host thread; a context owns a HIP stream and a device arena
 * that is reused across calls.  Every pointer argument may be host memory or
 * device memory (hipMalloc / torch CUDA tensors): the library detects which
 * and stages host buffers through the arena.  Functions never exit or prompt;
 * they return a pg_status and leave a message in pg_last_error().
 *
 * Each entry point names the reference interface it replaces.  Reference path
 * prefixes: U/ = src/BNAPlatform-win64-cuda7.0-20151118/src/,
 *           W/ = src/BNAPlatform-weighted-network-win64-cuda7.0-20151118/src/.
 */
#ifndef PAGANI_H
#define PAGANI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PAGANI_ABI_VERSION 1

typedef enum pg_status {
    PG_OK = 0,
    PG_EINVAL = 1,        /* bad argument */
    PG_ENOMEM = 2,        /* device allocation failed */
    PG_EDEVICE = 3,       /* HIP runtime error */
    PG_EFORMAT = 4,       /* malformed CSR (row[n] != nnz, unsorted, ...) */
    PG_ETHRESH = 5,       /* threshold out of range / selection mismatch */
    PG_ENOTSUP = 6,       /* size beyond what this build handles */
    PG_ESTATE = 7         /* call order (e.g. pg_corr_csr before pg_corr_build) */
} pg_status;

typedef struct pg_ctx pg_ctx;

/* Create a context on HIP device `device` (replaces the device pick of
 * U/BC/CUBC/CUBC.cpp:17-38 and findCudaDevice in W modularity_GPU.cu:409). */
pg_status pg_create(pg_ctx **ctx, int device);
void pg_destroy(pg_ctx *ctx);
const char *pg_last_error(const pg_ctx *ctx);
const char *pg_status_string(pg_status s);
int pg_abi_version(void);
/* Wait for all work queued by this context. */
pg_status pg_synchronize(pg_ctx *ctx);
/* Stage timing of pg_corr_build / pg_corr_emit / pg_corr_select (the ms_*
 * fields of pg_corr_info).  No counterpart in the reference: each timing
 * event leaves the device idle for a few microseconds, so callers that do not
 * read the stage times (batch pipelines, timing loops) can drop them.
 * level 2 (default): every stage; 1: only ms_gram_emit (the emitting Gram
 * kernel); 0: none.  Fields of untimed stages are 0. */
pg_status pg_set_timing(pg_ctx *ctx, int32_t level);

/* ------------------------------------------------------------ construction */

typedef enum pg_prec {
    PG_PREC_FP32 = 0,     /* FP32 Gram on MFMA (v_mfma_f32_32x32x2_f32) */
    PG_PREC_BF16X3 = 1    /* split z = hi + lo (bf16); hi.hi + hi.lo + lo.hi on MFMA */
} pg_prec;

typedef enum pg_thr_kind {
    PG_THR_R = 0,         /* keep r > thr (weighted: r > thr - 1e-6) */
    PG_THR_SPARSITY = 1   /* keep the k largest r, k from the sparsity percent */
} pg_thr_kind;

typedef enum pg_layout {
    PG_BOLD_TIME_MAJOR = 0,   /* bold[l*N + i]: the reference's BOLD gather order */
    PG_BOLD_VOXEL_MAJOR = 1   /* bold[i*T + l] */
} pg_layout;

typedef struct pg_corr_opts {
    int32_t thr_kind;     /* pg_thr_kind */
    float thr;            /* r threshold, or sparsity percent in (0, 100] */
    int32_t weighted;     /* 1: CSR values = r (float of the FP64 r) */
    int32_t prec;         /* pg_prec */
    int32_t layout;       /* pg_layout of the bold argument */
} pg_corr_opts;

typedef struct pg_corr_info {
    int32_t n;            /* voxels */
    int32_t t;            /* time points */
    int64_t nnz;          /* 2 x undirected edges = row[n] */
    float r_thresh;       /* r threshold as the reference names the file */
    float s_thresh;       /* sparsity percent as the reference names the file */
    double t_star;        /* FP64 k-th largest r (sparsity mode), else thr */
    double band;          /* proven bound on |r_fast - r_fp64| used for rechecks */
    int64_t band_pairs;   /* pairs recomputed in FP64 */
    int64_t band_flips;   /* rechecked pairs whose fast-path decision was wrong */
    int64_t gram_passes;  /* Gram sweeps run (1 r-mode, 2 sparsity, +1 on overflow) */
    double ms_zscore, ms_gram, ms_select, ms_csr;  /* device time per stage */
    double ms_gram_hist;  /* histogram Gram pass kernel alone (sparsity mode) */
    double ms_gram_emit;  /* emitting Gram pass kernel alone (last launch) */
    double band_max_err;  /* largest |r_fast - r_fp64| over the rechecked pairs; the
                             build fails (PG_EDEVICE) if it exceeds band / 1.5, the
                             bound before its safety factor */
} pg_corr_info;

/* Voxel time series -> thresholded CSR adjacency, on the device.
 * Replaces CorMat_gpu (U/CorMat/CUCorMat/CUCORMAT/CorMat.cu:18) together with
 * post_block (U CorMat.cpp:775), FormFullGraph (U CorMat.cpp:633,
 * W CorMat.cpp:647), FormFullGraph_s (U CorMat.cpp:693) and FormCSRGraph
 * (U CorMat.cpp:839, W CorMat.cpp:853).  The CSR stays in the context until
 * the next pg_corr_build; read it with pg_corr_csr or pg_corr_csr_view. */
pg_status pg_corr_build(pg_ctx *ctx, const float *bold, int32_t n, int32_t t,
                        const pg_corr_opts *opts, pg_corr_info *info);

/* Several thresholds from one Gram pass (U CorMat.cpp:300-346 computes r
 * once and applies every threshold of its command line).  pg_corr_emit runs
 * the z-score and ONE emitting Gram pass that keeps, with its fast r~, every
 * pair the loosest threshold opts->thr can keep (r-mode: the smallest r;
 * sparsity: the largest percent).  pg_corr_select then forms the CSR of one
 * threshold of the same kind at least as strict -- the FP64 recheck of the
 * kept pairs near it and the CSR build, no Gram pass -- with the same edge set
 * as pg_corr_build, and leaves it for pg_corr_csr / pg_corr_csr_view.  The
 * kept pairs stay until the next pg_corr_emit or pg_corr_build; a looser
 * threshold than the emitted one is PG_ETHRESH.  pg_corr_emit's info holds
 * the z-score and Gram figures and, in band_pairs, the number of kept pairs. */
pg_status pg_corr_emit(pg_ctx *ctx, const float *bold, int32_t n, int32_t t,
                       const pg_corr_opts *opts, pg_corr_info *info);
pg_status pg_corr_select(pg_ctx *ctx, float thr, pg_corr_info *info);

/* Copy the last CSR out: row[n+1], col[nnz], val[nnz] (val may be NULL; it is
 * only written for weighted builds).  Host or device destinations. */
pg_status pg_corr_csr(pg_ctx *ctx, int32_t *row, int32_t *col, float *val);

/* Device pointers of the last CSR, owned by the context (val NULL unless weighted). */
pg_status pg_corr_csr_view(pg_ctx *ctx, const int32_t **row, const int32_t **col,
                           const float **val);

/* FP64 z-scored series of the last build, voxel-major [n][t] (device pointer). */
pg_status pg_corr_z64_view(pg_ctx *ctx, const double **z64);

/* Dense correlation matrix of one or more subjects: subject averaging and the
 * .cormat output (U CorMat.cpp:236-262, 380-620, 775-838; W CorMat.cpp:386-600).
 * The matrix is the packed upper triangle, row-major over i < j, M1 = n(n-1)/2
 * floats, held by the context.  pg_cormat_begin zeroes it; pg_cormat_add adds
 * one subject's (float) r per pair (r = the FP64 Pearson r, clamped to [-1,1]),
 * or its Fisher z when fisher != 0 (U CorMat.cpp:35-42); every subject has the
 * same n, t may differ.  pg_cormat_finish divides by the subject count and,
 * for Fisher, maps back (U CorMat.cpp:507-517); a single subject without
 * Fisher is its own r matrix.  pg_cormat_read copies values [offset,
 * offset+count) out (host or device).  pg_cormat_threshold applies a threshold
 * to the finished matrix as the reference does for averaged networks
 * (FormFullGraph: a > thr, no 1e-6 offset; FormFullGraph_s: the k largest, ties
 * at the k-th value in row-major order) and leaves the CSR for pg_corr_csr /
 * pg_corr_csr_view (values = matrix entries when weighted). */
pg_status pg_cormat_begin(pg_ctx *ctx, int32_t n, int32_t fisher);
pg_status pg_cormat_add(pg_ctx *ctx, const float *bold, int32_t t, int32_t layout);
pg_status pg_cormat_finish(pg_ctx *ctx);
pg_status pg_cormat_read(pg_ctx *ctx, int64_t offset, int64_t count, float *dst);
pg_status pg_cormat_threshold(pg_ctx *ctx, int32_t thr_kind, float thr, int32_t weighted,
                              pg_corr_info *info);

/* ------------------------------------------------------------------ cohort */

/* Edges (undirected) the sparsity rule keeps for n voxels at `sparsity`
 * percent, with the reference's float arithmetic (U CorMat.cpp:311-313);
 * a subject's CSR has 2k column entries. */
int64_t pg_sparsity_pairs(int32_t n, float sparsity);

/* Bits of pg_subject.status: why a subject was rebuilt outside the graph. */
#define PG_COHORT_WINDOW 1u       /* selection window below the kept pairs */
#define PG_COHORT_SEGMENTS 2u     /* an edge segment overflowed */
#define PG_COHORT_CANDIDATES 4u   /* the candidate list overflowed */
#define PG_COHORT_SELECTION 8u    /* the exact selection did not hold */
#define PG_COHORT_CSR 16u         /* row[n] != 2k */
#define PG_COHORT_DEPTH 32u       /* a BFS deeper than the captured levels */
#define PG_COHORT_OUTSIDE 64u     /* too small for the single pass (n, k or supertiles) or T > 360 */

typedef struct pg_subject {
    const float *bold;       /* t x n FP32 series, time-major (bold[l*n + i]), device memory */
    int32_t n, t;
    float sparsity;          /* percent in (0, 100] */
    const int32_t *sources;  /* BC sources (device or host), nsrc of them; ignored when bc is NULL */
    int64_t nsrc;
    int32_t *row;            /* outputs, device memory: row[n + 1] */
    int32_t *col;            /* col[2 * pg_sparsity_pairs(n, sparsity)] */
    int32_t *deg;            /* deg[n], or NULL */
    float *cp;               /* cp[n] (binary clustering), or NULL */
    double *bc;              /* bc[n] (BC from the sources), or NULL */
    uint32_t status;         /* out: 0 = built in the graph, else PG_COHORT_* bits of the
                                checks that sent it through the ordinary entry points */
    float ms;                /* out: device time of the subject inside the last graph launch */
} pg_subject;

typedef struct pg_cohort_info {
    int32_t graph_nodes;     /* nodes of the captured graphs */
    int32_t graphs;          /* captured graphs (consecutive subjects, about 2048 nodes each at most) */
    int32_t reruns;          /* subjects rebuilt outside the graph */
    double ms_build;         /* host time to plan, capture and instantiate */
    double ms_graph;         /* device time of the last graph launch */
    double ms_reruns;        /* host time of the rebuilds */
} pg_cohort_info;

/* A cohort of subjects of uneven n and t in captured HIP graphs: per subject
 * the BF16x3 sparsity construction (the single pass of pg_corr_build:
 * z-score, sampled histogram, emitting Gram pass, window, FP64 recheck, exact
 * selection, CSR) and then degree, binary clustering and BC from its sources,
 * all enqueued without host round trips on workspaces sized for the largest
 * subject.  Consecutive subjects are captured into one graph until it holds
 * about 2048 kernel nodes (PAGANI_COHORT_GRAPH_NODES), so a cohort is one
 * graph or a few, launched back to back `launches` times (>= 1; every launch
 * recomputes the same results).  Each subject's device-side checks are read
 * after the last launch; a subject whose single-pass assumptions did not hold
 * (or whose BFS went deeper than the captured levels) is rebuilt with
 * pg_corr_build, pg_degree, pg_cp_binary and pg_bc, so the outputs are always
 * those of the ordinary entry points.  The reference runs subjects one after
 * another in a host loop (U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:190-378,
 * then one tool per metric); there is no single reference function. */
pg_status pg_cohort_run(pg_ctx *ctx, pg_subject *subjects, int32_t nsub, int32_t launches,
                        pg_cohort_info *info);

/* ----------------------------------------------------------------- metrics */
/* CSR arguments: row[n+1] (row[0] = 0, row[n] = nnz), col[nnz] ascending within
 * each row, symmetric, no self loops; val[nnz] FP32 weights where used. */

/* Degree (U/Degree/Degree/Degree.cpp:84-91).  deg: exact int32 (may be NULL);
 * deg_ref: the float the reference writes, (float)R[k+1] - R[k] evaluated in
 * float (may be NULL). */
pg_status pg_degree(pg_ctx *ctx, const int32_t *row, int32_t n, int32_t *deg, float *deg_ref);

/* Strength: FP32 running sum of each CSR row (W/Degree/Degree/Degree.cpp:84-88). */
pg_status pg_strength(pg_ctx *ctx, const int32_t *row, const float *val, int32_t n, float *s);

/* Triangles through each vertex by sorted-adjacency intersection (the count
 * U/Cp/Cp/Cp.cpp:102-108 forms with a dense lookup). */
pg_status pg_triangles(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                       int64_t nnz, int64_t *tri);

/* Binary clustering coefficient, Cp = (double)T/k/(k-1)*2 as float
 * (replaces Cp, U/Cp/Cp/Cp.cpp:21-115).  mean: FP64 mean of cp (may be NULL). */
pg_status pg_cp_binary(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                       int64_t nnz, float *cp, double *mean);

/* pg_cp_binary without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, fixed-memory serving loops).  row, col and cp must be device
 * memory (PG_EINVAL otherwise: staging host buffers would allocate); mean may
 * be host or device memory or NULL.  n <= PG_CP_STATIC_MAX_N (PG_ENOTSUP
 * above it).  The per-vertex triangle counts live in one fixed array per GPU
 * of 8 * PG_CP_STATIC_MAX_N bytes (3.9 MB), held while the library is loaded
 * and shared by every call on that GPU: calls that overlap on one GPU (two
 * contexts, two threads) race. */
#define PG_CP_STATIC_MAX_N 491520
pg_status pg_cp_binary_static(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                              int64_t nnz, float *cp, double *mean);

typedef enum pg_cp_kind { PG_CP_BARRAT = 1, PG_CP_ONNELA = 2 } pg_cp_kind;

/* Weighted clustering in FP64: Barrat (W/Cp/Cp/Cp_1.cpp:24-131) or Onnela
 * (W/Cp/Cp/Cp_2.cpp:23-129, W/Cp_GPU/CUCP/Cp.cu:111).  wmax_mode 0 normalises
 * Onnela weights by the largest weight, 1 by the largest of val[0..n-1]
 * (the reference CPU tool).  mean may be NULL. */
pg_status pg_cp_weighted(pg_ctx *ctx, const int32_t *row, const int32_t *col,
                         const float *val, int32_t n, int64_t nnz, int32_t kind,
                         int32_t wmax_mode, double *cp, double *mean);

/* Brandes betweenness, exact: uint64 path counts, FP64 dependencies, summed
 * over ordered pairs and unnormalised like the reference.  sources == NULL
 * means every vertex.  Replaces CUBC (U/BC/CUBC/CUBC.cpp:13) and
 * Betweenness_GPU_node_pred (U/BC/CUBC/BC_GPU.h:9). */
pg_status pg_bc(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                int64_t nnz, const int32_t *sources, int64_t nsrc, double *bc);

/* pg_bc with the reference's output type: the FP64 sums rounded to float,
 * for callers of CUBC(float *BC, int *row, int *col, int N, int nnz)
 * (U/BC/CUBC/CUBC.cpp:13), which fills N floats. */
pg_status pg_bc_float(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                      int64_t nnz, const int32_t *sources, int64_t nsrc, float *bc);

/* BFS hop levels for nsrc sources: level[s*n + v], -1 where unreachable
 * (the BFS of U/BFS_MulCPU/BFS_MulCPU/APSP_BFS.cpp:150-192). */
pg_status pg_bfs(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                 int64_t nnz, const int32_t *sources, int64_t nsrc, int32_t *level);

/* Binary nodal efficiency and characteristic path length from all-pairs BFS
 * hop counts (U/BFS_MulCPU/BFS_MulCPU/APSP_BFS.cpp:26-192, replaces APSP_BFS):
 * e_i = (FP32 sum over j != i, in increasing j, of 1.0f / d_ij) / (n - 1),
 * unreachable pairs skipped; Lp = 1 / (S / n / (n - 1)) / 2 with S the FP64
 * sum of 1.0f / d_ij over reachable i < j.  The graph must be undirected
 * (symmetric CSR).  isolated counts e_i == 0; unconnected is 1 when some pair
 * is unreachable.  eff, lp, isolated, unconnected may be NULL. */
pg_status pg_bfs_efficiency(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                            int64_t nnz, float *eff, double *lp, int32_t *isolated,
                            int32_t *unconnected);

/* Blocked Floyd-Warshall closure of a dense row-major n x n FP32 distance
 * matrix, in place (+inf = no edge).  Replaces cuAPSP (W/Lp/CUBFW_Lp/CUBFW_Lp/
 * CUBFW.cpp:19) and the tile kernels of CUBFW_ph1/2/3.cu.  block: tile edge
 * 64 or 128 (0 = default 64); results are bit-identical to the blocked algorithm of
 * W BFW.cpp:40-108 run with the same tile edge. */
pg_status pg_fw(pg_ctx *ctx, float *d, int32_t n, int32_t block);

/* Dense distance matrix from a CSR (W/Lp/CUBFW_Lp/CUBFW_Lp/FormBlock.cpp:45-75):
 * kind 0: 1/w (the reference), 1: 1 - w, 2: 1 (hops); +inf where no edge, 0 on
 * the diagonal.  d is n x n row-major. */
pg_status pg_fw_init(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val,
                     int32_t n, int64_t nnz, int32_t kind, float *d);

/* Nodal efficiency and characteristic path length from a closed distance
 * matrix (W/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp.cpp:70-113): for u != v with
 * d_uv >= 0.1, e_u = sum of 1/d_uv (FP32) / (n-1), set to 0 when below 1/n;
 * Lp = n (n-1) / sum of 1/d_uv (FP64 sum).  eff (n floats) and lp may be NULL;
 * isolated (may be NULL) counts the zeroed e_u. */
pg_status pg_fw_efficiency(pg_ctx *ctx, const float *d, int32_t n, float *eff, double *lp,
                           int32_t *isolated);

typedef enum pg_mod_kmode {
    PG_MOD_STRENGTH = 0,  /* k = strength, 2m = sum of weights (binary if val NULL) */
    PG_MOD_BINARY_K = 1   /* k = degree, 2m = nnz, A weighted: the reference's tool */
} pg_mod_kmode;

/* Newman leading-eigenvector modularity by recursive bisection (replaces
 * Partition_GPU, W/Newman_Modularity_GPU/CUmodularity/modularity_GPU.cu:405).
 * labels: 1..K.  q, lambda0 (round-0 leading eigenvalue), eig0 (round-0
 * vector, n doubles) and rounds may be NULL. */
pg_status pg_modularity(pg_ctx *ctx, const int32_t *row, const int32_t *col,
                        const float *val, int32_t n, int64_t nnz, int32_t kmode,
                        int32_t *labels, double *q, double *lambda0, double *eig0,
                        int32_t *rounds);

typedef enum pg_mod_stop {
    PG_MOD_STOP_VECTOR = 0,   /* the reference: |v - v0| or |v + v0| <= 1e-6 */
    PG_MOD_STOP_RAYLEIGH = 1  /* relative change of the Rayleigh quotient <= tol */
} pg_mod_stop;

typedef struct pg_mod_opts {
    int32_t kmode;        /* pg_mod_kmode */
    int32_t stop;         /* pg_mod_stop */
    double tol;           /* PG_MOD_STOP_RAYLEIGH: |rho_k - rho_(k-1)| <= tol |rho_k|
                             (0 means 1e-8) */
    int32_t max_iter;     /* power iterations per pass (0 means 10000, the reference's cap) */
    int32_t progress;     /* 1: one line per bisection level on stderr */
    int32_t max_depth;    /* 0: bisect until no group splits; d > 0: groups d
                             bisections deep are final (labels in round order) */
} pg_mod_opts;

/* pg_modularity with a choice of stopping rule.  With PG_MOD_STOP_RAYLEIGH the
 * quotient rho_k = x.(B + beta I)x / x.x of each iteration (FP64) is tracked
 * and a group's eigenvalue is the final quotient instead of +-|(B + beta I)x|;
 * the bisection (shift rule, split rule, label order) is the reference's. */
pg_status pg_modularity_ex(pg_ctx *ctx, const int32_t *row, const int32_t *col,
                           const float *val, int32_t n, int64_t nnz, const pg_mod_opts *opts,
                           int32_t *labels, double *q, double *lambda0, double *eig0,
                           int32_t *rounds);

/* Eigenvector centrality by the power iteration of CUEC (U/CUEC/CUEC/
 * CUEC_unweighted.cu:61-196; weighted: W/CUEC/weighted-net.cu:55-203): start
 * vector 0..n-1; per iteration x = u / |u|, u = A x; stop when |u - u_prev| or
 * |u + u_prev| <= 1e-6 (u_prev before normalising) or after 10000 iterations.
 * A is weighted by val (FP32 weights, FP64 arithmetic), binary when val is
 * NULL.  vhat = x of the last iteration (what the unweighted tool writes),
 * avec = A x of the last iteration (what the weighted tool writes), lambda =
 * |u| of the last iteration (the printed eigenvalue).  Any output may be NULL. */
pg_status pg_eigenvector_centrality(pg_ctx *ctx, const int32_t *row, const int32_t *col,
                                    const float *val, int32_t n, int64_t nnz, double *vhat,
                                    double *avec, double *lambda, int32_t *iters);

/* Participation coefficient (W/PC_CPU/PC_CPU/main.cpp:96-160): with s_i the
 * FP32 strength (CSR order) and s_ij the FP32 weight from i into module j,
 * pc_i = 1 - sum over j = 1..K of (s_ij / s_i)^2 (0 when s_i == 0), pc_norm =
 * pc / ((K-1)/K) (pc when K == 1); K = largest label.  labels (n, >= 1) are the
 * module of each vertex, e.g. from pg_modularity; val NULL means unit weights.
 * pc_norm and nmod may be NULL. */
pg_status pg_participation(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val,
                           int32_t n, int64_t nnz, const int32_t *labels, float *pc, float *pc_norm,
                           int32_t *nmod);

/* Degree-preserving random rewiring by double edge swaps, the random
 * reference network of Cp (gamma), CUBFS_Lp / CUBFW_Lp (lambda) and the
 * Maslov tool: replaces Maslov (U/BFS_MulCPU/BFS_MulCPU/Maslov.cpp:64) and
 * Maslov_weighted_1 (W/Maslov/Maslov/Maslov.cpp:106).  The input is a
 * symmetric CSR without self-loops, columns strictly ascending in each row
 * (PG_EFORMAT otherwise).  attempts swap attempts are made (0: 2 * nnz/2, the
 * reference's count); an attempt draws two edges, gives up when they share a
 * vertex, picks 1-3,2-4 or 1-4,3-2 by a fair coin and gives up when either new
 * edge exists.  The row offsets are unchanged (every vertex keeps its degree),
 * so only out_col (nnz) is written, rows sorted by column; with val and out_val
 * both non-NULL each edge keeps its weight in out_val (nnz).  The reference
 * seeds with the wall clock; here the result is a function of seed alone.
 * swaps (may be NULL) receives the number of swaps made. */
pg_status pg_rewire(pg_ctx *ctx, const int32_t *row, const int32_t *col, const float *val, int32_t n,
                    int64_t nnz, uint64_t seed, int64_t attempts, int32_t *out_col, float *out_val,
                    int64_t *swaps);

/* ---------------------------------------------------------------- autotuner */

typedef struct pg_tune_result {
    int32_t value;        /* the chosen launch shape */
    int32_t candidates;   /* shapes timed now (0: the cached value was used) */
    double ms;            /* time of the chosen shape */
    double ms_default;    /* time of the default shape (when timed) */
} pg_tune_result;

/* Autotuner.  The reference hard-codes its launch shapes (Batch_size =
 * 1024*3, U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:238; the BC grid from free
 * memory, U/BC/CUBC/CUBC.cpp:45-64; sizeBlock = 96*32,
 * W/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp.cpp:15).  Each call below times the
 * candidate shapes on the caller's input through the ordinary entry point and
 * keeps the fastest in a cache keyed by the device architecture and the
 * half-octave size buckets (file: $PAGANI_TUNE_CACHE, else
 * ~/.cache/pagani/tune.txt; PAGANI_TUNE=off ignores it, =refresh searches
 * again); later calls on inputs of the same bucket use it.  A cached value is
 * returned without a search.  With PAGANI_AUTOTUNE=1, pg_corr_build, pg_bc
 * and pg_fw (block 0) search by themselves on first use.  Results never
 * depend on the shape: every shape gives the same edge set, levels and
 * distances, and BC to FP64 summation order.
 *  - pg_autotune_corr: the supertile edge of the Gram walk (8, 4, 16 tiles),
 *    timed on the emitting Gram pass of pg_corr_build(opts); key (n, t).
 *  - pg_autotune_bc: 64-source BFS batches per launch (1..16) of pg_bc and
 *    pg_bfs, timed on pg_bc over the given sources; key (n, nnz).
 *  - pg_autotune_fw: the Floyd-Warshall tile (64, 128), timed on pg_fw over a
 *    copy of d (d is not changed); key n. */
pg_status pg_autotune_corr(pg_ctx *ctx, const float *bold, int32_t n, int32_t t,
                           const pg_corr_opts *opts, pg_tune_result *res);
pg_status pg_autotune_bc(pg_ctx *ctx, const int32_t *row, const int32_t *col, int32_t n,
                         int64_t nnz, const int32_t *sources, int64_t nsrc, pg_tune_result *res);
pg_status pg_autotune_fw(pg_ctx *ctx, const float *d, int32_t n, pg_tune_result *res);

#ifdef __cplusplus
}
#endif
#endif /* PAGANI_H */
