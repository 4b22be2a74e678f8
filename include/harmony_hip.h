/*
 * libharmony_hip -- C ABI of the MI355X (gfx950) Harmony integration path.
 *
 * Every entry point replaces a native call that rapids_singlecell's Python
 * driver makes through nanobind; the replaced interface is cited above each
 * declaration (paths relative to src/rapids_singlecell/ in the reference).
 *
 * Conventions
 *   - Plain pointers and sizes only.  Array pointers are device pointers
 *     (hipMalloc / torch-ROCm memory), C-contiguous, row-major.  Pointers typed
 *     `void*` hold float or double data as selected by `dtype`.
 *   - The caller owns all memory, including the workspace whose size the
 *     *_workspace_bytes queries return.  Nothing is freed by the library.
 *   - `stream` is a hipStream_t (NULL = default stream).  Work is enqueued on
 *     it; entry points that must report a host value synchronise it.
 *   - Return value: HZ_OK or a negative HZ_ERR_* code; hz_last_error()
 *     describes the last failure of the calling thread.
 *   - Calls that share a workspace must not overlap.
 */
#ifndef HARMONY_HIP_H
#define HARMONY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HZ_OK 0
#define HZ_ERR_ARG (-1)
#define HZ_ERR_HIP (-2)
#define HZ_ERR_LAUNCH (-3)
#define HZ_ERR_UNSUPPORTED (-4)
#define HZ_ERR_WORKSPACE (-5)

/* element size selects the arithmetic type */
#define HZ_F32 4
#define HZ_F64 8

#define HZ_PREC_NATIVE 0     /* GEMMs in the storage type */
#define HZ_PREC_SPLIT_BF16 1 /* FP32 storage, split-BF16 matrix-core products */

#define HZ_CORR_BATCHED 0
#define HZ_CORR_FAST 1
#define HZ_CORR_ORIGINAL 2

/* Problem shape shared by the job-level entry points.
 *
 * One covariate (the reference's case): n_batches levels, cats[i] is the
 * level of cell i, and O, E, Pr_b, theta, N_b and lambda have n_batches rows.
 *
 * Several covariates (stacked one-hot Phi; no counterpart in the reference,
 * which joins several keys into one categorical): every cell has one level
 * of each of n_covariates covariates.  Cells with the same levels form a
 * group; n_batches is then the number of groups, cats[i] the group of cell i
 * (cat_offsets / cell_indices group the cells by group), and group_levels
 * (device, n_batches x n_covariates int32) holds each group's level ids in
 * 0 .. n_levels-1 (covariate c's levels numbered after those of covariates
 * 0 .. c-1; n_levels0 = levels of covariate 0).  O, E, Pr_b, theta, N_b and
 * lambda then have n_levels rows. */
typedef struct hz_dims {
  int n_cells;
  int n_pcs;
  int n_clusters;
  int n_batches;
  int block_size; /* int(n_cells * block_proportion), >= 1 */
  int dtype;      /* HZ_F32 or HZ_F64 */
  int precision;  /* HZ_PREC_* */
  int n_covariates;              /* 0 or 1: one covariate; >= 2: stacked one-hot */
  int n_levels;                  /* n_covariates >= 2: total levels */
  int n_levels0;                 /* n_covariates >= 2: levels of covariate 0 */
  const int32_t* group_levels;   /* n_covariates >= 2: n_batches x n_covariates */
} hz_dims;

int hz_version(void);
const char* hz_last_error(void);

/* Releases what the library created for the calling thread: the side stream
 * (and its events) that builds block orders beside the clustering loop, and
 * the pinned ring of convergence flags, on every device the thread used.  It
 * waits for their pending work first.  Call it before a thread that ran
 * clustering loops exits, and once on the main thread before the process
 * exits; the next call on the thread re-creates them.  No counterpart in the
 * reference (its driver keeps no per-thread native state). */
int hz_release_thread(void);

/* Opt-in HIP-event timing of the library's own launches (benchmark
 * instrumentation; no counterpart in the reference).  While enabled, each
 * timed launch is bracketed by events on its stream; hz_profile_collect waits
 * for them, writes the summed milliseconds and launch counts per kind
 * (HZ_PROF_*) and resets the counters. */
#define HZ_PROF_ROW 0        /* row kernel, one launch per block */
#define HZ_PROF_CENTROID 1   /* centroid pass */
#define HZ_PROF_YREDUCE 2    /* centroid reduction + block-0 setup */
#define HZ_PROF_ORDER 3      /* block order pipeline (several launches) */
#define HZ_PROF_CORRECTION 4 /* whole correction */
#define HZ_PROF_PHI 5        /* correction: Phi pass */
#define HZ_PROF_APPLY 6      /* correction: apply + renormalise */
#define HZ_PROF_SWEEP 7      /* all blocks of one clustering iteration, one launch */
#define HZ_PROF_KINDS 8
int hz_profile_enable(int on);
int hz_profile_collect(double* ms_out, long long* launches_out, int n_kinds);

/* Process-wide tuning knobs (0 = automatic choice), read when a workspace
 * layout is computed, so a workspace must be sized and used under the same
 * settings.  No counterpart in the reference (its closest analogue is the
 * colsum_algo selection, _harmony/_helper.py:401-504). */
#define HZ_TUNE_ROW_CHUNK 0   /* rows per chunk of the row kernel (multiple of 16) */
#define HZ_TUNE_CENTROID_WG 1 /* workgroups of the centroid pass */
#define HZ_TUNE_SWEEP 2       /* 0 = block sweep in one launch when it fits, 1 = one launch per block */
#define HZ_TUNE_SWEEP_WG 3    /* workgroups of the block sweep (capped at what is co-resident) */
#define HZ_TUNE_SWEEP_PLAIN 4 /* 1 = launch the block sweep as a plain dispatch of the same,
                                 co-resident size (for profilers that mishandle cooperative
                                 launches); 0 = cooperative launch */
#define HZ_TUNE_SPIN_LIMIT 5 /* polls a block-sweep workgroup makes while waiting for the
                                 others' block sums before it gives up and the call returns
                                 HZ_ERR_LAUNCH (0 = 65536, about 65 ms) */
#define HZ_TUNE_BLOCK_UPDATE 6 /* per-block path, update between blocks: 0 = automatic (a separate
                                  launch above 512 levels x clusters), 1 = in the row launch's
                                  last workgroup, 2 = a separate launch spread over the clusters
                                  (identical results) */
#define HZ_TUNE_KEYS 8
int hz_set_tuning(int key, int value);
int hz_get_tuning(int key);

/* Workspace for hz_init_clusters / hz_clustering_loop / hz_correction.
 * Replaces _allocate_clustering_workspace (_harmony/__init__.py:356-384),
 * get_cub_sort_temp_bytes (_cuda/harmony/clustering/clustering.cu:48-59) and
 * the per-call workspace of _correction_batched (_harmony/__init__.py:648-656).
 * Returns 0 for an unsupported shape (see hz_last_error). */
size_t hz_workspace_bytes(const hz_dims* dims);

/* L2 row normalisation, dst = src * min(rsqrt(sum src^2), 1e12).
 * Replaces _harmony_normalize_cuda.l2_row_normalize
 * (_cuda/harmony/normalize/normalize.cu:35-68, kernels_normalize.cuh:6-54). */
int hz_l2_normalize(int dtype, const void* src, void* dst, long long n_rows, int n_cols,
                    void* stream);

/* Key-sorted cell order of one clustering iteration: order[r] = cell of rank r
 * for keys pcg(i ^ seed).  Replaces pcg_hash_kernel + iota_kernel + CUB
 * SortPairs (_cuda/harmony/clustering/clustering.cu:292-305,
 * kernels_clustering.cuh:106-126). */
size_t hz_block_order_workspace_bytes(int n_cells);
int hz_block_order(int n_cells, uint32_t seed, int32_t* order, void* ws, size_t ws_bytes,
                   void* stream);

/* hz_block_order without a workspace argument and without any allocation:
 * the scratch is a static device array of 32 MiB per GPU (the library's own,
 * held until the process ends), so n_cells is limited to what
 * hz_block_order_workspace_bytes fits in 32 MiB (about 1.6 million cells);
 * larger inputs return HZ_ERR_WORKSPACE before any launch.  Every call on one
 * GPU shares that scratch: calls that overlap on one GPU (other threads or
 * streams) race, so the caller serialises them.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). */
int hz_block_order_noalloc(int n_cells, uint32_t seed, int32_t* order, void* stream);

/* Block membership of n_iter consecutive clustering iterations (seeds seed ..
 * seed+n_iter-1), built by the same kernels and in the same batched launches
 * as hz_clustering_loop builds them: block_of_cell[y * n_cells + i] =
 * rank(pcg(i ^ (seed + y))) / block_size, i.e. the block of cell i in the
 * key-sorted order; order_out (optional, n_iter x n_cells) receives the
 * grouped order (block-major, then batch).  Synchronises the stream.  Returns
 * HZ_ERR_UNSUPPORTED if a block-boundary bucket overflowed its key list.
 * Test binding of the membership the clustering loop uses in place of the
 * reference's full sort + block slicing (_cuda/harmony/clustering/
 * clustering.cu:292-319: idx_list_alt[j*block_size .. (j+1)*block_size)). */
size_t hz_block_membership_workspace_bytes(int n_cells, int block_size, int n_batches,
                                           int n_iter);
int hz_block_membership(int n_cells, int block_size, int n_batches, const int32_t* cats,
                        uint32_t seed, int n_iter, int32_t* block_of_cell, int32_t* order_out,
                        void* ws, size_t ws_bytes, void* stream);

/* Initial R, E, O and objective from centroids Y0 (n_clusters x n_pcs).
 * Replaces the body of _initialize_centroids after k-means
 * (_harmony/__init__.py:323-351): _calc_R (_fuses.py:6-8), L1 normalise,
 * _outer_cp, _scatter_add_cp and compute_objective.  obj_out: device scalar. */
int hz_init_clusters(const hz_dims* dims, const void* Z_norm, const void* Y0,
                     const int32_t* cats, const int32_t* cat_offsets,
                     const int32_t* cell_indices, const void* Pr_b, const void* theta,
                     double sigma, int stabilized, void* R, void* E, void* O, void* obj_out,
                     void* ws, size_t ws_bytes, void* stream);

/* One Harmony iteration's clustering step; R, E, O are updated in place and
 * the final objective is written to the device scalar last_obj; *n_iter
 * receives the number of clustering iterations executed.
 * Replaces _harmony_clustering_cuda.clustering_loop
 * (_cuda/harmony/clustering/clustering.cu:241-393, binding :397-469). */
int hz_clustering_loop(const hz_dims* dims, const void* Z_norm, void* R, void* E, void* O,
                       const void* Pr_b, const int32_t* cats, const void* theta, double sigma,
                       double tol, int max_iter, uint32_t seed, int stabilized, void* last_obj,
                       int* n_iter, void* ws, size_t ws_bytes, void* stream);

/* One job of hz_clustering_loop_batched: the arguments of hz_clustering_loop
 * for one Harmony instance (its own shape, state and workspace). */
typedef struct hz_job {
  hz_dims dims;
  const void* Z_norm;
  void* R;
  void* E;
  void* O;
  const void* Pr_b;
  const int32_t* cats;
  const void* theta;
  double sigma;
  double tol;
  int max_iter;
  uint32_t seed;
  int stabilized;
  void* last_obj; /* device scalar */
  int n_iter;     /* out: clustering iterations executed */
  void* ws;
  size_t ws_bytes;
} hz_job;

/* Clustering step of several independent Harmony instances at once: per
 * clustering iteration, every job's centroid pass and block orders, then ONE
 * cooperative launch in which each job's block sweep runs on its own share
 * of the device's workgroups (shares proportional to the jobs' cell counts),
 * so that the waits of one job's block exchange are filled by the other
 * jobs' rows.  Each job's R, E, O, last_obj and n_iter are exactly those of
 * hz_clustering_loop on that job alone (fixed-point sums: the result does not
 * depend on the share).  All jobs must be float32 with one covariate, run
 * the block sweep (block_size >= n_cells / 64, n_clusters <= 256), share the
 * kernel shape (precision, ceil(n_clusters / 16) class and n_pcs class) and
 * max_iter; otherwise HZ_ERR_UNSUPPORTED and nothing is launched.  bws is a
 * device workspace of hz_batched_workspace_bytes bytes (per-iteration sweep
 * arguments).  No counterpart in the reference (it runs one instance per
 * call); the single-job interface it generalises is
 * _cuda/harmony/clustering/clustering.cu:397-469. */
size_t hz_batched_workspace_bytes(int n_jobs, int max_iter);
int hz_clustering_loop_batched(hz_job* jobs, int n_jobs, void* bws, size_t bws_bytes,
                               void* stream);
/* *eligible = 1 when a job of these dimensions can run in
 * hz_clustering_loop_batched on the current device (float32, one covariate,
 * at most 256 clusters and 64 blocks, and a block sweep whose LDS fits), else
 * 0.  Jobs that are not eligible run through hz_clustering_loop.  No reference
 * counterpart (the reference integrates one dataset per call). */
int hz_batched_eligible(const hz_dims* dims, int* eligible);

/* Ridge correction from the ORIGINAL X: lambda_kb (written to lambda_kb_out,
 * n_batches x n_clusters), Z_hat and its L2-normalised rows Z_norm.
 * Replaces _compute_lambda_kb (_harmony/__init__.py:454-480) and
 * _harmony_correction_batched_cuda.correction_batched
 * (_cuda/harmony/correction/correction_batched.cu:18-130, binding :135-164),
 * _harmony_correction_cuda.correction_fast (correction_fast.cu:23-109) and
 * _correction_original (_harmony/__init__.py:532-570), plus the following
 * _normalize_cp(Z_hat) (_harmony/__init__.py:276). */
int hz_correction(const hz_dims* dims, const void* X, const void* R, const void* O,
                  const void* E, const void* N_b, const int32_t* cat_offsets,
                  const int32_t* cell_indices, int method, int dynamic_lambda, double alpha,
                  int has_threshold, double threshold, double ridge_lambda,
                  void* lambda_kb_out, void* Z_hat, void* Z_norm, void* ws, size_t ws_bytes,
                  void* stream);

/* Three-term objective of given R and similarities; result copied to *obj_host.
 * Replaces _harmony_clustering_cuda.compute_objective
 * (_cuda/harmony/clustering/clustering.cu:188-237, binding :472-489). */
size_t hz_objective_workspace_bytes(void);
int hz_objective(int dtype, const void* R, const void* sim, const void* O, const void* E,
                 const void* theta, double sigma, int n_cells, int n_clusters, int n_batches,
                 int stabilized, double* obj_host, void* ws, size_t ws_bytes, void* stream);

/* Penalty ((E+1)/(O+E+1))^theta (stabilized) or ((E+1)/(O+1))^theta.
 * Replaces _harmony_pen_cuda.penalty (_cuda/harmony/pen/pen.cu:13-25,
 * kernels_pen.cuh:8-20). */
int hz_penalty(int dtype, const void* E, const void* O, const void* theta, void* pen,
               int n_batches, int n_clusters, int stabilized, void* stream);

/* R_out[r] = L1(exp(term (1 - sim[idx[r]])) * pen[cats[r]]).
 * Replaces the test binding _harmony_pen_cuda.fused_pen_norm_int
 * (_cuda/harmony/pen/pen.cu:102-106, kernels_pen.cuh:27-82). */
int hz_fused_pen_norm(int dtype, const void* sim, const void* pen, const int32_t* cats,
                      const int32_t* idx, void* R_out, double term, int n_rows, int n_cols,
                      void* stream);

/* lambda_kb (n_batches x n_clusters). Replaces _compute_lambda_kb
 * (_harmony/__init__.py:454-480). */
int hz_lambda_kb(int dtype, const void* E, const void* O, const void* N_b, int n_batches,
                 int n_clusters, int dynamic_lambda, double alpha, int has_threshold,
                 double threshold, double ridge_lambda, void* out, void* stream);

/* Closed-form inverses, n_clusters x (n_batches+1)^2.  Replaces the test
 * binding _harmony_correction_cuda.compute_inv_mat
 * (_cuda/harmony/correction/correction_fast.cu:142-159,
 * kernels_correction_fast.cuh:24-123). */
int hz_inv_mats(int dtype, const void* O, const void* lambda_kb, void* inv, int n_batches,
                int n_clusters, void* stream);

/* Seeded Lloyd k-means on Z_norm producing the initial centroids Y_out
 * (k x d).  Replaces cuML KMeans(init="k-means||", n_init=1, max_iter=25)
 * (_harmony/__init__.py:315-323) with this library's own initialiser: seeds
 * are the first k cells of the pcg order for `seed`. */
size_t hz_kmeans_workspace_bytes(int dtype, long long n_cells, int n_pcs, int k);
int hz_kmeans(int dtype, const void* Z_norm, long long n_cells, int n_pcs, int k, uint32_t seed,
              int max_iter, void* Y_out, void* ws, size_t ws_bytes, void* stream);

/* Exact k nearest neighbours (brute force) of every row of X (n_rows x
 * n_cols, float32, device) among all rows of X, nearest first: indices and
 * distances are n_rows x k (the row itself is among them).  Metrics:
 * HZ_KNN_EUCLIDEAN, HZ_KNN_SQEUCLIDEAN (expanded form |x|^2 + |y|^2 - 2 x.y,
 * clamped at 0) and HZ_KNN_COSINE (1 - x.y of unit rows).  k <= 32,
 * n_cols <= 128; neighbours are ordered by (distance, index).  Replaces
 * cuML NearestNeighbors(algorithm="brute").kneighbors
 * (preprocessing/_neighbors/_algorithms/_brute.py:14-38), called by
 * pp.neighbors (preprocessing/_neighbors/__init__.py:200-207). */
#define HZ_KNN_EUCLIDEAN 0
#define HZ_KNN_SQEUCLIDEAN 1
#define HZ_KNN_COSINE 2
size_t hz_knn_workspace_bytes(long long n_rows, int n_cols, int metric);
int hz_knn(int metric, const float* X, long long n_rows, int n_cols, int k, int32_t* indices,
           float* distances, void* ws, size_t ws_bytes, void* stream);

/* UMAP fuzzy-set membership weights of a kNN graph (n_rows x k, nearest
 * first, as hz_knn returns it): per row the smooth-kNN normalisation (rho =
 * nearest non-zero distance, sigma by bisection so that the memberships of
 * neighbours 1..k-1 sum to log2(k), floored at 1e-3 of the row's mean
 * distance, or of mean_distance -- the mean of all n_rows x k distances --
 * when rho = 0), then w = exp(-max(0, d - rho) / sigma), 0 for the row
 * itself.  sigmas and rhos (n_rows each) are optional outputs.  Replaces the
 * membership half of cuML's fuzzy_simplicial_set as called by
 * _get_connectivities_umap (preprocessing/_neighbors/_neighbors.py:134-159);
 * the symmetrisation w + w^T - w*w^T is done by the caller on the sparse
 * graph. */
int hz_umap_weights(const float* distances, const int32_t* indices, long long n_rows, int k,
                    double mean_distance, float* weights, float* sigmas, float* rhos,
                    void* stream);

/* Dense PCA with zero centring, the producer of the Harmony input
 * (obsm["X_pca"]).  hz_pca_covariance: column means (mean, n_cols, FP64) and
 * the covariance of the centred X (cov, n_cols x n_cols, FP64, divided by
 * n_rows - 1) of a float32 X (n_rows x n_cols); hz_pca_project: X_pca =
 * (X - mean) components (n_cols x n_comps float32, n_comps <= 64) into out
 * (n_rows x n_comps).  The eigendecomposition of cov is the caller's.  Both
 * take a workspace of hz_pca_workspace_bytes(n_rows, n_cols) bytes.
 * Replaces the dense zero-centred branch of pp.pca (cuML PCA, svd_solver
 * "full"/"covariance_eigh"; preprocessing/_pca.py:71-330). */
size_t hz_pca_workspace_bytes(long long n_rows, int n_cols);
int hz_pca_covariance(const float* X, long long n_rows, int n_cols, double* mean, double* cov,
                      void* ws, size_t ws_bytes, void* stream);
int hz_pca_project(const float* X, long long n_rows, int n_cols, const double* mean,
                   const float* components, int n_comps, float* out, void* ws, size_t ws_bytes,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HARMONY_HIP_H */
