/* C ABI of libart_hip.so: the MI355X (gfx950) hot path of Adaptive Regression
 * Trees.  Plain pointers and sizes only; every device pointer is HIP device
 * memory, every call is enqueued on `stream` and never synchronizes unless its
 * comment says so.  Return values: ART_OK (0) or a negative ART_ERR_* code;
 * art_last_error() returns the message of the calling thread's last failure.
 *
 * Per-system numerical outcomes are NOT errors: they are written to int32
 * status arrays with the reference's codes (art/_cuda_batched_cholesky.py:52-55):
 *   0 ok, +k first non-positive Cholesky pivot (one-based), -1 flagged input
 *   (bad gate or weight), -2 non-finite assembly/factor, -3 non-finite solve;
 *   failed coefficients are NaN.
 */
#ifndef ART_HIP_H
#define ART_HIP_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ART_ABI_VERSION 1

#define ART_OK 0
#define ART_ERR_ARG (-1)
#define ART_ERR_HIP (-2)
#define ART_ERR_ALLOC (-3)

#define ART_PRECISION_FP64 0  /* FP64 Gram + FP64 Cholesky ("plain") */
#define ART_PRECISION_MIXED 1 /* low-precision MFMA Gram + FP64 refinement */
#define ART_PRECISION_FP64_REFINED 2 /* FP64 Gram + Cholesky + FP64 data-space refinement */

#define ART_DEFAULT_CHUNK_ROWS 2048      /* equal-size node plans */
#define ART_DEFAULT_LEAF_CHUNK_ROWS 4096 /* uneven leaf plans (launched largest chunk first) */
#define ART_MIXED_REFINE_TOL 1e-4 /* accept |delta| <= tol*|beta| after one refinement */
#define ART_MIXED_REFINE_STEPS 1  /* further refinement steps on still-flagged leaves */
#define ART_SOLVE_LDS_BYTES 65536

/* Feature basis of the leaf models (art/_cuda_features.py:13-29): degree 1
 * or 2, optional bias, bias first or (degree 1 only) last; p columns. */
typedef struct {
  int d;
  int p;
  int degree;
  int bias;
  int bias_last;
} art_basis;

/* Opaque resident workspace: chunk tables and partial-Gram buffers. */
typedef struct art_ridge_plan art_ridge_plan;

const char* art_last_error(void);
int art_version(void);

/* Number of feature columns of the basis, or -1 for an unsupported basis.
 * Replaces FeaturePlan.__init__'s p (art/_cuda_features.py:13-29,56-64). */
int art_feature_count(int d, int degree, int include_bias, int bias_last);

/* Phi[r, f] = feature f of X[r, :] for rows r < rows; X (rows, d) and
 * out (rows, p) row-major float64.  Column order: optional bias, x_0..x_{d-1},
 * then x_i*x_j for i <= j with i outermost; bias_last moves the bias after the
 * linear terms (degree 1 only).
 * Replaces FeaturePlan.build / polynomial_features
 * (art/_cuda_features.py:66-85, art/_cuda_features.cu:5-20). */
int art_features_build(const double* X, int64_t rows, int d, int degree, int include_bias,
                       int bias_last, double* out, hipStream_t stream);

/* flags[g] = 1 when rows [g*rows_per_group, (g+1)*rows_per_group) of `a`
 * (cols values each) or of `y` (one value per row, may be NULL) hold a
 * non-finite value, else 0.  Replaces batch_check_data
 * (art/_cuda_batched_cholesky_kernels.cu:7-13, art/_cuda_batched_cholesky.py:177-183). */
int art_check_finite(const double* a, const double* y, int64_t groups, int64_t rows_per_group,
                     int cols, int* flags, hipStream_t stream);

/* Variable-length leaves.  offsets (K+1, host memory) delimit leaf j's rows
 * [offsets[j], offsets[j+1]).  Allocates device buffers (synchronous).
 * chunk_rows: rows per Gram work item (multiple of 16; 0 =
 * ART_DEFAULT_LEAF_CHUNK_ROWS).  Work items are launched largest first. */
int art_leaf_plan_create(const int64_t* offsets, int K, int d, int degree, int include_bias,
                         int bias_last, int chunk_rows, art_ridge_plan** plan);

/* One weighted ridge fit per leaf with features built from X on the fly:
 *   beta_j = argmin sum_r max(w_r, floor) (y_r - phi(x_r).beta)^2 + lambda_j |beta|^2,
 *   lambda_j = ridge * trace(G_j) / p  (0 when ridge == 0).
 * X (N, d), y (N,), w (N,) float64 device; beta (K, p), status (K,),
 * lambda (K,) device outputs.  A negative or non-finite weight flags its
 * leaf (-1).  Replaces the per-leaf solve_weighted_ridge(solver="normal")
 * (art/models.py:218-343) fed by PolynomialRidgeModel.prepare_design
 * (art/models.py:602-620). */
int art_leaf_ridge_fit(art_ridge_plan* plan, const double* X, const double* y, const double* w,
                       double ridge, double weight_floor, int precision, double* beta,
                       int* status, double* lambda, hipStream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 fit of art_leaf_ridge_fit (the plan's default
 * chunking and kernels, so the same results as a plan fit) with no plan and
 * no allocation.  offsets (K+1) is host memory; basis gives d, p, degree,
 * bias and bias_last (p must match them; p <= 95, d <= 24).  Every work
 * buffer (chunk and leaf tables, per-chunk partial systems, weight flags)
 * lives in one static device array of ART_STATIC_SCRATCH_BYTES per GPU,
 * held until the process ends; a problem needing more is rejected with
 * ART_ERR_ARG.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (different streams, or threads) race; calls on one
 * stream are ordered.  Needs about 12 KB per 4096-row chunk at p <= 47
 * (e.g. ~1300 chunks, ~5M rows, for the quadratic basis in d = 8). */
#define ART_STATIC_SCRATCH_BYTES (16u << 20)
int art_leaf_ridge_fit_static(const int64_t* offsets, int K, art_basis basis, const double* X,
                              const double* y, const double* w, double ridge,
                              double weight_floor, double* beta, int* status, double* lambda,
                              hipStream_t stream);

/* ART_PRECISION_MIXED (p <= 64): split-bf16 MFMA Gram with exact FP64
 * diagonal/right-hand side, one FP64 data-space refinement pass, and the exact
 * FP64 path for every leaf whose refinement correction exceeds the tolerance
 * (or whose mixed factorization fails).  Number of leaves that took the FP64
 * path in the last mixed fit (synchronizes the device): */
int art_leaf_fallback_count(art_ridge_plan* plan, int* count);

/* ART_PRECISION_FP64_REFINED (p <= 64, d <= 16): the FP64 fit, its factor
 * kept, then FP64 data-space refinement steps beta += (L L^T)^-1 (Phi^T W
 * (y - Phi beta) - lambda beta) until |delta| <= 1e-8 |beta| (at most
 * 1 + refine_steps steps).  For systems whose Cholesky succeeds this
 * converges to the least-squares solution of the augmented weighted design,
 * the solution the reference reaches by its QR fallback for ill-conditioned
 * Grams (art/models.py:298-357).
 *
 * rcond (K,) device output: 1 / (|A|_1 |A^-1|_1) of each leaf's regularized
 * system A = G + lambda I from the factor of the last mixed or refined fit
 * (exact up to rounding; LAPACK dpocon, art/models.py:288-296, estimates
 * it); 0 for failed leaves (status != 0).  Replaces the reference's dpocon
 * call. */
int art_leaf_rcond(art_ridge_plan* plan, const int* status, double* rcond, hipStream_t stream);

/* Acceptance tolerance of the mixed refinement (default ART_MIXED_REFINE_TOL). */
int art_leaf_set_refine_tol(art_ridge_plan* plan, double tol);

/* Refinement steps after the first, run only on leaves the previous step
 * flagged (default ART_MIXED_REFINE_STEPS, 0..16).  A later step accepts a
 * leaf when its correction contracts (rho < 1/2) and the estimated error
 * left, rho |delta| / (1 - rho), is at most tol^2 |beta|; leaves still
 * flagged take the exact FP64 path.  No reference counterpart. */
int art_leaf_set_refine_steps(art_ridge_plan* plan, int steps);

/* Equal-size gated nodes (the BatchedCholeskyWorkspace contract): capacity K
 * nodes of n rows and p features.  Allocates device buffers (synchronous).
 * Replaces BatchedCholeskyWorkspace.allocate (art/_cuda_batched_cholesky.py:68-76). */
int art_gated_plan_create(int K, int64_t n, int p, int chunk_rows, art_ridge_plan** plan);

/* Left/right fits for `count` active nodes: side 0 weights max(1-pi, floor),
 * side 1 max(pi, floor).  phi (K, n, p), y (K, n), pi (K, n) float64 device;
 * ids: the active slot indices, both a device copy and a host copy (count
 * entries each).  Outputs are side-major: beta (2, K, p), status (2, K),
 * lambda (2, K); slots not in ids keep their previous contents.
 * Replaces BatchedCholeskyWorkspace.fit (art/_cuda_batched_cholesky.py:273-332). */
int art_gated_ridge_fit(art_ridge_plan* plan, const double* phi, const double* y, const double* pi,
                        const int* ids_device, const int* ids_host, int count, double ridge,
                        double weight_floor, double* beta, int* status, double* lambda,
                        hipStream_t stream);

/* Plan shape: feature count p, Gram work items, 16-column tiles. */
int art_plan_info(const art_ridge_plan* plan, int* p, int* nchunks, int* tiles);

/* Kernel timing for benchmarks: keep `slots` rings of HIP events recorded on
 * each fit's stream around the Gram kernel and the solve kernel (0 disables).
 * No reference counterpart; the reference times phases with CUDA events in
 * _Timers (art/_cuda_optimizer.py:76-107). */
int art_plan_timing_enable(art_ridge_plan* plan, int slots);

/* Waits for and returns up to `max` recorded (gram_ms, solve_ms) pairs;
 * returns how many were written (or a negative error). */
int art_plan_timing_read(art_ridge_plan* plan, float* gram_ms, float* solve_ms, int max);

/* ---------------------------------------------------------------------
 * Soft-oblique split optimizer.  One problem = one node x one start theta0.
 * Device arrays: X (K, n, d), y (K, n), temperature (K,), theta0 (P, d+1),
 * node (P,) int32 -> outputs theta_out (P, d+1) (the final, unit-w split),
 * beta_out (P, 2, p) (hard-refit left/right coefficients), hard (P, 8)
 * [MSE, MSE_L, MSE_R, n_right, lambda_L, lambda_R, bad, min_side_fallback],
 * info (P, 8) int32 [stop or error code, converged, n_iters,
 * fallback_iteration, physical candidates, logical candidates, error,
 * final partition stage ran],
 * history (P, 5, max_iters+1) [soft loss, |projected grad|, accepted step,
 * backtracks, line-search loss], scalars (P, 6) [initial |grad|, gradient
 * tolerance, final alpha, temperature, lambda_L and lambda_R of the last
 * soft fit].  H is the logical backtrack batch (candidates examined per
 * batch before the earliest Armijo pass is taken).
 * Codes: 0 max_iters, 1 gradient_tolerance, 2 non_descent_direction,
 * 3 line_search_failed, 4 max_line_search_failures, 5 alpha_min_reached,
 * 6 cholesky_failure, 7 min_side_points, 8 nonfinite_split.
 * mode 0: Armijo + AdaptiveAlpha (the reference loop); mode 1: annealed
 * fixed step (anneal_iters iterations, T geometric from t0 to t1, step);
 * mode 2: parent fit only (unweighted fit of all rows; hard[0] = its MSE). */
typedef struct {
  const double* X;
  const double* y;
  const double* temperature;
  const double* theta0;
  const int* node;
  int n;
  int d;
  art_basis basis;
  double ridge, floor, grad_atol, grad_rtol, alpha0, rho, armijo_c, alpha_min, alpha_max, grow,
      recovery;
  int max_iters, max_backtracks, heavy, max_failures, adaptive, required, fallback, hard_refit, H,
      mode, anneal_iters;
  double t0, t1, step;
  double* theta_out;
  double* beta_out;
  double* hard;
  int* info;
  double* history;
  double* scalars;
  /* Uneven node sizes (optional): node slot k owns rows
   * [row_offsets[k], row_offsets[k+1]) of X (rows, d) and y (rows,); n is
   * then unused and each problem needs
   * max(min_side_points, ceil(min_side_fraction * n_k)) points per side
   * (art/splitters.py:750-752) instead of `required`.  NULL: n rows each. */
  const int64_t* row_offsets;
  int min_side_points;
  double min_side_fraction;
  /* Child-system precision of the annealed soft fits (art_split_anneal_batched
   * only; 0 elsewhere): ART_PRECISION_FP64, or ART_PRECISION_MIXED for the
   * split-bf16 MFMA product of the 32 child systems (bf16 hi + mid + lo of
   * each side weight and packed product, six products on
   * v_mfma_f32_16x16x32_bf16, FP32 accumulation per 32 rows into FP64).
   * Gates, predictions, gradients, solves and the final hard refit stay
   * FP64. */
  int child_precision;
} art_split_args;

/* Runs every problem of a cohort to completion in one launch (p <= 32,
 * d <= 31, H >= 1).  Replaces the host-driven loop of
 * BatchedOptimizerWorkspace._run/search/advance/_finish
 * (art/_cuda_batched_optimizer.py:331-526) and, for one node, the CPU
 * SoftObliqueSplitter._optimize_from_initialization (art/splitters.py:365-690). */
int art_split_optimize(const art_split_args* args, int problems, hipStream_t stream);

/* art_split_optimize with each problem's rows split over `parts` workgroups
 * (1 <= parts <= 8): the same passes, every pass's row sums added over the
 * parts (in part order, so all parts hold the same node state and make the
 * same decisions), for cohorts too small to fill the GPU with one workgroup
 * per problem.  Same outputs and controls as art_split_optimize (whose
 * reference counterparts it shares: art/_cuda_batched_optimizer.py:331-526,
 * art/splitters.py:365-690); the sums differ from it in rounding only.
 * count: `problems` ints and failed: one int, both cleared by the call; buf:
 * problems * art_split_part_doubles(p, parts) doubles of exchange space.
 * The grid (problems * parts workgroups) goes through a cooperative launch
 * and must be resident at once; the call returns ART_ERR_HIP, with nothing
 * run, when the runtime refuses it.  *failed = 1 after the launch means a
 * part waited too long for the others (the results are then invalid).
 * parts = 1 is art_split_optimize. */
int art_split_optimize_parts(const art_split_args* args, int problems, int parts, int* count,
                             double* buf, int* failed, hipStream_t stream);

/* Exchange doubles per problem of art_split_optimize_parts (-1 if invalid). */
int art_split_part_doubles(int p, int parts);

/* The most parts (<= 8) for which `problems` problems of this launch's
 * kernel (args' basis picks it) fit on the current device at once (1 when
 * none). */
int art_split_max_parts(const art_split_args* args, int problems);

/* Annealed mode only (mode and Armijo fields ignored), restart-batched:
 * problems come in blocks of 16 consecutive problems that share one node
 * (node[16b] is the block's node; pad a node's starts with duplicates), one
 * workgroup per block streams the node's rows once per pass for all 16
 * starts.  Each side's system is its own floored weighted Gram
 * (max(1 - pi, floor) and max(pi, floor)), as in the two weighted fits of
 * art/objectives.py:237-303, so the results equal art_split_optimize mode 1
 * per start.  Needs d <= 16 and at most 17 basis features.  Same
 * outputs as art_split_optimize mode 1 (history row 0 = soft loss per
 * iteration). */
int art_split_anneal_batched(const art_split_args* args, int problems, hipStream_t stream);

/* Random starts theta (problems, d+1) for nodes node[q] of X (K, n, d):
 * w = unit Gaussian direction (FP64 Box-Muller from Philox4x32-10, key seed,
 * counter (pair, q, counter_lo, counter_hi)), z = median over the node's rows
 * of x.w (mean of the two middle values for even n).  Replaces the
 * reference's cuRAND initialization kernels initial_directions /
 * initial_scores / initial_offsets with cp.median
 * (art/_cuda_batched_optimizer.cu:25-45, art/_cuda_batched_optimizer.py:291-305). */
int art_split_init(const double* X, int n, int d, const int* node, int problems, uint64_t seed,
                   uint64_t counter, double* theta, hipStream_t stream);

/* art_split_init for nodes of uneven size: node slot k owns rows
 * [row_offsets[k], row_offsets[k+1]) of X (rows, d); row_offsets is a device
 * array.  keys (problems,), optional: problem q then draws with key keys[q]
 * and counter (pair, 0, 0, 0), independent of its position in the call;
 * NULL gives art_split_init's stream. */
int art_split_init_ranges(const double* X, const int64_t* row_offsets, int d, const int* node,
                          int problems, uint64_t seed, uint64_t counter, const uint64_t* keys,
                          double* theta, hipStream_t stream);

/* Resident systems for rank-k refits.  art_leaf_gram writes each leaf's
 * FP64 augmented system [Phi | y]^T W [Phi | y] (the fit's Gram and
 * right-hand side, no ridge) to grams (K, art_gram_size(p)) in tile layout
 * -- tile pair (t1 <= t2, t1 outermost) element (a, b) is
 * G[16 t1 + a][16 t2 + b] of the (p+1)-column augmented matrix -- and ORs
 * the leaf's invalid-weight flags into bad (K,).  Systems add: the system of
 * old rows plus the system of new rows is the system of all rows.
 * art_gram_solve fits every resident system as art_leaf_ridge_fit's FP64
 * path does (trace-scaled ridge, Cholesky, the same status codes).  No
 * reference counterpart: the reference refits leaves from their rows
 * (art/builder.py:812-883). */
int art_leaf_gram(art_ridge_plan* plan, const double* X, const double* y, const double* w,
                  double weight_floor, double* grams, int* bad, hipStream_t stream);
int art_gram_size(int p);
/* art_leaf_gram with a precision and an optional leaf selection.
 * precision ART_PRECISION_FP64 is art_leaf_gram's system; ART_PRECISION_MIXED
 * (p <= 64, d <= 16) writes the split-bf16 MFMA system of the mixed fit in
 * the same layout: entry (i, j) within 4e-7 of sqrt(G_ii G_jj) of the FP64
 * system (the FP32 rounding of each scaled feature value), ~1e-8 on large
 * leaves.
 * mask (K,) int32 device array or NULL (every leaf): a leaf with mask[j] == 0
 * is skipped and keeps what grams and bad held.  No reference counterpart:
 * the serving loop forms split nodes' statistics at mixed precision and the
 * systems it keeps for refits in FP64. */
int art_leaf_gram_ex(art_ridge_plan* plan, const double* X, const double* y, const double* w,
                     double weight_floor, int precision, const int* mask, double* grams,
                     int* bad, hipStream_t stream);
int art_gram_solve(const double* grams, const int* bad, int K, art_basis basis, double ridge,
                   double* beta, int* status, double* lambda, hipStream_t stream);

/* Frees a plan's device buffers (synchronizes the device).
 * Replaces BatchedCholeskyWorkspace.close (art/_cuda_batched_cholesky.py:339-352). */
void art_plan_destroy(art_ridge_plan* plan);

/* ---------------------------------------------------------------------
 * Hit-and-run sampling of P convex polytopes {x : A x <= b} in d <= 32
 * dimensions, `chains` chains per polytope, lane per chain.
 * Device arrays: A (rows, d) and b (rows,) hold every polytope's constraints
 * back to back, polytope q owning rows [row_offsets[q], row_offsets[q+1])
 * (row_offsets (P+1,) int64); x0 (P, chains, d) feasible starts; shape
 * (d, d) row-major or NULL (direction d <- shape . d, the covariance shaping
 * V diag(sqrt(lambda)) of art/sampling.py:496-537,865-875).  Optional axis
 * bounds box_lo, box_hi (P, d) (+-inf where absent) add the rows
 * x_j <= box_hi[q, j] and -x_j <= -box_lo[q, j] to polytope q; they give the
 * same chord as listing those rows in A, for less work.  A polytope may then
 * have no rows in A.
 * Schedule: chain c records after moves first, first + stride, ...;
 * K_c = n / chains + (c < n % chains) records; record k is sample
 * k * chains + c of samples (P, n, d).  The batched sampler is
 * first = burn_in + max(1, thinning), stride = max(1, thinning)
 * (art/sampling.py:618-803); the scalar sampler is chains = 1,
 * first = burn_in + 1, stride = thinning (art/sampling.py:445-494).
 * Random stream: directions == NULL draws Philox4x32-10 (key = seed,
 * counter (move * W + w, chain, polytope, 0), W = ceil((2 ceil(d/2) + 2) / 4))
 * FP32 Box-Muller directions and a 53-bit chord uniform; otherwise
 * directions (P, steps, chains, d) and uniforms (P, steps, chains) are read
 * as given (steps >= the longest chain's moves).
 * Replaces BatchedHitAndRunSampler.sample / _advance_batch
 * (art/sampling.py:618-962) and the Numba kernels sample_precomputed_chains
 * (art/_numba_sampling.py:43-167). */
typedef struct {
  const double* A;
  const double* b;
  const int64_t* row_offsets;
  const double* x0;
  const double* shape;
  const double* box_lo;
  const double* box_hi;
  const double* directions;
  const double* uniforms;
  double* samples;
  uint64_t seed;
  int P;
  int d;
  int chains;
  int n;
  int first;
  int stride;
  int steps;
  /* Per-polytope streams and shapes (both optional): seeds (P,) gives
   * polytope q the key seeds[q] and counter (move * W + w, chain, 0, 0), the
   * stream a one-polytope call with seed seeds[q] draws; shape_stride > 0
   * reads polytope q's shaping matrix at shape + q * shape_stride. */
  const uint64_t* seeds;
  int shape_stride;
  /* 1 + the largest row count of any polytope in A, or 0 when not known.
   * With one chain per polytope and a known bound, a wave splits each block
   * of 64 moves of the chain (same samples, shorter critical path). */
  int row_bound;
} art_har_args;

int art_har_sample(const art_har_args* args, hipStream_t stream);

/* out[i] = Philox4x32-10(counters[i], keys[i]) for i < count: counters (count, 4),
 * keys (count, 2), out (count, 4) uint32 device arrays.  The random stream
 * of art_har_sample is defined in terms of this function.  It replaces the
 * cuRAND Philox generator the reference uses for its random starts
 * (art/_cuda_batched_optimizer.py:158-160,291-305). */
int art_philox4x32_10(const uint32_t* counters, const uint32_t* keys, int64_t count,
                      uint32_t* out, hipStream_t stream);


/* ---------------------------------------------------------------------
 * Tree inference: out[r] = prediction of the flattened tree at X[r, :]
 * (X (n, d) row-major).  Split node k: w = node_w[k*d .. k*d+d), z =
 * node_z[k], children node_child[2k] (left) and node_child[2k+1] (right);
 * a link c < 0 names leaf -1-c.  root is the root link (a leaf when < 0,
 * then the split arrays may be NULL).  Leaf l's model is
 * leaf_coef[l*p .. l*p+p) on the basis (d <= 32, degree 1 or 2).  A point
 * goes right when w.x - z >= 0.  Replaces RegressionTree.predict and
 * SplitNode.route (art/tree.py:45-50,124-165). */
int art_tree_predict(const double* X, int64_t n, const double* node_w, const double* node_z,
                     const int* node_child, int root, const double* leaf_coef, art_basis basis,
                     double* out, hipStream_t stream);

/* leaf[r] = index of the leaf that X[r, :] reaches in the flattened tree of
 * art_tree_predict (the same routing rule and leaf numbering). */
int art_tree_leaf(const double* X, int64_t n, int d, const double* node_w, const double* node_z,
                  const int* node_child, int root, int* leaf, hipStream_t stream);

/* out[s] = median over the points of set s of the distance to their nearest
 * other point (mode 0), or median over all point pairs of their distance
 * (mode 1).  Set s holds rows [offsets[s], offsets[s+1]) of X (rows, d);
 * offsets lives on the device and host_offsets is its host copy (2 to 4096
 * points per set).  An even count averages the two middle distances, as
 * numpy.median does.  Replaces median_nearest_neighbor_distance and
 * median_pairwise_distance (art/temperature.py:73-118). */
int art_knn_median(const double* X, const int64_t* offsets, const int64_t* host_offsets, int sets,
                   int d, int mode, double* out, hipStream_t stream);

/* ---- Serving loop: data-driven tree fits (no counterpart in the reference,
 * whose builder queries an oracle; these move a request's own rows).  ---- */

/* sse[k] = y'Wy - 2 beta'h + beta'G beta of system k of art_leaf_gram's
 * layout (K systems of p coefficients) for coefficients beta (K, p);
 * wsum[k] (optional) = G[bias_col][bias_col], the weight sum when bias_col
 * is the constant column (-1: none, wsum = 0). */
int art_gram_sse(const double* grams, int K, int p, int bias_col, const double* beta, double* sse,
                 double* wsum, hipStream_t stream);

/* Row sample of K row ranges: output row i of range k (rows
 * [dst_offsets[k], dst_offsets[k+1])) copies input row
 * src_begin[k] + floor((i - dst_offsets[k]) n_k / m_k), where
 * n_k = src_end[k] - src_begin[k] and m_k <= n_k is the output range size.
 * Offsets are device arrays; total = dst_offsets[K].  w and wo are both NULL
 * or both set. */
int art_gather_strided(const double* X, const double* y, const double* w, int d, int K,
                       const int64_t* src_begin, const int64_t* src_end,
                       const int64_t* dst_offsets, int64_t total, double* Xo, double* yo,
                       double* wo, hipStream_t stream);

/* Stable partition of K row ranges [src_begin[k], src_end[k]) of
 * (X, y, w) by hyperplanes: a row goes right when node_w[k].x - node_z[k] >= 0
 * (art_tree_leaf's rule).  Range k's left rows land at dst_base[k] + 0.. and
 * its right rows after them, each side in input order; node_right[k] gets
 * the right count.  Ranges are cut into blocks of art_partition_block_rows()
 * rows: blk_offsets (K+1) gives range k blocks blk_offsets[k] ..
 * blk_offsets[k+1] (ceil(n_k / rows) of them), blocks = blk_offsets[K].
 * flags (one byte per input row) and blk_right (blocks ints) are scratch.
 * All arrays are device arrays; outputs must not overlap inputs. */
int art_partition_rows(const double* X, const double* y, const double* w, int d, int K,
                       const int64_t* src_begin, const int64_t* src_end,
                       const int64_t* blk_offsets, int64_t blocks,
                       const int64_t* dst_base, const double* node_w, const double* node_z,
                       unsigned char* flags, int* blk_right, int64_t* node_right, double* Xo,
                       double* yo, double* wo, hipStream_t stream);
int art_partition_block_rows(void);

#ifdef __cplusplus
}
#endif

#endif /* ART_HIP_H */
