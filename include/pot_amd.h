/* pot_amd.h -- C ABI of the MI355X (gfx950) entropic Sinkhorn library.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t
 * (passed as void*), runs stream-ordered on that stream and returns an int:
 * >= 0 a solver status, < 0 an error.  All buffers are owned by the caller;
 * scratch space is the caller's `workspace` (its size comes from the matching
 * *_workspace_size function).  Weights, potentials and error histories are
 * float64 device arrays whatever the cost dtype.
 *
 * The reference (POT 0.9.7.dev0) reaches this arithmetic through its backend
 * operators (ot/backend.py:1306-1307 logsumexp, nx.dot / nx.einsum); its own
 * native FFI precedent is EMD_wrap (ot/lp/EMD_wrapper.cpp:209, declared in
 * ot/lp/EMD.h): caller-allocated float64 outputs plus an int status that the
 * Python layer maps to warnings (ot/lp/emd_wrap.pyx:30-43).  The conventions
 * below follow that precedent.
 */
#ifndef POT_AMD_H
#define POT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status and error codes ------------------------------------------ */
#define POTAMD_CONVERGED 0    /* err < stopThr at a check iteration */
#define POTAMD_MAX_ITER 1     /* numItermax reached (the reference warns) */
#define POTAMD_NUMERICAL 2    /* scaling domain: 0 / nan / inf, state reverted */
#define POTAMD_NEED_EXACT 3   /* log domain: a column mass left the fp32 range */
#define POTAMD_EINVAL (-1)    /* invalid argument */
#define POTAMD_EHIP (-2)      /* HIP runtime error */
#define POTAMD_EUNSUPPORTED (-3) /* shape outside the compiled kernel set */
#define POTAMD_EWORKSPACE (-4)   /* workspace too small */

#define POTAMD_F32 0
#define POTAMD_F64 1
#define POTAMD_LOG 0      /* log-domain potentials (sinkhorn_log, log batch) */
#define POTAMD_SCALING 1  /* scaling vectors (sinkhorn_knopp, plain batch) */
#define POTAMD_COL_FIRST 0 /* v update first: ot.sinkhorn / sinkhorn_log / knopp */
#define POTAMD_ROW_FIRST 1 /* u update first: ot.batch projections */

/* Library version (major*10000 + minor*100 + patch). */
int potamd_version(void);

/* Human-readable text of a status / error code. */
const char* potamd_status_string(int code);

/* Bytes of workspace needed by potamd_sinkhorn_dense. */
size_t potamd_dense_workspace_size(int dtype, int batch, int n, int m, int rows_per_block);

/* Dense-cost Sinkhorn on `batch` problems of shape n x m sharing one reg.
 *
 * Replaces the iteration loops of
 *   ot.bregman.sinkhorn_log       ot/bregman/_sinkhorn.py:690-918  (LOG, COL_FIRST)
 *   ot.bregman.sinkhorn_knopp     ot/bregman/_sinkhorn.py:487-687  (SCALING, COL_FIRST)
 *   ot.batch.bregman_log_projection_batch  ot/batch/_utils.py:165-290 (LOG, ROW_FIRST)
 *   ot.batch.bregman_projection_batch      ot/batch/_utils.py:46-162  (SCALING, ROW_FIRST)
 * with the reference's check cadence (every 10th iteration), strict
 * `err < stopThr`, and the batch's single stop on the largest error.
 *
 * M: cost (kernel_given = 0; element exponent -M/reg) or the kernel itself
 *    (kernel_given = 1: log kernel in LOG, K in SCALING), dtype `dtype`,
 *    problem p row i at M + p*bstride + i*ldm.
 * a, b: float64 weights, batch x n and batch x m.
 * u, v: float64, batch x n / batch x m.  On entry the initial state
 *    (LOG: log potentials, zeros for a cold start; the COL_FIRST solve uses
 *    only u, v is recomputed first.  SCALING: scalings).  On exit the state
 *    the reference returns (log_u/log_v, or u/v).
 * err_norm: how a batch combines its per-problem errors: 0 = maximum (the
 *    ot.batch stop, ot/batch/_utils.py:273-278), 1 = Frobenius norm (the
 *    multi-histogram knopp, ot/bregman/_sinkhorn.py:646-652, which runs the
 *    histograms as a batch sharing M through bstride = 0).
 * err_hist: float64 device array of err_cap entries, the reference's
 *    log["err"] (one value per check).
 * threads, cols_per_thread, rows_per_block, prefetch (rows loaded ahead, 1
 *    or 2): tile shape, 0 = automatic.
 * host_ctrl: pinned host memory of >= 512 bytes used for asynchronous stop
 *    polling, or NULL (then each check synchronises).
 * info (host, 4 ints): [0] niter (the reference's `ii`), [1] number of err
 *    values, [2] status, [3] iteration of a numerical stop.
 */
int potamd_sinkhorn_dense(int dtype, int domain, int order, int kernel_given,
                          const void* M, int64_t ldm, int64_t bstride, int batch, int n, int m,
                          double reg, const double* a, const double* b, double* u, double* v,
                          int max_iter, double stop_thr, int err_norm, double* err_hist,
                          int err_cap, int threads, int cols_per_thread, int rows_per_block,
                          int prefetch, void* workspace, size_t ws_bytes, void* host_ctrl,
                          void* stream, int* info);

/* The same dense solve with each problem stopping on its own check, as a
 * lone call per problem would (a batch of sinkhorn_log problems, e.g. the
 * histograms of a 2-D b sharing M through bstride = 0).  Replaces the
 * per-column loop of the multi-histogram sinkhorn_log
 * (ot/bregman/_sinkhorn.py:816-854: one sinkhorn_log call per column of b,
 * each with its own stop at :884-899) with one batched solve.
 * niters, statuses: device int[batch], each problem's `ii` at its stop and
 * its status (POTAMD_CONVERGED or POTAMD_MAX_ITER).  u, v receive every
 * problem's state at its own stop.  The return value is POTAMD_CONVERGED
 * when every problem converged, POTAMD_MAX_ITER when one reached max_iter,
 * POTAMD_NEED_EXACT as for potamd_sinkhorn_dense (then u, v are not written).
 * Workspace: potamd_dense_workspace_size.  No error history is kept.
 */
int potamd_sinkhorn_dense_each(int dtype, int domain, int order, int kernel_given,
                               const void* M, int64_t ldm, int64_t bstride, int batch, int n,
                               int m, double reg, const double* a, const double* b, double* u,
                               double* v, int max_iter, double stop_thr, int threads,
                               int cols_per_thread, int rows_per_block, int prefetch,
                               void* workspace, size_t ws_bytes, void* host_ctrl, void* stream,
                               int* niters, int* statuses, int* info);

/* Bytes of workspace needed by potamd_plan_dense. */
size_t potamd_plan_workspace_size(int batch, int n);

/* Plan and per-problem sums from potentials (fp64 element arithmetic).
 *
 * Replaces the returns of sinkhorn_log (ot/bregman/_sinkhorn.py:915-918),
 * sinkhorn_knopp (:683-687), the loss of sinkhorn2 (:436-438), the batch
 * plan / log plan (ot/batch/_utils.py:160, :286-290), loss_linear_batch
 * (ot/batch/_linear.py:123-147) and the entropy / KL values of solve_batch
 * (_linear.py:338-345).
 * P / logP: outputs in `dtype`, contiguous n x m per problem with problem
 *    stride pstride; either may be NULL.
 * sums: float64 device, batch x 4: sum M*P, sum P log(P+1e-16),
 *    sum P log(P/(a b)+1e-16) (needs a, b; else 0), sum P.  May be NULL.
 */
int potamd_plan_dense(int dtype, int domain, int kernel_given, const void* M, int64_t ldm,
                      int64_t bstride, int batch, int n, int m, double reg, const double* u,
                      const double* v, const double* a, const double* b, void* P, void* logP,
                      int64_t pstride, double* sums, void* workspace, size_t ws_bytes,
                      void* stream);

/* Batched (squared) Euclidean cost matrices.
 * Replaces ot.dist / ot.utils.euclidean_distances (ot/utils.py:270-307,
 * :465-468; clamp = 1, zero_diag when X is Y) and
 * ot.batch.dist_euclidean_batch (ot/batch/_linear.py:54-85; clamp = 0).
 * X: batch x n x d, Y: batch x m x d, M: batch x n x ldm (same dtype).
 */
int potamd_sqeuclidean(int dtype, const void* X, const void* Y, int batch, int n, int m, int d,
                       int64_t xstride, int64_t ystride, void* M, int64_t ldm, int64_t mstride,
                       int squared, int clamp, int zero_diag, void* stream);

/* Ragged batch of squared-Euclidean costs (float32), clamped at 0, with the
 * per-problem maximum (the `M / M.max()` normalisation of the workloads).
 * Problem p: points X[xoff[p] .. +ns[p]) and Y[yoff[p] .. +ms[p]) of
 * dimension d (row-major), cost at M + moff[p] with row stride ldms[p];
 * maxout[p] receives max_ij M (may be NULL).  Descriptor arrays live on the
 * device; max_n / max_m bound the shapes.  Generalises ot.dist (ot/utils.py:
 * 270-307) to many problems in one launch.
 */
int potamd_sqeuclidean_ragged(int batch, int max_n, int max_m, int d, const float* X,
                              const float* Y, const int64_t* xoff, const int64_t* yoff,
                              const int* ns, const int* ms, float* M, const int64_t* moff,
                              const int* ldms, float* maxout, void* stream);

/* Ragged batch of independent sinkhorn_log problems, one per workgroup.
 *
 * Each problem p runs exactly the loop of ot.bregman.sinkhorn_log
 * (ot/bregman/_sinkhorn.py:880-906) on its own cost with its own reg and
 * initial u, stopping on its own check (the batch of the reference,
 * ot/batch/_utils.py:165-290, needs equal shapes, one reg and one stop; this
 * entry point generalises it for many uneven problems).
 * M: packed float32 costs; problem p at M + moff[p], n_p x m_p with row stride
 *    ldms[p].  All descriptor arrays (moff, ns, ms, ldms, aoff, boff, regs,
 *    order) are device arrays of `batch` entries; max_n >= every n_p and
 *    max_m >= every m_p (both <= 1024).
 * a, u: problem p's rows at offset aoff[p]; b, v: columns at boff[p].
 *    u holds u0 on entry (zeros = the reference's cold start).
 * check_every: iterations between marginal checks (0 = 10, the reference's
 *    cadence; 1 stops a problem at the first iteration that meets stopThr --
 *    the error is formed every sweep anyway).
 * order: launch order of the problems (e.g. most work first) or NULL.
 * vec_aligned: 1 if M is 16-byte aligned and every moff, ldms is a multiple
 *    of 4 (vector loads), else 0.
 * fallback_thr: > 0 enables the warm-start fallback: a problem whose first
 *    checked error exceeds it restarts from u0 = 0 (status gets bit 4).
 * niter, status: per-problem outputs (device int arrays); err_last / err_first:
 *    last and first checked marginal errors (device, may be NULL).
 * waves: wavefronts per problem (4 or 8; 0 = default) in the low byte, rows
 *    prefetched per wave (1 or 2, 0 = 1; 2 needs max_m > 512) in the next byte,
 *    register cap in waves per SIMD in the third (0 = default, 4 = two 8-wave
 *    workgroups per CU (8 waves, 1 row prefetched only), 255 = no cap); bit 24
 *    starts each problem's row sweep at a problem-dependent row.
 */
int potamd_sinkhorn_log_batch(int dtype, int batch, int max_n, int max_m, const void* M,
                              const int64_t* moff, const int* ns, const int* ms, const int* ldms,
                              const int64_t* aoff, const int64_t* boff, const double* regs,
                              const double* a, const double* b, double* u, double* v,
                              int max_iter, int check_every, double stop_thr,
                              double fallback_thr, const int* order, int vec_aligned,
                              int* niter, int* status,
                              double* err_last, double* err_first, int waves, void* stream);

/* The same batch of sinkhorn_log problems given as point clouds: the squared
 * Euclidean cost |x_r - y_c|^2 (ot.dist, ot/utils.py:270-307) is computed on
 * the fly on MFMA and never stored; with `normalize` it is divided by its
 * maximum (ot.dist(...) / M.max(), found on the device and returned in
 * cmax).  Replaces, like potamd_sinkhorn_log_batch, a per-problem
 * ot.bregman.sinkhorn_log call (ot/bregman/_sinkhorn.py:690-918) with the
 * reference's stop rule, or the ot.batch solvers (ot/batch/_linear.py:
 * 236-357) on uneven sizes.
 * X: (sum n, d) float32 row-major, problem p's points from row xoff[p];
 *    Y likewise with yoff.  a, u share X's offsets, b, v share Y's.
 * n, m <= 1024, d <= 16.  regs: reg of each problem (on the normalised cost
 *    when normalize != 0).  normalize: 0 the raw cost, 1 the cost divided by
 *    its maximum (found here, written to cmax), 2 divided by the given cmax[p].
 *    u: in u0 (natural log; 0 for the cold start), out log u; v: out log v.
 * max_iter, check_every, stop_thr, fallback_thr, order, niter, status,
 *    err_last, err_first: as for potamd_sinkhorn_log_batch; status 2 means
 *    the problem was not solved (cost maximum 0, or reg so far below the
 *    cost scale that the scaled coordinates leave the f16 range), with NaN
 *    potentials.
 * cmax: per-problem cost maximum (device double[B]): written with normalize
 *    0 (1.0) or 1, read with 2; may be NULL unless normalize == 2.
 * waves: wavefronts per problem, 8 or 16 (0 = default 8).
 * queue: device int[2], zero before the first launch; with it each
 *    workgroup takes the next problem of `order` when it starts (global
 *    longest-first scheduling whatever the block-to-CU assignment), and the
 *    kernel leaves it zero again, so a buffer can be reused by later calls
 *    and graph replays.  Calls that overlap on one GPU need separate queues.
 *    NULL: workgroup i solves problem order[i].
 */
int potamd_sinkhorn_log_batch_points(int batch, int max_n, int max_m, int d, const float* X,
                                     const float* Y, const int64_t* xoff, const int64_t* yoff,
                                     const int* ns, const int* ms, const double* regs,
                                     int normalize, const double* a, const double* b, double* u,
                                     double* v, int max_iter, int check_every, double stop_thr,
                                     double fallback_thr, const int* order, int* niter,
                                     int* status, double* err_last, double* err_first,
                                     double* cmax, int waves, int* queue, void* stream);

/* The same batch with large problems split over several workgroups, for
 * batches too small to fill the GPU with one workgroup per problem (the
 * slowest problem then sets the time).  No counterpart in the reference
 * (ot/bregman/_sinkhorn.py:880-906 per problem, as above).
 * items, order: the queue, `items` entries, each p | k << 24 (problem p, part
 *    k); the parts of a split problem must be taken together, so list them
 *    next to each other.
 * split: device int[B], 0 for a whole problem, else parts | slot << 4 with
 *    parts 2..8 (each part needs at least one 32-row block: n >= 32 parts)
 *    and slot < nsplit, one slot per split problem.
 * split_ws: nsplit exchange slots (potamd_points_split_workspace_size bytes,
 *    256-byte aligned), zero before the first launch; the kernel leaves the
 *    counters zero again, like the queue.  The parts combine their column
 *    reductions there in part order, so every part takes the same decisions;
 *    a part that waits too long for the others gives up and the problem
 *    reports status 2.
 * queue: required; 8 wavefronts per part.  Other arguments as above.
 */
size_t potamd_points_split_workspace_size(int nsplit);
int potamd_sinkhorn_log_batch_points_split(int batch, int max_n, int max_m, int d, const float* X,
                                           const float* Y, const int64_t* xoff,
                                           const int64_t* yoff, const int* ns, const int* ms,
                                           const double* regs, int normalize, const double* a,
                                           const double* b, double* u, double* v, int max_iter,
                                           int check_every, double stop_thr, double fallback_thr,
                                           int items, const int* order, const int* split,
                                           int nsplit, void* split_ws, size_t split_ws_bytes,
                                           int* niter, int* status, double* err_last,
                                           double* err_first, double* cmax, int* queue,
                                           void* stream);

/* No-allocation variant of potamd_sinkhorn_log_batch_points_split.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The work-queue counters and the exchange slots
 * live in a fixed device array of the library (potamd_noalloc_scratch_bytes
 * bytes per GPU, held until the process ends); the call clears the parts it
 * uses with a fill kernel on `stream`, then solves, and allocates nothing.
 * nsplit: 0 .. POTAMD_NOALLOC_MAX_SPLIT (more returns POTAMD_EWORKSPACE);
 *    with 0 every split entry must be 0 (a problem with a slot reports
 *    status 2, as a malformed entry does).  Other arguments as for
 *    potamd_sinkhorn_log_batch_points_split; the results are the same.
 * Calls that overlap on one GPU share the array and race: serialise them
 * (one stream, or events).
 */
#define POTAMD_NOALLOC_MAX_SPLIT 256
size_t potamd_noalloc_scratch_bytes(void);
int potamd_sinkhorn_log_batch_points_split_noalloc(
    int batch, int max_n, int max_m, int d, const float* X, const float* Y, const int64_t* xoff,
    const int64_t* yoff, const int* ns, const int* ms, const double* regs, int normalize,
    const double* a, const double* b, double* u, double* v, int max_iter, int check_every,
    double stop_thr, double fallback_thr, int items, const int* order, const int* split,
    int nsplit, int* niter, int* status, double* err_last, double* err_first, double* cmax,
    void* stream);

/* Per-problem maximum of |x_r - y_c|^2 (the M.max() of ot.dist(x, y) /
 * M.max()), the same device search as potamd_sinkhorn_log_batch_points with
 * normalize = 1 (MFMA pass, then the best pair of each lane evaluated in fp64
 * from the points), without the solve.  cmax: device double[B]; a problem
 * with n or m outside [1, 1024] gets NaN.
 */
int potamd_ragged_points_cmax(int batch, int max_n, int max_m, int d, const float* X,
                              const float* Y, const int64_t* xoff, const int64_t* yoff,
                              const int* ns, const int* ms, double* cmax, void* stream);

/* Batched Minkowski / KL cost matrices, M[b] (batch x n x m, row-major, the
 * dtype of X and Y; X: batch x n x d, Y: batch x m x d, contiguous):
 *   metric 0: M_bij = (sum_k |x_bik - y_bjk|^p)^(q/p)   ot.batch.dist_lp_batch
 *             (ot/batch/_linear.py:20-51); p, q > 0
 *   metric 1: M_bij = sum_k y log(y + eps) - y log(x + eps)   dist_kl_batch
 *             (:88-120, logits_X = False)
 *   metric 2: M_bij = sum_k y log(y + eps) - y x   (logits_X = True, with the
 *             sum over k of the docstring formula; the reference's own
 *             logits branch, :115-116, omits that sum).
 */
int potamd_pairwise_batch(int dtype, int metric, const void* X, const void* Y, int batch, int n,
                          int m, int d, double p, double q, double eps, void* M, void* stream);

/* Strided batched GEMM, C_b = alpha op(A_b) op(B_b) + beta C_b (C read only
 * when beta != 0), dtype POTAMD_F32 or POTAMD_F64 for all three.  Element
 * (i, k) of op(A_b) is A[b sA + i rsA + k csA], element (k, j) of op(B_b) is
 * B[b sB + k rsB + j csB] (so transposes and channel-last tensors are read in
 * place), element (i, j) of C_b is C[b sC + i ldc + j].  Replaces the
 * einsum products of the batched Gromov-Wasserstein tensor product
 * (ot/batch/_quadratic.py:535-572, hC1 T hC2^T per channel) and of the
 * batched cost backward.
 */
int potamd_gemm_batched(int dtype, int batch, int M, int N, int K, double alpha, const void* A,
                        int64_t sA, int64_t rsA, int64_t csA, const void* B, int64_t sB,
                        int64_t rsB, int64_t csB, double beta, void* C, int64_t sC,
                        int64_t ldc, void* stream);

/* ---- learned warm start (no counterpart in the reference) ------------- */

/* bf16 MFMA GEMM with fused epilogue, the linear layers of the warm-start
 * network (forward Y = X W, backward dX = dY W^T, dW = X^T dY):
 *   C[M x N] = alpha * op(A)[M x K] op(B)[K x N]  (+ bias[N])
 *              (+ rowvec[i / seg_len][N]) (leaky relu, slope 0.01)
 *              (* (mask[i][j] > 0 ? 1 : 0.01))
 * op(A)[i][k] = transA ? A[k*lda + i] : A[i*lda + k]; op(B) likewise with
 * op(B)[k][j] = transB ? B[j*ldb + k] : B[k*ldb + j].  A, B, mask bf16.
 * out_kind: 0 bf16 store, 1 fp32 store, 2 fp32 atomic add (split-K over
 * `splits` workgroups; bias / rowvec must then be NULL).
 */
int potamd_gemm_bf16(int M, int N, int K, const void* A, int64_t lda, int transA,
                     const void* B, int64_t ldb, int transB, void* C, int64_t ldc,
                     int out_kind, float alpha, const float* bias, const float* rowvec,
                     int seg_len, int relu, const void* mask, int64_t ldmask, int splits,
                     void* stream);

/* Small kernels of the warm-start network's forward, backward and Adam step,
 * selected by `op` (argument meaning per op in pot_amd/warmstart/net.py):
 *  0 build inputs, 1 segment mean, 2 vector x matrix, 3 vector x matrix^T,
 *  4 output head, 5 centred weighted loss + gradient, 6 head backward,
 *  7 head weight gradient, 8 column sums, 9 segment sums, 10 pooling
 *  backward, 11 outer-product accumulation, 12 Adam step, 13 fp32 -> bf16,
 *  14 scaled squared row norms.
 * hp: 7 floats of hyper-parameters (Adam) or NULL.
 */
int potamd_mlp_op(int op, int rows, int seg_len, int P, int k, int n, int64_t count, int flag,
                  double scalar, const float* hp, const void* x, const void* y, const void* z,
                  const void* w, void* out, void* out2, void* stream);

/* Fused inference forward of one two-layer chain of the warm-start network
 * (bf16 MFMA, fp32 accumulation, leaky ReLU slope 0.01; hidden width 256,
 * input width 32):  A1 = act(inp W1 + bias1), A2 = act(A1 W2 + b2) (A2
 * rounded to bf16), then
 *   mode 0: out[s][:] = mean of A2 over the rows of segment s (bias1 = b1[256]);
 *   mode 1: out[r] = A2[r] . w3 + b3[0] + qn[r] (bias1 = rowvec[segment][256],
 *           qn may be NULL).
 * inp: rows x 32 bf16 input rows (16-byte aligned), or NULL to build them
 *    from pts (rows x 16 fp32, 16-byte aligned), wts (per-row weight, fp64),
 *    regs (fp64) and cmax (fp32) per segment as
 *    [x * rsqrt(cmax), seg_len * w - 1, (log reg + 4.147366) / 1.151293, 0...].
 * W1: 32 x 256 and W2: 256 x 256 bf16, row-major.
 * seg_len must be a multiple of 64 and divide rows.
 */
int potamd_mlp_fused(int mode, int rows, int seg_len, const void* inp, const float* pts,
                     const double* wts, const double* regs, const float* cmax, const void* W1,
                     const float* bias1, const void* W2, const float* b2, const float* w3,
                     const float* b3, const float* qn, float* out, void* stream);

/* ---- Online cost: squared Euclidean cost of two point clouds, never stored ----
 *
 * Replaces the loop of sinkhorn_log (ot/bregman/_sinkhorn.py:880-906) for
 * M = dist(X, Y) / scale (ot/utils.py:400-466, euclidean_distances :270-307)
 * without materialising M; the reference's own lazy path is
 * empirical_sinkhorn(isLazy=True) (ot/bregman/_empirical.py:193-270).
 * Precision of the x . y^T term: */
#define POTAMD_ONLINE_F32 0    /* fp32 MFMA */
#define POTAMD_ONLINE_BF16 1   /* bf16 MFMA, fp32 accumulation */
#define POTAMD_ONLINE_BF16X3 2 /* bf16 hi + lo split, three MFMA products */
#define POTAMD_ONLINE_FP8 3    /* fp8 e4m3 MFMA, per-point scale */
#define POTAMD_ONLINE_F64 4    /* fp64 vector ALU (reference precision) */

/* Workspace of potamd_sinkhorn_online; 0 for unsupported arguments (the MFMA
 * precisions take d <= 128). */
size_t potamd_online_workspace_size(int precision, int n, int m, int d);

/* sinkhorn_log on the online cost, column update first, marginal error
 * ||P^T 1 - b||_2 every 10 iterations, strict `err < stop_thr`, exactly as
 * _sinkhorn.py:880-899.  X (n x d, row stride ldx) and Y (m x d) are device
 * points of `dtype` (POTAMD_F64 only with POTAMD_ONLINE_F64); a, b, u (in:
 * u0, out: log_u), v (out: log_v) fp64 device, natural logs.  cost_scale > 0 divides the cost by it; 0 divides by the
 * maximum cost, computed on the device and returned in *cost_scale_out
 * (host pointer, may be NULL).  splits: C-side splits per pass (0 = auto,
 * at most 64).  semantics 0 is the loop above; 1 is the loop of
 * empirical_sinkhorn(isLazy=True) (ot/bregman/_empirical.py:218-256) seen
 * from its second point set (call it with X = X_t, Y = X_s, a = b_t, b = a_s,
 * u0 = g0): the error is the 1-norm, checked when (ii + 1) % 10 == 0, with a
 * non-strict `err <= stop_thr`.  info as potamd_sinkhorn_dense.  Returns the
 * status. */
int potamd_sinkhorn_online(int precision, int dtype, const void* X, int64_t ldx, const void* Y,
                           int64_t ldy, int n, int m, int d, double reg, double cost_scale,
                           const double* a,
                           const double* b, double* u, double* v, int max_iter, double stop_thr,
                           int semantics, double* err_hist, int err_cap, int splits,
                           void* workspace,
                           size_t ws_bytes, void* host_ctrl, void* stream, int* info,
                           double* cost_scale_out);

/* Workspace of potamd_online_plan. */
size_t potamd_online_plan_workspace_size(int n, int m);

/* Plan, cost and marginals of the online cost in fp64 (_sinkhorn.py:915-918
 * and the cost of sinkhorn2, :436): P_ij = exp(-M_ij / reg + u_i + v_j) with
 * M_ij from the points (fp32 or fp64, `dtype`) in fp64 (clamped at 0, as ot/utils.py:299).
 * P (n x ldp fp64) may be NULL.  sums (device, 4 doubles) = [<P, M>, sum P,
 * ||P 1 - a||_2, ||P^T 1 - b||_2]. */
int potamd_online_plan(int dtype, const void* X, int64_t ldx, const void* Y, int64_t ldy, int n,
                       int m, int d, double reg, double cost_scale, const double* u, const double* v,
                       const double* a, const double* b, double* P, int64_t ldp, double* sums,
                       void* workspace, size_t ws_bytes, void* stream);

/* ---- Greedy Sinkhorn -------------------------------------------------- */

/* Workspace of potamd_greenkhorn: K = exp(-M / reg) and its transpose in
 * fp64 (16 n m bytes) and the two violation vectors; 0 outside the supported
 * sizes (n, m <= 2097120). */
size_t potamd_greenkhorn_workspace_size(int n, int m);

/* ot.bregman.greenkhorn (ot/bregman/_sinkhorn.py:921-1102): the greedy loop
 * of :1058-1086 (argmax of |viol| and |viol_2|, first index on ties, NaN
 * largest; the larger one's row or column scaling updated, the violations
 * updated incrementally; stop = max of the two |violations| read before the
 * update, stop <= stop_thr ends the loop after it), in fp64 whatever the
 * cost's dtype, in one launch.
 * u, v: fp64 device scalings, in: the start (1/n, 1/m, or exp of a warm
 *    start), out: the final ones.
 * P: fp64 device n x m, or NULL: the plan u_i K_ij v_j of the final scalings.
 * info (host, 4 ints): [0] the reference's `ii` at exit (-1 when
 *    max_iter < 1), [1] 1 when the stop rule ended the loop, [2] status.
 * Returns POTAMD_CONVERGED, or POTAMD_MAX_ITER (the reference warns).
 * Synchronises the stream. */
int potamd_greenkhorn(int dtype, const void* M, int64_t ldm, int n, int m, double reg,
                      const double* a, const double* b, double* u, double* v, int max_iter,
                      double stop_thr, double* P, void* workspace, size_t ws_bytes, void* stream,
                      int* info);

/* ---- Log-stabilised Sinkhorn ------------------------------------------ */

/* Workspace of potamd_sinkhorn_stabilized (holds K in the cost's dtype). */
size_t potamd_stabilized_workspace_size(int dtype, int n, int m);

/* ot.bregman.sinkhorn_stabilized (ot/bregman/_sinkhorn.py:1105-1367), one
 * histogram: u, v start at 1/dim; v = b / K^T u, u = a / K v with
 * K = exp(-(M - alpha - beta) / reg); when max|u| or max|v| exceeds tau,
 * alpha += reg log u, beta += reg log v, u, v restart at 1/dim and K is
 * rebuilt; every print_period-th iteration err = ||Gamma^T 1 - b||_2
 * (err_hist); err <= stop_thr stops; a NaN restores the previous u, v and
 * returns POTAMD_NUMERICAL.  alpha, beta (fp64 device, in: warm start) hold
 * the absorbed potentials on exit and u, v (fp64 device, out) the scalings,
 * so log u = alpha / reg + log(u) as in the reference's log.  marg_sq (device,
 * 2 doubles, may be NULL) receives ||Gamma^T 1 - b||^2 and ||Gamma 1 - a||^2
 * of the final plan (the check of sinkhorn_epsilon_scaling, :1502-1505).
 * info as potamd_sinkhorn_dense. */
int potamd_sinkhorn_stabilized(int dtype, const void* M, int64_t ldm, int n, int m, double reg,
                               const double* a, const double* b, double* alpha, double* beta,
                               double* u, double* v, double tau, int max_iter, double stop_thr,
                               int print_period, double* err_hist, int err_cap, double* marg_sq,
                               void* workspace, size_t ws_bytes, void* host_ctrl, void* stream,
                               int* info);

/* The several-histogram branch of ot.bregman.sinkhorn_stabilized (a 2-D b,
 * ot/bregman/_sinkhorn.py:1230-1367): H target histograms against one M.
 * b: H x m histogram-major; u: H x n, v: H x m (out: the final scalings);
 * alpha (n), beta (m): in the warm start, out the absorbed potentials.
 * Per iteration v = b / (K^T u), u = a / (K v) for every histogram; when
 * max|u| or max|v| exceeds tau, alpha += reg max_h log u_h (row-wise),
 * beta += reg max log v (one maximum over all of v) and K is rebuilt, with
 * u and v kept (the reference resets them only for one histogram); every
 * print_period-th iteration err = (max|u - u'| / max(max|u|, max|u'|, 1) +
 * the same for v) / 2 is logged; err <= stop_thr stops; a NaN restores the
 * previous u, v and stops (status POTAMD_NUMERICAL).  info as for
 * potamd_sinkhorn_stabilized.  Workspace:
 * potamd_stabilized_multi_workspace_size.
 */
size_t potamd_stabilized_multi_workspace_size(int dtype, int n, int m, int H);
int potamd_sinkhorn_stabilized_multi(int dtype, const void* M, int64_t ldm, int n, int m, int H,
                                     double reg, const double* a, const double* b,
                                     double* alpha, double* beta, double* u, double* v,
                                     double tau, int max_iter, double stop_thr, int print_period,
                                     double* err_hist, int err_cap, void* workspace,
                                     size_t ws_bytes, void* host_ctrl, void* stream, int* info);

#ifdef __cplusplus
}
#endif

#endif /* POT_AMD_H */
