// Internal C++ interface between the kernel files and the C-ABI layer.
#pragma once

#include "common.h"

namespace potamd {

struct RowSweepArgs {
  const void* M;      // B problems of n x ldm elements (TM)
  int64_t ldm;        // row stride in elements
  int64_t bstride;    // problem stride in elements
  int n, m, rpb;      // rows, columns, rows per workgroup
  float s_hi, s_lo;   // fp32 split of s2 (fp32 costs)
  double s2;          // s2 in full (fp64 costs)
  int kernel_given;   // scaling domain: M already holds K
  int row_first;      // emit row-marginal error partials
  int iter;           // iteration index (flag slot parity)
  const double* a;    // B x n
  const double* loga2;
  const double* u_old;  // B x n
  double* u_new;        // B x n
  const double* v;      // B x m (current column potential / scaling)
  void* colpart;        // B x G x m partial column sums (float or double)
  double* rowerr;       // B x G partial row-marginal err^2
  int* bad;             // [2] numerical-guard slots (scaling domain)
  const Ctrl* ctrl;
  const int* pdone;     // per-problem stop: pdone[p] != 0 skips problem p (or null)
};

struct MergeArgs {
  int n, m, G, B;
  int iter, max_iter;
  int order, domain;
  const double* b;      // B x m
  const double* logb2;  // B x m
  const double* v_cur;  // B x m
  double* v_next;       // B x m
  const void* colpart;  // B x G x m (float or double)
  const double* rowerr; // B x G
  double* errpart;      // B x gridDim.x
  int* bad;             // [2]
  Ctrl* ctrl;
  double* err_hist;
  int err_cap;
  double stop_thr;
  int prologue;         // scaling prologue: v1 = b / (K^T u0), no check
  int err_norm;         // batch error: 0 max over problems, 1 Frobenius, 2 each
                        // problem stops on its own error (per-problem state in pstate)
  int* pstate;          // err_norm 2: [B] done, [B] niter, [B] status
};

struct ColLseArgs {
  const void* M;
  int64_t ldm, bstride;
  int n, m, rpb;
  double s2;
  const double* u2;   // B x n
  double* pmax;       // B x G x m
  double* psum;
};

struct PlanArgs {
  const void* M;
  int64_t ldm, bstride;
  int n, m, rpb;
  int domain, kernel_given;
  double reg;
  const double* u;   // B x n natural-log potentials (log) or scalings
  const double* v;   // B x m
  const double* a;   // B x n (KL sums; may be null)
  const double* b;   // B x m (KL sums; may be null)
  void* P;           // plan out (dtype of M, row stride m), may be null
  void* logP;        // log plan out (log domain), may be null
  int64_t pstride;   // problem stride of P / logP
  double* partial;   // B x blocks x 4
};

struct DistArgs {
  const void* X;
  const void* Y;
  void* M;
  int n, m, d;
  int64_t xstride, ystride, mstride, ldm;
  int squared, clamp, zero_diag;
};

struct BatchArgs {
  const void* M;         // packed costs (float)
  const int64_t* moff;   // [B] element offset of problem p's cost
  const int* ns;         // [B] rows (<= 1024)
  const int* ms;         // [B] columns (<= 1024)
  const int* ldms;       // [B] row stride
  const int64_t* aoff;   // [B] offset of problem p in a / u (length n_p)
  const int64_t* boff;   // [B] offset of problem p in b / v (length m_p)
  const double* regs;    // [B]
  const double* a;
  const double* b;
  double* u;             // in: u0 (natural log), out: log_u
  double* v;             // out: log_v
  const int* order;      // [B] launch order (largest first) or null
  int max_iter;
  int rotate;            // start each problem's row sweep at a problem-dependent row
  int check_every;       // iterations between marginal checks (10 in the reference)
  double stop_thr;
  double fallback_thr;   // > 0: restart cold when the first check exceeds it
  int* niter;            // [B]
  int* status;           // [B]
  double* err_last;      // [B] last checked error (may be null)
  double* err_first;     // [B] error at the first check (may be null)
};

int batch_cpt(int max_m);
int batch_waves_default();
int batch_occupancy_default(int waves, int pf);
hipError_t launch_batch_log(int cpt, int vec_ok, int waves, int pf, int occ,
                            const BatchArgs& a, int B, hipStream_t s);

struct RaggedDistArgs {
  const float* X;        // packed points, problem p rows at xoff[p]
  const float* Y;
  const int64_t* xoff;   // [B] point offsets (rows)
  const int64_t* yoff;
  const int* ns;
  const int* ms;
  int d;
  float* M;
  const int64_t* moff;
  const int* ldms;
  float* maxout;         // [B] per-problem maximum cost (may be null)
};

hipError_t launch_sqeuclid_ragged(const RaggedDistArgs& a, int B, int max_n, int max_m,
                                  hipStream_t s);

struct GemmArgs {
  int M, N, K;
  const void* A;   // bf16
  int64_t lda;
  int transA;      // op(A)[i][k] = transA ? A[k*lda + i] : A[i*lda + k]
  const void* B;   // bf16
  int64_t ldb;
  int transB;      // op(B)[k][j] = transB ? B[j*ldb + k] : B[k*ldb + j]
  void* C;
  int64_t ldc;
  int out_kind;    // 0 bf16 store, 1 fp32 store, 2 fp32 atomic add
  float alpha;
  const float* bias;    // [N] or null
  const float* rowvec;  // [M / seg_len][N] or null: row i adds rowvec[i / seg_len]
  int seg_len;
  int relu;
  const void* mask;     // bf16 [M][ldmask] or null: zero where mask <= 0
  int64_t ldmask;
};

// Warm-start network input features: the weight column enters as
// seg_len * w - 1 (0 for uniform weights) and log reg as
// (log reg - log 0.0158) / 1.1513, i.e. [-1, 1] over reg in [5e-3, 5e-2].
constexpr double kLogRegCenter = -4.147366;
constexpr double kLogRegScale = 1.151293;

enum MlpOpCode : int {
  kOpBuildInputs = 0,
  kOpSegmentMean = 1,
  kOpVecMat = 2,
  kOpVecMatT = 3,
  kOpHead = 4,
  kOpLoss = 5,
  kOpHeadBack = 6,
  kOpHeadWGrad = 7,
  kOpColSum = 8,
  kOpSegmentSum = 9,
  kOpPoolBack = 10,
  kOpOuterAcc = 11,
  kOpAdam = 12,
  kOpToBf16 = 13,
  kOpRowNorm = 14,
};

struct MlpOp {
  int op, rows, seg_len, P, k, n;
  int64_t count;
  int flag;
  double scalar;
  float hp[7];
  const void* x;
  const void* y;
  const void* z;
  const void* w;
  void* out;
  void* out2;
};

hipError_t launch_gemm_bf16(const GemmArgs& g, int splits, hipStream_t s);
hipError_t launch_mlp_op(const MlpOp& o, hipStream_t s);

int plan_blocks(int n);
hipError_t launch_plan(int dtype, PlanArgs a, int B, double* sums, hipStream_t s);
hipError_t launch_sqeuclid(int dtype, const DistArgs& a, int B, hipStream_t s);

bool row_config_valid(int dtype, int T, int cpt, int pf);
void row_config_default(int dtype, int m, int* T, int* cpt, int* pf);
hipError_t launch_row_sweep(int dtype, int vec_ok, int T, int cpt, int pf, int domain, int mode,
                            const RowSweepArgs& a, int G, int B, hipStream_t s);
hipError_t launch_col_merge(int dtype, const MergeArgs& a, int B, hipStream_t s);
hipError_t launch_col_lse(int dtype, const ColLseArgs& a, int G, int B, const double* logb2,
                          double* v_out, hipStream_t s);
hipError_t launch_log2(int64_t count, const double* x, double* y, hipStream_t s);
hipError_t launch_scale_copy(int64_t count, const double* x, double f, double* y, hipStream_t s);
hipError_t launch_fill(int64_t count, double val, double* y, hipStream_t s);
hipError_t launch_ctrl_init(Ctrl* c, int* bad, hipStream_t s);
hipError_t launch_pstate_init(int* pstate, int B, hipStream_t s);
hipError_t launch_gather_final(int B, int n, int m, const int* pstate, double* const* U,
                               double* const* V, double f, double* u, double* v,
                               int* niters, int* statuses, hipStream_t s);

// Point-cloud batch solver (batch_points.hip).  The split exchange slots
// are kPointsSplitSlot bytes each; a slot's counters live in its first
// kPointsSplitCounters bytes.
constexpr size_t kPointsSplitSlot = 66048;
constexpr size_t kPointsSplitCounters = 256;
int points_validate(int batch, int max_n, int max_m, int d, const float* X, const float* Y,
                    const int64_t* xoff, const int64_t* yoff, const int* ns, const int* ms,
                    const double* regs, int normalize, const double* a, const double* b,
                    const double* u, const double* v, int max_iter, int check_every,
                    const int* niter, const int* status, const double* cmax, int waves,
                    bool queued);
int points_launch(int batch, int max_n, int max_m, int d, const float* X, const float* Y,
                  const int64_t* xoff, const int64_t* yoff, const int* ns, const int* ms,
                  const double* regs, int normalize, const double* a, const double* b, double* u,
                  double* v, int max_iter, int check_every, double stop_thr,
                  double fallback_thr, const int* order, int* niter, int* status,
                  double* err_last, double* err_first, double* cmax, int waves, int* queue,
                  int items, const int* split, char* split_ws, int nsplit, void* stream);

}  // namespace potamd
