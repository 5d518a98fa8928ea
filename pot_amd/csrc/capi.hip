// C ABI (include/pot_amd.h): argument checks, workspace carving and the
// host-side iteration driver of the dense solver.
//
// The driver enqueues the per-iteration kernels asynchronously.  The merge
// kernel of a check iteration takes the stop decision on the device and sets
// Ctrl::done, after which every later kernel returns at once; the host polls
// that flag through an asynchronous copy into pinned memory, at most
// kRing check intervals behind the GPU, so the GPU is never idle waiting for
// the host and no iteration is enqueued long after the stop.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "../../include/pot_amd.h"
#include "host.h"
#include "internal.h"

using namespace potamd;

int potamd::device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  return cus;
}

namespace {

// Stop checks in flight: the host runs at most kRing checks (10 iterations
// each) ahead of the device; the iterations it queues past the device's stop
// exit at once but still cost their launches.
constexpr int kRing = 2;
constexpr int kColLseRows = 256;

int default_rpb(int n) {
  // About two workgroups per CU, at least 4 rows each (the column partials
  // then cost <= 1/4 of the cost matrix traffic), at most kMaxRowsPerBlock.
  const int target = 2 * device_cus();
  int rpb = (n + target - 1) / target;
  return std::min(256, std::max(4, rpb));
}

struct DenseLayout {
  double* U[2];
  double* V[2];
  double* loga2;
  double* logb2;
  void* colpart;
  double* rowerr;
  double* errpart;
  double* pmax;
  double* psum;
  Ctrl* ctrl;
  int* bad;
  int* pstate;  // per-problem stop: [B] done, [B] niter, [B] status
  size_t bytes;
};

DenseLayout carve_dense(void* ws, int dtype, int B, int n, int m, int rpb) {
  Carver c(ws);
  DenseLayout L;
  const int G = (n + rpb - 1) / rpb;
  const int G2 = (n + kColLseRows - 1) / kColLseRows;
  const size_t es = dtype == POTAMD_F32 ? 4 : 8;
  L.U[0] = c.take<double>((size_t)B * n);
  L.U[1] = c.take<double>((size_t)B * n);
  L.V[0] = c.take<double>((size_t)B * m);
  L.V[1] = c.take<double>((size_t)B * m);
  L.loga2 = c.take<double>((size_t)B * n);
  L.logb2 = c.take<double>((size_t)B * m);
  L.colpart = c.take<char>((size_t)B * G * m * es);
  L.rowerr = c.take<double>((size_t)B * G);
  L.errpart = c.take<double>((size_t)B * ((m + 63) / 64));  // one per merge block
  L.pmax = c.take<double>((size_t)B * G2 * m);
  L.psum = c.take<double>((size_t)B * G2 * m);
  L.ctrl = c.take<Ctrl>(1);
  L.bad = c.take<int>(2);
  L.pstate = c.take<int>((size_t)3 * B);
  L.bytes = c.off;
  return L;
}

void split_f32(double x, float* hi, float* lo) {
  *hi = (float)x;
  *lo = (float)(x - (double)*hi);
}

}  // namespace

extern "C" int potamd_version(void) { return 100; }

extern "C" const char* potamd_status_string(int code) {
  switch (code) {
    case POTAMD_CONVERGED: return "converged";
    case POTAMD_MAX_ITER: return "maximum number of iterations reached";
    case POTAMD_NUMERICAL: return "numerical errors (zero, nan or inf scaling)";
    case POTAMD_NEED_EXACT: return "column mass below the fp32 range";
    case POTAMD_EINVAL: return "invalid argument";
    case POTAMD_EHIP: return "HIP runtime error";
    case POTAMD_EUNSUPPORTED: return "shape not supported by the compiled kernels";
    case POTAMD_EWORKSPACE: return "workspace too small";
    default: return "unknown status";
  }
}

extern "C" size_t potamd_dense_workspace_size(int dtype, int batch, int n, int m,
                                              int rows_per_block) {
  if (batch < 1 || n < 1 || m < 1) return 0;
  const int rpb = rows_per_block > 0 ? rows_per_block : default_rpb(n);
  return carve_dense(nullptr, dtype, batch, n, m, rpb).bytes;
}

namespace {

// The dense solve behind potamd_sinkhorn_dense (err_norm 0 or 1: one stop
// for the batch) and potamd_sinkhorn_dense_each (err_norm 2: each problem
// stops on its own check; niters / statuses receive the per-problem results).
int dense_impl(int dtype, int domain, int order, int kernel_given, const void* M, int64_t ldm,
               int64_t bstride, int batch, int n, int m, double reg, const double* a,
               const double* b, double* u, double* v, int max_iter, double stop_thr,
               int err_norm, double* err_hist, int err_cap, int threads, int cols_per_thread,
               int rows_per_block, int prefetch, void* workspace, size_t ws_bytes,
               void* host_ctrl, void* stream, int* info, int* niters, int* statuses) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || (domain != POTAMD_LOG && domain != POTAMD_SCALING) ||
      (order != POTAMD_COL_FIRST && order != POTAMD_ROW_FIRST) || !M || !a || !b || !u || !v ||
      !info || batch < 1 || n < 1 || m < 1 || ldm < m || bstride < 0 ||
      err_norm < 0 || err_norm > 2 ||
      max_iter < 1 || (!kernel_given && !(reg > 0.0) && domain == POTAMD_LOG) ||
      (err_cap > 0 && !err_hist))
    return POTAMD_EINVAL;
  const bool each = err_norm == 2;
  const int rpb = rows_per_block > 0 ? rows_per_block : default_rpb(n);
  if (rpb > 256) return POTAMD_EINVAL;
  int T = threads, cpt = cols_per_thread, pf = prefetch;
  if (T == 0 || cpt == 0) {
    int pf_default = 1;
    row_config_default(dtype, m, &T, &cpt, &pf_default);
    if (pf == 0) pf = pf_default;
  }
  if (pf == 0) pf = 1;
  if (T == 0 || !row_config_valid(dtype, T, cpt, pf) || (int64_t)T * cpt < m)
    return POTAMD_EUNSUPPORTED;
  const size_t es = dtype == POTAMD_F32 ? 4 : 8;
  const int vec = dtype == POTAMD_F32 ? 4 : 2;
  const int vec_ok = ((uintptr_t)M % 16 == 0) && (ldm % vec == 0) && (bstride % vec == 0) &&
                     (m % vec == 0);
  (void)es;
  DenseLayout L = carve_dense(workspace, dtype, batch, n, m, rpb);
  if (!workspace || ws_bytes < L.bytes) return POTAMD_EWORKSPACE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int G = (n + rpb - 1) / rpb;
  const int64_t Bn = (int64_t)batch * n, Bm = (int64_t)batch * m;
  const bool logd = domain == POTAMD_LOG;

  // s2: element exponent scale in base 2.
  double s2;
  if (kernel_given)
    s2 = logd ? kLog2e : 1.0;
  else
    s2 = -kLog2e / reg;
  float s_hi, s_lo;
  split_f32(s2, &s_hi, &s_lo);

  // Set-up: log2 weights, initial state, control block.
  HIP_TRY(launch_ctrl_init(L.ctrl, L.bad, s));
  if (each) HIP_TRY(launch_pstate_init(L.pstate, batch, s));
  HIP_TRY(launch_log2(Bn, a, L.loga2, s));
  HIP_TRY(launch_log2(Bm, b, L.logb2, s));
  HIP_TRY(launch_scale_copy(Bn, u, logd ? kLog2e : 1.0, L.U[0], s));
  HIP_TRY(launch_scale_copy(Bm, v, logd ? kLog2e : 1.0, L.V[0], s));

  RowSweepArgs ra;
  memset(&ra, 0, sizeof(ra));
  ra.M = M;
  ra.ldm = ldm;
  ra.bstride = bstride;
  ra.n = n;
  ra.m = m;
  ra.rpb = rpb;
  ra.s_hi = s_hi;
  ra.s_lo = s_lo;
  ra.s2 = s2;
  ra.kernel_given = kernel_given;
  ra.row_first = order == POTAMD_ROW_FIRST;
  ra.a = a;
  ra.loga2 = L.loga2;
  ra.colpart = L.colpart;
  ra.rowerr = L.rowerr;
  ra.bad = L.bad;
  ra.ctrl = L.ctrl;
  ra.pdone = each ? L.pstate : nullptr;

  MergeArgs ma;
  memset(&ma, 0, sizeof(ma));
  ma.n = n;
  ma.m = m;
  ma.G = G;
  ma.B = batch;
  ma.max_iter = max_iter;
  ma.order = order;
  ma.domain = domain;
  ma.b = b;
  ma.logb2 = L.logb2;
  ma.colpart = L.colpart;
  ma.rowerr = L.rowerr;
  ma.errpart = L.errpart;
  ma.bad = L.bad;
  ma.ctrl = L.ctrl;
  ma.err_hist = err_hist;
  ma.err_cap = err_cap;
  ma.stop_thr = stop_thr;
  ma.err_norm = err_norm;
  ma.pstate = each ? L.pstate : nullptr;

  // Prologue of the column-first order: v_1 from u_0.
  if (order == POTAMD_COL_FIRST) {
    if (logd) {
      ColLseArgs ca;
      ca.M = M;
      ca.ldm = ldm;
      ca.bstride = bstride;
      ca.n = n;
      ca.m = m;
      ca.rpb = kColLseRows;
      ca.s2 = s2;
      ca.u2 = L.U[0];
      ca.pmax = L.pmax;
      ca.psum = L.psum;
      const int G2 = (n + kColLseRows - 1) / kColLseRows;
      HIP_TRY(launch_col_lse(dtype, ca, G2, batch, L.logb2, L.V[1], s));
    } else {
      ra.iter = -1;
      ra.u_old = L.U[0];
      ra.u_new = L.U[1];
      ra.v = L.V[0];
      HIP_TRY(launch_row_sweep(dtype, vec_ok, T, cpt, pf, domain, 1, ra, G, batch, s));
      ma.iter = -1;
      ma.prologue = 1;
      ma.v_cur = L.V[0];
      ma.v_next = L.V[1];
      HIP_TRY(launch_col_merge(dtype, ma, batch, s));
      ma.prologue = 0;
    }
  }

  Ctrl* hc = static_cast<Ctrl*>(host_ctrl);
  Ctrl local;
  hipEvent_t ev[kRing];
  bool ev_used[kRing] = {};
  for (int k = 0; k < kRing; ++k) HIP_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
  int ring_pos = 0;
  int ring_head = 0;  // oldest check whose copy the host has not read
  bool stopped_early = false;
  int rc = 0;
  const int vo = order == POTAMD_COL_FIRST ? 1 : 0;

  for (int ii = 0; ii < max_iter && !stopped_early; ++ii) {
    // Checks the device has already passed end the loop without a host wait.
    while (hc && ring_head < ring_pos) {
      const int k = ring_head % kRing;
      const hipError_t q = hipEventQuery(ev[k]);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) {
        rc = POTAMD_EHIP;
        break;
      }
      ev_used[k] = false;
      ++ring_head;
      if (hc[k].done) stopped_early = true;
    }
    if (rc || stopped_early) break;
    ra.iter = ii;
    ra.u_old = L.U[ii & 1];
    ra.u_new = L.U[(ii + 1) & 1];
    ra.v = L.V[(ii + vo) & 1];
    if (launch_row_sweep(dtype, vec_ok, T, cpt, pf, domain, 0, ra, G, batch, s) != hipSuccess) {
      rc = POTAMD_EHIP;
      break;
    }
    ma.iter = ii;
    ma.v_cur = L.V[(ii + vo) & 1];
    ma.v_next = L.V[(ii + vo + 1) & 1];
    if (launch_col_merge(dtype, ma, batch, s) != hipSuccess) {
      rc = POTAMD_EHIP;
      break;
    }
    const bool check = order == POTAMD_COL_FIRST ? (ii % 10 == 0) : (ii >= 1 && (ii - 1) % 10 == 0);
    if (check && ii + 1 < max_iter) {
      const int k = ring_pos % kRing;
      if (ev_used[k]) {
        if (hipEventSynchronize(ev[k]) != hipSuccess) { rc = POTAMD_EHIP; break; }
        ev_used[k] = false;
        ++ring_head;
        const Ctrl* snap = hc ? &hc[k] : &local;
        if (snap->done) stopped_early = true;
      }
      if (!stopped_early) {
        if (hc) {
          if (hipMemcpyAsync(&hc[k], L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess) {
            rc = POTAMD_EHIP;
            break;
          }
          if (hipEventRecord(ev[k], s) != hipSuccess) { rc = POTAMD_EHIP; break; }
          ev_used[k] = true;
        } else {
          if (hipMemcpyAsync(&local, L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess) {
            rc = POTAMD_EHIP;
            break;
          }
          if (local.done) stopped_early = true;
        }
        ++ring_pos;
      }
    }
  }
  for (int k = 0; k < kRing; ++k) hipEventDestroy(ev[k]);
  if (rc) return rc;

  HIP_TRY(hipMemcpyAsync(&local, L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  int status = local.status;
  int niter = local.niter;
  if (!local.done) {  // cannot happen: the last merge always decides
    status = POTAMD_MAX_ITER;
    niter = max_iter - 1;
  }
  int fu, fv;
  if (status == POTAMD_NUMERICAL) {
    fu = niter & 1;
    fv = niter & 1;
  } else {
    fu = (niter + 1) & 1;
    fv = (niter + 1) & 1;
  }
  if (status != POTAMD_NEED_EXACT) {
    if (each) {
      HIP_TRY(launch_gather_final(batch, n, m, L.pstate, L.U, L.V, logd ? kLn2 : 1.0, u, v,
                                  niters, statuses, s));
    } else {
      HIP_TRY(launch_scale_copy(Bn, L.U[fu], logd ? kLn2 : 1.0, u, s));
      HIP_TRY(launch_scale_copy(Bm, L.V[fv], logd ? kLn2 : 1.0, v, s));
    }
  }
  info[0] = niter;
  info[1] = local.n_err;
  info[2] = status;
  info[3] = local.bad_iter;
  return status;
}

}  // namespace

extern "C" int potamd_sinkhorn_dense(int dtype, int domain, int order, int kernel_given,
                                     const void* M, int64_t ldm, int64_t bstride, int batch,
                                     int n, int m, double reg, const double* a, const double* b,
                                     double* u, double* v, int max_iter, double stop_thr,
                                     int err_norm, double* err_hist, int err_cap, int threads,
                                     int cols_per_thread, int rows_per_block, int prefetch,
                                     void* workspace,
                                     size_t ws_bytes, void* host_ctrl, void* stream, int* info) {
  if (err_norm != 0 && err_norm != 1) return POTAMD_EINVAL;
  return dense_impl(dtype, domain, order, kernel_given, M, ldm, bstride, batch, n, m, reg, a, b,
                    u, v, max_iter, stop_thr, err_norm, err_hist, err_cap, threads,
                    cols_per_thread, rows_per_block, prefetch, workspace, ws_bytes, host_ctrl,
                    stream, info, nullptr, nullptr);
}

extern "C" int potamd_sinkhorn_dense_each(int dtype, int domain, int order, int kernel_given,
                                          const void* M, int64_t ldm, int64_t bstride,
                                          int batch, int n, int m, double reg, const double* a,
                                          const double* b, double* u, double* v, int max_iter,
                                          double stop_thr, int threads, int cols_per_thread,
                                          int rows_per_block, int prefetch, void* workspace,
                                          size_t ws_bytes, void* host_ctrl, void* stream,
                                          int* niters, int* statuses, int* info) {
  if (!niters || !statuses) return POTAMD_EINVAL;
  return dense_impl(dtype, domain, order, kernel_given, M, ldm, bstride, batch, n, m, reg, a, b,
                    u, v, max_iter, stop_thr, 2, nullptr, 0, threads, cols_per_thread,
                    rows_per_block, prefetch, workspace, ws_bytes, host_ctrl, stream, info,
                    niters, statuses);
}

extern "C" size_t potamd_plan_workspace_size(int batch, int n) {
  if (batch < 1 || n < 1) return 0;
  return align_up((size_t)batch * plan_blocks(n) * 4 * sizeof(double));
}

extern "C" int potamd_plan_dense(int dtype, int domain, int kernel_given, const void* M,
                                 int64_t ldm, int64_t bstride, int batch, int n, int m,
                                 double reg, const double* u, const double* v, const double* a,
                                 const double* b, void* P, void* logP, int64_t pstride,
                                 double* sums, void* workspace, size_t ws_bytes, void* stream) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || !M || !u || !v || batch < 1 || n < 1 ||
      m < 1 || ldm < m || (!kernel_given && domain == POTAMD_LOG && !(reg > 0.0)) ||
      ((P || logP) && batch > 1 && pstride < (int64_t)n * m))
    return POTAMD_EINVAL;
  if (!workspace || ws_bytes < potamd_plan_workspace_size(batch, n)) return POTAMD_EWORKSPACE;
  PlanArgs pa;
  memset(&pa, 0, sizeof(pa));
  pa.M = M;
  pa.ldm = ldm;
  pa.bstride = bstride;
  pa.n = n;
  pa.m = m;
  pa.domain = domain;
  pa.kernel_given = kernel_given;
  pa.reg = reg;
  pa.u = u;
  pa.v = v;
  pa.a = (a && b) ? a : nullptr;
  pa.b = (a && b) ? b : nullptr;
  pa.P = P;
  pa.logP = logP;
  pa.pstride = pstride;
  pa.partial = static_cast<double*>(workspace);
  HIP_TRY(launch_plan(dtype, pa, batch, sums, static_cast<hipStream_t>(stream)));
  return 0;
}

extern "C" int potamd_sqeuclidean(int dtype, const void* X, const void* Y, int batch, int n,
                                  int m, int d, int64_t xstride, int64_t ystride, void* M,
                                  int64_t ldm, int64_t mstride, int squared, int clamp,
                                  int zero_diag, void* stream) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || !X || !Y || !M || batch < 1 || n < 1 ||
      m < 1 || d < 1 || ldm < m)
    return POTAMD_EINVAL;
  DistArgs da;
  da.X = X;
  da.Y = Y;
  da.M = M;
  da.n = n;
  da.m = m;
  da.d = d;
  da.xstride = xstride;
  da.ystride = ystride;
  da.mstride = mstride;
  da.ldm = ldm;
  da.squared = squared;
  da.clamp = clamp;
  da.zero_diag = zero_diag;
  HIP_TRY(launch_sqeuclid(dtype, da, batch, static_cast<hipStream_t>(stream)));
  return 0;
}

extern "C" int potamd_sinkhorn_log_batch(int dtype, int batch, int max_n, int max_m,
                                         const void* M, const int64_t* moff, const int* ns,
                                         const int* ms, const int* ldms, const int64_t* aoff,
                                         const int64_t* boff, const double* regs,
                                         const double* a, const double* b, double* u,
                                         double* v, int max_iter, int check_every,
                                         double stop_thr, double fallback_thr, const int* order,
                                         int vec_aligned,
                                         int* niter, int* status, double* err_last,
                                         double* err_first, int waves, void* stream) {
  if (batch < 1 || max_n < 1 || max_m < 1 || !M || !moff || !ns || !ms || !ldms || !aoff ||
      !boff || !regs || !a || !b || !u || !v || max_iter < 1 || !niter || !status ||
      check_every < 0)
    return POTAMD_EINVAL;
  if (dtype != POTAMD_F32) return POTAMD_EUNSUPPORTED;
  const int cpt = batch_cpt(max_m);
  if (cpt == 0 || max_n > 1024) return POTAMD_EUNSUPPORTED;
  BatchArgs ba;
  memset(&ba, 0, sizeof(ba));
  ba.M = M;
  ba.moff = moff;
  ba.ns = ns;
  ba.ms = ms;
  ba.ldms = ldms;
  ba.aoff = aoff;
  ba.boff = boff;
  ba.regs = regs;
  ba.a = a;
  ba.b = b;
  ba.u = u;
  ba.v = v;
  ba.order = order;
  ba.max_iter = max_iter;
  ba.check_every = check_every > 0 ? check_every : 10;
  ba.stop_thr = stop_thr;
  ba.fallback_thr = fallback_thr;
  ba.niter = niter;
  ba.status = status;
  ba.err_last = err_last;
  ba.err_first = err_first;
  const int vec_ok = vec_aligned && ((uintptr_t)M % 16 == 0);
  // waves: low byte = wavefronts per problem (4 or 8), next byte = rows
  // prefetched per wave (1 or 2; 2 only with 1024-column tiles), third byte =
  // register cap in waves per SIMD (0 = library default, 4 = two 8-wave
  // workgroups per CU, 255 = none), bit 24 = rotate each problem's row order.
  int pf = (waves >> 8) & 0xff;
  int occ = (waves >> 16) & 0xff;
  ba.rotate = (waves >> 24) & 1;
  waves &= 0xff;
  if (waves == 0) waves = batch_waves_default();
  if (pf == 0) pf = 1;
  if (occ == 0) occ = batch_occupancy_default(waves, pf);
  if (occ == 255) occ = 0;
  if ((waves != 4 && waves != 8) || (pf != 1 && pf != 2) || (pf == 2 && cpt != 16) ||
      (occ != 0 && occ != 4) || (occ == 4 && (waves != 8 || pf != 1)))
    return POTAMD_EINVAL;
  HIP_TRY(launch_batch_log(cpt, vec_ok, waves, pf, occ, ba, batch,
                           static_cast<hipStream_t>(stream)));
  return 0;
}

extern "C" int potamd_sqeuclidean_ragged(int batch, int max_n, int max_m, int d, const float* X,
                                         const float* Y, const int64_t* xoff,
                                         const int64_t* yoff, const int* ns, const int* ms,
                                         float* M, const int64_t* moff, const int* ldms,
                                         float* maxout, void* stream) {
  if (batch < 1 || max_n < 1 || max_m < 1 || d < 1 || !X || !Y || !xoff || !yoff || !ns || !ms ||
      !M || !moff || !ldms)
    return POTAMD_EINVAL;
  RaggedDistArgs ra;
  ra.X = X;
  ra.Y = Y;
  ra.xoff = xoff;
  ra.yoff = yoff;
  ra.ns = ns;
  ra.ms = ms;
  ra.d = d;
  ra.M = M;
  ra.moff = moff;
  ra.ldms = ldms;
  ra.maxout = maxout;
  HIP_TRY(launch_sqeuclid_ragged(ra, batch, max_n, max_m, static_cast<hipStream_t>(stream)));
  return 0;
}

extern "C" int potamd_gemm_bf16(int M, int N, int K, const void* A, int64_t lda, int transA,
                                const void* B, int64_t ldb, int transB, void* C, int64_t ldc,
                                int out_kind, float alpha, const float* bias, const float* rowvec,
                                int seg_len, int relu, const void* mask, int64_t ldmask,
                                int splits, void* stream) {
  if (M < 1 || N < 1 || K < 1 || !A || !B || !C || out_kind < 0 || out_kind > 2 || splits < 1 ||
      (rowvec && seg_len < 1) || (splits > 1 && (out_kind != 2 || bias || rowvec)))
    return POTAMD_EINVAL;
  if ((!transA && lda < K) || (transA && lda < M) || (!transB && ldb < N) || (transB && ldb < K) ||
      ldc < N || (mask && ldmask < N))
    return POTAMD_EINVAL;
  GemmArgs g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.lda = lda;
  g.transA = transA;
  g.B = B;
  g.ldb = ldb;
  g.transB = transB;
  g.C = C;
  g.ldc = ldc;
  g.out_kind = out_kind;
  g.alpha = alpha;
  g.bias = bias;
  g.rowvec = rowvec;
  g.seg_len = seg_len > 0 ? seg_len : 1;
  g.relu = relu;
  g.mask = mask;
  g.ldmask = ldmask;
  HIP_TRY(launch_gemm_bf16(g, splits, static_cast<hipStream_t>(stream)));
  return 0;
}

extern "C" int potamd_mlp_op(int op, int rows, int seg_len, int P, int k, int n, int64_t count,
                             int flag, double scalar, const float* hp, const void* x,
                             const void* y, const void* z, const void* w, void* out, void* out2,
                             void* stream) {
  if (op < 0 || op > kOpRowNorm || !out) return POTAMD_EINVAL;
  if ((op == kOpColSum || op == kOpHeadWGrad) && (k < 1 || k > 1024)) return POTAMD_EINVAL;
  if ((op == kOpSegmentMean || op == kOpSegmentSum || op == kOpPoolBack || op == kOpLoss ||
       op == kOpBuildInputs || op == kOpRowNorm) && seg_len < 1)
    return POTAMD_EINVAL;
  MlpOp o;
  o.op = op;
  o.rows = rows;
  o.seg_len = seg_len;
  o.P = P;
  o.k = k;
  o.n = n;
  o.count = count;
  o.flag = flag;
  o.scalar = scalar;
  for (int q = 0; q < 7; ++q) o.hp[q] = hp ? hp[q] : 0.f;
  o.x = x;
  o.y = y;
  o.z = z;
  o.w = w;
  o.out = out;
  o.out2 = out2;
  HIP_TRY(launch_mlp_op(o, static_cast<hipStream_t>(stream)));
  return 0;
}
