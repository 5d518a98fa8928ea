// Log-stabilised Sinkhorn (ot.bregman.sinkhorn_stabilized,
// ot/bregman/_sinkhorn.py:1105-1367) on a materialised cost, for one target
// histogram or several (a 2-D b: the scalings become n x H and m x H, the
// absorption and the error follow the reference's several-histogram branch).
//
// Per iteration, exactly as the reference:
//   v = b / (K^T u);  u = a / (K v)                          (:1268-1270)
//   if max|u| > tau or max|v| > tau: alpha += reg log u, beta += reg log v,
//      u = 1/dim_a, v = 1/dim_b, K = exp(-(M - alpha - beta) / reg)   (:1273-1286)
//   every print_period-th iteration: err = ||sum_i Gamma_ij - b||_2 with
//      Gamma = K u v, appended to the log                    (:1288-1302)
//   err <= stopThr stops; a NaN in u or v restores the previous u, v and
//      stops with the "Numerical errors" warning             (:1309-1318)
// with K = exp(-(M - alpha - beta) / reg) stored in the cost's dtype.
//
// All decisions are taken on the device by the last block of a reduction
// (ticket), so the host enqueues iterations without waiting; every later
// kernel returns at once after the stop.  The kernel rebuild runs every
// iteration but returns immediately unless the absorption flag is set.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "../../include/pot_amd.h"
#include "host.h"
#include "internal.h"

namespace potamd {
namespace {

constexpr int kColRows = 256;  // rows per block of the column pass

struct StabCtrl {
  Ctrl c;          // done, status, niter, bad_iter, counter, n_err, err
  int absorb;      // this iteration absorbs into alpha, beta
  int anynan;      // this iteration produced a NaN in u or v
  unsigned counter2;
  int pad;
  double maxu, maxv;
};

template <typename TM>
__global__ __launch_bounds__(256) void stab_build_k(const TM* M, int64_t ldm, int n, int m,
                                                    double reg, const double* alpha,
                                                    const double* beta, TM* K,
                                                    const StabCtrl* sc) {
  if (sc) {
    if (sc->c.done || !sc->absorb) return;
  }
  const int64_t total = (int64_t)n * m;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e / m), j = (int)(e % m);
    const double x = -((double)M[(int64_t)i * ldm + j] - alpha[i] - beta[j]) / reg;
    K[(int64_t)i * m + j] = (TM)exp(x);
  }
}

// part[g][j] = sum_{i in block g} K_ij w_i  (fp64)
// (histogram h = blockIdx.z: w + h n, part + h G m)
template <typename TM>
__global__ __launch_bounds__(256) void stab_col_partial(const TM* K, int n, int m,
                                                        const double* w, double* part,
                                                        const StabCtrl* sc) {
  if (sc && sc->c.done) return;
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int g = blockIdx.y;
  const int h = blockIdx.z;
  if (j >= m) return;
  w += (int64_t)h * n;
  const int r0 = g * kColRows, r1 = min(n, r0 + kColRows);
  double acc = 0.0;
  for (int i = r0; i < r1; ++i) acc = fma((double)K[(int64_t)i * m + j], w[i], acc);
  part[((int64_t)h * gridDim.y + g) * m + j] = acc;
}

// v_j = b_j / sum_g part[g][j]
__global__ __launch_bounds__(256) void stab_col_update(const double* part, int G, int m,
                                                       const double* b, double* v,
                                                       const StabCtrl* sc) {
  if (sc->c.done) return;
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int h = blockIdx.y;  // histogram
  if (j >= m) return;
  part += (int64_t)h * G * m;
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += part[(int64_t)g * m + j];
  v[(int64_t)h * m + j] = b[(int64_t)h * m + j] / s;
}

// u_i = a_i / sum_j K_ij v_j, one block per row
template <typename TM>
__global__ __launch_bounds__(256) void stab_row(const TM* K, int n, int m, const double* v,
                                                const double* a, double* u, const StabCtrl* sc) {
  if (sc->c.done) return;
  const int i = blockIdx.x;
  const int h = blockIdx.y;  // histogram (a is shared)
  v += (int64_t)h * m;
  double acc = 0.0;
  for (int j = threadIdx.x; j < m; j += 256) acc = fma((double)K[(int64_t)i * m + j], v[j], acc);
  __shared__ double red[4];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) u[(int64_t)h * n + i] = a[i] / (red[0] + red[1] + red[2] + red[3]);
}

__global__ void stab_copy(int64_t n, const double* x, double* y, const StabCtrl* sc) {
  if (sc && sc->c.done) return;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// max|u|, max|v| and NaN over u, v; the last block sets the absorption flag
// (a NaN makes the reference's maximum NaN and its comparison false).
__global__ __launch_bounds__(256) void stab_check(const double* u, int64_t n, const double* v,
                                                  int64_t m, double tau, double* blk,
                                                  StabCtrl* sc) {
  if (sc->c.done) return;
  double mu = 0.0, mv = 0.0;
  int nan = 0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const double x = fabs(u[i]);
    if (x != x) nan = 1; else mu = fmax(mu, x);
  }
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < m; j += stride) {
    const double x = fabs(v[j]);
    if (x != x) nan = 1; else mv = fmax(mv, x);
  }
  mu = wave_max(mu);
  mv = wave_max(mv);
  __shared__ double r0[4], r1[4];
  if ((threadIdx.x & 63) == 0) {
    r0[threadIdx.x >> 6] = mu;
    r1[threadIdx.x >> 6] = mv;
  }
  nan = __syncthreads_or(nan);
  if (threadIdx.x == 0) {
    blk[3 * blockIdx.x] = fmax(fmax(r0[0], r0[1]), fmax(r0[2], r0[3]));
    blk[3 * blockIdx.x + 1] = fmax(fmax(r1[0], r1[1]), fmax(r1[2], r1[3]));
    blk[3 * blockIdx.x + 2] = nan;
  }
  __threadfence();
  __shared__ int last;
  if (threadIdx.x == 0) last = atomicAdd(&sc->counter2, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last || threadIdx.x != 0) return;
  __threadfence();
  double MU = 0.0, MV = 0.0;
  int any = 0;
  for (int k = 0; k < (int)gridDim.x; ++k) {
    MU = fmax(MU, blk[3 * k]);
    MV = fmax(MV, blk[3 * k + 1]);
    any |= blk[3 * k + 2] != 0.0;
  }
  sc->counter2 = 0;
  sc->maxu = MU;
  sc->maxv = MV;
  sc->anynan = any;
  sc->absorb = !any && (MU > tau || MV > tau);
}

// alpha += reg log u, beta += reg log v, u = 1/n, v = 1/m (when absorbing)
__global__ void stab_absorb(double* alpha, double* u, int n, double* beta, double* v, int m,
                            double reg, const StabCtrl* sc) {
  if (sc->c.done || !sc->absorb) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) {
    alpha[k] += reg * log(u[k]);
    u[k] = 1.0 / n;
  }
  if (k < m) {
    beta[k] += reg * log(v[k]);
    v[k] = 1.0 / m;
  }
}

// err = ||v_j sum_i K_ij u_i - b_j||_2 from the column partials; the last
// block logs it and stops on err <= stop_thr, else on a NaN restores the
// previous u, v (done by stab_revert) and stops as numerical.
__global__ __launch_bounds__(256) void stab_err(const double* part, int G, int m,
                                                const double* b, const double* v, int check,
                                                int iter, int max_iter, double stop_thr,
                                                double* err_hist, int err_cap, double* blk,
                                                StabCtrl* sc) {
  if (sc->c.done) return;
  double e2 = 0.0;
  if (check) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < m) {
      double s = 0.0;
      for (int g = 0; g < G; ++g) s += part[(int64_t)g * m + j];
      const double d = v[j] * s - b[j];
      e2 = d * d;
    }
  }
  __shared__ double red[4];
  e2 = wave_sum(e2);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) blk[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
  __threadfence();
  __shared__ int last;
  if (threadIdx.x == 0) last = atomicAdd(&sc->c.counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last || threadIdx.x != 0) return;
  __threadfence();
  sc->c.counter = 0;
  if (check) {
    double acc = 0.0;
    for (int k = 0; k < (int)gridDim.x; ++k) acc += blk[k];
    const double err = sqrt(acc);
    sc->c.err = err;
    const int k = sc->c.n_err;
    if (k < err_cap) err_hist[k] = err;
    sc->c.n_err = k + 1;
  }
  // the reference keeps the last err between checks (initially 1)
  if (sc->c.err <= stop_thr) {
    sc->c.done = 1;
    sc->c.status = kConverged;
    sc->c.niter = iter;
    return;
  }
  if (sc->anynan) {
    sc->c.done = 2;  // stab_revert restores u, v, then sets done = 1
    sc->c.status = kNumerical;
    sc->c.niter = iter;
    sc->c.bad_iter = iter;
    return;
  }
  if (iter >= max_iter - 1) {
    sc->c.done = 1;
    sc->c.status = kMaxIter;
    sc->c.niter = iter;
  }
}

// Several histograms (:1275-1283 with n_hists): alpha_i += reg max_h log u_hi,
// beta_j += reg max log v (one maximum over all of v), and u, v keep their
// values (the reference resets them only for a single histogram).
__global__ void stab_absorb_multi(double* alpha, const double* u, int n, int H, double* beta,
                                  int m, double reg, const StabCtrl* sc) {
  if (sc->c.done || !sc->absorb) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) {
    double mx = u[k];
    for (int h = 1; h < H; ++h) mx = fmax(mx, u[(int64_t)h * n + k]);
    alpha[k] += reg * log(mx);
  }
  if (k < m) beta[k] += reg * log(sc->maxv);  // v > 0 here, so max log v = log max |v|
}

// Several histograms: the check of :1290-1296, the relative changes
// err = (max|u - u'| / max(max|u|, max|u'|, 1) + the same for v) / 2, then
// the stop / NaN / max-iteration decisions of stab_err.
__global__ __launch_bounds__(256) void stab_err_multi(const double* u, const double* up,
                                                      int64_t nu, const double* v,
                                                      const double* vp, int64_t nv, int check,
                                                      int iter, int max_iter, double stop_thr,
                                                      double* err_hist, int err_cap,
                                                      double* blk, StabCtrl* sc) {
  if (sc->c.done) return;
  double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // |du|, |u|, |u'|, |dv|, |v|, |v'|
  if (check) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nu; i += stride) {
      r[0] = fmax(r[0], fabs(u[i] - up[i]));
      r[1] = fmax(r[1], fabs(u[i]));
      r[2] = fmax(r[2], fabs(up[i]));
    }
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nv; j += stride) {
      r[3] = fmax(r[3], fabs(v[j] - vp[j]));
      r[4] = fmax(r[4], fabs(v[j]));
      r[5] = fmax(r[5], fabs(vp[j]));
    }
  }
  __shared__ double red[6][4];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double x = wave_max(r[q]);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q)
      blk[6 * blockIdx.x + q] = fmax(fmax(red[q][0], red[q][1]), fmax(red[q][2], red[q][3]));
  }
  __threadfence();
  __shared__ int last;
  if (threadIdx.x == 0) last = atomicAdd(&sc->c.counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last || threadIdx.x != 0) return;
  __threadfence();
  sc->c.counter = 0;
  if (check) {
    double R[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < (int)gridDim.x; ++k)
      for (int q = 0; q < 6; ++q) R[q] = fmax(R[q], blk[6 * k + q]);
    const double eu = R[0] / fmax(fmax(R[1], R[2]), 1.0);
    const double evv = R[3] / fmax(fmax(R[4], R[5]), 1.0);
    const double err = 0.5 * (eu + evv);
    sc->c.err = err;
    const int k = sc->c.n_err;
    if (k < err_cap) err_hist[k] = err;
    sc->c.n_err = k + 1;
  }
  if (sc->c.err <= stop_thr) {
    sc->c.done = 1;
    sc->c.status = kConverged;
    sc->c.niter = iter;
    return;
  }
  if (sc->anynan) {
    sc->c.done = 2;
    sc->c.status = kNumerical;
    sc->c.niter = iter;
    sc->c.bad_iter = iter;
    return;
  }
  if (iter >= max_iter - 1) {
    sc->c.done = 1;
    sc->c.status = kMaxIter;
    sc->c.niter = iter;
  }
}

__global__ void stab_revert(double* u, const double* up, int64_t n, double* v, const double* vp,
                            int64_t m, StabCtrl* sc) {
  if (sc->c.done != 2) return;
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < n) u[k] = up[k];
  if (k < m) v[k] = vp[k];
}

__global__ void stab_finish(StabCtrl* sc) {
  if (sc->c.done == 2) sc->c.done = 1;
}

__global__ void stab_init(StabCtrl* sc, double* u, int64_t n, double uval, double* v, int64_t m,
                          double vval) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k == 0) {
    sc->c.done = 0;
    sc->c.status = 0;
    sc->c.niter = -1;
    sc->c.bad_iter = -1;
    sc->c.counter = 0;
    sc->c.n_err = 0;
    sc->c.need_exact = 0;
    sc->c.pad1 = 0;
    sc->c.err = 1.0;  // the reference's initial err
    sc->absorb = 0;
    sc->anynan = 0;
    sc->counter2 = 0;
    sc->pad = 0;
    sc->maxu = 0.0;
    sc->maxv = 0.0;
  }
  if (k < n) u[k] = uval;
  if (k < m) v[k] = vval;
}

// Final marginal errors of Gamma = K u v: sq[0] = ||Gamma^T 1 - b||^2,
// sq[1] = ||Gamma 1 - a||^2 (the epsilon-scaling check, :1502-1505).
template <typename TM>
__global__ __launch_bounds__(256) void stab_row_marg(const TM* K, int n, int m, const double* u,
                                                     const double* v, const double* a,
                                                     double* rowerr) {
  const int i = blockIdx.x;
  double acc = 0.0;
  for (int j = threadIdx.x; j < m; j += 256) acc = fma((double)K[(int64_t)i * m + j], v[j], acc);
  __shared__ double red[4];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double d = u[i] * (red[0] + red[1] + red[2] + red[3]) - a[i];
    rowerr[i] = d * d;
  }
}

__global__ void stab_marg_final(const double* part, int G, int m, const double* b,
                                const double* v, const double* rowerr, int n, double* sq) {
  // single block
  double ec = 0.0, er = 0.0;
  for (int j = threadIdx.x; j < m; j += 256) {
    double s = 0.0;
    for (int g = 0; g < G; ++g) s += part[(int64_t)g * m + j];
    const double d = v[j] * s - b[j];
    ec += d * d;
  }
  for (int i = threadIdx.x; i < n; i += 256) er += rowerr[i];
  __shared__ double r0[4], r1[4];
  ec = wave_sum(ec);
  er = wave_sum(er);
  if ((threadIdx.x & 63) == 0) {
    r0[threadIdx.x >> 6] = ec;
    r1[threadIdx.x >> 6] = er;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    sq[0] = r0[0] + r0[1] + r0[2] + r0[3];
    sq[1] = r1[0] + r1[1] + r1[2] + r1[3];
  }
}

struct StabLayout {
  void* K;
  double* part;
  double* up;
  double* vp;
  double* blk;
  double* rowerr;
  StabCtrl* sc;
  size_t bytes;
};

StabLayout carve_stab(void* ws, int dtype, int n, int m, int H) {
  Carver c(ws);
  StabLayout L;
  const size_t es = dtype == POTAMD_F32 ? 4 : 8;
  const int G = (n + kColRows - 1) / kColRows;
  L.K = c.take<char>((size_t)n * m * es);
  L.part = c.take<double>((size_t)G * m * H);
  L.up = c.take<double>((size_t)n * H);
  L.vp = c.take<double>((size_t)m * H);
  L.blk = c.take<double>(std::max(6 * 1024, (m + 255) / 256 + 8));
  L.rowerr = c.take<double>(n);
  L.sc = c.take<StabCtrl>(1);
  L.bytes = c.off;
  return L;
}

// H histograms (u: H x n, v and b: H x m, histogram-major); `multi` gives
// the reference's several-histogram branch (a 2-D b, even of one column).
template <typename TM>
int run_stabilized(const TM* M, int64_t ldm, int n, int m, int H, bool multi, double reg,
                   const double* a, const double* b, double* alpha, double* beta, double* u,
                   double* v, double tau, int max_iter, double stop_thr, int print_period,
                   double* err_hist, int err_cap, double* marg_sq, const StabLayout& L,
                   void* host_ctrl, hipStream_t s, int* info) {
  TM* K = static_cast<TM*>(L.K);
  const int G = (n + kColRows - 1) / kColRows;
  const int64_t nu = (int64_t)n * H, nv = (int64_t)m * H;
  const int64_t nm = std::max(nu, nv);
  const dim3 g1((unsigned)((nm + 255) / 256));
  const dim3 g1s((unsigned)((std::max(n, m) + 255) / 256));
  const int kblocks = (int)std::min<int64_t>(4096, ((int64_t)n * m + 255) / 256);
  const int cblocks = (int)std::min<int64_t>(1024, (nm + 255) / 256);
  hipLaunchKernelGGL(stab_init, g1, dim3(256), 0, s, L.sc, u, nu, 1.0 / n, v, nv, 1.0 / m);
  hipLaunchKernelGGL((stab_build_k<TM>), dim3(kblocks), dim3(256), 0, s, M, ldm, n, m, reg,
                     alpha, beta, K, (const StabCtrl*)nullptr);
  HIP_TRY(hipGetLastError());

  constexpr int kRing = 2;  // stop checks in flight: the host queues at most 2 past the device
  Ctrl* hc = static_cast<Ctrl*>(host_ctrl);
  hipEvent_t ev[kRing];
  bool ev_used[kRing] = {};
  for (int k = 0; k < kRing; ++k) HIP_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
  int ring_pos = 0;
  int ring_head = 0;  // oldest check whose copy the host has not read
  bool stopped = false;
  int rc = 0;
  StabCtrl local;
  for (int ii = 0; ii < max_iter && !stopped; ++ii) {
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
      if (hc[k].done) stopped = true;
    }
    if (rc || stopped) break;
    hipLaunchKernelGGL(stab_copy, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, s, nu, u,
                       L.up, L.sc);
    hipLaunchKernelGGL(stab_copy, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, nv, v,
                       L.vp, L.sc);
    hipLaunchKernelGGL((stab_col_partial<TM>), dim3((m + 255) / 256, G, H), dim3(256), 0, s, K,
                       n, m, u, L.part, L.sc);
    hipLaunchKernelGGL(stab_col_update, dim3((m + 255) / 256, H), dim3(256), 0, s, L.part, G, m,
                       b, v, L.sc);
    hipLaunchKernelGGL((stab_row<TM>), dim3(n, H), dim3(256), 0, s, K, n, m, v, a, u, L.sc);
    hipLaunchKernelGGL(stab_check, dim3(cblocks), dim3(256), 0, s, u, nu, v, nv, tau, L.blk,
                       L.sc);
    if (multi)
      hipLaunchKernelGGL(stab_absorb_multi, g1s, dim3(256), 0, s, alpha, u, n, H, beta, m, reg,
                         L.sc);
    else
      hipLaunchKernelGGL(stab_absorb, g1s, dim3(256), 0, s, alpha, u, n, beta, v, m, reg, L.sc);
    hipLaunchKernelGGL((stab_build_k<TM>), dim3(kblocks), dim3(256), 0, s, M, ldm, n, m, reg,
                       alpha, beta, K, L.sc);
    const int check = ii % print_period == 0;
    if (multi) {
      hipLaunchKernelGGL(stab_err_multi, dim3(cblocks), dim3(256), 0, s, u, L.up, nu, v, L.vp,
                         nv, check, ii, max_iter, stop_thr, err_hist, err_cap, L.blk, L.sc);
    } else {
      if (check)
        hipLaunchKernelGGL((stab_col_partial<TM>), dim3((m + 255) / 256, G), dim3(256), 0, s, K,
                           n, m, u, L.part, L.sc);
      hipLaunchKernelGGL(stab_err, dim3((m + 255) / 256), dim3(256), 0, s, L.part, G, m, b, v,
                         check, ii, max_iter, stop_thr, err_hist, err_cap, L.blk, L.sc);
    }
    hipLaunchKernelGGL(stab_revert, g1, dim3(256), 0, s, u, L.up, nu, v, L.vp, nv, L.sc);
    hipLaunchKernelGGL(stab_finish, dim3(1), dim3(1), 0, s, L.sc);
    if (hipGetLastError() != hipSuccess) {
      rc = POTAMD_EHIP;
      break;
    }
    if (check && ii + 1 < max_iter) {
      const int k = ring_pos % kRing;
      if (ev_used[k]) {
        if (hipEventSynchronize(ev[k]) != hipSuccess) { rc = POTAMD_EHIP; break; }
        ev_used[k] = false;
        ++ring_head;
        const Ctrl* snap = hc ? &hc[k] : &local.c;
        if (snap->done) stopped = true;
      }
      if (!stopped) {
        if (hc) {
          if (hipMemcpyAsync(&hc[k], L.sc, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipEventRecord(ev[k], s) != hipSuccess) {
            rc = POTAMD_EHIP;
            break;
          }
          ev_used[k] = true;
        } else {
          if (hipMemcpyAsync(&local, L.sc, sizeof(StabCtrl), hipMemcpyDeviceToHost, s) !=
                  hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess) {
            rc = POTAMD_EHIP;
            break;
          }
          if (local.c.done) stopped = true;
        }
        ++ring_pos;
      }
    }
  }
  for (int k = 0; k < kRing; ++k) (void)hipEventDestroy(ev[k]);
  if (rc) return rc;
  if (marg_sq && !multi) {
    // the stop flag is set by now; rebuild the partials without it
    hipLaunchKernelGGL((stab_col_partial<TM>), dim3((m + 255) / 256, G), dim3(256), 0, s, K, n,
                       m, u, L.part, (const StabCtrl*)nullptr);
    hipLaunchKernelGGL((stab_row_marg<TM>), dim3(n), dim3(256), 0, s, K, n, m, u, v, a,
                       L.rowerr);
    hipLaunchKernelGGL(stab_marg_final, dim3(1), dim3(256), 0, s, L.part, G, m, b, v, L.rowerr,
                       n, marg_sq);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipMemcpyAsync(&local, L.sc, sizeof(StabCtrl), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  int status = local.c.status, niter = local.c.niter;
  if (!local.c.done) {
    status = POTAMD_MAX_ITER;
    niter = max_iter - 1;
  }
  info[0] = niter;
  info[1] = local.c.n_err;
  info[2] = status;
  info[3] = local.c.bad_iter;
  return status;
}

}  // namespace
}  // namespace potamd

using namespace potamd;

extern "C" size_t potamd_stabilized_workspace_size(int dtype, int n, int m) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || n < 1 || m < 1) return 0;
  return carve_stab(nullptr, dtype, n, m, 1).bytes;
}

extern "C" size_t potamd_stabilized_multi_workspace_size(int dtype, int n, int m, int H) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || n < 1 || m < 1 || H < 1) return 0;
  return carve_stab(nullptr, dtype, n, m, H).bytes;
}

extern "C" int potamd_sinkhorn_stabilized_multi(int dtype, const void* M, int64_t ldm, int n,
                                                int m, int H, double reg, const double* a,
                                                const double* b, double* alpha, double* beta,
                                                double* u, double* v, double tau, int max_iter,
                                                double stop_thr, int print_period,
                                                double* err_hist, int err_cap, void* workspace,
                                                size_t ws_bytes, void* host_ctrl, void* stream,
                                                int* info) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || !M || !a || !b || !alpha || !beta || !u ||
      !v || !info || n < 1 || m < 1 || H < 1 || H > 65535 || ldm < m || !(reg > 0.0) ||
      max_iter < 1 || print_period < 1 || (err_cap > 0 && !err_hist))
    return POTAMD_EINVAL;
  StabLayout L = carve_stab(workspace, dtype, n, m, H);
  if (!workspace || ws_bytes < L.bytes) return POTAMD_EWORKSPACE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == POTAMD_F32)
    return run_stabilized<float>(static_cast<const float*>(M), ldm, n, m, H, true, reg, a, b,
                                 alpha, beta, u, v, tau, max_iter, stop_thr, print_period,
                                 err_hist, err_cap, nullptr, L, host_ctrl, s, info);
  return run_stabilized<double>(static_cast<const double*>(M), ldm, n, m, H, true, reg, a, b,
                                alpha, beta, u, v, tau, max_iter, stop_thr, print_period,
                                err_hist, err_cap, nullptr, L, host_ctrl, s, info);
}

extern "C" int potamd_sinkhorn_stabilized(int dtype, const void* M, int64_t ldm, int n, int m,
                                          double reg, const double* a, const double* b,
                                          double* alpha, double* beta, double* u, double* v,
                                          double tau, int max_iter, double stop_thr,
                                          int print_period, double* err_hist, int err_cap,
                                          double* marg_sq, void* workspace, size_t ws_bytes,
                                          void* host_ctrl, void* stream, int* info) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || !M || !a || !b || !alpha || !beta || !u ||
      !v || !info || n < 1 || m < 1 || ldm < m || !(reg > 0.0) || max_iter < 1 ||
      print_period < 1 || (err_cap > 0 && !err_hist))
    return POTAMD_EINVAL;
  StabLayout L = carve_stab(workspace, dtype, n, m, 1);
  if (!workspace || ws_bytes < L.bytes) return POTAMD_EWORKSPACE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == POTAMD_F32)
    return run_stabilized<float>(static_cast<const float*>(M), ldm, n, m, 1, false, reg, a, b,
                                 alpha, beta, u, v, tau, max_iter, stop_thr, print_period,
                                 err_hist, err_cap, marg_sq, L, host_ctrl, s, info);
  return run_stabilized<double>(static_cast<const double*>(M), ldm, n, m, 1, false, reg, a, b,
                                alpha, beta, u, v, tau, max_iter, stop_thr, print_period,
                                err_hist, err_cap, marg_sq, L, host_ctrl, s, info);
}
