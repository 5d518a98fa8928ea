// Ragged batch of small sinkhorn_log problems, one workgroup per problem.
//
// Each problem runs the reference's sinkhorn_log loop (ot/bregman/_sinkhorn.py:
// 880-906: v first, then u, the column-marginal check every 10th iteration,
// strict `err < stopThr`, niter = ii) entirely inside one workgroup, so a
// problem stops on its own and no iteration needs a kernel boundary or a
// host round trip.  Problems may differ in n, m, reg and initial u (the
// learned warm start writes u0 here).
//
// Layout: the workgroup has W wavefronts; wave w handles rows w, w+W, ...
// and each of its 64 lanes owns CPT columns (4 consecutive per 16-byte
// load).  A row's log-sum-exp is a wave-local reduction (no barrier); the
// column sums of the next half-step accumulate in registers and are merged
// across the W waves through LDS once per sweep.  As in the dense solver,
// the row exponent is shifted by the previous row potential, so one exp2
// per element suffices and rows whose shifted sum leaves [2^-60, 2^60] are
// redone with their maximum as the shift.
#include "internal.h"

namespace potamd {

constexpr int kBatchWaves = 4;
constexpr int kBatchThreads = kBatchWaves * kWave;

template <int CPT, int VEC>
__global__ __launch_bounds__(kBatchThreads) void batch_log_kernel(BatchArgs A) {
  constexpr int W = kBatchWaves;
  constexpr int NCH = CPT / VEC;
  constexpr int MMAX = kWave * CPT;
  const int p = A.order ? A.order[blockIdx.x] : blockIdx.x;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  const int n = A.ns[p], m = A.ms[p], ldm = A.ldms[p];
  const float* Mp = static_cast<const float*>(A.M) + A.moff[p];
  const double* ap = A.a + A.aoff[p];
  const double* bp = A.b + A.boff[p];
  double* up = A.u + A.aoff[p];
  double* vp = A.v + A.boff[p];
  double* u2 = A.u2 + A.aoff[p];
  double* la2 = A.loga2 + A.aoff[p];
  double* sbuf = A.sbuf + A.aoff[p];
  const double reg = A.regs[p];
  const double s2d = -1.4426950408889634 / reg;
  const float s_hi = (float)s2d;
  const float s_lo = (float)(s2d - (double)s_hi);

  __shared__ double vsh[MMAX];          // column potentials, base 2
  __shared__ float part[W][MMAX];       // per-wave column partials
  __shared__ double red[W];
  __shared__ int flag;

  // Set-up: base-2 state, log2 a.
  for (int i = tid; i < n; i += kBatchThreads) {
    u2[i] = up[i] * 1.4426950408889634;
    la2[i] = log2(ap[i]);
  }
  __syncthreads();

  auto col = [&](int q) { return ((q / VEC) * kWave + lane) * VEC + (q % VEC); };
  auto load_row = [&](int i, float* dst) {
    const float* row = Mp + (int64_t)i * ldm;
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const int j0 = (k * kWave + lane) * VEC;
      if (j0 < m) {
        if constexpr (VEC == 4) {
          typedef float f4 __attribute__((ext_vector_type(4)));
          f4 x = __builtin_nontemporal_load(reinterpret_cast<const f4*>(row + j0));
          dst[k * 4 + 0] = x.x;
          dst[k * 4 + 1] = x.y;
          dst[k * 4 + 2] = x.z;
          dst[k * 4 + 3] = x.w;
        } else {
          dst[k] = __builtin_nontemporal_load(row + j0);
        }
      } else {
#pragma unroll
        for (int c = 0; c < VEC; ++c) dst[k * VEC + c] = 0.f;
      }
    }
  };

  // Prologue: v_1 = log b - LSE_i(Mr + u_0), online (max, sum) per column
  // over this wave's rows, merged across waves as per-wave LSE values.
  {
    float mx[CPT], sm[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      mx[q] = -__builtin_inff();
      sm[q] = 0.f;
    }
    float cur[CPT];
    for (int i = wid; i < n; i += W) {
      load_row(i, cur);
      const float ui = (float)u2[i];
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const float x = scaled_exponent(cur[q], s_hi, s_lo, ui);
        const float d = x - mx[q];
        const float e = exp2_fast(-fabsf(d));
        if (d > 0.f) {
          sm[q] = __builtin_fmaf(sm[q], e, 1.f);
          mx[q] = x;
        } else {
          sm[q] += e;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      part[wid][col(q)] = (sm[q] > 0.f) ? mx[q] + __builtin_amdgcn_logf(sm[q]) : -__builtin_inff();
    __syncthreads();
    for (int j = tid; j < m; j += kBatchThreads) {
      float g = -__builtin_inff();
#pragma unroll
      for (int w = 0; w < W; ++w) g = fmaxf(g, part[w][j]);
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < W; ++w) s += exp2((double)part[w][j] - (double)g);
      const double bj = bp[j];
      vsh[j] = bj > 0.0 ? log2(bj) - ((double)g + log2(s)) : -__builtin_inf();
    }
    for (int j = m + tid; j < MMAX; j += kBatchThreads) vsh[j] = -__builtin_inf();
    __syncthreads();
  }

  int status = kMaxIter, niter = A.max_iter - 1;
  double err_last = -1.0, err_first = -1.0;
  float cur[CPT], nxt[CPT];
  for (int ii = 0; ii < A.max_iter; ++ii) {
    // ---- row sweep: u from v, column sums of P(u_new, v) ----
    float acc[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) acc[q] = 0.f;
    int i = wid;
    if (i < n) load_row(i, cur);
    for (; i < n; i += W) {
      if (i + W < n) load_row(i + W, nxt);
      const double a_i = ap[i];
      if (a_i > 0.0) {
        double c2 = la2[i] - u2[i];
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
          const float t = (float)(vsh[col(q)] - c2);
          cur[q] = exp2_fast(scaled_exponent(cur[q], s_hi, s_lo, t));
          s += cur[q];
        }
        double S = (double)wave_sum(s);
        double seff = S;
        if (!in_safe_range(S)) {
          load_row(i, cur);
          float mxr = -__builtin_inff();
#pragma unroll
          for (int q = 0; q < CPT; ++q) {
            const float x = scaled_exponent(cur[q], s_hi, s_lo, (float)(vsh[col(q)] - c2));
            mxr = fmaxf(mxr, x);
          }
          const double gmx = (double)wave_max(mxr);
          if (__builtin_isfinite(gmx)) c2 += gmx;
          s = 0.f;
#pragma unroll
          for (int q = 0; q < CPT; ++q) {
            const float t = (float)(vsh[col(q)] - c2);
            cur[q] = exp2_fast(scaled_exponent(cur[q], s_hi, s_lo, t));
            s += cur[q];
          }
          S = (double)wave_sum(s);
          seff = S * exp2(__builtin_isfinite(gmx) ? gmx : 0.0);
        }
        if (lane == 0) sbuf[i] = seff;
        const float kf = (float)a_i * __builtin_amdgcn_rcpf((float)S);
#pragma unroll
        for (int q = 0; q < CPT; ++q) acc[q] = __builtin_fmaf(cur[q], kf, acc[q]);
      } else if (lane == 0) {
        sbuf[i] = 0.0;
      }
#pragma unroll
      for (int q = 0; q < CPT; ++q) cur[q] = nxt[q];
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q) part[wid][col(q)] = acc[q];
    __syncthreads();

    // ---- row potentials, column sums, marginal error ----
    for (int r = tid; r < n; r += kBatchThreads) {
      const double S = sbuf[r];
      u2[r] = ap[r] > 0.0 ? u2[r] - log2(S) : -__builtin_inf();
    }
    double e2 = 0.0;
    const bool check = (ii % 10 == 0);
    for (int j = tid; j < m; j += kBatchThreads) {
      float cs = 0.f;
#pragma unroll
      for (int w = 0; w < W; ++w) cs += part[w][j];
      const double d = (double)cs - bp[j];
      e2 += d * d;
      part[0][j] = cs;
    }
    e2 = wave_sum(e2);
    if (lane == 0) red[wid] = e2;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < W; ++w) t += red[w];
      const double err = sqrt(t);
      int stop = 0;
      if (check) {
        err_last = err;
        if (ii == 0) err_first = err;
        if (err < A.stop_thr) stop = 1;
      }
      flag = stop;
    }
    __syncthreads();
    if (flag) {
      status = kConverged;
      niter = ii;
      break;
    }
    if (ii == A.max_iter - 1) break;
    // ---- column update v_j += log2 b_j - log2 colsum_j ----
    for (int j = tid; j < m; j += kBatchThreads) {
      const double bj = bp[j];
      const double cs = (double)part[0][j];
      vsh[j] = bj > 0.0 ? vsh[j] + log2(bj) - log2(cs) : -__builtin_inf();
    }
    __syncthreads();
  }

  // ---- outputs (natural-log potentials) ----
  for (int i = tid; i < n; i += kBatchThreads) up[i] = u2[i] * 0.6931471805599453;
  for (int j = tid; j < m; j += kBatchThreads) vp[j] = vsh[j] * 0.6931471805599453;
  if (tid == 0) {
    A.niter[p] = niter;
    A.status[p] = status;
    if (A.err_last) A.err_last[p] = err_last;
    if (A.err_first) A.err_first[p] = err_first;
  }
}

int batch_cpt(int max_m) {
  if (max_m <= 256) return 4;
  if (max_m <= 512) return 8;
  if (max_m <= 1024) return 16;
  return 0;
}

hipError_t launch_batch_log(int cpt, int vec_ok, const BatchArgs& a, int B, hipStream_t s) {
  dim3 grid(B), block(kBatchThreads);
#define POTAMD_BATCH_CASE(C)                                                      \
  if (cpt == C) {                                                                 \
    if (vec_ok)                                                                   \
      hipLaunchKernelGGL((batch_log_kernel<C, 4>), grid, block, 0, s, a);         \
    else                                                                          \
      hipLaunchKernelGGL((batch_log_kernel<C, 1>), grid, block, 0, s, a);         \
    return hipGetLastError();                                                     \
  }
  POTAMD_BATCH_CASE(4)
  POTAMD_BATCH_CASE(8)
  POTAMD_BATCH_CASE(16)
#undef POTAMD_BATCH_CASE
  return hipErrorInvalidConfiguration;
}

}  // namespace potamd
