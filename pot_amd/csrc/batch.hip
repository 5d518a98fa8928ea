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
// across the W waves through LDS once per sweep.  Everything a row needs
// besides its cost (shift c_i, weight a_i, the column potentials) is in LDS,
// stored lane-major so each LDS read is conflict-free.  As in the dense
// solver, the row exponent is shifted by the previous row potential, so one
// exp2 per element suffices; a row whose shifted sum leaves [2^-60, 2^60] is
// redone with its maximum as the shift.
//
// Warm-start fallback: when fallback_thr > 0 and the first check of a
// problem started from a given u0 exceeds it, the problem restarts from
// u0 = 0 (the cold start) and reports status | 4.
#include "internal.h"

namespace potamd {

constexpr int kBatchMaxN = 1024;

// OCC > 0 caps the registers so that OCC waves fit per SIMD (two 8-wave
// workgroups per CU at OCC = 4); the few values this pushes to scratch are
// per-problem and per-iteration scalars, none is touched in the row loop.
template <int CPT, int VEC, int W, int PF, int OCC>
__global__ __launch_bounds__(W * kWave)
__attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, OCC > 0 ? OCC : 8)))
void batch_log_kernel(BatchArgs A) {
  constexpr int T = W * kWave;
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
  const double reg = A.regs[p];
  const double s2d = -1.4426950408889634 / reg;
  const float s_hi = (float)s2d;
  const float s_lo = (float)(s2d - (double)s_hi);

  // Lane-major slot of column j: lane (j / VEC) % 64, register q.
  auto slot = [&](int j) {
    const int q = (j / (VEC * kWave)) * VEC + (j % VEC);
    return q * kWave + (j / VEC) % kWave;
  };
  __shared__ double vsh[MMAX];          // column potentials (base 2), lane-major
  __shared__ float part[W][MMAX];       // per-wave column partials, lane-major
  __shared__ double u2[kBatchMaxN];     // row potentials (base 2)
  __shared__ double c2s[kBatchMaxN];    // row shifts log2 a_i - u2_i
  __shared__ float af[kBatchMaxN];      // row weights
  __shared__ float sh[kBatchMaxN];      // shifted row sums (fp32 mantissa)
  __shared__ double red[W];
  __shared__ int flag;

  for (int i = tid; i < n; i += T) {
    u2[i] = up[i] * 1.4426950408889634;
    af[i] = (float)ap[i];
  }
  __syncthreads();

  // Every load is unconditional: the compiler's wait counts assume the fewest
  // loads any path may have issued, so a load under a condition (or whose
  // value is consumed inside one) is waited for at once.  A lane past the last
  // column loads the row's first columns; mask_row zeroes those values (and
  // the row padding past m, which may hold anything) where they are used.
  auto load_row = [&](int i, float* dst) {
    const float* row = Mp + (int64_t)i * ldm;
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const int j0 = (k * kWave + lane) * VEC;
      const int jl = j0 < m ? j0 : 0;
      if constexpr (VEC == 4) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4 x = __builtin_nontemporal_load(reinterpret_cast<const f4*>(row + jl));
        dst[k * 4 + 0] = x.x;
        dst[k * 4 + 1] = x.y;
        dst[k * 4 + 2] = x.z;
        dst[k * 4 + 3] = x.w;
      } else {
        dst[k] = __builtin_nontemporal_load(row + jl);
      }
    }
  };
  const bool padded = m < MMAX;
  auto mask_row = [&](float* cur) {
    if (padded) {
#pragma unroll
      for (int k = 0; k < NCH; ++k)
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
          const int j = (k * kWave + lane) * VEC + c;
          cur[k * VEC + c] = j < m ? cur[k * VEC + c] : 0.f;
        }
    }
  };
  // This wave's rows wid, wid + W, ... streamed through PF + 1 buffers: while
  // row k is processed, the next PF rows are in flight.  The loop is unrolled
  // over the buffers so they rotate without register copies (a copy would
  // wait for the prefetch); a prefetch past the wave's last row reads that
  // row again.  body(i, cur) gets the row's logical index and its values.
  constexpr int NB = PF + 1;
  auto for_rows = [&](auto phys, auto body) {
    const int nl = wid < n ? (n - wid + W - 1) / W : 0;
    if (nl == 0) return;
    float buf[NB][CPT];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      load_row(phys(wid + (b < nl ? b : nl - 1) * W), buf[b]);
      __builtin_amdgcn_sched_barrier(0);  // issue the buffers in order
    }
    for (int k0 = 0; k0 < nl; k0 += NB) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int k = k0 + b;
        if (k < nl) {
          mask_row(buf[b]);
          body(wid + k * W, buf[b]);
        }
        load_row(phys(wid + (k + NB < nl ? k + NB : nl - 1) * W), buf[b]);
      }
    }
  };

  int status = kMaxIter, niter = A.max_iter - 1, restarted = 0;
  double err_last = -1.0, err_first = -1.0;
  const bool warm = A.fallback_thr > 0.0;

restart:
  // ---- prologue: v_1 = log b - LSE_i(Mr + u_0) per column ----
  {
    float mx[CPT], sm[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      mx[q] = -__builtin_inff();
      sm[q] = 0.f;
    }
    for_rows([](int r) { return r; }, [&](int i, float* cur) {
      const float ui = (float)u2[i];
#pragma unroll
      for (int q = 0; q < CPT; ++q) {
        const float x = scaled_exponent(cur[q], s_hi, s_lo, ui);
        const float d = x - mx[q];
        const float e = exp2_fast(-fabsf(d));
        const bool up_ = d > 0.f;
        sm[q] = up_ ? __builtin_fmaf(sm[q], e, 1.f) : sm[q] + e;
        mx[q] = up_ ? x : mx[q];
      }
    });
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      part[wid][q * kWave + lane] =
          (sm[q] > 0.f) ? mx[q] + __builtin_amdgcn_logf(sm[q]) : -__builtin_inff();
    __syncthreads();
    for (int j = tid; j < m; j += T) {
      const int sj = slot(j);
      float g = -__builtin_inff();
#pragma unroll
      for (int w = 0; w < W; ++w) g = fmaxf(g, part[w][sj]);
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < W; ++w) s += exp2((double)part[w][sj] - (double)g);
      const double bj = bp[j];
      vsh[sj] = bj > 0.0 ? log2(bj) - ((double)g + log2(s)) : -__builtin_inf();
    }
    for (int j = m + tid; j < MMAX; j += T) vsh[slot(j)] = -__builtin_inf();
    for (int i = tid; i < n; i += T) c2s[i] = log2(ap[i]) - u2[i];
    __syncthreads();
  }

  for (int ii = 0; ii < A.max_iter; ++ii) {
    // ---- row sweep: u from v, column sums of P(u_new, v) ----
    float acc[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) acc[q] = 0.f;
    // Row i of the sweep is physical row pr(i): with A.rotate, problem p
    // starts its sweep at row rot, so equal-size problems streamed side by
    // side do not hit the same address offsets at the same time.
    const int rot = (A.rotate && n > 0) ? (int)(((int64_t)p * 337) % n) : 0;
    auto pr = [&](int r) { const int q = r + rot; return q < n ? q : q - n; };
    for_rows(pr, [&](int i, float* cur) {
      const int ip = pr(i);
      const float a_i = af[ip];
      if (a_i > 0.f) {
        double c2 = c2s[ip];
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
          const float t = (float)(vsh[q * kWave + lane] - c2);
          cur[q] = exp2_fast(scaled_exponent(cur[q], s_hi, s_lo, t));
          s += cur[q];
        }
        float S = wave_sum_dpp(s);
        if (!in_safe_range((double)S)) {
          load_row(ip, cur);
          mask_row(cur);
          float mxr = -__builtin_inff();
#pragma unroll
          for (int q = 0; q < CPT; ++q) {
            const float x =
                scaled_exponent(cur[q], s_hi, s_lo, (float)(vsh[q * kWave + lane] - c2));
            mxr = fmaxf(mxr, x);
          }
          const float gmx = wave_max(mxr);
          if (__builtin_isfinite(gmx)) c2 += (double)gmx;
          s = 0.f;
#pragma unroll
          for (int q = 0; q < CPT; ++q) {
            const float t = (float)(vsh[q * kWave + lane] - c2);
            cur[q] = exp2_fast(scaled_exponent(cur[q], s_hi, s_lo, t));
            s += cur[q];
          }
          S = wave_sum_dpp(s);
          if (lane == 0) c2s[ip] = c2;  // the shift actually used
        }
        if (lane == 0) sh[ip] = S;
        const float kf = a_i * __builtin_amdgcn_rcpf(S);
#pragma unroll
        for (int q = 0; q < CPT; ++q) acc[q] = __builtin_fmaf(cur[q], kf, acc[q]);
      }
    });
#pragma unroll
    for (int q = 0; q < CPT; ++q) part[wid][q * kWave + lane] = acc[q];
    __syncthreads();

    // ---- row potentials, column sums, marginal error ----
    // u2_new = log2 a - c2_used - log2 S ; next shift c2 = log2 a - u2_new = c2_used + log2 S.
    for (int r = tid; r < n; r += T) {
      if (af[r] > 0.f) {
        const double c = c2s[r] + log2((double)sh[r]);
        c2s[r] = c;
        u2[r] = log2(ap[r]) - c;
      } else {
        u2[r] = -__builtin_inf();
      }
    }
    double e2 = 0.0;
    for (int j = tid; j < m; j += T) {
      const int sj = slot(j);
      float cs = 0.f;
#pragma unroll
      for (int w = 0; w < W; ++w) cs += part[w][sj];
      const double d = (double)cs - bp[j];
      e2 += d * d;
      part[0][sj] = cs;
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
      if (ii % A.check_every == 0) {
        err_last = err;
        if (ii == 0) {
          err_first = err;
          if (warm && !restarted && err > A.fallback_thr) stop = 2;
        }
        if (stop == 0 && err < A.stop_thr) stop = 1;
      }
      flag = stop;
    }
    __syncthreads();
    if (flag == 2) {
      // Warm start rejected: restart from the cold start u0 = 0.
      restarted = 1;
      for (int r = tid; r < n; r += T) u2[r] = 0.0;
      __syncthreads();
      goto restart;
    }
    if (flag == 1) {
      status = kConverged;
      niter = ii;
      break;
    }
    if (ii == A.max_iter - 1) break;
    // ---- column update v_j += log2 b_j - log2 colsum_j ----
    for (int j = tid; j < m; j += T) {
      const int sj = slot(j);
      const double bj = bp[j];
      const double cs = (double)part[0][sj];
      vsh[sj] = bj > 0.0 ? vsh[sj] + log2(bj) - log2(cs) : -__builtin_inf();
    }
    __syncthreads();
  }

  // ---- outputs (natural-log potentials) ----
  for (int i = tid; i < n; i += T) up[i] = u2[i] * 0.6931471805599453;
  for (int j = tid; j < m; j += T) vp[j] = vsh[slot(j)] * 0.6931471805599453;
  if (tid == 0) {
    A.niter[p] = niter;
    A.status[p] = status | (restarted ? 4 : 0);
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

int batch_waves_default() { return 8; }

int batch_occupancy_default(int waves, int pf) { return (waves == 8 && pf == 1) ? 4 : 0; }

hipError_t launch_batch_log(int cpt, int vec_ok, int waves, int pf, int occ,
                            const BatchArgs& a, int B, hipStream_t s) {
  dim3 grid(B);
#define POTAMD_BATCH_CASE(C, WV, PP, OC)                                                          \
  if (cpt == C && waves == WV && pf == PP && occ == OC) {                                        \
    if (vec_ok)                                                                                  \
      hipLaunchKernelGGL((batch_log_kernel<C, 4, WV, PP, OC>), grid, dim3(WV * kWave), 0, s, a); \
    else                                                                                         \
      hipLaunchKernelGGL((batch_log_kernel<C, 1, WV, PP, OC>), grid, dim3(WV * kWave), 0, s, a); \
    return hipGetLastError();                                                                    \
  }
  POTAMD_BATCH_CASE(4, 4, 1, 0)
  POTAMD_BATCH_CASE(8, 4, 1, 0)
  POTAMD_BATCH_CASE(16, 4, 1, 0)
  POTAMD_BATCH_CASE(4, 8, 1, 0)
  POTAMD_BATCH_CASE(8, 8, 1, 0)
  POTAMD_BATCH_CASE(16, 8, 1, 0)
  POTAMD_BATCH_CASE(16, 4, 2, 0)
  POTAMD_BATCH_CASE(16, 8, 2, 0)
  POTAMD_BATCH_CASE(4, 8, 1, 4)
  POTAMD_BATCH_CASE(8, 8, 1, 4)
  POTAMD_BATCH_CASE(16, 8, 1, 4)
#undef POTAMD_BATCH_CASE
  return hipErrorInvalidConfiguration;
}

}  // namespace potamd
