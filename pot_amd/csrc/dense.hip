// Dense-cost Sinkhorn for one problem (or a batch of equal-shape problems):
// one HBM sweep of the cost per iteration.
//
// The reference alternates two full passes over the n x m cost per iteration
// (ot/bregman/_sinkhorn.py:881-882 for sinkhorn_log, :629-631 for
// sinkhorn_knopp, ot/batch/_utils.py:269-270 and :142-143 for the batch).
// Here a workgroup owns whole rows (its threads cover all m columns), so it
// can finish the row update of a row and, from the same registers, add that
// row's contribution to the column sums of the next half-step:
//
//   row_sweep   u_i    <- row update from v                 (exact row LSE)
//               col_j  += P_ij(u_new, v) over the block's rows  -> partials
//   col_merge   colsum_j = sum of partials; marginal error = |colsum - b|;
//               v_j    <- v_j + log b_j - log colsum_j      (exact in exact math)
//
// so each iteration reads M once and the reference's every-10th-iteration
// marginal check costs no extra pass.  The row LSE is shifted by the previous
// row potential (c_i = log a_i - u_i): the shifted row sum is then the
// previous plan's row mass over a_i, bounded in [min b, 1/a_i], so one exp2
// per element suffices; if a sum leaves [2^-60, 2^60] the row is redone with
// its exact maximum as the shift.
#include "internal.h"

namespace potamd {


template <int W, typename R>
__device__ __forceinline__ R block_sum(R x, R* red, int& parity) {
  x = wave_sum_fast(x);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) red[parity * W + wid] = x;
  lds_barrier();
  R s = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) s += red[parity * W + w];
  parity ^= 1;
  return s;
}

template <int W, typename R>
__device__ __forceinline__ R block_max(R x, R* red, int& parity) {
  x = wave_max(x);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) red[parity * W + wid] = x;
  lds_barrier();
  R s = red[parity * W];
#pragma unroll
  for (int w = 1; w < W; ++w) s = s > red[parity * W + w] ? s : red[parity * W + w];
  parity ^= 1;
  return s;
}

// Vector of VEC elements loaded with one 16-byte (or scalar) access.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename TM, int VEC> struct VecT;
template <> struct VecT<float, 4> { using type = f32x4; };
template <> struct VecT<float, 1> { using type = float; };
template <> struct VecT<double, 2> { using type = f64x2; };
template <> struct VecT<double, 1> { using type = double; };

template <typename TM, int VEC>
__device__ __forceinline__ void load_vec(const TM* p, TM* out) {
  using V = typename VecT<TM, VEC>::type;
  V x = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
  const TM* e = reinterpret_cast<const TM*>(&x);
#pragma unroll
  for (int c = 0; c < VEC; ++c) out[c] = e[c];
}

// Row sweep.  MODE 0: row update + column accumulation.  MODE 1 (scaling
// domain only): column accumulation with u_old as the row scale, no update
// (the knopp prologue K^T u0).
//
// Inside the row loop only element arithmetic runs; the per-row shift c_i
// and shifted sum S_i go to LDS and the fp64 potential update
// u_i = log2 a_i - c_i - log2 S_i is done after the loop, so the transient
// registers of the fp64 log2 do not add to the loop's live set.
constexpr int kMaxRowsPerBlock = 256;

template <typename TM, int T, int CPT, int VEC, int DOMAIN, int MODE, int PF>
__global__ __launch_bounds__(T) void row_sweep_kernel(RowSweepArgs A) {
  using R = typename Real<TM>::type;
  constexpr int W = T / kWave;
  constexpr int NCH = CPT / VEC;  // vector chunks per thread
  if (A.ctrl->done) return;
  const int p = blockIdx.y;
  if (A.pdone && A.pdone[p]) return;  // this problem has stopped on its own
  const int g = blockIdx.x;
  const int G = gridDim.x;
  const int tid = threadIdx.x;
  const int n = A.n, m = A.m;
  const TM* Mp = static_cast<const TM*>(A.M) + (int64_t)p * A.bstride;
  const double* ap = A.a + (int64_t)p * n;
  const double* la2 = A.loga2 + (int64_t)p * n;
  const double* uo = A.u_old + (int64_t)p * n;
  double* un = A.u_new + (int64_t)p * n;
  const double* vp = A.v + (int64_t)p * m;

  __shared__ R red[2 * W];
  __shared__ double row_c[kMaxRowsPerBlock];
  __shared__ double row_s[kMaxRowsPerBlock];
  __shared__ double red_d[W];
  // Column potentials (log domain) or scalings (scaling domain), fp64, kept
  // in LDS: read once per element, at a small fraction of LDS bandwidth,
  // which leaves the registers to the row data and its prefetch.
  __shared__ double vsh[T * CPT];
  int parity = 0;

  R acc[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) acc[q] = 0;
  for (int j = tid; j < T * CPT; j += T) {
    const bool ok = j < m;
    if (DOMAIN == kLog)
      vsh[j] = ok ? vp[j] : -__builtin_inf();
    else
      vsh[j] = ok ? vp[j] : 0.0;
  }
  __syncthreads();
  auto vcol = [&](int q) -> double {
    return vsh[((q / VEC) * T + tid) * VEC + (q % VEC)];
  };

  // Rows are dealt round-robin: workgroup g owns rows g, g + G, g + 2G, ...
  // so at any moment the G workgroups stream neighbouring rows (one
  // contiguous region) instead of G regions a row block apart.
  const int nloc = g < n ? (n - g + G - 1) / G : 0;
  auto row_of = [&](int k) { return g + k * G; };

  // The per-row scalars of this workgroup's rows are staged in LDS before the
  // row loop: a global load of them inside the loop would be issued after the
  // row prefetch, and its wait (the vector memory counter drains in order)
  // would also wait for the prefetch.
  __shared__ double row_a[kMaxRowsPerBlock];
  __shared__ double row_la[kMaxRowsPerBlock];
  __shared__ double row_uo[kMaxRowsPerBlock];
  for (int r = tid; r < nloc; r += T) {
    const int i = row_of(r);
    row_a[r] = ap[i];
    row_la[r] = DOMAIN == kLog ? la2[i] : 0.0;
    row_uo[r] = uo[i];
  }
  __syncthreads();

  // PF + 1 row buffers: while row k is processed, rows k+1 .. k+PF are in
  // flight.  The loop is unrolled over the buffers so they rotate without
  // register copies (a copy would wait for the prefetch).  Every load is
  // unconditional, because the compiler's wait counts assume the fewest
  // loads any path may have issued: a lane past the last column loads
  // column 0 of the row (its value is replaced by 0 where it is used), and a
  // prefetch past the workgroup's last row reads that row again.
  constexpr int NB = PF + 1;
  const bool padded = m < T * CPT;
  R buf[NB][CPT];
  auto load_row = [&](int i, R* dst) {
    const TM* row = Mp + (int64_t)i * A.ldm;
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const int j0 = (k * T + tid) * VEC;
      TM tmp[VEC];
      load_vec<TM, VEC>(row + (j0 < m ? j0 : 0), tmp);
#pragma unroll
      for (int c = 0; c < VEC; ++c) dst[k * VEC + c] = (R)tmp[c];
    }
  };
  auto mask_row = [&](R* cur) {
    if (padded) {
#pragma unroll
      for (int k = 0; k < NCH; ++k) {
        const bool ok = (k * T + tid) * VEC < m;
#pragma unroll
        for (int c = 0; c < VEC; ++c) cur[k * VEC + c] = ok ? cur[k * VEC + c] : (R)0;
      }
    }
  };
  if (nloc > 0) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      load_row(row_of(b < nloc ? b : nloc - 1), buf[b]);
      // Issue the buffers in order: the waits inside the loop count on it.
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  const R s_hi = (R)(sizeof(TM) == 4 ? (double)A.s_hi : A.s2);
  const R s_lo = (R)(sizeof(TM) == 4 ? (double)A.s_lo : 0.0);

  auto process = [&](int k, R* cur) {
    const int i = row_of(k);
    mask_row(cur);
    const double a_i = row_a[k];
    if (DOMAIN == kLog) {
      if (a_i == 0.0) {
        if (tid == 0) { row_c[k] = __builtin_nan(""); row_s[k] = 0.0; }
      } else {
        double c2 = row_la[k] - row_uo[k];
        if (!__builtin_isfinite(c2)) c2 = 0.0;
        R s = 0;
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
          const R t = (R)(vcol(q) - c2);
          cur[q] = exp2_fast(scaled_exponent(cur[q], s_hi, s_lo, t));
          s += cur[q];
        }
        double S = (double)block_sum<W>(s, red, parity);
        if (!in_safe_range(S)) {
          // Exact path: reload the row and shift by its maximum.
          load_row(i, cur);
          mask_row(cur);
          R mx = -(R)__builtin_inf();
#pragma unroll
          for (int q = 0; q < CPT; ++q) {
            const R x = scaled_exponent(cur[q], s_hi, s_lo, (R)(vcol(q) - c2));
            mx = x > mx ? x : mx;
          }
          const double gmx = (double)block_max<W>(mx, red, parity);
          if (__builtin_isfinite(gmx)) c2 += gmx;
          s = 0;
#pragma unroll
          for (int q = 0; q < CPT; ++q) {
            const R t = (R)(vcol(q) - c2);
            cur[q] = exp2_fast(scaled_exponent(cur[q], s_hi, s_lo, t));
            s += cur[q];
          }
          S = (double)block_sum<W>(s, red, parity);
        }
        if (tid == 0) { row_c[k] = c2; row_s[k] = S; }
        R kf;
        if (sizeof(R) == 4)
          kf = (R)((float)a_i * __builtin_amdgcn_rcpf((float)S));
        else
          kf = (R)(a_i / S);
#pragma unroll
        for (int q = 0; q < CPT; ++q) acc[q] = __builtin_fma(cur[q], kf, acc[q]);
      }
    } else {  // scaling domain
#pragma unroll
      for (int q = 0; q < CPT; ++q)
        if (!A.kernel_given) cur[q] = exp2_fast(scaled_exponent(cur[q], s_hi, s_lo, (R)0));
      if (MODE == 1) {
        const R f = (R)row_uo[k];
#pragma unroll
        for (int q = 0; q < CPT; ++q) acc[q] = __builtin_fma(cur[q], f, acc[q]);
      } else {
        R s = 0;
#pragma unroll
        for (int q = 0; q < CPT; ++q) s = __builtin_fma(cur[q], (R)vcol(q), s);
        const double S = (double)block_sum<W>(s, red, parity);
        if (tid == 0) row_s[k] = S;
        R f;
        if (sizeof(R) == 4)
          f = (R)((float)a_i * __builtin_amdgcn_rcpf((float)S));
        else
          f = (R)(a_i / S);
#pragma unroll
        for (int q = 0; q < CPT; ++q) acc[q] = __builtin_fma(cur[q], f, acc[q]);
      }
    }
  };

  for (int k0 = 0; k0 < nloc; k0 += NB) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int k = k0 + b;
      if (k < nloc) process(k, buf[b]);
      load_row(row_of(k + NB < nloc ? k + NB : nloc - 1), buf[b]);
    }
  }

  // Column partials of this row block.
  R* cp = static_cast<R*>(A.colpart) + ((int64_t)p * G + g) * m;
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
#pragma unroll
    for (int c = 0; c < VEC; ++c) {
      const int j = (k * T + tid) * VEC + c;
      if (j < m) cp[j] = acc[k * VEC + c];
    }
  }
  if (MODE == 1) return;

  // fp64 row potentials and row-marginal error of the previous plan.
  __syncthreads();
  double rerr = 0.0;
  int bad = 0;
  for (int r = tid; r < nloc; r += T) {
    const int i = row_of(r);
    const double a_i = ap[i];
    const double S = row_s[r];
    if (DOMAIN == kLog) {
      const double c2 = row_c[r];
      if (a_i == 0.0) {
        un[i] = -__builtin_inf();
      } else {
        un[i] = la2[i] - c2 - log2(S);
        if (A.row_first) {
          const double rs = exp2(uo[i] + c2) * S;
          rerr += (rs - a_i) * (rs - a_i);
        }
      }
    } else {
      const double f = a_i / S;
      un[i] = f;
      if (!__builtin_isfinite(f)) bad = 1;
      if (A.row_first) {
        const double rs = uo[i] * S;
        rerr += (rs - a_i) * (rs - a_i);
      }
    }
  }
  if (A.row_first) {
    rerr = wave_sum(rerr);
    if ((tid & (kWave - 1)) == 0) red_d[tid / kWave] = rerr;
  }
  bad = __syncthreads_or(bad);
  if (tid == 0) {
    if (A.row_first) {
      double t = 0.0;
      for (int w = 0; w < W; ++w) t += red_d[w];
      A.rowerr[(int64_t)p * G + g] = t;
    }
    if (bad) A.bad[A.iter & 1] = 1;
  }
}

// Column merge + marginal check partials + column update; the stop decision
// follows in col_decide_kernel.  A block owns kMergeCols columns; its 256 threads are 16
// column quads (4 adjacent columns, one 16-byte load) by 16 groups of row
// blocks, so each thread keeps all its loads in flight; the 16 group sums
// meet in LDS and are added in a fixed order (deterministic result).
constexpr int kMergeCols = 64;

template <typename TP>
__global__ __launch_bounds__(256) void col_merge_kernel(MergeArgs A) {
  Ctrl* ctrl = A.ctrl;
  if (ctrl->done) return;
  constexpr int NQ = kMergeCols / 4;  // column quads per block
  constexpr int NG = 256 / NQ;        // groups of row blocks
  __shared__ double gsum[NG][kMergeCols];
  const int p = blockIdx.y;
  if (A.pstate && A.pstate[p]) return;  // this problem has stopped on its own
  const int tid = threadIdx.x;
  const int m = A.m;
  const bool col_first = A.order == kColFirst;
  const bool check_col = !A.prologue && col_first && (A.iter % 10 == 0);
  const bool check_row = !A.prologue && !col_first && A.iter >= 1 && ((A.iter - 1) % 10 == 0);
  const int next_slot = (A.iter + 1) & 1;

  {
    const int qd = tid % NQ, grp = tid / NQ;
    const int j0 = blockIdx.x * kMergeCols + qd * 4;
    const int per = (A.G + NG - 1) / NG;
    const int g0 = grp * per;
    const int g1 = min(A.G, g0 + per);
    const TP* cp = static_cast<const TP*>(A.colpart) + (int64_t)p * A.G * m;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    bool vec = false;
    if constexpr (sizeof(TP) == 4)
      vec = (m & 3) == 0 && (reinterpret_cast<uintptr_t>(A.colpart) & 15) == 0;
    if (vec) {
      // m % 4 == 0 and j0 % 4 == 0: the quad is whole when it starts inside.
      if (j0 < m) {
#pragma unroll 8
        for (int g = g0; g < g1; ++g) {
          const f32x4 x = *reinterpret_cast<const f32x4*>(cp + (int64_t)g * m + j0);
          part[0] += (double)x[0];
          part[1] += (double)x[1];
          part[2] += (double)x[2];
          part[3] += (double)x[3];
        }
      }
    } else {
      for (int g = g0; g < g1; ++g) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (j0 + c < m) part[c] += (double)cp[(int64_t)g * m + j0 + c];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) gsum[grp][qd * 4 + c] = part[c];
  }
  __syncthreads();

  const int cl = tid;  // column of this block handled by threads 0..kMergeCols-1
  const int j = blockIdx.x * kMergeCols + cl;
  double e2 = 0.0;
  int bad = 0, need_exact = 0;
  if (tid < kMergeCols && j < m) {
    double cs = 0.0;
#pragma unroll
    for (int g = 0; g < NG; ++g) cs += gsum[g][cl];
    const double bj = A.b[(int64_t)p * m + j];
    const double vc = A.v_cur[(int64_t)p * m + j];
    double vn;
    if (A.domain == kLog) {
      if (check_col) e2 = (cs - bj) * (cs - bj);
      if (bj == 0.0) {
        vn = -__builtin_inf();
      } else {
        vn = vc + A.logb2[(int64_t)p * m + j] - log2(cs);
        // Column mass below the fp32 normal range cannot be trusted.
        const double floor_cs = sizeof(TP) == 4 ? 1e-36 : 1e-300;
        if (!(cs >= floor_cs) || !__builtin_isfinite(vn)) need_exact = 1;
      }
    } else {
      if (check_col) {
        const double mj = vc * cs;
        e2 = (mj - bj) * (mj - bj);
      }
      vn = bj / cs;
      if (cs == 0.0 || !__builtin_isfinite(vn)) bad = 1;
    }
    A.v_next[(int64_t)p * m + j] = vn;
  }
  // Block reduction of e2.
  __shared__ double red[4];
  double s = wave_sum(e2);
  if ((tid & 63) == 0) red[tid >> 6] = s;
  const int any_bad = __syncthreads_or(bad);
  const int any_exact = __syncthreads_or(need_exact);
  if (tid == 0) {
    A.errpart[(int64_t)p * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    if (any_bad) A.bad[next_slot] = 1;
    if (any_exact) atomicOr(&ctrl->need_exact, 1);
  }
}

// Stop decision of one iteration, one workgroup, launched after col_merge:
// the kernel boundary makes every block's partials and flags visible, where a
// last-block ticket inside col_merge needed a device-scope fence in every
// block (an L2 write-back and invalidate per block).  The error partials are
// reduced in a fixed order (lane-strided sums, then a butterfly per wave, then
// the waves in order), so the result is deterministic.
__global__ __launch_bounds__(256) void col_decide_kernel(MergeArgs A) {
  Ctrl* ctrl = A.ctrl;
  if (ctrl->done || A.prologue) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int it = A.iter;
  const bool col_first = A.order == kColFirst;
  const bool check_col = col_first && (it % 10 == 0);
  const bool check_row = !col_first && it >= 1 && ((it - 1) % 10 == 0);
  const int nblk = (A.m + kMergeCols - 1) / kMergeCols;
  if (A.err_norm == 2) {
    // Each problem stops on its own check, exactly as a lone solve would:
    // a wave per problem (problems w, w + 4, ...), then the solve stops when
    // none is left running or at the last iteration.
    __shared__ int running;
    __shared__ int need_exact;
    if (tid == 0) {
      running = 0;
      need_exact = ctrl->need_exact;
    }
    __syncthreads();
    if (need_exact) {
      if (tid == 0) {
        ctrl->done = 1;
        ctrl->status = kNeedExact;
        ctrl->niter = it;
      }
      return;
    }
    int* pdone = A.pstate;
    int* pniter = A.pstate + A.B;
    int* pstat = A.pstate + 2 * A.B;
    const bool last = it >= A.max_iter - 1;
    for (int q = w; q < A.B; q += 4) {
      if (pdone[q]) continue;  // uniform per wave
      bool stop = false;
      if (check_col || check_row) {
        const double* src = check_col ? A.errpart + (int64_t)q * nblk : A.rowerr + (int64_t)q * A.G;
        const int cnt = check_col ? nblk : A.G;
        double acc = 0.0;
        for (int b = lane; b < cnt; b += 64) acc += src[b];
        acc = wave_sum(acc);
        stop = sqrt(acc) < A.stop_thr;
      }
      if (lane == 0) {
        if (stop) {
          pdone[q] = 1;
          pniter[q] = check_col ? it : it - 1;
          pstat[q] = kConverged;
        } else if (last) {
          pdone[q] = 1;
          pniter[q] = it;
          pstat[q] = kMaxIter;
        } else {
          atomicAdd(&running, 1);
        }
      }
    }
    __syncthreads();
    if (tid == 0 && running == 0) {
      ctrl->done = 1;
      ctrl->status = last ? kMaxIter : kConverged;
      ctrl->niter = it;
    }
    return;
  }
  __shared__ double wpart[4];
  if (check_col || check_row) {
    double tot = 0.0;
    for (int q = w; q < A.B; q += 4) {
      const double* src = check_col ? A.errpart + (int64_t)q * nblk : A.rowerr + (int64_t)q * A.G;
      const int cnt = check_col ? nblk : A.G;
      double acc = 0.0;
      for (int b = lane; b < cnt; b += 64) acc += src[b];
      acc = wave_sum(acc);
      if (A.err_norm == 1) {
        tot += acc;  // Frobenius norm over the batch
      } else {
        const double e = sqrt(acc);
        tot = e > tot ? e : tot;  // maximum over the batch
      }
    }
    if (lane == 0) wpart[w] = tot;
  }
  __syncthreads();
  if (tid != 0) return;
  const int cur_slot = it & 1;
  // Numerical guard of the scaling domain (ot/bregman/_sinkhorn.py:633-645).
  if (A.domain == kScaling && A.bad[cur_slot]) {
    ctrl->done = 1;
    ctrl->status = kNumerical;
    ctrl->niter = it;
    ctrl->bad_iter = it;
    return;
  }
  A.bad[cur_slot] = 0;
  if (ctrl->need_exact) {
    ctrl->done = 1;
    ctrl->status = kNeedExact;
    ctrl->niter = it;
    return;
  }
  if (check_col || check_row) {
    double err;
    if (A.err_norm == 1) {
      err = sqrt(((wpart[0] + wpart[1]) + wpart[2]) + wpart[3]);
    } else {
      err = wpart[0];
      for (int k = 1; k < 4; ++k) err = wpart[k] > err ? wpart[k] : err;
    }
    ctrl->err = err;
    const int k = ctrl->n_err;
    if (k < A.err_cap) A.err_hist[k] = err;
    ctrl->n_err = k + 1;
    if (err < A.stop_thr) {
      ctrl->done = 1;
      ctrl->status = kConverged;
      ctrl->niter = check_col ? it : it - 1;
      return;
    }
  }
  if (it >= A.max_iter - 1) {
    ctrl->done = 1;
    ctrl->status = kMaxIter;
    ctrl->niter = it;
  }
}

// Exact column log-sum-exp in base 2 (prologue of the log domain,
// v1 = log b - LSE_i(Mr + u0), ot/bregman/_sinkhorn.py:881 at ii = 0).
// One column per thread, rows split over gridDim.y blocks; online
// (max, sum) in fp64.

template <typename TM>
__global__ __launch_bounds__(256) void col_lse_partial_kernel(ColLseArgs A) {
  const int p = blockIdx.z, g = blockIdx.y, G = gridDim.y;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= A.m) return;
  const TM* Mp = static_cast<const TM*>(A.M) + (int64_t)p * A.bstride;
  const double* u = A.u2 + (int64_t)p * A.n;
  const int r0 = g * A.rpb, r1 = min(A.n, r0 + A.rpb);
  double mx = -__builtin_inf(), sm = 0.0;
  for (int i = r0; i < r1; ++i) {
    const double x = (double)Mp[(int64_t)i * A.ldm + j] * A.s2 + u[i];
    if (x == -__builtin_inf()) continue;
    if (x > mx) {
      sm = sm * exp2(mx - x) + 1.0;
      mx = x;
    } else {
      sm += exp2(x - mx);
    }
  }
  const int64_t o = ((int64_t)p * G + g) * A.m + j;
  A.pmax[o] = mx;
  A.psum[o] = sm;
}

__global__ __launch_bounds__(256) void col_lse_merge_kernel(int m, int G, const double* pmax,
                                                            const double* psum,
                                                            const double* logb2, double* v_out) {
  const int p = blockIdx.y;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= m) return;
  double mx = -__builtin_inf();
  for (int g = 0; g < G; ++g) {
    const double x = pmax[((int64_t)p * G + g) * m + j];
    mx = x > mx ? x : mx;
  }
  double sm = 0.0;
  if (mx != -__builtin_inf()) {
    for (int g = 0; g < G; ++g) {
      const int64_t o = ((int64_t)p * G + g) * m + j;
      sm += psum[o] * exp2(pmax[o] - mx);
    }
  }
  const double lse = mx + log2(sm);
  v_out[(int64_t)p * m + j] = logb2[(int64_t)p * m + j] - lse;
}

// ---------------------------------------------------------------------------
// Small elementwise helpers of the solve set-up / tear-down.
__global__ void log2_kernel(int64_t count, const double* x, double* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) y[i] = log2(x[i]);
}
__global__ void scale_copy_kernel(int64_t count, const double* x, double f, double* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) y[i] = x ? x[i] * f : 0.0;
}
__global__ void fill_kernel(int64_t count, double val, double* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) y[i] = val;
}
__global__ void pstate_init_kernel(int* pstate, int B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B) {
    pstate[i] = 0;
    pstate[B + i] = -1;
    pstate[2 * B + i] = kMaxIter;
  }
}
// Final potentials of a per-problem solve: problem p's state sits in the
// ping-pong buffer of its own stop, (niter_p + 1) & 1.
__global__ void gather_final_kernel(int B, int n, int m, const int* pstate, const double* U0,
                                    const double* U1, const double* V0, const double* V1,
                                    double f, double* u, double* v, int* niters, int* statuses) {
  const int p = blockIdx.y;
  const int slot = (pstate[B + p] + 1) & 1;
  const double* Us = slot ? U1 : U0;
  const double* Vs = slot ? V1 : V0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    u[(int64_t)p * n + i] = Us[(int64_t)p * n + i] * f;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x)
    v[(int64_t)p * m + j] = Vs[(int64_t)p * m + j] * f;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (niters) niters[p] = pstate[B + p];
    if (statuses) statuses[p] = pstate[2 * B + p];
  }
}
__global__ void ctrl_init_kernel(Ctrl* c, int* bad) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    c->done = 0;
    c->status = 0;
    c->niter = -1;
    c->bad_iter = -1;
    c->counter = 0;
    c->n_err = 0;
    c->need_exact = 0;
    c->pad1 = 0;
    c->err = 0.0;
    if (bad) {
      bad[0] = 0;
      bad[1] = 0;
    }
  }
}

}  // namespace potamd

// ---------------------------------------------------------------------------
// Host launchers (internal C++ API, declared in internal.h).

namespace potamd {

namespace {

template <typename TM, int T, int CPT, int VEC, int PF>
hipError_t launch_row_t(int domain, int mode, const RowSweepArgs& a, dim3 grid, hipStream_t s) {
  if (domain == kLog) {
    hipLaunchKernelGGL((row_sweep_kernel<TM, T, CPT, VEC, kLog, 0, PF>), grid, dim3(T), 0, s, a);
  } else if (mode == 0) {
    hipLaunchKernelGGL((row_sweep_kernel<TM, T, CPT, VEC, kScaling, 0, PF>), grid, dim3(T), 0, s,
                       a);
  } else {
    hipLaunchKernelGGL((row_sweep_kernel<TM, T, CPT, VEC, kScaling, 1, PF>), grid, dim3(T), 0, s,
                       a);
  }
  return hipGetLastError();
}

template <typename TM, int VEC>
hipError_t launch_row_v(int T, int cpt, int pf, int domain, int mode, const RowSweepArgs& a,
                        dim3 grid, hipStream_t s) {
#define POTAMD_ROW_CASE(TT, CC, PP)                                                   \
  if (T == TT && cpt == CC && pf == PP)                                              \
    return launch_row_t<TM, TT, CC, VEC, PP>(domain, mode, a, grid, s);
  if constexpr (sizeof(TM) == 4) {
    POTAMD_ROW_CASE(64, 4, 1)
    POTAMD_ROW_CASE(256, 4, 1)
    POTAMD_ROW_CASE(256, 8, 1)
    POTAMD_ROW_CASE(512, 8, 1)
    POTAMD_ROW_CASE(1024, 8, 1)
    POTAMD_ROW_CASE(1024, 16, 1)
    POTAMD_ROW_CASE(1024, 8, 2)
    POTAMD_ROW_CASE(1024, 16, 2)
    POTAMD_ROW_CASE(512, 32, 1)
    POTAMD_ROW_CASE(512, 32, 2)
  } else {
    POTAMD_ROW_CASE(64, 2, 1)
    POTAMD_ROW_CASE(256, 2, 1)
    POTAMD_ROW_CASE(256, 4, 1)
    POTAMD_ROW_CASE(512, 4, 1)
    POTAMD_ROW_CASE(1024, 4, 1)
    POTAMD_ROW_CASE(1024, 8, 1)
  }
#undef POTAMD_ROW_CASE
  return hipErrorInvalidConfiguration;
}

}  // namespace

namespace {
struct RowCfg {
  int T, cpt, pf;
};
// Tile shapes compiled per dtype; the default for m is the first entry that
// covers it (the autotuner may pick any entry that covers m).
const RowCfg kRowF32[] = {{64, 4, 1},    {256, 4, 1},   {256, 8, 1},  {512, 8, 1},
                          {1024, 8, 1},  {1024, 16, 1}, {1024, 8, 2}, {1024, 16, 2},
                          {512, 32, 1},  {512, 32, 2}};
const RowCfg kRowF64[] = {{64, 2, 1},  {256, 2, 1},  {256, 4, 1},
                          {512, 4, 1}, {1024, 4, 1}, {1024, 8, 1}};
}  // namespace

bool row_config_valid(int dtype, int T, int cpt, int pf) {
  const RowCfg* tab = dtype == 0 ? kRowF32 : kRowF64;
  const int cnt = dtype == 0 ? sizeof(kRowF32) / sizeof(RowCfg) : sizeof(kRowF64) / sizeof(RowCfg);
  for (int q = 0; q < cnt; ++q)
    if (tab[q].T == T && tab[q].cpt == cpt && tab[q].pf == pf) return true;
  return false;
}

void row_config_default(int dtype, int m, int* T, int* cpt, int* pf) {
  const RowCfg* tab = dtype == 0 ? kRowF32 : kRowF64;
  const int cnt = dtype == 0 ? sizeof(kRowF32) / sizeof(RowCfg) : sizeof(kRowF64) / sizeof(RowCfg);
  *T = 0;
  *cpt = 0;
  *pf = 1;
  for (int q = 0; q < cnt; ++q) {
    if ((int64_t)tab[q].T * tab[q].cpt >= m) {
      *T = tab[q].T;
      *cpt = tab[q].cpt;
      *pf = tab[q].pf;
      return;
    }
  }
}

hipError_t launch_row_sweep(int dtype, int vec_ok, int T, int cpt, int pf, int domain, int mode,
                            const RowSweepArgs& a, int G, int B, hipStream_t s) {
  dim3 grid(G, B);
  if (dtype == 0) {
    return vec_ok ? launch_row_v<float, 4>(T, cpt, pf, domain, mode, a, grid, s)
                  : launch_row_v<float, 1>(T, cpt, pf, domain, mode, a, grid, s);
  }
  return vec_ok ? launch_row_v<double, 2>(T, cpt, pf, domain, mode, a, grid, s)
                : launch_row_v<double, 1>(T, cpt, pf, domain, mode, a, grid, s);
}

hipError_t launch_col_merge(int dtype, const MergeArgs& a, int B, hipStream_t s) {
  dim3 grid((a.m + kMergeCols - 1) / kMergeCols, B);
  if (dtype == 0)
    hipLaunchKernelGGL(col_merge_kernel<float>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(col_merge_kernel<double>, grid, dim3(256), 0, s, a);
  if (!a.prologue) hipLaunchKernelGGL(col_decide_kernel, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_col_lse(int dtype, const ColLseArgs& a, int G, int B, const double* logb2,
                          double* v_out, hipStream_t s) {
  dim3 grid((a.m + 255) / 256, G, B);
  if (dtype == 0)
    hipLaunchKernelGGL(col_lse_partial_kernel<float>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(col_lse_partial_kernel<double>, grid, dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(col_lse_merge_kernel, dim3((a.m + 255) / 256, B), dim3(256), 0, s, a.m, G,
                     a.pmax, a.psum, logb2, v_out);
  return hipGetLastError();
}

static inline unsigned nblocks(int64_t count, int t) { return (unsigned)((count + t - 1) / t); }

hipError_t launch_log2(int64_t count, const double* x, double* y, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(log2_kernel, dim3(nblocks(count, 256)), dim3(256), 0, s, count, x, y);
  return hipGetLastError();
}
hipError_t launch_pstate_init(int* pstate, int B, hipStream_t s) {
  hipLaunchKernelGGL(pstate_init_kernel, dim3((B + 255) / 256), dim3(256), 0, s, pstate, B);
  return hipGetLastError();
}

hipError_t launch_gather_final(int B, int n, int m, const int* pstate, double* const* U,
                               double* const* V, double f, double* u, double* v, int* niters,
                               int* statuses, hipStream_t s) {
  const int mx = n > m ? n : m;
  const int gx = (mx + 255) / 256 < 64 ? (mx + 255) / 256 : 64;
  hipLaunchKernelGGL(gather_final_kernel, dim3(gx, B), dim3(256), 0, s, B, n, m, pstate, U[0],
                     U[1], V[0], V[1], f, u, v, niters, statuses);
  return hipGetLastError();
}

hipError_t launch_scale_copy(int64_t count, const double* x, double f, double* y, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(scale_copy_kernel, dim3(nblocks(count, 256)), dim3(256), 0, s, count, x, f,
                     y);
  return hipGetLastError();
}
hipError_t launch_fill(int64_t count, double val, double* y, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(count, 256)), dim3(256), 0, s, count, val, y);
  return hipGetLastError();
}
hipError_t launch_ctrl_init(Ctrl* c, int* bad, hipStream_t s) {
  hipLaunchKernelGGL(ctrl_init_kernel, dim3(1), dim3(64), 0, s, c, bad);
  return hipGetLastError();
}

}  // namespace potamd
