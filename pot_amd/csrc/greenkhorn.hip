// Greedy Sinkhorn (ot.bregman.greenkhorn, ot/bregman/_sinkhorn.py:921-1102)
// on a materialised cost, the whole loop in one workgroup.
//
// Per iteration, as the reference (:1058-1086):
//   i1 = argmax |viol|, i2 = argmax |viol_2| (the first index of the
//   largest, a NaN counting as largest, as numpy's argmax);
//   stop = max(|viol[i1]|, |viol_2[i2]|) (NaN propagates, as nx.maximum);
//   if |viol[i1]| > |viol_2[i2]|, the row update
//      new_u = a[i1] / (K[i1,:] . v);  viol[i1] = (new_u K[i1,:]) . v - a[i1];
//      viol_2 += (K[i1,:] (new_u - u[i1])) v;  u[i1] = new_u
//   else the column update
//      new_v = b[i2] / (K[:,i2] . u);  viol += ((new_v - v[i2]) K[:,i2]) u;
//      viol_2[i2] = new_v (K[:,i2] . u) - b[i2];  v[i2] = new_v
//   and stop <= stopThr ends the loop, after that update (as the reference).
// The start: viol = rowsum(G) - a, viol_2 = colsum(G) - b with G = (u K) v
// (:1048-1051).  Everything is fp64; K = exp(-M / reg) is stored twice, as K
// and K^T, so a row update and a column update both read contiguous memory.
// The plan is formed after the loop as (u_i K_ij) v_j: the reference keeps G
// up to date row by row and column by column, and each entry it last wrote
// equals that product of the final u and v.
//
// The loop is sequential by construction (each update moves the next argmax),
// so one workgroup runs it (512 threads up to 2048 points a side, 1024
// above): per iteration one pass over the violations for both argmaxes, then
// two passes over a row of K (or K^T) with block reductions, 2 barriers in
// all.  Up to 4096 points a side the scalings and violations stay in LDS for
// the whole loop.
#include <limits.h>
#include <math.h>

#include "../../include/pot_amd.h"
#include "common.h"
#include "host.h"

namespace potamd {
namespace {

constexpr int kGridMax = 65535 * 32;  // rows (or columns) the tile grid covers
constexpr int kGLds = 4096;           // n, m up to this: the loop's state in LDS (128 KB)
constexpr int kGSmall = 2048;         // n, m up to this: a 512-thread loop (64 KB of LDS)

struct GkOut {
  int niter;      // the reference's ii at exit
  int stopped;    // 1: stop <= stopThr ended the loop
  int pad[2];
  double stop;    // the last stop value
};

struct GkLayout {
  double* K;
  double* KT;
  double* viol;
  double* viol2;
  GkOut* out;
  size_t bytes;
};

GkLayout carve_gk(void* base, int n, int m) {
  Carver c(base);
  GkLayout L;
  L.K = c.take<double>((size_t)n * m);
  L.KT = c.take<double>((size_t)n * m);
  L.viol = c.take<double>(n);
  L.viol2 = c.take<double>(m);
  L.out = c.take<GkOut>(1);
  L.bytes = c.off;
  return L;
}

// K = exp(-M / reg) and K^T, one 32 x 32 tile per workgroup through LDS so
// that both stores are coalesced.
template <typename TM>
__global__ __launch_bounds__(256) void gk_build_kernel(const TM* M, int64_t ldm, int n, int m,
                                                       double reg, double* K, double* KT) {
  __shared__ double tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int r = r0 + ty + k, c = c0 + tx;
    double x = 0.0;
    if (r < n && c < m) {
      x = exp(-((double)M[(int64_t)r * ldm + c] / reg));
      K[(int64_t)r * m + c] = x;
    }
    tile[ty + k][tx] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int c = c0 + ty + k, r = r0 + tx;
    if (c < m && r < n) KT[(int64_t)c * n + r] = tile[tx][ty + k];
  }
}

// Sum over a workgroup of `threads` threads, the same value (same order) in
// every thread; red holds one slot per wave.
template <int THREADS>
__device__ __forceinline__ double block_sum(double x, double* red) {
  constexpr int W = THREADS / kWave;
  x = wave_sum_dpp_d(x);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = x;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int w = 1; w < W; ++w) s += red[w];
  return s;
}

// The same sum with the per-wave partials added on DPP by one wave's lanes
// (the order differs from block_sum's; every thread gets the same value).
template <int W>
__device__ __forceinline__ double block_sum_dpp(double x, double* red) {
  const int lane = threadIdx.x & (kWave - 1);
  x = wave_sum_dpp_d(x);
  if (lane == 0) red[threadIdx.x / kWave] = x;
  __syncthreads();
  return wave_sum_dpp_d(lane < W ? red[lane] : 0.0);
}

// The starting violations: block r < n sums row r of G = (u K) v (minus
// a_r), block n + c sums column c (from row c of K^T, minus b_c).
__global__ __launch_bounds__(256) void gk_init_kernel(const double* K, const double* KT, int n,
                                                      int m, const double* a, const double* b,
                                                      const double* u, const double* v,
                                                      double* viol, double* viol2) {
  __shared__ double red[256 / kWave];
  const int row = blockIdx.x;
  const bool col = row >= n;
  const int r = col ? row - n : row;
  double s = 0.0;
  if (col) {
    const double* src = KT + (int64_t)r * n;
    const double vr = v[r];
    for (int k = threadIdx.x; k < n; k += 256) s += (u[k] * src[k]) * vr;
  } else {
    const double* src = K + (int64_t)r * m;
    const double ur = u[r];
    for (int k = threadIdx.x; k < m; k += 256) s += (ur * src[k]) * v[k];
  }
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) {
    if (col)
      viol2[r] = s - b[r];
    else
      viol[r] = s - a[r];
  }
}

// (x, i) ranks above (bx, bi): a NaN above any number, then the larger value,
// then the smaller index -- numpy's argmax order.
__device__ __forceinline__ bool gk_above(double x, int i, double bx, int bi) {
  const bool xn = __builtin_isnan(x), bn = __builtin_isnan(bx);
  if (xn != bn) return xn;
  if (xn) return i < bi;
  return x > bx || (x == bx && i < bi);
}

// The same order through DPP (no LDS round trips): the largest key -- the
// bits of a non-negative double order like its value, a NaN above +inf, and
// a lane with no candidate (-inf) counts as +0 with index INT_MAX -- then
// the smallest index holding it.  Every lane gets the same pair.
template <int CTRL>
__device__ __forceinline__ unsigned long long gk_dpp_u64(unsigned long long x) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, 0xF, 0xF, false);
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ __forceinline__ unsigned long long gk_readlane_u64(unsigned long long x, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ void gk_wave_argmax(double& bx, int& bi) {
  const unsigned long long key =
      (bx >= 0.0 || __builtin_isnan(bx)) ? (unsigned long long)__double_as_longlong(bx) : 0ull;
  unsigned long long k = key;
  k = max(k, gk_dpp_u64<0xB1>(k));   // quad_perm [1,0,3,2]
  k = max(k, gk_dpp_u64<0x4E>(k));   // quad_perm [2,3,0,1]
  k = max(k, gk_dpp_u64<0x141>(k));  // row_half_mirror
  k = max(k, gk_dpp_u64<0x140>(k));  // row_mirror
  k = max(max(gk_readlane_u64(k, 0), gk_readlane_u64(k, 16)),
          max(gk_readlane_u64(k, 32), gk_readlane_u64(k, 48)));
  int i = key == k ? bi : INT_MAX;
  i = min(i, __builtin_amdgcn_update_dpp(0, i, 0xB1, 0xF, 0xF, false));
  i = min(i, __builtin_amdgcn_update_dpp(0, i, 0x4E, 0xF, 0xF, false));
  i = min(i, __builtin_amdgcn_update_dpp(0, i, 0x141, 0xF, 0xF, false));
  i = min(i, __builtin_amdgcn_update_dpp(0, i, 0x140, 0xF, 0xF, false));
  i = min(min(__builtin_amdgcn_readlane(i, 0), __builtin_amdgcn_readlane(i, 16)),
          min(__builtin_amdgcn_readlane(i, 32), __builtin_amdgcn_readlane(i, 48)));
  bx = __longlong_as_double((long long)k);
  bi = i;
}

// The loop.  C > 0: u, v and both violation vectors live in LDS (n, m <= C)
// and each thread keeps its share of the updated row (or column) of K in
// registers between the two passes over it, so an iteration touches HBM only
// to read that row once; C == 0: the same loop with the state in HBM, for
// larger problems.
//
// Two barriers per iteration: one after the argmax scan, one in the first
// dot product.  Thread t owns entries t, t + T, ... of both violation
// vectors: it alone scans and updates them, so the updates of one iteration
// reach the next scan without a barrier.  A row update's own violation
// viol[i1] = (new_u K[i1,:]) . v - a[i1] needs a second block sum; its wave
// partials are left in rd2, and the next iteration sums them after its argmax
// barrier (the owner skips i1 in its scan and the value joins the per-wave
// pairs in the block argmax, a maximum under a total order, so the result is
// the same).  u[i1] (v[i2]) is read before the dot-product barrier
// and rewritten by thread 0 after it.
template <int C, int T>
__global__ __launch_bounds__(T) void gk_loop_kernel(const double* K, const double* KT, int n,
                                                      int m, const double* a, const double* b,
                                                      double* u, double* v, double* viol,
                                                      double* viol2, int max_iter,
                                                      double stop_thr, GkOut* out) {
  constexpr int W = T / kWave;  // waves
  static_assert(W < kWave, "one lane per wave, and one more, in the cross-wave merge");
  __shared__ double rv1[W], rv2[W], rd1[W], rd2[W];
  __shared__ int ri1[W], ri2[W];
  constexpr int CS = C > 0 ? C : 1;
  constexpr int R = C > 0 ? C / T : 1;  // row elements per thread kept in registers
  __shared__ double su[CS], sv[CS], sq[CS], sq2[CS];
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid & (kWave - 1);
  double* U = u;
  double* V = v;
  double* Q = viol;
  double* Q2 = viol2;
  if constexpr (C > 0) {
    for (int k = tid; k < n; k += T) {
      su[k] = u[k];
      sq[k] = viol[k];
    }
    for (int k = tid; k < m; k += T) {
      sv[k] = v[k];
      sq2[k] = viol2[k];
    }
    __syncthreads();
    U = su;
    V = sv;
    Q = sq;
    Q2 = sq2;
  }
  int ii = 0, stopped = 0;
  double stop = 0.0;
  int pr = -1;       // the previous iteration's row update, its violation pending
  double pa = 0.0;   // a[pr]
  for (; ii < max_iter; ++ii) {
    // ---- the two argmaxes ----
    double b1 = -__builtin_inf(), b2 = -__builtin_inf();
    int i1 = INT_MAX, i2 = INT_MAX;
    for (int k = tid; k < n; k += T) {
      const double x = fabs(Q[k]);
      if (k != pr && gk_above(x, k, b1, i1)) {
        b1 = x;
        i1 = k;
      }
    }
    for (int k = tid; k < m; k += T) {
      const double x = fabs(Q2[k]);
      if (gk_above(x, k, b2, i2)) {
        b2 = x;
        i2 = k;
      }
    }
    gk_wave_argmax(b1, i1);
    gk_wave_argmax(b2, i2);
    if (lane == 0) {
      rv1[wave] = b1;
      ri1[wave] = i1;
      rv2[wave] = b2;
      ri2[wave] = i2;
    }
    __syncthreads();
    // lane w < W takes wave w's pairs, lane W the pending row's
    b1 = b2 = -__builtin_inf();
    i1 = i2 = INT_MAX;
    if (lane < W) {
      b1 = rv1[lane];
      i1 = ri1[lane];
      b2 = rv2[lane];
      i2 = ri2[lane];
    }
    if (pr >= 0) {
      const double q = wave_sum_dpp_d(lane < W ? rd2[lane] : 0.0) - pa;
      if (tid == pr % T) Q[pr] = q;
      if (lane == W) {
        b1 = fabs(q);
        i1 = pr;
      }
      pr = -1;
    }
    gk_wave_argmax(b1, i1);
    gk_wave_argmax(b2, i2);
    // (every thread holds the same pairs; n, m >= 1 give indexes in range)
    stop = (__builtin_isnan(b1) || __builtin_isnan(b2)) ? __builtin_nan("") : fmax(b1, b2);
    const bool row = b1 > b2;
    if (row) {
      // ---- row i1 ----
      const double* Kr = K + (int64_t)i1 * m;
      const double ai = a[i1];
      const double u_old = U[i1];
      double s = 0.0, s2 = 0.0;
      double dot1, new_u, du;
      if constexpr (C > 0) {
        double kr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int j = tid + r * T;
          kr[r] = j < m ? Kr[j] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int j = tid + r * T;
          if (j < m) s += kr[r] * V[j];
        }
        dot1 = block_sum_dpp<W>(s, rd1);
        new_u = ai / dot1;
        du = new_u - u_old;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int j = tid + r * T;
          if (j < m) {
            const double vj = V[j];
            Q2[j] += (kr[r] * du) * vj;
            s2 += (new_u * kr[r]) * vj;
          }
        }
      } else {
        for (int j = tid; j < m; j += T) s += Kr[j] * V[j];
        dot1 = block_sum_dpp<W>(s, rd1);
        new_u = ai / dot1;
        du = new_u - u_old;
        for (int j = tid; j < m; j += T) {
          const double kj = Kr[j], vj = V[j];
          Q2[j] += (kj * du) * vj;
          s2 += (new_u * kj) * vj;
        }
      }
      s2 = wave_sum_dpp_d(s2);
      if (lane == 0) rd2[wave] = s2;
      if (tid == 0) U[i1] = new_u;
      pr = i1;
      pa = ai;
    } else {
      // ---- column i2 ----
      const double* Kc = KT + (int64_t)i2 * n;
      const double bi = b[i2];
      const double v_old = V[i2];
      double s = 0.0;
      double dot1, new_v, dv;
      if constexpr (C > 0) {
        double kc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = tid + r * T;
          kc[r] = i < n ? Kc[i] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = tid + r * T;
          if (i < n) s += kc[r] * U[i];
        }
        dot1 = block_sum_dpp<W>(s, rd1);
        new_v = bi / dot1;
        dv = new_v - v_old;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = tid + r * T;
          if (i < n) Q[i] += (dv * kc[r]) * U[i];
        }
      } else {
        for (int i = tid; i < n; i += T) s += Kc[i] * U[i];
        dot1 = block_sum_dpp<W>(s, rd1);
        new_v = bi / dot1;
        dv = new_v - v_old;
        for (int i = tid; i < n; i += T) Q[i] += (dv * Kc[i]) * U[i];
      }
      if (tid == i2 % T) Q2[i2] = new_v * dot1 - bi;
      if (tid == 0) V[i2] = new_v;
    }
    if (stop <= stop_thr) {
      stopped = 1;
      break;
    }
  }
  if constexpr (C > 0) {
    __syncthreads();
    for (int k = tid; k < n; k += T) u[k] = su[k];
    for (int k = tid; k < m; k += T) v[k] = sv[k];
  }
  if (tid == 0) {
    out->niter = stopped ? ii : max_iter - 1;
    out->stopped = stopped;
    out->stop = stop;
  }
}

__global__ __launch_bounds__(256) void gk_plan_kernel(const double* K, int n, int m,
                                                      const double* u, const double* v,
                                                      double* P) {
  const int64_t total = (int64_t)n * m;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e / m), j = (int)(e - (int64_t)i * m);
    P[e] = (u[i] * K[e]) * v[j];
  }
}

}  // namespace
}  // namespace potamd

using namespace potamd;

extern "C" size_t potamd_greenkhorn_workspace_size(int n, int m) {
  if (n < 1 || m < 1 || n > kGridMax || m > kGridMax) return 0;
  return carve_gk(nullptr, n, m).bytes;
}

extern "C" int potamd_greenkhorn(int dtype, const void* M, int64_t ldm, int n, int m, double reg,
                                 const double* a, const double* b, double* u, double* v,
                                 int max_iter, double stop_thr, double* P, void* workspace,
                                 size_t ws_bytes, void* stream, int* info) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || !M || !a || !b || !u || !v || !info ||
      n < 1 || m < 1 || n > kGridMax || m > kGridMax || ldm < m || !(reg > 0.0))
    return POTAMD_EINVAL;
  GkLayout L = carve_gk(workspace, n, m);
  if (!workspace || ws_bytes < L.bytes) return POTAMD_EWORKSPACE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 tiles((m + 31) / 32, (n + 31) / 32);
  if (dtype == POTAMD_F32)
    hipLaunchKernelGGL(gk_build_kernel<float>, tiles, dim3(256), 0, s,
                       static_cast<const float*>(M), ldm, n, m, reg, L.K, L.KT);
  else
    hipLaunchKernelGGL(gk_build_kernel<double>, tiles, dim3(256), 0, s,
                       static_cast<const double*>(M), ldm, n, m, reg, L.K, L.KT);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(gk_init_kernel, dim3(n + m), dim3(256), 0, s, L.K, L.KT, n, m, a, b, u, v,
                     L.viol, L.viol2);
  HIP_TRY(hipGetLastError());
  GkOut res = {};
  if (max_iter > 0) {
    if (n <= kGSmall && m <= kGSmall)
      hipLaunchKernelGGL((gk_loop_kernel<kGSmall, 512>), dim3(1), dim3(512), 0, s, L.K, L.KT, n,
                         m, a, b, u, v, L.viol, L.viol2, max_iter, stop_thr, L.out);
    else if (n <= kGLds && m <= kGLds)
      hipLaunchKernelGGL((gk_loop_kernel<kGLds, 1024>), dim3(1), dim3(1024), 0, s, L.K, L.KT, n,
                         m, a, b, u, v, L.viol, L.viol2, max_iter, stop_thr, L.out);
    else
      hipLaunchKernelGGL((gk_loop_kernel<0, 1024>), dim3(1), dim3(1024), 0, s, L.K, L.KT, n, m,
                         a, b, u, v, L.viol, L.viol2, max_iter, stop_thr, L.out);
    HIP_TRY(hipGetLastError());
  }
  if (P) {
    const int64_t total = (int64_t)n * m;
    const int blocks = (int)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
    hipLaunchKernelGGL(gk_plan_kernel, dim3(blocks), dim3(256), 0, s, L.K, n, m, u, v, P);
    HIP_TRY(hipGetLastError());
  }
  if (max_iter > 0) HIP_TRY(hipMemcpyAsync(&res, L.out, sizeof(GkOut), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  info[0] = max_iter > 0 ? res.niter : -1;
  info[1] = max_iter > 0 ? res.stopped : 0;
  info[2] = res.stopped ? POTAMD_CONVERGED : POTAMD_MAX_ITER;
  info[3] = 0;
  return info[2];
}
