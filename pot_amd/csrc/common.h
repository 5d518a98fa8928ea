// Shared device helpers for the gfx950 Sinkhorn kernels.
//
// Conventions used by every kernel in this directory:
//  * potentials live in HBM as float64 in base-2 log units (u2 = u / ln 2),
//    so that the element exponent is an exp2, the native v_exp_f32;
//  * the cost scale s2 = -log2(e) / reg is carried as an fp32 hi/lo pair so
//    that M * s2 is exact to ~1e-14 relative before the single fp32 rounding;
//  * wavefronts are 64 lanes; reductions use cross-lane shuffles inside a
//    wave and LDS across waves.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace potamd {

constexpr int kWave = 64;

// Log / scaling domains of the solver.
enum Domain : int { kLog = 0, kScaling = 1 };
// Iteration order: column update first (ot.sinkhorn / sinkhorn_log /
// sinkhorn_knopp) or row update first (ot.batch projections).
enum Order : int { kColFirst = 0, kRowFirst = 1 };

// Control block shared by the kernels of one solve (device memory).
struct Ctrl {
  int done;        // 1 once the solve has stopped; later kernels return early
  int status;      // 0 converged, 1 max-iter, 2 numerical, 3 needs exact column pass
  int niter;       // the reference's `ii` at the stop
  int bad_iter;    // scaling domain: iteration whose update produced 0/nan/inf
  unsigned counter;  // last-block-done ticket of the merge kernel (self-resetting)
  int n_err;       // number of err values written
  int need_exact;  // log domain: a column mass fell below the trusted range
  int pad1;
  double err;      // last marginal error
};

enum Status : int {
  kConverged = 0,
  kMaxIter = 1,
  kNumerical = 2,
  kNeedExact = 3,
};

__device__ __forceinline__ float exp2_fast(float x) {
  return __builtin_amdgcn_exp2f(x);
}
__device__ __forceinline__ double exp2_fast(double x) { return exp2(x); }

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}
// Wave sum through DPP (no LDS round trips): four row-local steps give every
// lane its 16-lane row sum, then the four row sums are read from lanes 0, 16,
// 32, 48 and added in a fixed order.  The result is wave-uniform and
// deterministic.  DPP controls (GFX9): quad_perm [1,0,3,2] = 0xB1,
// quad_perm [2,3,0,1] = 0x4E, row_half_mirror = 0x141, row_mirror = 0x140.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_sum_dpp(float x) {
  x += dpp_f<0xB1>(x);
  x += dpp_f<0x4E>(x);
  x += dpp_f<0x141>(x);
  x += dpp_f<0x140>(x);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 48));
  return (r0 + r1) + (r2 + r3);
}

// The same for a double (each DPP step moves both halves); every lane gets
// the same value.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double wave_sum_dpp_d(double x) {
  x += dpp_d<0xB1>(x);
  x += dpp_d<0x4E>(x);
  x += dpp_d<0x141>(x);
  x += dpp_d<0x140>(x);
  return (readlane_d(x, 0) + readlane_d(x, 16)) + (readlane_d(x, 32) + readlane_d(x, 48));
}

__device__ __forceinline__ float wave_sum_fast(float x) { return wave_sum_dpp(x); }
__device__ __forceinline__ double wave_sum_fast(double x) { return wave_sum(x); }

__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, kWave));
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, kWave));
  return x;
}

// Workgroup barrier that orders LDS traffic only: waits for this wave's LDS
// operations, then s_barrier.  Outstanding global loads stay in flight, so a
// row prefetched before the barrier is not drained by it.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Element exponent in base 2: x2 = M * s2 + t with s2 = s_hi + s_lo.
__device__ __forceinline__ float scaled_exponent(float m, float s_hi, float s_lo,
                                                 float t) {
  return __builtin_fmaf(m, s_lo, __builtin_fmaf(m, s_hi, t));
}
__device__ __forceinline__ double scaled_exponent(double m, double s_hi, double s_lo,
                                                  double t) {
  return fma(m, s_hi, t) + m * s_lo;
}

// Real type used for element arithmetic: fp32 for fp32 costs, fp64 for fp64.
template <typename TM> struct Real { using type = float; };
template <> struct Real<double> { using type = double; };

__device__ __forceinline__ bool in_safe_range(double s) {
  // 2^-60 .. 2^60: entries that underflowed are < 2^-66 of the sum, nothing
  // overflowed (exp2 of fp32 saturates at 2^128).
  return s >= 8.673617379884035e-19 && s <= 1.152921504606847e18;
}

}  // namespace potamd
