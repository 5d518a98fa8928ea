// Transport plan, its cost and entropy terms from converged potentials.
//
// P_ij = exp(-M_ij / reg + u_i + v_j) (log domain, ot/bregman/_sinkhorn.py:915-918,
// ot/batch/_utils.py:286-290) or P_ij = u_i K_ij v_j (scaling domain,
// _sinkhorn.py:683-687, _utils.py:160).  Every element is formed in fp64 from
// the fp64 potentials and the input cost, then stored in the cost's dtype;
// the per-problem sums are fp64 with a fixed reduction order:
//   sums[0] = sum M * P                  (sinkhorn2 loss, value_linear)
//   sums[1] = sum P * log(P + 1e-16)     (-entropy_batch, _utils.py:4-22)
//   sums[2] = sum P * log(P / (a b) + 1e-16)  (KL value, _linear.py:342-345)
//   sums[3] = sum P
#include "internal.h"

namespace potamd {

constexpr int kPlanThreads = 256;
constexpr int kPlanSums = 4;

template <typename TM>
__global__ __launch_bounds__(kPlanThreads) void plan_kernel(PlanArgs A) {
  const int p = blockIdx.y;
  const int n = A.n, m = A.m;
  const TM* Mp = static_cast<const TM*>(A.M) + (int64_t)p * A.bstride;
  const double* u = A.u + (int64_t)p * n;
  const double* v = A.v + (int64_t)p * m;
  TM* P = A.P ? static_cast<TM*>(A.P) + (int64_t)p * A.pstride : nullptr;
  TM* L = A.logP ? static_cast<TM*>(A.logP) + (int64_t)p * A.pstride : nullptr;
  const double inv_reg = A.kernel_given ? 0.0 : 1.0 / A.reg;
  double s_cost = 0, s_ent = 0, s_kl = 0, s_mass = 0;
  const int r0 = blockIdx.x * A.rpb, r1 = min(n, r0 + A.rpb);
  for (int i = r0; i < r1; ++i) {
    const double ui = u[i];
    const double ai = A.a ? A.a[(int64_t)p * n + i] : 1.0;
    for (int j = threadIdx.x; j < m; j += kPlanThreads) {
      const double mij = (double)Mp[(int64_t)i * A.ldm + j];
      double pij, xij = 0.0;
      if (A.domain == kLog) {
        xij = (A.kernel_given ? mij : -mij * inv_reg) + ui + v[j];
        pij = exp(xij);
      } else {
        const double kij = A.kernel_given ? mij : exp(-mij * inv_reg);
        pij = ui * kij * v[j];
      }
      if (P) P[(int64_t)i * m + j] = (TM)pij;
      if (L) L[(int64_t)i * m + j] = (TM)xij;
      s_cost += mij * pij;
      s_ent += pij * log(pij + 1e-16);
      if (A.b) {
        const double bj = A.b[(int64_t)p * m + j];
        s_kl += pij * log(pij / (ai * bj) + 1e-16);
      }
      s_mass += pij;
    }
  }
  __shared__ double red[kPlanSums][kPlanThreads / kWave];
  double vals[kPlanSums] = {s_cost, s_ent, s_kl, s_mass};
#pragma unroll
  for (int q = 0; q < kPlanSums; ++q) {
    const double w = wave_sum(vals[q]);
    if ((threadIdx.x & (kWave - 1)) == 0) red[q][threadIdx.x / kWave] = w;
  }
  __syncthreads();
  if (threadIdx.x < kPlanSums) {
    double t = 0;
    for (int w = 0; w < kPlanThreads / kWave; ++w) t += red[threadIdx.x][w];
    A.partial[((int64_t)p * gridDim.x + blockIdx.x) * kPlanSums + threadIdx.x] = t;
  }
}

__global__ void plan_sums_kernel(int nblk, const double* partial, double* sums) {
  const int p = blockIdx.x;
  const int q = threadIdx.x;
  if (q >= kPlanSums) return;
  double t = 0;
  for (int b = 0; b < nblk; ++b) t += partial[((int64_t)p * nblk + b) * kPlanSums + q];
  sums[(int64_t)p * kPlanSums + q] = t;
}

int plan_blocks(int n) { return (n + 7) / 8; }

hipError_t launch_plan(int dtype, PlanArgs a, int B, double* sums, hipStream_t s) {
  const int nblk = plan_blocks(a.n);
  a.rpb = 8;
  dim3 grid(nblk, B);
  if (dtype == 0)
    hipLaunchKernelGGL(plan_kernel<float>, grid, dim3(kPlanThreads), 0, s, a);
  else
    hipLaunchKernelGGL(plan_kernel<double>, grid, dim3(kPlanThreads), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !sums) return e;
  hipLaunchKernelGGL(plan_sums_kernel, dim3(B), dim3(64), 0, s, nblk, a.partial, sums);
  return hipGetLastError();
}

}  // namespace potamd
