// Batched dense operations of the ot.batch API that are not Sinkhorn sweeps:
//
// * pairwise_kernel: the Minkowski and KL cost matrices of
//   ot.batch.dist_lp_batch (ot/batch/_linear.py:20-51) and dist_kl_batch
//   (:88-120), one 64 x 64 output tile per workgroup with the two point
//   tiles staged through LDS (the squared Euclidean cost lives in dist.hip);
// * gemm_batched_kernel: C_b = alpha op(A_b) op(B_b) + beta C_b with
//   arbitrary element strides, fp32 / fp64 accumulation in the input dtype.
//   It carries the Gromov-Wasserstein tensor product (hC1 T hC2^T per
//   feature channel, ot/batch/_quadratic.py:535-572, whose channel-last
//   layout is read in place through the strides) and the backward pass of
//   the squared Euclidean cost.
//
// Both are small next to the Sinkhorn sweeps (they run once per solve or
// once per outer Gromov step); they are written for correctness and full
// occupancy rather than MFMA peak: 256 threads, 4 x 4 outputs per thread.
#include "../../include/pot_amd.h"
#include "host.h"
#include "internal.h"

namespace potamd {

namespace {

constexpr int kTile = 64;
constexpr int kChunk = 16;

template <typename T>
__device__ inline T abs_pow(T x, T p) {
  x = x < (T)0 ? -x : x;
  if (p == (T)2) return x * x;
  if (p == (T)1) return x;
  return x == (T)0 ? (T)0 : (T)pow(x, p);
}

// metric 0: (sum_d |x - y|^p)^(q / p); 1: sum_d y log(y + eps) - y log(x + eps);
// 2: sum_d y log(y + eps) - y x (x holds logits).
template <typename T, int METRIC>
__global__ __launch_bounds__(256) void pairwise_kernel(const T* __restrict__ X,
                                                       const T* __restrict__ Y, int B, int n,
                                                       int m, int d, T p, T q, T eps,
                                                       T* __restrict__ M) {
  __shared__ T xs[kChunk][kTile + 1];
  __shared__ T ys[kChunk][kTile + 1];
  __shared__ T yl[kChunk][kTile + 1];  // y log(y + eps) (KL only)
  const int i0 = blockIdx.y * kTile, j0 = blockIdx.x * kTile;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  for (int b = blockIdx.z; b < B; b += gridDim.z) {
    const T* Xb = X + (int64_t)b * n * d;
    const T* Yb = Y + (int64_t)b * m * d;
    T acc[4][4] = {}, ent[4] = {};
    for (int k0 = 0; k0 < d; k0 += kChunk) {
      for (int e = threadIdx.x; e < kChunk * kTile; e += 256) {
        const int r = e / kChunk, k = e % kChunk;
        const int i = i0 + r, j = j0 + r, kk = k0 + k;
        T xv = (i < n && kk < d) ? Xb[(int64_t)i * d + kk] : (T)0;
        T yv = (j < m && kk < d) ? Yb[(int64_t)j * d + kk] : (T)0;
        if (METRIC == 1) xv = (T)log(xv + eps);
        xs[k][r] = xv;
        ys[k][r] = yv;
        if (METRIC != 0) yl[k][r] = yv * (T)log(yv + eps);
      }
      __syncthreads();
      const int kn = min(kChunk, d - k0);
      for (int k = 0; k < kn; ++k) {
        T xv[4], yv[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          xv[t] = xs[k][ty + 16 * t];
          yv[t] = ys[k][tx + 16 * t];
          if (METRIC != 0) ent[t] += yl[k][tx + 16 * t];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            if (METRIC == 0) {
              acc[r][c] += abs_pow(xv[r] - yv[c], p);
            } else {
              // rounded product, then the add, in the order of the entropy
              // sum: KL(y || x) of a row with itself is then exactly 0
#pragma clang fp contract(off)
              acc[r][c] = acc[r][c] + yv[c] * xv[r];
            }
          }
      }
      __syncthreads();
    }
    T* Mb = M + (int64_t)b * n * m;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + ty + 16 * r;
      if (i >= n) continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = j0 + tx + 16 * c;
        if (j >= m) continue;
        T val;
        if (METRIC == 0)
          val = (q == p) ? acc[r][c] : (T)pow(acc[r][c], q / p);
        else
          val = ent[c] - acc[r][c];
        Mb[(int64_t)i * m + j] = val;
      }
    }
  }
}

struct GemmBatchArgs {
  int B, M, N, K;
  const void* A;
  int64_t sA, rsA, csA;  // element (i, k) of op(A_b) at A + b sA + i rsA + k csA
  const void* Bm;
  int64_t sB, rsB, csB;  // element (k, j) of op(B_b) at B + b sB + k rsB + j csB
  void* C;
  int64_t sC, ldc;       // element (i, j) of C_b at C + b sC + i ldc + j
  double alpha, beta;
};

template <typename T>
__global__ __launch_bounds__(256) void gemm_batched_kernel(GemmBatchArgs g) {
  __shared__ T as[kChunk][kTile + 1];
  __shared__ T bs[kChunk][kTile + 1];
  const int i0 = blockIdx.y * kTile, j0 = blockIdx.x * kTile;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const T* A = static_cast<const T*>(g.A);
  const T* Bm = static_cast<const T*>(g.Bm);
  T* C = static_cast<T*>(g.C);
  // Thread-to-element maps that keep the loads coalesced: along k when k is
  // the contiguous index of the operand, along the row (column) otherwise.
  const bool a_k_fast = g.csA == 1;
  const bool b_j_fast = g.csB == 1;
  for (int b = blockIdx.z; b < g.B; b += gridDim.z) {
    const T* Ab = A + (int64_t)b * g.sA;
    const T* Bb = Bm + (int64_t)b * g.sB;
    T acc[4][4] = {};
    for (int k0 = 0; k0 < g.K; k0 += kChunk) {
      for (int e = threadIdx.x; e < kChunk * kTile; e += 256) {
        int r, k;
        if (a_k_fast) {
          r = e / kChunk;
          k = e % kChunk;
        } else {
          r = e % kTile;
          k = e / kTile;
        }
        const int i = i0 + r, kk = k0 + k;
        as[k][r] = (i < g.M && kk < g.K) ? Ab[(int64_t)i * g.rsA + (int64_t)kk * g.csA] : (T)0;
        int c, kb;
        if (b_j_fast) {
          c = e % kTile;
          kb = e / kTile;
        } else {
          c = e / kChunk;
          kb = e % kChunk;
        }
        const int j = j0 + c, kk2 = k0 + kb;
        bs[kb][c] = (j < g.N && kk2 < g.K) ? Bb[(int64_t)kk2 * g.rsB + (int64_t)j * g.csB] : (T)0;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kChunk; ++k) {
        T av[4], bv[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          av[t] = as[k][ty + 16 * t];
          bv[t] = bs[k][tx + 16 * t];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[r][c] = fma(av[r], bv[c], acc[r][c]);
      }
      __syncthreads();
    }
    T* Cb = C + (int64_t)b * g.sC;
    const T alpha = (T)g.alpha, beta = (T)g.beta;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + ty + 16 * r;
      if (i >= g.M) continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = j0 + tx + 16 * c;
        if (j >= g.N) continue;
        T* dst = Cb + (int64_t)i * g.ldc + j;
        // beta == 0 never reads C, so an uninitialised C stays harmless
        *dst = beta == (T)0 ? alpha * acc[r][c] : fma(alpha, acc[r][c], beta * *dst);
      }
    }
  }
}

dim3 tile_grid(int rows, int cols, int B) {
  return dim3((cols + kTile - 1) / kTile, (rows + kTile - 1) / kTile, B < 65535 ? B : 65535);
}

}  // namespace

}  // namespace potamd

using namespace potamd;

extern "C" int potamd_pairwise_batch(int dtype, int metric, const void* X, const void* Y,
                                     int batch, int n, int m, int d, double p, double q,
                                     double eps, void* M, void* stream) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || metric < 0 || metric > 2 || !X || !Y ||
      !M || batch < 1 || n < 1 || m < 1 || d < 1)
    return POTAMD_EINVAL;
  if (metric == 0 && !(p > 0.0 && q > 0.0)) return POTAMD_EINVAL;
  if (metric != 0 && !(eps >= 0.0)) return POTAMD_EINVAL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid = tile_grid(n, m, batch);
#define POTAMD_PAIRWISE(T, MET)                                                               \
  hipLaunchKernelGGL((pairwise_kernel<T, MET>), grid, dim3(256), 0, s,                        \
                     static_cast<const T*>(X), static_cast<const T*>(Y), batch, n, m, d,      \
                     (T)p, (T)q, (T)eps, static_cast<T*>(M))
  if (dtype == POTAMD_F32) {
    if (metric == 0) POTAMD_PAIRWISE(float, 0);
    else if (metric == 1) POTAMD_PAIRWISE(float, 1);
    else POTAMD_PAIRWISE(float, 2);
  } else {
    if (metric == 0) POTAMD_PAIRWISE(double, 0);
    else if (metric == 1) POTAMD_PAIRWISE(double, 1);
    else POTAMD_PAIRWISE(double, 2);
  }
#undef POTAMD_PAIRWISE
  HIP_TRY(hipGetLastError());
  return 0;
}

extern "C" int potamd_gemm_batched(int dtype, int batch, int M, int N, int K, double alpha,
                                   const void* A, int64_t sA, int64_t rsA, int64_t csA,
                                   const void* B, int64_t sB, int64_t rsB, int64_t csB,
                                   double beta, void* C, int64_t sC, int64_t ldc, void* stream) {
  if ((dtype != POTAMD_F32 && dtype != POTAMD_F64) || !A || !B || !C || batch < 1 || M < 1 ||
      N < 1 || K < 1 || ldc < N)
    return POTAMD_EINVAL;
  GemmBatchArgs g{batch, M, N, K, A, sA, rsA, csA, B, sB, rsB, csB, C, sC, ldc, alpha, beta};
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid = tile_grid(M, N, batch);
  if (dtype == POTAMD_F32)
    hipLaunchKernelGGL(gemm_batched_kernel<float>, grid, dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL(gemm_batched_kernel<double>, grid, dim3(256), 0, s, g);
  HIP_TRY(hipGetLastError());
  return 0;
}
