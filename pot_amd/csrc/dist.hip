// Batched squared-Euclidean / Euclidean cost matrices.
//
// M_bij = |x_bi|^2 + |y_bj|^2 - 2 x_bi . y_bj, the formula of ot.dist
// (ot/utils.py:292-299, clamped at 0, zero diagonal when X is Y) and of
// ot.batch.dist_euclidean_batch (ot/batch/_linear.py:79-84, not clamped).
// 64x64 output tiles, x and y tiles staged through LDS; accumulation in the
// input dtype.
#include "internal.h"

namespace potamd {

constexpr int kDistTile = 64;
constexpr int kDistK = 16;

template <typename T>
__global__ __launch_bounds__(256) void sqeuclid_kernel(DistArgs A) {
  const int p = blockIdx.z;
  const T* X = static_cast<const T*>(A.X) + (int64_t)p * A.xstride;
  const T* Y = static_cast<const T*>(A.Y) + (int64_t)p * A.ystride;
  T* M = static_cast<T*>(A.M) + (int64_t)p * A.mstride;
  const int i0 = blockIdx.y * kDistTile, j0 = blockIdx.x * kDistTile;
  __shared__ T xs[kDistK][kDistTile + 1];
  __shared__ T ys[kDistK][kDistTile + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;  // 4x4 outputs per thread
  T acc[4][4] = {}, xn[4] = {}, yn[4] = {};
  for (int k0 = 0; k0 < A.d; k0 += kDistK) {
    for (int e = threadIdx.x; e < kDistK * kDistTile; e += 256) {
      const int r = e / kDistK, k = e % kDistK;
      const int i = i0 + r, j = j0 + r, kk = k0 + k;
      xs[k][r] = (i < A.n && kk < A.d) ? X[(int64_t)i * A.d + kk] : (T)0;
      ys[k][r] = (j < A.m && kk < A.d) ? Y[(int64_t)j * A.d + kk] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kDistK; ++k) {
      T xv[4], yv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        xv[q] = xs[k][ty + 16 * q];
        yv[q] = ys[k][tx + 16 * q];
        xn[q] = fma(xv[q], xv[q], xn[q]);
        yn[q] = fma(yv[q], yv[q], yn[q]);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fma(xv[a], yv[b], acc[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ty + 16 * a;
    if (i >= A.n) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int j = j0 + tx + 16 * b;
      if (j >= A.m) continue;
      T c = (T)-2 * acc[a][b] + xn[a] + yn[b];
      if (A.clamp) c = c > (T)0 ? c : (T)0;
      if (A.squared == 0) c = sqrt(c);
      if (A.zero_diag && i == j) c = 0;
      M[(int64_t)i * A.ldm + j] = c;
    }
  }
}

hipError_t launch_sqeuclid(int dtype, const DistArgs& a, int B, hipStream_t s) {
  dim3 grid((a.m + kDistTile - 1) / kDistTile, (a.n + kDistTile - 1) / kDistTile, B);
  if (dtype == 0)
    hipLaunchKernelGGL(sqeuclid_kernel<float>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(sqeuclid_kernel<double>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace potamd
