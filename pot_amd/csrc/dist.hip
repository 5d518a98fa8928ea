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

// Ragged variant: problem p has n_p x m_p costs from points at xoff / yoff,
// written at moff[p] with row stride ldms[p]; the per-problem maximum goes to
// maxout[p] (atomicMax on the bit pattern, valid for the non-negative clamped
// costs).  Tiles outside a problem exit at once.
__global__ __launch_bounds__(256) void sqeuclid_ragged_kernel(RaggedDistArgs A) {
  const int p = blockIdx.z;
  const int n = A.ns[p], m = A.ms[p], d = A.d;
  const int i0 = blockIdx.y * kDistTile, j0 = blockIdx.x * kDistTile;
  if (i0 >= n || j0 >= m) return;
  const float* X = A.X + A.xoff[p] * d;
  const float* Y = A.Y + A.yoff[p] * d;
  float* M = A.M + A.moff[p];
  const int ldm = A.ldms[p];
  __shared__ float xs[kDistK][kDistTile + 1];
  __shared__ float ys[kDistK][kDistTile + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  float acc[4][4] = {}, xn[4] = {}, yn[4] = {};
  for (int k0 = 0; k0 < d; k0 += kDistK) {
    for (int e = threadIdx.x; e < kDistK * kDistTile; e += 256) {
      const int r = e / kDistK, k = e % kDistK;
      const int i = i0 + r, j = j0 + r, kk = k0 + k;
      xs[k][r] = (i < n && kk < d) ? X[(int64_t)i * d + kk] : 0.f;
      ys[k][r] = (j < m && kk < d) ? Y[(int64_t)j * d + kk] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kDistK; ++k) {
      float xv[4], yv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        xv[q] = xs[k][ty + 16 * q];
        yv[q] = ys[k][tx + 16 * q];
        xn[q] = fmaf(xv[q], xv[q], xn[q]);
        yn[q] = fmaf(yv[q], yv[q], yn[q]);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fmaf(xv[a], yv[b], acc[a][b]);
    }
    __syncthreads();
  }
  float mx = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ty + 16 * a;
    if (i >= n) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int j = j0 + tx + 16 * b;
      if (j >= m) continue;
      float c = -2.f * acc[a][b] + xn[a] + yn[b];
      c = c > 0.f ? c : 0.f;
      M[(int64_t)i * ldm + j] = c;
      mx = fmaxf(mx, c);
    }
  }
  // Zero the row padding [m, ldm) so packed rows are fully defined.
  if (j0 + kDistTile >= m && ldm > m) {
    for (int e = threadIdx.x; e < kDistTile * (ldm - m); e += 256) {
      const int r = e / (ldm - m), c = m + e % (ldm - m);
      if (i0 + r < n) M[(int64_t)(i0 + r) * ldm + c] = 0.f;
    }
  }
  mx = wave_max(mx);
  if ((threadIdx.x & (kWave - 1)) == 0 && A.maxout)
    atomicMax(reinterpret_cast<int*>(A.maxout + p), __float_as_int(mx));
}

__global__ void zero_f32_kernel(int count, float* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) x[i] = 0.f;
}

hipError_t launch_sqeuclid_ragged(const RaggedDistArgs& a, int B, int max_n, int max_m,
                                  hipStream_t s) {
  if (a.maxout) {
    hipLaunchKernelGGL(zero_f32_kernel, dim3((B + 255) / 256), dim3(256), 0, s, B, a.maxout);
  }
  dim3 grid((max_m + kDistTile - 1) / kDistTile, (max_n + kDistTile - 1) / kDistTile, B);
  hipLaunchKernelGGL(sqeuclid_ragged_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
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
