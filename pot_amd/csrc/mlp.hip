// Kernels of the warm-start network: a bf16 MFMA GEMM with fused epilogues
// and the small reduction / elementwise kernels of its forward, backward and
// Adam step.  No counterpart in the reference.
//
// GEMM: C[M x N] (+)= op(A)[M x K] * op(B)[K x N], bf16 operands, fp32
// accumulation on v_mfma_f32_32x32x8bf16_1k.  A 256-thread workgroup owns a
// 64 x 64 output tile; wave w computes the 32 x 32 quarter (w / 2, w % 2).
// K advances in slices of 32 staged through LDS with both operands stored
// k-contiguous (As[m][k], Bt[n][k]), so every MFMA operand is one 8-byte LDS
// read; transposed operands are transposed while being staged.
#include "internal.h"

namespace potamd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}
__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16) | ((u & 0xffff) ? 0x40 : 0);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int kTM = 64, kTN = 64, kTK = 32, kPad = 8;
// Activation of every hidden layer: leaky ReLU (no unit can die).  A stored
// activation has the sign of its pre-activation, so backward masks read it.
constexpr float kLeak = 0.01f;

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ uint16_t As[kTM][kTK + kPad];
  __shared__ uint16_t Bt[kTN][kTK + kPad];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
  const int kchunk = (g.K + gridDim.z - 1) / gridDim.z;
  const int kbeg = blockIdx.z * kchunk, kend = min(g.K, kbeg + kchunk);
  const uint16_t* A = static_cast<const uint16_t*>(g.A);
  const uint16_t* B = static_cast<const uint16_t*>(g.B);
  f32x16 acc = {};
  const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
  for (int k0 = kbeg; k0 < kend; k0 += kTK) {
    // Stage op(A)[m0:m0+64, k0:k0+32] and op(B)[k0:k0+32, n0:n0+64].
    for (int e = tid; e < kTM * kTK; e += 256) {
      int i, k;
      if (!g.transA) { i = e / kTK; k = e % kTK; } else { k = e / kTM; i = e % kTM; }
      const int gi = m0 + i, gk = k0 + k;
      uint16_t v = 0;
      if (gi < g.M && gk < kend)
        v = g.transA ? A[(int64_t)gk * g.lda + gi] : A[(int64_t)gi * g.lda + gk];
      As[i][k] = v;
    }
    for (int e = tid; e < kTN * kTK; e += 256) {
      int j, k;
      if (g.transB) { j = e / kTK; k = e % kTK; } else { k = e / kTN; j = e % kTN; }
      const int gj = n0 + j, gk = k0 + k;
      uint16_t v = 0;
      if (gj < g.N && gk < kend)
        v = g.transB ? B[(int64_t)gj * g.ldb + gk] : B[(int64_t)gk * g.ldb + gj];
      Bt[j][k] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 8) {
      const int kq = kk + 4 * (lane >> 5);
      bf16x4 a = *reinterpret_cast<const bf16x4*>(&As[wm + (lane & 31)][kq]);
      bf16x4 b = *reinterpret_cast<const bf16x4*>(&Bt[wn + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // Epilogue: lane holds column wn + lane%32, rows wm + 8*q + 4*(lane/32) + r.
  const int j = n0 + wn + (lane & 31);
  if (j >= g.N) return;
  const float bj = g.bias ? g.bias[j] : 0.f;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = m0 + wm + 8 * q + 4 * (lane >> 5) + r;
      if (i >= g.M) continue;
      float x = acc[4 * q + r] * g.alpha + bj;
      if (g.rowvec) x += g.rowvec[(int64_t)(i / g.seg_len) * g.N + j];
      if (g.relu) x = x > 0.f ? x : kLeak * x;
      if (g.mask) x = bf2f(static_cast<const uint16_t*>(g.mask)[(int64_t)i * g.ldmask + j]) > 0.f ? x : kLeak * x;
      const int64_t o = (int64_t)i * g.ldc + j;
      if (g.out_kind == 0) {
        static_cast<uint16_t*>(g.C)[o] = f2bf(x);
      } else if (g.out_kind == 1) {
        static_cast<float*>(g.C)[o] = x;
      } else {
        atomicAdd(static_cast<float*>(g.C) + o, x);
      }
    }
  }
}

hipError_t launch_gemm_bf16(const GemmArgs& g, int splits, hipStream_t s) {
  dim3 grid((g.N + kTN - 1) / kTN, (g.M + kTM - 1) / kTM, splits);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, g);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Network input rows.  Encoder row j of problem p: [y_j * scale_p, m_p b_j,
// log reg_p, 0...] padded to `width`; source row i: [x_i * scale_p, 0...].
__global__ void build_inputs_kernel(int rows, int seg_len, int dx, int width, const float* pts,
                                    const double* b, const double* regs, const float* cmax,
                                    int with_extra, uint16_t* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * width) return;
  const int r = (int)(e / width), c = (int)(e % width);
  const int p = r / seg_len;
  const float scale = rsqrtf(cmax ? cmax[p] : 1.f);
  float v = 0.f;
  if (c < dx) {
    v = pts[(int64_t)r * dx + c] * scale;
  } else if (with_extra && c == dx) {
    v = (float)(b[r] * seg_len - 1.0);
  } else if (with_extra && c == dx + 1) {
    v = (float)((log(regs[p]) - kLogRegCenter) / kLogRegScale);
  }
  out[e] = f2bf(v);
}

// z[p][h] = mean over the rows of segment p of H[r][h] (bf16 in, fp32 out).
__global__ void segment_mean_kernel(int seg_len, int h, const uint16_t* H, float* z) {
  const int p = blockIdx.x;
  for (int c = threadIdx.x; c < h; c += blockDim.x) {
    float s = 0.f;
    for (int r = 0; r < seg_len; ++r) s += bf2f(H[((int64_t)p * seg_len + r) * h + c]);
    z[(int64_t)p * h + c] = s / seg_len;
  }
}

// out[p][n] = sum_k in[p][k] W[k][n] (+ bias[n], or bias[p][n] when bias_per_row) --
// small fp32 products with
// the per-problem vectors (W fp32, row-major, k x n).
// kVecP problems per workgroup: their input rows pass through LDS kVecK
// entries at a time, so each W element read serves kVecP problems.  Each
// output is accumulated in q order, as one problem at a time would.
constexpr int kVecP = 8;
constexpr int kVecK = 256;
__global__ __launch_bounds__(256) void vec_mat_kernel(int P, int k, int n, const float* in,
                                                      const float* W, const float* bias,
                                                      int bias_per_row, float* out) {
  __shared__ __attribute__((aligned(16))) float xs[kVecP][kVecK];
  const int p0 = blockIdx.x * kVecP;
  for (int c0 = 0; c0 < n; c0 += blockDim.x) {
    const int c = c0 + threadIdx.x;
    float s[kVecP];
#pragma unroll
    for (int i = 0; i < kVecP; ++i) {
      const int p = p0 + i < P ? p0 + i : P - 1;
      s[i] = (bias && c < n) ? bias[(bias_per_row ? (int64_t)p * n : 0) + c] : 0.f;
    }
    for (int k0 = 0; k0 < k; k0 += kVecK) {
      const int kc = k - k0 < kVecK ? k - k0 : kVecK;
      __syncthreads();  // the previous chunk's readers are done
      for (int e = threadIdx.x; e < kVecP * kVecK; e += blockDim.x) {
        const int i = e / kVecK, q = e % kVecK;
        const int p = p0 + i < P ? p0 + i : P - 1;
        xs[i][q] = q < kc ? in[(int64_t)p * k + k0 + q] : 0.f;
      }
      __syncthreads();
      if (c < n) {
        int q = 0;
        // four q at a time: their W loads issue together, the inputs are
        // read from LDS as float4 (q and kVecK are multiples of 4)
        for (; q + 4 <= kc; q += 4) {
          float w[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) w[t] = W[(int64_t)(k0 + q + t) * n + c];
#pragma unroll
          for (int i = 0; i < kVecP; ++i) {
            const float4 x = *reinterpret_cast<const float4*>(&xs[i][q]);
            s[i] = fmaf(x.x, w[0], s[i]);
            s[i] = fmaf(x.y, w[1], s[i]);
            s[i] = fmaf(x.z, w[2], s[i]);
            s[i] = fmaf(x.w, w[3], s[i]);
          }
        }
        for (; q < kc; ++q) {
          const float w = W[(int64_t)(k0 + q) * n + c];
#pragma unroll
          for (int i = 0; i < kVecP; ++i) s[i] = fmaf(xs[i][q], w, s[i]);
        }
      }
    }
    if (c < n) {
#pragma unroll
      for (int i = 0; i < kVecP; ++i)
        if (p0 + i < P) out[(int64_t)(p0 + i) * n + c] = s[i];
    }
  }
}

// out[p][q] = sum_c in[p][c] W[q][c]   (vector times W^T)
__global__ void vec_matT_kernel(int P, int k, int n, const float* in, const float* W, float* out) {
  const int p = blockIdx.x;
  for (int q = threadIdx.x; q < k; q += blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < n; ++c) s = fmaf(in[(int64_t)p * n + c], W[(int64_t)q * n + c], s);
    out[(int64_t)p * k + q] = s;
  }
}

// f[r] = sum_h G[r][h] w[h] + b0 (+ qn[r]) (one wave per row).
__global__ void head_kernel(int rows, int h, const uint16_t* G, const float* w, const float* b0,
                            const float* qn, float* f) {
  const int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  float s = 0.f;
  for (int c = lane; c < h; c += 64) s = fmaf(bf2f(G[(int64_t)r * h + c]), w[c], s);
  s = wave_sum(s);
  if (lane == 0) f[r] = s + b0[0] + (qn ? qn[r] : 0.f);
}

// qn[r] = |pts[r] * scale_p|^2 over the first dx coordinates (fp32).
__global__ void row_norm_kernel(int rows, int seg_len, int dx, const float* pts,
                                const float* cmax, float* qn) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int p = r / seg_len;
  const float sc2 = 1.f / (cmax ? cmax[p] : 1.f);
  float s = 0.f;
  const float* row = pts + (int64_t)r * dx;
  if ((dx & 3) == 0 && (reinterpret_cast<uintptr_t>(pts) & 15) == 0) {
    // 16-byte loads, the same order of accumulation
    for (int c = 0; c < dx; c += 4) {
      const float4 v = *reinterpret_cast<const float4*>(row + c);
      s = fmaf(v.x, v.x, s);
      s = fmaf(v.y, v.y, s);
      s = fmaf(v.z, v.z, s);
      s = fmaf(v.w, v.w, s);
    }
  } else {
    for (int c = 0; c < dx; ++c) s = fmaf(row[c], row[c], s);
  }
  qn[r] = s * sc2;
}

// Centred weighted MSE per problem and its gradient with respect to f.
// loss_out[p] = mean_i (fc - tc)^2 (r0/reg_p)^2 / P ; df[r] = d loss / d f[r].
__global__ void loss_kernel(int seg_len, int P, const float* f, const float* t,
                            const double* regs, double r0, float* df, float* loss_out) {
  const int p = blockIdx.x;
  __shared__ float red[2][16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, W = blockDim.x / 64;
  const float* fp = f + (int64_t)p * seg_len;
  const float* tp = t + (int64_t)p * seg_len;
  float sf = 0.f, st = 0.f;
  for (int i = tid; i < seg_len; i += blockDim.x) { sf += fp[i]; st += tp[i]; }
  sf = wave_sum(sf);
  st = wave_sum(st);
  if (lane == 0) { red[0][wid] = sf; red[1][wid] = st; }
  __syncthreads();
  float mf = 0.f, mt = 0.f;
  for (int w = 0; w < W; ++w) { mf += red[0][w]; mt += red[1][w]; }
  mf /= seg_len;
  mt /= seg_len;
  __syncthreads();
  const float wgt = (float)((r0 / regs[p]) * (r0 / regs[p]));
  float l = 0.f, sd = 0.f;
  for (int i = tid; i < seg_len; i += blockDim.x) {
    const float d = (fp[i] - mf) - (tp[i] - mt);
    l += d * d;
    const float g = 2.f * d / seg_len * wgt / P;
    df[(int64_t)p * seg_len + i] = g;
    sd += g;
  }
  l = wave_sum(l);
  sd = wave_sum(sd);
  if (lane == 0) { red[0][wid] = l; red[1][wid] = sd; }
  __syncthreads();
  float L = 0.f, SD = 0.f;
  for (int w = 0; w < W; ++w) { L += red[0][w]; SD += red[1][w]; }
  SD /= seg_len;
  for (int i = tid; i < seg_len; i += blockDim.x) df[(int64_t)p * seg_len + i] -= SD;
  if (tid == 0 && loss_out) loss_out[p] = L / seg_len * wgt / P;
}

// dB[r][h] = df[r] * w[h] * leak'(G[r][h])  (bf16 out)
__global__ void head_back_kernel(int rows, int h, const float* df, const float* w,
                                 const uint16_t* G, uint16_t* dB) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * h) return;
  const int r = (int)(e / h), c = (int)(e % h);
  dB[e] = f2bf(df[r] * w[c] * (bf2f(G[e]) > 0.f ? 1.f : kLeak));
}

// gw[h] += sum_r G[r][h] df[r] ; gb += sum_r df[r]   (atomic per block)
__global__ void head_wgrad_kernel(int rows, int h, const uint16_t* G, const float* df, float* gw,
                                  float* gb) {
  const int c = threadIdx.x;  // blockDim.x == h (<= 1024)
  const int r0 = blockIdx.x * 64, r1 = min(rows, r0 + 64);
  float s = 0.f, sb = 0.f;
  for (int r = r0; r < r1; ++r) {
    s = fmaf(bf2f(G[(int64_t)r * h + c]), df[r], s);
    sb += df[r];
  }
  atomicAdd(gw + c, s);
  if (c == 0) atomicAdd(gb, sb);
}

// Column sums (bias gradients): out[c] += sum_r X[r][c], bf16 X.
__global__ void colsum_kernel(int rows, int h, const uint16_t* X, float* out) {
  const int c = threadIdx.x;
  const int r0 = blockIdx.x * 64, r1 = min(rows, r0 + 64);
  float s = 0.f;
  for (int r = r0; r < r1; ++r) s += bf2f(X[(int64_t)r * h + c]);
  atomicAdd(out + c, s);
}

// Segment sums: out[p][c] = sum over segment p of X[r][c].
__global__ void segment_sum_kernel(int seg_len, int h, const uint16_t* X, float* out) {
  const int p = blockIdx.x;
  for (int c = threadIdx.x; c < h; c += blockDim.x) {
    float s = 0.f;
    for (int r = 0; r < seg_len; ++r) s += bf2f(X[((int64_t)p * seg_len + r) * h + c]);
    out[(int64_t)p * h + c] = s;
  }
}

// dA[r][c] = dz[seg(r)][c] / seg_len * leak'(H[r][c])   (bf16 out)
__global__ void pool_back_kernel(int rows, int seg_len, int h, const float* dz, const uint16_t* H,
                                 uint16_t* dA) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * h) return;
  const int r = (int)(e / h), c = (int)(e % h);
  const int p = r / seg_len;
  dA[e] = f2bf(dz[(int64_t)p * h + c] / seg_len * (bf2f(H[e]) > 0.f ? 1.f : kLeak));
}

// gW[q][c] += sum_p z[p][q] dc[p][c]  (outer products of per-problem vectors)
__global__ void outer_acc_kernel(int P, int k, int n, const float* z, const float* dc, float* gW) {
  const int q = blockIdx.x;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    float s = 0.f;
    for (int p = 0; p < P; ++p) s = fmaf(z[(int64_t)p * k + q], dc[(int64_t)p * n + c], s);
    gW[(int64_t)q * n + c] += s;
  }
}

// Adam on fp32 master weights; writes the bf16 copy used by the GEMMs.
__global__ void adam_kernel(int64_t count, float* w, const float* g, float* m1, float* m2,
                            float lr, float b1, float b2, float eps, float bc1, float bc2,
                            float wd, uint16_t* w_bf16) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const float gi = g[e] + wd * w[e];
  const float a = b1 * m1[e] + (1.f - b1) * gi;
  const float b = b2 * m2[e] + (1.f - b2) * gi * gi;
  m1[e] = a;
  m2[e] = b;
  w[e] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
  if (w_bf16) w_bf16[e] = f2bf(w[e]);
}

__global__ void to_bf16_kernel(int64_t count, const float* x, uint16_t* y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) y[e] = f2bf(x[e]);
}

static inline unsigned nb(int64_t c, int t) { return (unsigned)((c + t - 1) / t); }

hipError_t launch_mlp_op(const MlpOp& o, hipStream_t s) {
  switch (o.op) {
    case kOpBuildInputs:
      hipLaunchKernelGGL(build_inputs_kernel, dim3(nb((int64_t)o.rows * o.k, 256)), dim3(256), 0,
                         s, o.rows, o.seg_len, o.n, o.k, static_cast<const float*>(o.x),
                         static_cast<const double*>(o.y), static_cast<const double*>(o.z),
                         static_cast<const float*>(o.w), o.flag,
                         static_cast<uint16_t*>(o.out));
      break;
    case kOpSegmentMean:
      hipLaunchKernelGGL(segment_mean_kernel, dim3(o.P), dim3(256), 0, s, o.seg_len, o.k,
                         static_cast<const uint16_t*>(o.x), static_cast<float*>(o.out));
      break;
    case kOpVecMat:
      hipLaunchKernelGGL(vec_mat_kernel, dim3((o.P + kVecP - 1) / kVecP), dim3(256), 0, s,
                         o.P, o.k, o.n,
                         static_cast<const float*>(o.x), static_cast<const float*>(o.w),
                         static_cast<const float*>(o.y), o.flag, static_cast<float*>(o.out));
      break;
    case kOpVecMatT:
      hipLaunchKernelGGL(vec_matT_kernel, dim3(o.P), dim3(256), 0, s, o.P, o.k, o.n,
                         static_cast<const float*>(o.x), static_cast<const float*>(o.w),
                         static_cast<float*>(o.out));
      break;
    case kOpHead:
      hipLaunchKernelGGL(head_kernel, dim3(nb(o.rows, 4)), dim3(256), 0, s, o.rows, o.k,
                         static_cast<const uint16_t*>(o.x), static_cast<const float*>(o.w),
                         static_cast<const float*>(o.y), static_cast<const float*>(o.z),
                         static_cast<float*>(o.out));
      break;
    case kOpLoss:
      hipLaunchKernelGGL(loss_kernel, dim3(o.P), dim3(256), 0, s, o.seg_len, o.P,
                         static_cast<const float*>(o.x), static_cast<const float*>(o.y),
                         static_cast<const double*>(o.z), o.scalar,
                         static_cast<float*>(o.out), static_cast<float*>(o.out2));
      break;
    case kOpHeadBack:
      hipLaunchKernelGGL(head_back_kernel, dim3(nb((int64_t)o.rows * o.k, 256)), dim3(256), 0, s,
                         o.rows, o.k, static_cast<const float*>(o.x),
                         static_cast<const float*>(o.w), static_cast<const uint16_t*>(o.y),
                         static_cast<uint16_t*>(o.out));
      break;
    case kOpHeadWGrad:
      hipLaunchKernelGGL(head_wgrad_kernel, dim3(nb(o.rows, 64)), dim3(o.k), 0, s, o.rows, o.k,
                         static_cast<const uint16_t*>(o.x), static_cast<const float*>(o.y),
                         static_cast<float*>(o.out), static_cast<float*>(o.out2));
      break;
    case kOpColSum:
      hipLaunchKernelGGL(colsum_kernel, dim3(nb(o.rows, 64)), dim3(o.k), 0, s, o.rows, o.k,
                         static_cast<const uint16_t*>(o.x), static_cast<float*>(o.out));
      break;
    case kOpSegmentSum:
      hipLaunchKernelGGL(segment_sum_kernel, dim3(o.P), dim3(256), 0, s, o.seg_len, o.k,
                         static_cast<const uint16_t*>(o.x), static_cast<float*>(o.out));
      break;
    case kOpPoolBack:
      hipLaunchKernelGGL(pool_back_kernel, dim3(nb((int64_t)o.rows * o.k, 256)), dim3(256), 0, s,
                         o.rows, o.seg_len, o.k, static_cast<const float*>(o.x),
                         static_cast<const uint16_t*>(o.y), static_cast<uint16_t*>(o.out));
      break;
    case kOpOuterAcc:
      hipLaunchKernelGGL(outer_acc_kernel, dim3(o.k), dim3(256), 0, s, o.P, o.k, o.n,
                         static_cast<const float*>(o.x), static_cast<const float*>(o.y),
                         static_cast<float*>(o.out));
      break;
    case kOpAdam:
      hipLaunchKernelGGL(adam_kernel, dim3(nb(o.count, 256)), dim3(256), 0, s, o.count,
                         static_cast<float*>(o.out), static_cast<const float*>(o.x),
                         static_cast<float*>(const_cast<void*>(o.y)),
                         static_cast<float*>(const_cast<void*>(o.z)), o.hp[0], o.hp[1],
                         o.hp[2], o.hp[3], o.hp[4], o.hp[5], o.hp[6],
                         static_cast<uint16_t*>(o.out2));
      break;
    case kOpRowNorm:
      hipLaunchKernelGGL(row_norm_kernel, dim3(nb(o.rows, 256)), dim3(256), 0, s, o.rows,
                         o.seg_len, o.n, static_cast<const float*>(o.x),
                         static_cast<const float*>(o.w), static_cast<float*>(o.out));
      break;
    case kOpToBf16:
      hipLaunchKernelGGL(to_bf16_kernel, dim3(nb(o.count, 256)), dim3(256), 0, s, o.count,
                         static_cast<const float*>(o.x), static_cast<uint16_t*>(o.out));
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace potamd
