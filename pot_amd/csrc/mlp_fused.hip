// Fused two-layer forward of the warm-start network for inference: the
// 256-wide hidden activations never leave the chip.  No counterpart in the
// reference.
//
// Both modes run, on blocks of 64 input rows (32 bf16 features each),
//     A1 = act(inp W1 + bias1)           (64 x 256, bf16 in LDS)
//     A2 = act(A1 W2 + b2)               (64 x 256, rounded to bf16)
// with act the leaky ReLU (slope 0.01), and then
//  * mode 0 (encoder): the mean of A2 over each segment of seg_len rows
//    (bias1 = b1[256]); out = (rows / seg_len) x 256 means;
//  * mode 1 (head): out[r] = A2[r] . w3 + b3 + qn[r], with bias1 the
//    per-segment vector rowvec[r / seg_len][256].
// These are the layers that pot_amd/warmstart/net.py otherwise runs as bf16
// GEMMs with the same epilogues (bias / broadcast, leaky ReLU, bf16 store),
// a segment mean and a head kernel.
//
// The input rows are either given (bf16, rows x 32) or built here from the
// fp32 points (rows x 16), the per-row weights and the per-segment reg and
// cost maximum, with the arithmetic of the network's input-row kernel:
// [x * rsqrt(cmax), seg_len * w - 1, (log reg - c) / s, 0 ...].
//
// Layout: a 256-thread workgroup (one wave per SIMD) walks whole segments;
// wave w owns hidden columns 64 w .. 64 w + 63 of both layers and keeps its
// slices of W1 (32 x 64) and W2 (256 x 64) in registers for the whole
// launch.  Layer 1 is computed transposed (D[j][r]) so each lane holds four
// consecutive columns of one row and writes them to LDS with one 8-byte
// store; layer 2 reads A1 from LDS (16 bytes per lane per MFMA K step).  In
// mode 0 layer 2 gives D[r][j], so a lane owns a column and the segment sum
// is a per-lane accumulation; in mode 1 it gives D[j][r] (operands swapped),
// so a lane owns a row and the head's dot product is per lane.  A1 is double
// buffered in LDS: one barrier per block (two in mode 1).  The next block's
// inputs are fetched while the current one is computed.
#include <string.h>

#include "../../include/pot_amd.h"
#include "host.h"
#include "internal.h"

namespace potamd {
namespace {

constexpr int kFH = 256;             // hidden width
constexpr int kFK = 32;              // input width (one K slice pair)
constexpr int kFD = 16;              // point coordinates
constexpr int kFRows = 64;           // rows per block
constexpr int kFStride = kFH + 8;    // bf16 per LDS row (16-byte pad against bank conflicts)
constexpr int kFW = 4;               // waves per workgroup
constexpr float kFLeak = 0.01f;

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

struct FusedMlpArgs {
  int mode;
  int rows, seg_len;
  const uint16_t* inp;   // rows x 32 bf16, or null: built from the points
  const float* pts;      // rows x 16 fp32
  const double* wts;     // rows (a or b)
  const double* regs;    // rows / seg_len
  const float* cmax;     // rows / seg_len
  const uint16_t* W1;    // 32 x 256 bf16, row-major [k][j]
  const float* bias1;    // mode 0: b1[256]; mode 1: rowvec[rows / seg_len][256]
  const uint16_t* W2;    // 256 x 256 bf16, row-major [k][j]
  const float* b2;       // [256]
  const float* w3;       // mode 1: [256]
  const float* b3;       // mode 1: [1]
  const float* qn;       // mode 1: [rows] or null
  float* out;
};

__device__ __forceinline__ float bf_to_f(unsigned h) { return __uint_as_float(h << 16); }
// Two floats to packed bf16 (round to nearest even, v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned pack_bf2(float lo, float hi) {
  const bf16x2 v = {(__bf16)lo, (__bf16)hi};
  return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ float round_bf(float x) { return (float)(__bf16)x; }
// Leaky ReLU with a slope below 1: max(x, slope x) (two VALU operations).
__device__ __forceinline__ float leaky(float x) { return fmaxf(x, kFLeak * x); }

__device__ __forceinline__ f32x16 mfma16(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Eight bf16 of column j, rows k0 .. k0+7, of a row-major [k][256] matrix.
__device__ __forceinline__ u32x4 load_col8(const uint16_t* W, int k0, int j) {
  u32x4 v;
#pragma unroll
  for (int e = 0; e < 4; ++e)
    v[e] = (unsigned)W[(k0 + 2 * e) * kFH + j] | ((unsigned)W[(k0 + 2 * e + 1) * kFH + j] << 16);
  return v;
}

// The layer-1 B operands of one block: row r = r0 + 32 t + l32, input
// features 16 s + 8 h .. +7.  Raw holds what the loads return (fetched one
// block ahead); make() turns it into the bf16 operands.
template <int IN>
struct BlockIn;

template <>
struct BlockIn<0> {  // given bf16 rows
  u32x4 raw[2][2];
  __device__ void fetch(const FusedMlpArgs& A, int64_t r0, int h, int l32) {
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s = 0; s < 2; ++s)
        raw[t][s] = *reinterpret_cast<const u32x4*>(A.inp + (r0 + 32 * t + l32) * kFK + 16 * s +
                                                    8 * h);
  }
  __device__ void make(const FusedMlpArgs&, int, float, float, u32x4 (&x)[2][2]) const {
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int s = 0; s < 2; ++s) x[t][s] = raw[t][s];
  }
};

template <>
struct BlockIn<1> {  // built from the points
  f32x4 c[2][2];     // [t][half of the lane's 8 coordinates]
  double w[2];
  __device__ void fetch(const FusedMlpArgs& A, int64_t r0, int h, int l32) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int64_t r = r0 + 32 * t + l32;
      const f32x4* p = reinterpret_cast<const f32x4*>(A.pts + r * kFD + 8 * h);
      c[t][0] = p[0];
      c[t][1] = p[1];
      w[t] = A.wts[r];
    }
  }
  // scale, lr: the segment's point scale and normalised log reg.
  __device__ void make(const FusedMlpArgs& A, int h, float scale, float lr,
                       u32x4 (&x)[2][2]) const {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        x[t][0][2 * q] = pack_bf2(c[t][q][0] * scale, c[t][q][1] * scale);
        x[t][0][2 * q + 1] = pack_bf2(c[t][q][2] * scale, c[t][q][3] * scale);
      }
      const float wf = (float)(w[t] * A.seg_len - 1.0);
      x[t][1][0] = h == 0 ? pack_bf2(wf, lr) : 0u;
      x[t][1][1] = 0u;
      x[t][1][2] = 0u;
      x[t][1][3] = 0u;
    }
  }
};

template <int MODE, int IN>
__global__ __launch_bounds__(kFW * kWave) __attribute__((amdgpu_waves_per_eu(1, 1)))
void mlp_fused_kernel(FusedMlpArgs A) {
  __shared__ __attribute__((aligned(16))) uint16_t act[2][kFRows][kFStride];
  __shared__ float part[2][kFW][kFRows];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int h = lane >> 5, l32 = lane & 31;
  const int jbase = 64 * w;

  // Weight slices: column j = jbase + 32 c + l32, rows 16 s + 8 h .. +7.
  u32x4 w1r[2][2], w2r[2][16];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
#pragma unroll
    for (int s = 0; s < 2; ++s) w1r[c][s] = load_col8(A.W1, 16 * s + 8 * h, jbase + 32 * c + l32);
#pragma unroll
    for (int s = 0; s < 16; ++s) w2r[c][s] = load_col8(A.W2, 16 * s + 8 * h, jbase + 32 * c + l32);
  }
  // In mode 0 a lane owns column jbase + 32 c + l32 of layer 2.
  const float b2c[2] = {A.b2[jbase + l32], A.b2[jbase + 32 + l32]};
  // Mode 1: the layer-2 bias and the head weights of the lane's values
  // (column jbase + 32 c + 8 q + 4 h + i), kept in registers: a global load
  // inside the block loop would be waited for there, with one wave per SIMD.
  float b2v[2][4][4], w3v[2][4][4];
  if constexpr (MODE == 1) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int j = jbase + 32 * c + 8 * q + 4 * h + i;
          b2v[c][q][i] = A.b2[j];
          w3v[c][q][i] = A.w3[j];
        }
  }

  const int nseg = A.rows / A.seg_len;
  const int bps = A.seg_len / kFRows;
  BlockIn<IN> nxt;
  if (blockIdx.x < nseg) nxt.fetch(A, (int64_t)blockIdx.x * A.seg_len, h, l32);
  int buf = 0;
  for (int p = blockIdx.x; p < nseg; p += gridDim.x) {
    float zacc[2] = {0.f, 0.f};
    // Per-segment values, once per segment: the layer-1 bias of the lane's
    // values and (inputs built from the points) the scale and log reg.
    const float* bias1 = MODE == 0 ? A.bias1 : A.bias1 + (int64_t)p * kFH;
    float b1v[2][4][4];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) b1v[c][q][i] = bias1[jbase + 32 * c + 8 * q + 4 * h + i];
    float seg_scale = 1.f, seg_lr = 0.f;
    if constexpr (IN == 1) {
      seg_scale = rsqrtf(A.cmax[p]);
      seg_lr = (float)((log(A.regs[p]) - kLogRegCenter) / kLogRegScale);
    }
    for (int blk = 0; blk < bps; ++blk) {
      const int64_t r0 = (int64_t)p * A.seg_len + (int64_t)blk * kFRows;
      u32x4 xin[2][2];
      nxt.make(A, h, seg_scale, seg_lr, xin);
      // Fetch the next block (this segment's next, or the next segment's
      // first; after the last block this one again, so the fetch is
      // unconditional and its wait is not pulled forward).
      {
        int64_t rn = r0;
        if (blk + 1 < bps)
          rn = r0 + kFRows;
        else if (p + (int)gridDim.x < nseg)
          rn = (int64_t)(p + gridDim.x) * A.seg_len;
        nxt.fetch(A, rn, h, l32);
      }
      // ---- layer 1, transposed: D[j][r], r = 32 t + l32 ----
#pragma unroll
      for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          f32x16 acc = {};
          acc = mfma16(w1r[c][0], xin[t][0], acc);
          acc = mfma16(w1r[c][1], xin[t][1], acc);
          // lane value 4 q + i: column jbase + 32 c + 8 q + 4 h + i, row 32 t + l32
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int j0 = jbase + 32 * c + 8 * q + 4 * h;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = leaky(acc[4 * q + i] + b1v[c][q][i]);
            u32x2 pk;
            pk[0] = pack_bf2(v[0], v[1]);
            pk[1] = pack_bf2(v[2], v[3]);
            *reinterpret_cast<u32x2*>(&act[buf][32 * t + l32][j0]) = pk;
          }
        }
      }
      __syncthreads();
      // ---- layer 2 ----
      f32x16 acc[2][2];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[c][t] = f32x16{};
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        u32x4 a[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
          a[t] = *reinterpret_cast<const u32x4*>(&act[buf][32 * t + l32][16 * s + 8 * h]);
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int t = 0; t < 2; ++t)
            acc[c][t] = MODE == 0 ? mfma16(a[t], w2r[c][s], acc[c][t])    // D[r][j]
                                  : mfma16(w2r[c][s], a[t], acc[c][t]);   // D[j][r]
      }
      if constexpr (MODE == 0) {
        // lane value 4 q + i: row 32 t + 8 q + 4 h + i, column jbase + 32 c + l32
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 16; ++v) zacc[c] += round_bf(leaky(acc[c][t][v] + b2c[c]));
      } else {
        // lane value 4 q + i: column jbase + 32 c + 8 q + 4 h + i, row 32 t + l32
        float fp[2] = {0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
              for (int t = 0; t < 2; ++t)
                fp[t] = __builtin_fmaf(round_bf(leaky(acc[c][t][4 * q + i] + b2v[c][q][i])),
                                       w3v[c][q][i], fp[t]);
            }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          fp[t] += __shfl_xor(fp[t], 32, kWave);
          if (h == 0) part[buf][w][32 * t + l32] = fp[t];
        }
        __syncthreads();
        if (tid < kFRows) {
          float f = ((part[buf][0][tid] + part[buf][1][tid]) + part[buf][2][tid]) +
                    part[buf][3][tid];
          f += A.b3[0] + (A.qn ? A.qn[r0 + tid] : 0.f);
          A.out[r0 + tid] = f;
        }
      }
      buf ^= 1;
    }
    if constexpr (MODE == 0) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float z = zacc[c] + __shfl_xor(zacc[c], 32, kWave);
        if (h == 0) A.out[(int64_t)p * kFH + jbase + 32 * c + l32] = z / (float)A.seg_len;
      }
    }
  }
}

}  // namespace
}  // namespace potamd

using namespace potamd;

extern "C" int potamd_mlp_fused(int mode, int rows, int seg_len, const void* inp,
                                const float* pts, const double* wts, const double* regs,
                                const float* cmax, const void* W1, const float* bias1,
                                const void* W2, const float* b2, const float* w3,
                                const float* b3, const float* qn, float* out, void* stream) {
  if (rows < 1 || seg_len < 1 || !W1 || !bias1 || !W2 || !b2 || !out ||
      (mode != 0 && mode != 1) || (mode == 1 && (!w3 || !b3)) ||
      (!inp && (!pts || !wts || !regs || !cmax)))
    return POTAMD_EINVAL;
  if (seg_len % kFRows != 0 || rows % seg_len != 0) return POTAMD_EUNSUPPORTED;
  // 16-byte vector loads of the input rows / points.
  if ((uintptr_t)(inp ? inp : (const void*)pts) % 16 != 0) return POTAMD_EINVAL;
  FusedMlpArgs fa;
  memset(&fa, 0, sizeof(fa));
  fa.mode = mode;
  fa.rows = rows;
  fa.seg_len = seg_len;
  fa.inp = static_cast<const uint16_t*>(inp);
  fa.pts = pts;
  fa.wts = wts;
  fa.regs = regs;
  fa.cmax = cmax;
  fa.W1 = static_cast<const uint16_t*>(W1);
  fa.bias1 = bias1;
  fa.W2 = static_cast<const uint16_t*>(W2);
  fa.b2 = b2;
  fa.w3 = w3;
  fa.b3 = b3;
  fa.qn = qn;
  fa.out = out;
  const int nseg = rows / seg_len;
  const int grid = nseg < device_cus() ? nseg : device_cus();
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 g(grid), b(kFW * kWave);
  if (mode == 0 && inp)
    hipLaunchKernelGGL((mlp_fused_kernel<0, 0>), g, b, 0, s, fa);
  else if (mode == 0)
    hipLaunchKernelGGL((mlp_fused_kernel<0, 1>), g, b, 0, s, fa);
  else if (inp)
    hipLaunchKernelGGL((mlp_fused_kernel<1, 0>), g, b, 0, s, fa);
  else
    hipLaunchKernelGGL((mlp_fused_kernel<1, 1>), g, b, 0, s, fa);
  HIP_TRY(hipGetLastError());
  return 0;
}
