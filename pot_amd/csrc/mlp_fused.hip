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
// a segment mean and a head kernel; the results agree up to fp32 rounding
// order (the biases enter first, as the accumulators' starting values) and,
// in the head, the split of w3 into bf16 hi + lo (2^-16 relative).
//
// The input rows are either given (bf16, rows x 32) or built here from the
// fp32 points (rows x 16), the per-row weights and the per-segment reg and
// cost maximum, with the arithmetic of the network's input-row kernel:
// [x * rsqrt(cmax), seg_len * w - 1, (log reg - c) / s, 0 ...].
//
// Layout: a 512-thread workgroup (two waves per SIMD, each under 256
// registers) walks whole segments; wave w owns hidden columns 32 w .. 32 w + 31
// of both layers and keeps its slices of W1 (32 x 32) and W2 (256 x 32) in
// registers for the whole launch.  Layer 1 is computed transposed (D[j][r])
// so each lane holds four consecutive columns of one row and writes them to
// LDS with one 8-byte store; layer 2 reads A1 from LDS (16 bytes per lane per
// MFMA K step).  The reductions after layer 2 also run on MFMA, on the
// packed bf16 A2 values: in mode 0 (D[r][j], a lane owns a column) the
// segment sum is ones x A2, accumulated over the segment's blocks; in mode 1
// (D[j][r], operands swapped, a lane owns a row) the head's dot product is
// w3 x A2 over the wave's columns, the waves' partials then added in LDS.
// A1 is double buffered in LDS: one barrier per block (two in mode 1).  The
// next block's inputs are fetched while the current one is computed.
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
constexpr int kFW = 8;               // waves per workgroup
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

// Two floats to packed bf16 (round to nearest even, v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned pack_bf2(float lo, float hi) {
  const bf16x2 v = {(__bf16)lo, (__bf16)hi};
  return __builtin_bit_cast(unsigned, v);
}
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

// bf16 of a float split as hi + lo (each a bf16), packed per pair.
__device__ __forceinline__ void split_bf2(float x0, float x1, unsigned& hi, unsigned& lo) {
  const bf16x2 h = {(__bf16)x0, (__bf16)x1};
  hi = __builtin_bit_cast(unsigned, h);
  lo = pack_bf2(x0 - (float)h[0], x1 - (float)h[1]);
}

// Head output of row r0 + i: the kFW wave partials of its dot product with w3,
// plus b3 and the optional per-row offset.
__device__ __forceinline__ void finish_rows(const FusedMlpArgs& A, const float (*pt)[kFRows],
                                            float b3, int64_t r0, int i) {
  const int64_t r = r0 + i;
  float f = 0.f;
#pragma unroll
  for (int ww = 0; ww < kFW; ++ww) f += pt[ww][i];
  f += b3 + (A.qn ? A.qn[r] : 0.f);
  A.out[r] = f;
}

template <int MODE, int IN>
__global__ __launch_bounds__(kFW * kWave) __attribute__((amdgpu_waves_per_eu(2, 2)))
void mlp_fused_kernel(FusedMlpArgs A) {
  __shared__ __attribute__((aligned(16))) uint16_t act[2][kFRows][kFStride];
  __shared__ float part[2][kFW][kFRows];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int h = lane >> 5, l32 = lane & 31;
  const int jbase = 32 * w;

  // Weight slices: column j = jbase + l32, rows 16 s + 8 h .. +7.
  u32x4 w1r[2], w2r[16];
#pragma unroll
  for (int s = 0; s < 2; ++s) w1r[s] = load_col8(A.W1, 16 * s + 8 * h, jbase + l32);
#pragma unroll
  for (int s = 0; s < 16; ++s) w2r[s] = load_col8(A.W2, 16 * s + 8 * h, jbase + l32);
  // The biases enter as the accumulators' starting values.  Layer 2: in mode
  // 0 every value of a lane is column jbase + l32; in mode 1 value 4 q + i is
  // column jbase + 8 q + 4 h + i.
  f32x16 b2init;
#pragma unroll
  for (int v = 0; v < 16; ++v)
    b2init[v] = MODE == 0 ? A.b2[jbase + l32] : A.b2[jbase + 8 * (v / 4) + 4 * h + (v % 4)];
  // Mode 0: the segment sum runs on MFMA against ones; mode 1: the head's dot
  // product runs on MFMA against w3 split into bf16 hi + lo.  Operand g
  // (columns jbase + 16 g .. +15) holds, in K slot 8 h + e, column
  // jbase + 16 g + 4 h + 8 (e / 4) + e % 4: the columns this lane holds.
  const u32x4 ones4 = {0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u};
  u32x4 w3hi[2], w3lo[2];
  if constexpr (MODE == 1) {
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        const int j = jbase + 16 * g + 4 * h + 8 * (e / 4) + e % 4;
        unsigned hi, lo;
        split_bf2(A.w3[j], A.w3[j + 1], hi, lo);
        w3hi[g][e / 2] = hi;
        w3lo[g][e / 2] = lo;
      }
  }
  const float b3 = MODE == 1 ? A.b3[0] : 0.f;

  const int nseg = A.rows / A.seg_len;
  const int bps = A.seg_len / kFRows;
  BlockIn<IN> nxt;
  if ((int)blockIdx.x < nseg) nxt.fetch(A, (int64_t)blockIdx.x * A.seg_len, h, l32);
  int buf = 0;
  // Mode 1: the row block whose wave partials sit in part[buf ^ 1], summed
  // after the next block's barrier (-1: none), so each block has one barrier.
  int64_t pend = -1;
  for (int p = blockIdx.x; p < nseg; p += gridDim.x) {
    f32x16 zsum = {};
    // Per-segment values, once per segment: the layer-1 bias of the lane's
    // values (value 4 q + i: column jbase + 8 q + 4 h + i) and (inputs built
    // from the points) the scale and log reg.
    const float* bias1 = MODE == 0 ? A.bias1 : A.bias1 + (int64_t)p * kFH;
    f32x16 b1init;
#pragma unroll
    for (int v = 0; v < 16; ++v) b1init[v] = bias1[jbase + 8 * (v / 4) + 4 * h + (v % 4)];
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
      for (int t = 0; t < 2; ++t) {
        f32x16 acc = mfma16(w1r[0], xin[t][0], b1init);
        acc = mfma16(w1r[1], xin[t][1], acc);
        // lane value 4 q + i: column jbase + 8 q + 4 h + i, row 32 t + l32
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          u32x2 pk;
          pk[0] = pack_bf2(leaky(acc[4 * q]), leaky(acc[4 * q + 1]));
          pk[1] = pack_bf2(leaky(acc[4 * q + 2]), leaky(acc[4 * q + 3]));
          *reinterpret_cast<u32x2*>(&act[buf][32 * t + l32][jbase + 8 * q + 4 * h]) = pk;
        }
      }
      __syncthreads();
      if constexpr (MODE == 1) {
        if (pend >= 0 && tid < kFRows) finish_rows(A, part[buf ^ 1], b3, pend, tid);
        pend = r0;
      }
      // ---- layer 2 ----
      f32x16 acc[2] = {b2init, b2init};
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        u32x4 a[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
          a[t] = *reinterpret_cast<const u32x4*>(&act[buf][32 * t + l32][16 * s + 8 * h]);
#pragma unroll
        for (int t = 0; t < 2; ++t)
          acc[t] = MODE == 0 ? mfma16(a[t], w2r[s], acc[t])    // D[r][j]
                             : mfma16(w2r[s], a[t], acc[t]);   // D[j][r]
      }
      // A2 = leaky(acc) rounded to bf16, packed per pair: y[t][g] holds
      // values 8 g .. 8 g + 7 of acc[t].
      u32x4 y[2][2];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            y[t][g][e] = pack_bf2(leaky(acc[t][8 * g + 2 * e]), leaky(acc[t][8 * g + 2 * e + 1]));
      if constexpr (MODE == 0) {
        // Mode 0 (D[r][j], a lane's values all column jbase + l32, rows
        // 32 t + 8 q + 4 h + i): column sums over the 64 rows, as ones x A2.
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int g = 0; g < 2; ++g) zsum = mfma16(ones4, y[t][g], zsum);
      } else {
        // Mode 1 (D[j][r], lane value 4 q + i: column jbase + 8 q + 4 h + i,
        // row 32 t + l32): the dot product with w3 over the wave's 32
        // columns, as w3 x A2 (hi and lo parts).
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          f32x16 d = {};
#pragma unroll
          for (int g = 0; g < 2; ++g) {
            d = mfma16(w3hi[g], y[t][g], d);
            d = mfma16(w3lo[g], y[t][g], d);
          }
          if (h == 0) part[buf][w][32 * t + l32] = d[0];
        }
      }
      buf ^= 1;
    }
    if constexpr (MODE == 0) {
      if (h == 0) A.out[(int64_t)p * kFH + jbase + l32] = zsum[0] / (float)A.seg_len;
    }
  }
  if constexpr (MODE == 1) {
    if (pend >= 0) {
      __syncthreads();
      if (tid < kFRows) finish_rows(A, part[buf ^ 1], b3, pend, tid);
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
