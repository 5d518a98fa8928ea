// Online-cost Sinkhorn: the squared Euclidean cost is never stored.
//
// sinkhorn_log (ot/bregman/_sinkhorn.py:880-899) over
//     M_rc = (|x_r|^2 + |y_c|^2 - 2 x_r . y_c) / cmax        (ot/utils.py:270-307)
// with cmax = max M (the configs divide by M.max()).  In base 2, with
// s2 = log2(e) / (reg * cmax), the exponent of an element of a row update is
//     t_rc = 2 s2 x_r . y_c + (pot_c - s2 |y_c|^2) - s2 |x_r|^2,
// so a half-iteration is a GEMM x . y^T with a log-sum-exp epilogue.  The row
// update and the column update are the same pass with the two point sets
// swapped ("R" is the side being updated, "C" the side reduced over).
//
// Pass kernel (one per precision of x . y^T):
//  * a workgroup owns 128 R points (4 waves x 32) and one split of the C
//    points, streamed through LDS 64 at a time (double buffered);
//  * each wave computes 32 x 32 tiles D[c][r] on 32x32 MFMA, with the R
//    operand held in registers for the whole pass and the C operand read
//    from LDS, so one lane owns one R point and reduces 16 C values per tile
//    without any cross-lane traffic;
//  * the operands are pre-scaled by sqrt(2 s2), and the accumulator starts
//    at (pot_c - s2 |y_c|^2) - shift_r (rounded once from fp64), so the MFMA
//    output is the base-2 exponent itself and the epilogue is one exp2 and
//    one add per element;
//  * shift_r is the previous pass's inner log-sum-exp of the same point;
//    a lane re-centres (and the wave recomputes that tile) only when the
//    tile's sum shows an element above about +16 or everything below -40
//    with nothing accumulated yet, which keeps the sum exact at the cost of
//    a rare extra tile and no per-element maximum;
//  * partial sums are combined in fp64 (16 fp32 terms per tile each).
// The merge kernel combines the split partials in fp64, writes the new
// potential, and on the column update computes the reference's marginal
// error ||P^T 1 - b||_2 (from the column sums the next update needs anyway)
// and takes the stop decision on the device (last-block ticket).
//
// Precisions of x . y^T: fp32 MFMA (32x32x2), bf16 MFMA (32x32x16), bf16
// hi+lo split with three products (compensated), fp8 e4m3 (32x32x16) with a
// per-point scale applied in the epilogue, and an fp64 vector-ALU pass that
// serves as the full-size reference.  Exponents, sums and potentials are
// fp32 / fp64 as described above in every variant.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "../../include/pot_amd.h"
#include "host.h"
#include "internal.h"

namespace potamd {
namespace {

constexpr int kRowsPerWG = 128;  // R points per workgroup of the MFMA pass
constexpr int kStep = 64;        // C points per LDS stage
constexpr int kPad = 128;        // packed operands are padded to this many points
constexpr float kFp8Max = 224.f; // largest |q| of a scaled fp8 operand

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <int P> struct Op;
template <> struct Op<POTAMD_ONLINE_F32> { static constexpr int es = 4; static constexpr bool lo = false; };
template <> struct Op<POTAMD_ONLINE_BF16> { static constexpr int es = 2; static constexpr bool lo = false; };
template <> struct Op<POTAMD_ONLINE_BF16X3> { static constexpr int es = 2; static constexpr bool lo = true; };
template <> struct Op<POTAMD_ONLINE_FP8> { static constexpr int es = 1; static constexpr bool lo = false; };
template <> struct Op<POTAMD_ONLINE_F64> { static constexpr int es = 4; static constexpr bool lo = false; };

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) {
  return __uint_as_float((unsigned int)h << 16);
}

// x = q[0] + q[1] + q[2] + q[3] to 2^-32 relative, each part a bf16, packed
// two per word; a non-finite x gives (x, 0, 0, 0).
__device__ __forceinline__ u32x2 split4_bf16(double x) {
  const bool fin = __builtin_isfinite(x);
  unsigned q[4];
  double r = x;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const unsigned short b = bf16_rne((float)r);
    q[i] = (i == 0 || fin) ? b : 0u;
    r -= (double)bf16_val(b);
  }
  return u32x2{q[0] | (q[1] << 16), q[2] | (q[3] << 16)};
}
constexpr unsigned kOnes = 0x3F803F80u;  // two bf16 1.0

// t <- exp2(t) elementwise; returns the sum of the 16 values as a pairwise
// tree.
__device__ __forceinline__ float exp2_tree_sum(f32x16& t) {
#pragma unroll
  for (int i = 0; i < 16; ++i) t[i] = exp2_fast(t[i]);
  float s[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) s[i] = t[2 * i] + t[2 * i + 1];
#pragma unroll
  for (int i = 0; i < 4; ++i) s[i] = s[2 * i] + s[2 * i + 1];
  return (s[0] + s[1]) + (s[2] + s[3]);
}

// A point coordinate as double; points are fp32 or fp64 (x64) in HBM.
__device__ __forceinline__ double coord(const void* p, int64_t idx, int x64) {
  return x64 ? static_cast<const double*>(p)[idx] : (double)static_cast<const float*>(p)[idx];
}

__device__ __forceinline__ double s2_of(const double* cscale, double reg) {
  return kLog2e / (reg * cscale[0]);
}

// ---------------------------------------------------------------------------
// Operand packing: one wave per point, 4 coordinates per lane (D <= 256).
// Writes the fp64 squared norm of the fp32 point and, for the MFMA
// precisions, the point scaled by sqrt(2 s2) (or unscaled) in the operand
// format.  Points and coordinates beyond (rows, d) are zero.
struct PackArgs {
  const void* X;
  int x64;
  int64_t ldx;
  int rows, d, D, rows_pad;
  const double* cscale;
  double reg;
  int scaled;
  void* pk;
  void* pk_lo;
  float* pk_scale;
  double* norms;
};

template <int P>
__global__ __launch_bounds__(256) void online_pack_kernel(PackArgs A) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= A.rows_pad) return;  // uniform per wave
  const double sc = A.scaled ? sqrt(2.0 * s2_of(A.cscale, A.reg)) : 1.0;
  double nrm = 0.0;
  if (row < A.rows)
    for (int e = lane; e < A.d; e += 64) {
      const double x = coord(A.X, (int64_t)row * A.ldx + e, A.x64);
      nrm = fma(x, x, nrm);
    }
  nrm = wave_sum(nrm);
  if (lane == 0) A.norms[row] = nrm;
  if constexpr (P == POTAMD_ONLINE_F64) return;
  const int e0 = lane * 4;
  float t[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int e = e0 + j;
    float x = 0.f;
    if (row < A.rows && e < A.d) x = (float)coord(A.X, (int64_t)row * A.ldx + e, A.x64);
    t[j] = (float)((double)x * sc);
  }
  if constexpr (P == POTAMD_ONLINE_F32) {
    if (e0 < A.D) {
      float* o = static_cast<float*>(A.pk) + (int64_t)row * A.D + e0;
      *reinterpret_cast<float4*>(o) = make_float4(t[0], t[1], t[2], t[3]);
    }
  } else if constexpr (P == POTAMD_ONLINE_BF16 || P == POTAMD_ONLINE_BF16X3) {
    if (e0 < A.D) {
      unsigned short hi[4], lo[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        hi[j] = bf16_rne(t[j]);
        lo[j] = bf16_rne(t[j] - bf16_val(hi[j]));
      }
      u32x2 wh = {(unsigned)hi[0] | ((unsigned)hi[1] << 16), (unsigned)hi[2] | ((unsigned)hi[3] << 16)};
      *reinterpret_cast<u32x2*>(static_cast<unsigned short*>(A.pk) + (int64_t)row * A.D + e0) = wh;
      if constexpr (P == POTAMD_ONLINE_BF16X3) {
        u32x2 wl = {(unsigned)lo[0] | ((unsigned)lo[1] << 16), (unsigned)lo[2] | ((unsigned)lo[3] << 16)};
        *reinterpret_cast<u32x2*>(static_cast<unsigned short*>(A.pk_lo) + (int64_t)row * A.D + e0) = wl;
      }
    }
  } else if constexpr (P == POTAMD_ONLINE_FP8) {
    float am = fmaxf(fmaxf(fabsf(t[0]), fabsf(t[1])), fmaxf(fabsf(t[2]), fabsf(t[3])));
    am = wave_max(am);
    const float sig = am > 0.f ? am / kFp8Max : 1.f;
    const float inv = 1.f / sig;
    if (e0 < A.D) {
      int w = __builtin_amdgcn_cvt_pk_fp8_f32(t[0] * inv, t[1] * inv, 0, false);
      w = __builtin_amdgcn_cvt_pk_fp8_f32(t[2] * inv, t[3] * inv, w, true);
      *reinterpret_cast<int*>(static_cast<unsigned char*>(A.pk) + (int64_t)row * A.D + e0) = w;
    }
    if (lane == 0) A.pk_scale[row] = sig;
  }
}

// ---------------------------------------------------------------------------
// MFMA pass.
struct PassArgs {
  const void* r_pk;
  const void* r_lo;
  const float* r_sc;
  const void* c_pk;
  const void* c_lo;
  const float* c_sc;
  const double* c_pot;   // C potential, base 2 (LSE mode)
  const double* c_norm;  // |c|^2, fp64
  const double* r_shift; // previous inner LSE of each R point (NaN = none)
  int nr, nc;
  int steps_per_split;
  const double* cscale;
  double reg;
  double* part;   // [splits][nr_pad]
  int* part_idx;  // max mode: arg max
  int nr_pad;
  const Ctrl* ctrl;
};

enum PassMode : int { kModeLse = 0, kModeMax = 1 };

template <int P>
__device__ __forceinline__ void mma_chunk(f32x16& acc, u32x4 a, u32x4 b) {
  if constexpr (P == POTAMD_ONLINE_F32) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a[j]), __uint_as_float(b[j]),
                                                 acc, 0, 0, 0);
  } else if constexpr (P == POTAMD_ONLINE_FP8) {
    const long a0 = (long)(((unsigned long)a[1] << 32) | a[0]);
    const long a1 = (long)(((unsigned long)a[3] << 32) | a[2]);
    const long b0 = (long)(((unsigned long)b[1] << 32) | b[0]);
    const long b1 = (long)(((unsigned long)b[3] << 32) | b[2]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a1, b1, acc, 0, 0, 0);
  } else {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                  __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
  }
}

// bf16x3 at d = 64 needs a few registers more than three waves per SIMD
// allow; the pass is latency-bound there, so three waves are asked for.
template <int P, int D, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    (P == POTAMD_ONLINE_BF16X3 && D == 64 && MODE == kModeLse) ? 3 : 1)))
void online_pass_kernel(PassArgs A) {
  if (A.ctrl && A.ctrl->done) return;
  constexpr int es = Op<P>::es;
  constexpr bool LO = Op<P>::lo;
  constexpr bool FP8 = P == POTAMD_ONLINE_FP8;
  constexpr int ROWB = D * es + 16;           // LDS bytes per C point (padded)
  constexpr int TILE = kStep * ROWB;
  constexpr int CPR = D * es / 16;            // 16-byte chunks per point
  constexpr int NCHUNK = kStep * CPR;         // chunks per stage
  constexpr int NPT = (NCHUNK + 255) / 256;   // chunks per thread
  constexpr int NB = D * es / 32;             // 16-byte chunks per lane (half a point)
  // bf16 operands: the potential term enters the accumulator through one
  // more bf16 MFMA (K slots [P1..P4, 1, 1, 1, 1] x [1, 1, 1, 1, -Q1..-Q4],
  // P the C point's term and Q the lane's shift, each split into four bf16
  // parts), issued after the dot product.
  constexpr bool KS = MODE == kModeLse && (P == POTAMD_ONLINE_BF16 || P == POTAMD_ONLINE_BF16X3);

  __shared__ __attribute__((aligned(16))) char sm_t[2][TILE];
  __shared__ __attribute__((aligned(16))) char sm_lo[LO ? 2 : 1][LO ? TILE : 16];
  __shared__ double sm_v[2][kStep];
  __shared__ float sm_sc[2][kStep];
  __shared__ u32x2 sm_kp[KS ? 2 : 1][KS ? kStep : 1];

  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int h = lane >> 5;
  const int l32 = lane & 31;
  const int r = blockIdx.x * kRowsPerWG + wave * 32 + l32;
  const int split = blockIdx.y;
  const int nsteps = (A.nc + kStep - 1) / kStep;
  const int step0 = split * A.steps_per_split;
  const int step1 = min(step0 + A.steps_per_split, nsteps);
  const double s2 = MODE == kModeLse ? s2_of(A.cscale, A.reg) : 0.0;

  // R operand: half a point per lane, for the whole pass.
  u32x4 rb[NB];
  u32x4 rl[LO ? NB : 1];
  {
    const char* rp = static_cast<const char*>(A.r_pk) + ((int64_t)r * D + h * (D / 2)) * es;
#pragma unroll
    for (int i = 0; i < NB; ++i) rb[i] = *reinterpret_cast<const u32x4*>(rp + 16 * i);
    if constexpr (LO) {
      const char* lp = static_cast<const char*>(A.r_lo) + ((int64_t)r * D + h * (D / 2)) * es;
#pragma unroll
      for (int i = 0; i < NB; ++i) rl[i] = *reinterpret_cast<const u32x4*>(lp + 16 * i);
    }
  }
  const float r_sig = FP8 ? A.r_sc[r] : 1.f;

  double cl = 0.0;     // shift of this lane's sum (base 2)
  bool fresh = true;   // no shift yet
  if (MODE == kModeLse && A.r_shift) {
    const double s0 = A.r_shift[r];
    if (__builtin_isfinite(s0)) {
      cl = s0;
      fresh = false;
    }
  }
  // K slots of the lane's shift (B operand): [1, 1, 1, 1, -Q1..-Q4].
  auto shift_slots = [&]() {
    const u32x2 q = split4_bf16(-cl);
    const u32x4 k = {kOnes, kOnes, q[0], q[1]};
    return k;
  };
  u32x4 rkq = {0u, 0u, 0u, 0u};
  if constexpr (KS) rkq = shift_slots();
  double ssum = 0.0;
  float best = -__builtin_inff();
  int best_c = 0;

  // Stage loader: C points [step*64, step*64 + 64) into registers.
  u32x4 st[NPT];
  u32x4 stl[LO ? NPT : 1];
  double st_v = 0.0;
  float st_sc = 1.f;
  u32x2 st_kp = {0u, 0u};
  auto load = [&](int step) {
    const int64_t c0 = (int64_t)step * kStep;
    const char* cp = static_cast<const char*>(A.c_pk) + c0 * D * es;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      const int q = tid + 256 * i;
      if (NCHUNK % 256 == 0 || q < NCHUNK)
        st[i] = *reinterpret_cast<const u32x4*>(cp + (int64_t)(q / CPR) * D * es + (q % CPR) * 16);
    }
    if constexpr (LO) {
      const char* lp = static_cast<const char*>(A.c_lo) + c0 * D * es;
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        const int q = tid + 256 * i;
        if (NCHUNK % 256 == 0 || q < NCHUNK)
          stl[i] = *reinterpret_cast<const u32x4*>(lp + (int64_t)(q / CPR) * D * es + (q % CPR) * 16);
      }
    }
    if (tid < kStep) {
      const int64_t c = c0 + tid;
      if (c < A.nc) {
        st_v = MODE == kModeLse ? A.c_pot[c] - s2 * A.c_norm[c] : A.c_norm[c];
      } else {
        st_v = -__builtin_inf();
      }
      if constexpr (FP8) st_sc = A.c_sc[c];
      if constexpr (KS) st_kp = split4_bf16(st_v);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      const int q = tid + 256 * i;
      if (NCHUNK % 256 == 0 || q < NCHUNK)
        *reinterpret_cast<u32x4*>(&sm_t[buf][(q / CPR) * ROWB + (q % CPR) * 16]) = st[i];
    }
    if constexpr (LO) {
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        const int q = tid + 256 * i;
        if (NCHUNK % 256 == 0 || q < NCHUNK)
          *reinterpret_cast<u32x4*>(&sm_lo[buf][(q / CPR) * ROWB + (q % CPR) * 16]) = stl[i];
      }
    }
    if (tid < kStep) {
      sm_v[buf][tid] = st_v;
      if constexpr (FP8) sm_sc[buf][tid] = st_sc;
      if constexpr (KS) sm_kp[buf][tid] = st_kp;
    }
  };

  // One 32 x 32 tile: C points sb*32 .. sb*32+31 of the stage against this
  // wave's 32 R points.  Lane value i corresponds to C point
  // sb*32 + 8*(i/4) + 4*h + (i%4).
  auto tile = [&](int buf, int sb, double shift) -> f32x16 {
    f32x16 acc;
    const double* vrow = &sm_v[buf][sb * 32 + 4 * h];
    if constexpr (MODE == kModeLse && !FP8 && !KS) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) acc[4 * q + rr] = (float)(vrow[8 * q + rr] - shift);
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    }
    const char* trow = &sm_t[buf][(sb * 32 + l32) * ROWB + h * (D / 2) * es];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const u32x4 a = *reinterpret_cast<const u32x4*>(trow + 16 * i);
      if constexpr (LO) {
        const char* lrow = &sm_lo[buf][(sb * 32 + l32) * ROWB + h * (D / 2) * es];
        const u32x4 al = *reinterpret_cast<const u32x4*>(lrow + 16 * i);
        mma_chunk<P>(acc, a, rl[i]);
        mma_chunk<P>(acc, al, rb[i]);
      }
      mma_chunk<P>(acc, a, rb[i]);
    }
    if constexpr (KS) {
      // Output value i of lane (r, h) is C point 8(i/4) + 4h + i%4 of the
      // tile, so the C points whose bit 2 is h take the K slots 8h..8h+7,
      // which meet the shift slots of lane (r, h): each lane's outputs get
      // its own shift.
      const u32x2 kp = sm_kp[buf][sb * 32 + l32];
      const bool mine = ((l32 >> 2) & 1) == h;
      const u32x4 ka = {mine ? kp[0] : 0u, mine ? kp[1] : 0u, mine ? kOnes : 0u, mine ? kOnes : 0u};
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ka),
                                                    __builtin_bit_cast(bf16x8, rkq), acc, 0, 0, 0);
    }
    if constexpr (MODE == kModeLse && FP8) {
      const float* srow = &sm_sc[buf][sb * 32 + 4 * h];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr)
          acc[4 * q + rr] = __builtin_fmaf(acc[4 * q + rr], r_sig * srow[8 * q + rr],
                                           (float)(vrow[8 * q + rr] - shift));
    }
    return acc;
  };

  if (step0 < step1) {
    load(step0);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  for (int step = step0; step < step1; ++step) {
    const bool more = step + 1 < step1;
    if (more) load(step + 1);
#pragma unroll
    for (int sb = 0; sb < 2; ++sb) {
      f32x16 t = tile(buf, sb, cl);
      if constexpr (MODE == kModeLse) {
        // exp2 in place (no extra registers), then a depth-4 tree sum, so
        // the 16 exp2 of a lane are independent of each other.
        float ts = exp2_tree_sum(t);
        // The tile sum itself flags a shift that is missing, too low (an
        // element above about +16: overflow risk) or too high while nothing
        // has been accumulated (everything below -40: underflow risk); only
        // then is the tile maximum formed (from the exponents, recomputed)
        // and the lane re-centred.
        const bool suspect = fresh || !(ts <= 0x1p20f) || (ts < 0x1p-40f && ssum < 0x1p-40);
        if (__ballot(suspect)) {
          t = tile(buf, sb, cl);
          float tmax = t[0];
#pragma unroll
          for (int i = 1; i < 16; ++i) tmax = fmaxf(tmax, t[i]);
          const bool need = suspect && tmax >= -3.0e38f &&
                            (fresh || tmax > 16.f || (tmax < -40.f && ssum < 0x1p-40));
          if (__ballot(need)) {
            if (need) {
              double delta = tmax;
              if (!fresh && ssum > 0.0) delta = fmax(delta, log2(ssum));
              ssum *= exp2(-delta);
              cl += delta;
              fresh = false;
              if constexpr (KS) rkq = shift_slots();
            }
            t = tile(buf, sb, cl);
            ts = exp2_tree_sum(t);
          }
        }
        ssum += (double)ts;
      } else {
        const double* vrow = &sm_v[buf][sb * 32 + 4 * h];
        const int cb = step * kStep + sb * 32 + 4 * h;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            const float val = __builtin_fmaf(-2.f, t[4 * q + rr], (float)vrow[8 * q + rr]);
            if (val > best) {
              best = val;
              best_c = cb + 8 * q + rr;
            }
          }
      }
    }
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // The two half-waves hold the same R point: combine, lanes 0..31 write.
  if constexpr (MODE == kModeLse) {
    const double cl_o = __shfl_xor(cl, 32, 64);
    const double s_o = __shfl_xor(ssum, 32, 64);
    double I;
    if (ssum > 0.0 && s_o > 0.0) {
      const double cm = fmax(cl, cl_o);
      I = cm + log2(ssum * exp2(cl - cm) + s_o * exp2(cl_o - cm));
    } else if (ssum > 0.0) {
      I = cl + log2(ssum);
    } else if (s_o > 0.0) {
      I = cl_o + log2(s_o);
    } else {
      I = -__builtin_inf();
    }
    if (h == 0) A.part[(int64_t)split * A.nr_pad + r] = I;
  } else {
    const float b_o = __shfl_xor(best, 32, 64);
    const int c_o = __shfl_xor(best_c, 32, 64);
    if (b_o > best || (b_o == best && c_o < best_c)) {
      best = b_o;
      best_c = c_o;
    }
    if (h == 0) {
      A.part[(int64_t)split * A.nr_pad + r] = best;
      A.part_idx[(int64_t)split * A.nr_pad + r] = best_c;
    }
  }
}

// ---------------------------------------------------------------------------
// fp64 pass (reference precision): 64 R points x 64 C points per stage,
// 4 x 4 per thread, exact online maximum per point.
struct F64PassArgs {
  const void* R;
  int64_t ldr;
  const void* C;
  int64_t ldc;
  int x64;
  int nr, nc, d;
  const double* c_pot;
  const double* c_norm;
  const double* cscale;
  double reg;
  int steps_per_split;
  double* part;
  int nr_pad;
  const Ctrl* ctrl;
};

__global__ __launch_bounds__(256) void online_f64_pass_kernel(F64PassArgs A) {
  if (A.ctrl && A.ctrl->done) return;
  __shared__ double Rs[64][33];
  __shared__ double Cs[64][33];
  __shared__ double vs[64];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int r0 = blockIdx.x * 64;
  const int split = blockIdx.y;
  const int nsteps = (A.nc + 63) / 64;
  const int step0 = split * A.steps_per_split;
  const int step1 = min(step0 + A.steps_per_split, nsteps);
  const double s2 = s2_of(A.cscale, A.reg);
  const double k2 = 2.0 * s2;
  double mx[4], sm[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    mx[i] = -__builtin_inf();
    sm[i] = 0.0;
  }
  for (int step = step0; step < step1; ++step) {
    const int c0 = step * 64;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < A.d; k0 += 32) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = tid + 256 * q;
        const int row = e >> 5, col = e & 31;
        const int k = k0 + col;
        Rs[row][col] = (r0 + row < A.nr && k < A.d) ? coord(A.R, (int64_t)(r0 + row) * A.ldr + k, A.x64) : 0.0;
        Cs[row][col] = (c0 + row < A.nc && k < A.d) ? coord(A.C, (int64_t)(c0 + row) * A.ldc + k, A.x64) : 0.0;
      }
      if (k0 == 0 && tid < 64) {
        const int c = c0 + tid;
        vs[tid] = c < A.nc ? A.c_pot[c] - s2 * A.c_norm[c] : -__builtin_inf();
      }
      __syncthreads();
#pragma unroll 8
      for (int kk = 0; kk < 32; ++kk) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Rs[ty * 4 + i][kk];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Cs[tx + 16 * j][kk];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double t[4];
      double tm = -__builtin_inf();
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        t[j] = fma(k2, acc[i][j], vs[tx + 16 * j]);
        tm = fmax(tm, t[j]);
      }
      if (tm > mx[i]) {
        sm[i] *= exp2(mx[i] - tm);
        mx[i] = tm;
      }
      if (mx[i] > -__builtin_inf()) {
#pragma unroll
        for (int j = 0; j < 4; ++j) sm[i] += exp2(t[j] - mx[i]);
      }
    }
  }
  // Combine the 16 threads of a row group (consecutive lanes).
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const double mo = __shfl_xor(mx[i], o, 64);
      const double so = __shfl_xor(sm[i], o, 64);
      const double mm = fmax(mx[i], mo);
      if (mm > -__builtin_inf()) {
        sm[i] = sm[i] * exp2(mx[i] - mm) + so * exp2(mo - mm);
        mx[i] = mm;
      }
    }
    if (tx == 0) {
      const int r = r0 + ty * 4 + i;
      A.part[(int64_t)split * A.nr_pad + r] = sm[i] > 0.0 ? mx[i] + log2(sm[i]) : -__builtin_inf();
    }
  }
}

// ---------------------------------------------------------------------------
// Cost maximum: per R point the best C of the max pass, re-evaluated exactly
// in fp64 from the fp32 points; then the maximum over points (last block).
struct MaxReduceArgs {
  const double* part;
  const int* part_idx;
  int S, nr, nr_pad;
  const void* X;
  int64_t ldx;
  const void* Y;
  int64_t ldy;
  int d;
  int x64;
  double* blockmax;
  Ctrl* ctrl;
  double* cscale;
};

__global__ __launch_bounds__(256) void online_max_reduce_kernel(MaxReduceArgs A) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  double v = -__builtin_inf();
  if (r < A.nr) {
    double bv = -__builtin_inf();
    int bc = 0;
    for (int s = 0; s < A.S; ++s) {
      const double pv = A.part[(int64_t)s * A.nr_pad + r];
      const int pc = A.part_idx[(int64_t)s * A.nr_pad + r];
      if (pv > bv || (pv == bv && pc < bc)) {
        bv = pv;
        bc = pc;
      }
    }
    double acc = 0.0;
    for (int k = 0; k < A.d; ++k) {
      const double df = coord(A.X, (int64_t)r * A.ldx + k, A.x64) -
                        coord(A.Y, (int64_t)bc * A.ldy + k, A.x64);
      acc = fma(df, df, acc);
    }
    v = acc;
  }
  __shared__ double red[4];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    A.blockmax[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  __threadfence();
  __shared__ int last;
  if (threadIdx.x == 0) last = atomicAdd(&A.ctrl->counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  double m = -__builtin_inf();
  for (int b = threadIdx.x; b < (int)gridDim.x; b += 256) m = fmax(m, A.blockmax[b]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    A.cscale[0] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    A.ctrl->counter = 0;
  }
}

// ---------------------------------------------------------------------------
// Merge of the split partials into the new potential of the R side.
struct OnlineMergeArgs {
  const double* part;
  int S, nr, nr_pad;
  const double* w;       // weights of the R side
  const double* logw2;   // log2 of the weights
  const double* norm;    // |r|^2
  const double* pot_cur; // current potential of the R side (error only)
  double* pot_new;
  double* shift;
  const double* cscale;
  double reg;
  int check, decide, iter, max_iter;
  int l1;         // error: 1 = sum |marginal - w| (empirical_sinkhorn), 0 = 2-norm
  int inclusive;  // stop: 1 = err <= stop_thr (empirical_sinkhorn), 0 = err < stop_thr
  double stop_thr;
  double* err_hist;
  int err_cap;
  double* errpart;
  Ctrl* ctrl;
};

// One wave per workgroup (the R side has 16384 points at the headline size:
// 256 workgroups fill the chip), the split partials read 8 at a time so their
// loads overlap.  The stop decision is online_decide_kernel's, launched after
// this kernel, so no block needs a device-scope fence.
constexpr int kMergeRows = 64;

__global__ __launch_bounds__(kMergeRows) void online_merge_kernel(OnlineMergeArgs A) {
  Ctrl* ctrl = A.ctrl;
  if (ctrl->done) return;
  const int tid = threadIdx.x;
  const int r = blockIdx.x * kMergeRows + tid;
  double e2 = 0.0;
  if (r < A.nr) {
    const double* pr = A.part + r;
    double mx = -__builtin_inf();
#pragma unroll 8
    for (int s = 0; s < A.S; ++s) mx = fmax(mx, pr[(int64_t)s * A.nr_pad]);
    double I = -__builtin_inf();
    if (mx > -__builtin_inf()) {
      double acc = 0.0;
#pragma unroll 8
      for (int s = 0; s < A.S; ++s) acc += exp2(pr[(int64_t)s * A.nr_pad] - mx);
      I = mx + log2(acc);
    }
    const double s2 = s2_of(A.cscale, A.reg);
    const double L2 = I - s2 * A.norm[r];
    const double wr = A.w[r];
    if (A.check) {
      const double cs = exp2(A.pot_cur[r] + L2);
      e2 = A.l1 ? fabs(cs - wr) : (cs - wr) * (cs - wr);
    }
    A.pot_new[r] = wr > 0.0 ? A.logw2[r] - L2 : -__builtin_inf();
    A.shift[r] = I;
  }
  if (!A.check) return;
  e2 = wave_sum(e2);
  if (tid == 0) A.errpart[blockIdx.x] = e2;
}

// Stop decision after a checking (or the last) merge: one workgroup; the
// kernel boundary makes the merge's partials visible.  The partials are
// summed in a fixed order (lane-strided, a butterfly per wave, the waves in
// order), so the result is deterministic.
__global__ __launch_bounds__(256) void online_decide_kernel(OnlineMergeArgs A, int nblk) {
  Ctrl* ctrl = A.ctrl;
  if (ctrl->done) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  __shared__ double wpart[4];
  if (A.check) {
    double acc = 0.0;
    for (int b = tid; b < nblk; b += 256) acc += A.errpart[b];
    acc = wave_sum(acc);
    if (lane == 0) wpart[w] = acc;
  }
  __syncthreads();
  if (tid != 0) return;
  const int it = A.iter;
  if (A.check) {
    const double acc = ((wpart[0] + wpart[1]) + wpart[2]) + wpart[3];
    const double err = A.l1 ? acc : sqrt(acc);
    ctrl->err = err;
    const int k = ctrl->n_err;
    if (k < A.err_cap) A.err_hist[k] = err;
    ctrl->n_err = k + 1;
    if (A.inclusive ? err <= A.stop_thr : err < A.stop_thr) {
      ctrl->done = 1;
      ctrl->status = kConverged;
      ctrl->niter = it;
      return;
    }
  }
  if (it >= A.max_iter - 1) {
    ctrl->done = 1;
    ctrl->status = kMaxIter;
    ctrl->niter = it;
  }
}

// ---------------------------------------------------------------------------
// Plan, cost and marginals in fp64: M_rc from the fp32 points in fp64,
// P_rc = exp(-M_rc / reg + u_r + v_c).  One workgroup per 64 rows sweeps all
// columns; column sums are fp64 atomics, row sums and cost are per block.
struct OnlinePlanArgs {
  const void* X;
  int64_t ldx;
  const void* Y;
  int64_t ldy;
  int x64;
  int n, m, d;
  double reg;
  const double* cscale;
  const double* u;   // natural log potentials
  const double* v;
  const double* xn;
  const double* yn;
  double* P;         // optional n x ldp
  int64_t ldp;
  double* rowsum;    // [n]
  double* colsum;    // [m], zeroed
  double* costpart;  // [blocks]
};

__global__ __launch_bounds__(256) void online_plan_kernel(OnlinePlanArgs A) {
  __shared__ double Rs[64][33];
  __shared__ double Cs[64][33];
  __shared__ double cred[16][64];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int r0 = blockIdx.x * 64;
  const double inv_c = 1.0 / A.cscale[0];
  double rs[4] = {0.0, 0.0, 0.0, 0.0};
  double cost = 0.0;
  double ur[4], xr[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + ty * 4 + i;
    ur[i] = r < A.n ? A.u[r] : -__builtin_inf();
    xr[i] = r < A.n ? A.xn[r] : 0.0;
  }
  for (int c0 = 0; c0 < A.m; c0 += 64) {
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < A.d; k0 += 32) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = tid + 256 * q;
        const int row = e >> 5, col = e & 31;
        const int k = k0 + col;
        Rs[row][col] = (r0 + row < A.n && k < A.d) ? coord(A.X, (int64_t)(r0 + row) * A.ldx + k, A.x64) : 0.0;
        Cs[row][col] = (c0 + row < A.m && k < A.d) ? coord(A.Y, (int64_t)(c0 + row) * A.ldy + k, A.x64) : 0.0;
      }
      __syncthreads();
#pragma unroll 8
      for (int kk = 0; kk < 32; ++kk) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Rs[ty * 4 + i][kk];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Cs[tx + 16 * j][kk];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
      }
    }
    double cs[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + tx + 16 * j;
      if (c >= A.m) continue;
      const double vc = A.v[c];
      const double yc = A.yn[c];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = r0 + ty * 4 + i;
        if (r >= A.n) continue;
        double M = xr[i] + yc - 2.0 * acc[i][j];
        M = (M > 0.0 ? M : 0.0) * inv_c;
        const double p = exp(-M / A.reg + ur[i] + vc);
        rs[i] += p;
        cs[j] += p;
        cost = fma(p, M, cost);
        if (A.P) A.P[(int64_t)r * A.ldp + c] = p;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) cred[ty][tx + 16 * j] = cs[j];
    __syncthreads();
    if (tid < 64 && c0 + tid < A.m) {
      double t = 0.0;
#pragma unroll
      for (int g = 0; g < 16; ++g) t += cred[g][tid];
      atomicAdd(&A.colsum[c0 + tid], t);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) rs[i] += __shfl_xor(rs[i], o, 64);
    const int r = r0 + ty * 4 + i;
    if (tx == 0 && r < A.n) A.rowsum[r] = rs[i];
  }
  __shared__ double red[4];
  cost = wave_sum(cost);
  if ((tid & 63) == 0) red[tid >> 6] = cost;
  __syncthreads();
  if (tid == 0) A.costpart[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// sums = [cost, total mass, ||rowsum - a||_2, ||colsum - b||_2]
__global__ __launch_bounds__(256) void online_plan_final_kernel(
    int n, int m, int nblk, const double* rowsum, const double* colsum, const double* costpart,
    const double* a, const double* b, double* sums) {
  const int tid = threadIdx.x;
  double c = 0.0, mass = 0.0, ea = 0.0, eb = 0.0;
  for (int i = tid; i < nblk; i += 256) c += costpart[i];
  for (int i = tid; i < n; i += 256) {
    mass += rowsum[i];
    const double d = rowsum[i] - a[i];
    ea += d * d;
  }
  for (int j = tid; j < m; j += 256) {
    const double d = colsum[j] - b[j];
    eb += d * d;
  }
  __shared__ double red[4][4];
  c = wave_sum(c);
  mass = wave_sum(mass);
  ea = wave_sum(ea);
  eb = wave_sum(eb);
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = c;
    red[1][tid >> 6] = mass;
    red[2][tid >> 6] = ea;
    red[3][tid >> 6] = eb;
  }
  __syncthreads();
  if (tid == 0) {
    double s[4];
    for (int k = 0; k < 4; ++k) s[k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
    sums[0] = s[0];
    sums[1] = s[1];
    sums[2] = sqrt(s[2]);
    sums[3] = sqrt(s[3]);
  }
}

__global__ void online_fill_kernel(int64_t count, double val, double* y) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < count) y[i] = val;
}

// ---------------------------------------------------------------------------
// Host side.

int round_up(int x, int k) { return (x + k - 1) / k * k; }

int padded_dim(int prec, int d) {
  return prec == POTAMD_ONLINE_F64 ? d : round_up(d, 32);
}

bool supported(int prec, int d) {
  if (prec < POTAMD_ONLINE_F32 || prec > POTAMD_ONLINE_F64) return false;
  if (prec == POTAMD_ONLINE_F64) return d >= 1;
  return d >= 1 && padded_dim(prec, d) <= 128;
}

struct OnlineLayout {
  char* pkX;
  char* pkY;
  char* loX;
  char* loY;
  float* scX;
  float* scY;
  double* xn;
  double* yn;
  double* U[2];
  double* V[2];
  double* loga2;
  double* logb2;
  double* shX;
  double* shY;
  double* part;
  int* pidx;
  double* errpart;
  double* blockmax;
  double* cscale;
  Ctrl* ctrl;
  size_t bytes;
};

int max_splits(int n, int m) {
  // Splits of the C side for a pass: about 2048 workgroups, at least 4 stages each.
  const int wgs = round_up(n, kPad) / kRowsPerWG;
  const int steps = (m + kStep - 1) / kStep;
  int s = (2048 + wgs - 1) / wgs;
  s = std::min(s, std::max(1, steps / 4));
  return std::max(1, s);
}

OnlineLayout carve_online(void* ws, int prec, int n, int m, int d) {
  Carver c(ws);
  OnlineLayout L;
  memset(&L, 0, sizeof(L));
  // Operand rows hold the widest format (fp32, also used by the max pass).
  const int Dm = round_up(d, 32);
  const size_t row_bytes = Dm <= 128 ? (size_t)Dm * 4 : 0;
  const int D = padded_dim(prec, d);
  const int np = round_up(n, kPad), mp = round_up(m, kPad);
  L.pkX = c.take<char>((size_t)np * row_bytes);
  L.pkY = c.take<char>((size_t)mp * row_bytes);
  const bool lo = prec == POTAMD_ONLINE_BF16X3;
  L.loX = c.take<char>(lo ? (size_t)np * D * 2 : 0);
  L.loY = c.take<char>(lo ? (size_t)mp * D * 2 : 0);
  L.scX = c.take<float>(np);
  L.scY = c.take<float>(mp);
  L.xn = c.take<double>(np);
  L.yn = c.take<double>(mp);
  L.U[0] = c.take<double>(np);
  L.U[1] = c.take<double>(np);
  L.V[0] = c.take<double>(mp);
  L.V[1] = c.take<double>(mp);
  L.loga2 = c.take<double>(np);
  L.logb2 = c.take<double>(mp);
  L.shX = c.take<double>(np);
  L.shY = c.take<double>(mp);
  // Partials: up to 64 splits per side (split counts are capped below).
  const size_t parts = (size_t)64 * std::max(np, mp);
  L.part = c.take<double>(parts);
  L.pidx = c.take<int>(parts);
  L.errpart = c.take<double>((std::max(n, m) + kMergeRows - 1) / kMergeRows);
  L.blockmax = c.take<double>((std::max(n, m) + 255) / 256);
  L.cscale = c.take<double>(1);
  L.ctrl = c.take<Ctrl>(1);
  L.bytes = c.off;
  return L;
}

template <int P>
hipError_t launch_pack_t(const PackArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(online_pack_kernel<P>, dim3((a.rows_pad + 3) / 4), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_pack(int prec, const PackArgs& a, hipStream_t s) {
  switch (prec) {
    case POTAMD_ONLINE_F32: return launch_pack_t<POTAMD_ONLINE_F32>(a, s);
    case POTAMD_ONLINE_BF16: return launch_pack_t<POTAMD_ONLINE_BF16>(a, s);
    case POTAMD_ONLINE_BF16X3: return launch_pack_t<POTAMD_ONLINE_BF16X3>(a, s);
    case POTAMD_ONLINE_FP8: return launch_pack_t<POTAMD_ONLINE_FP8>(a, s);
    default: return launch_pack_t<POTAMD_ONLINE_F64>(a, s);
  }
}

template <int P, int MODE>
hipError_t launch_pass_d(int D, const PassArgs& a, int S, hipStream_t s) {
  dim3 grid(a.nr_pad / kRowsPerWG, S);
  switch (D) {
    case 32: hipLaunchKernelGGL((online_pass_kernel<P, 32, MODE>), grid, dim3(256), 0, s, a); break;
    case 64: hipLaunchKernelGGL((online_pass_kernel<P, 64, MODE>), grid, dim3(256), 0, s, a); break;
    case 96: hipLaunchKernelGGL((online_pass_kernel<P, 96, MODE>), grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((online_pass_kernel<P, 128, MODE>), grid, dim3(256), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_pass(int prec, int D, const PassArgs& a, int S, hipStream_t s) {
  switch (prec) {
    case POTAMD_ONLINE_F32: return launch_pass_d<POTAMD_ONLINE_F32, kModeLse>(D, a, S, s);
    case POTAMD_ONLINE_BF16: return launch_pass_d<POTAMD_ONLINE_BF16, kModeLse>(D, a, S, s);
    case POTAMD_ONLINE_BF16X3: return launch_pass_d<POTAMD_ONLINE_BF16X3, kModeLse>(D, a, S, s);
    default: return launch_pass_d<POTAMD_ONLINE_FP8, kModeLse>(D, a, S, s);
  }
}

struct Side {
  const void* P;    // points (fp32, or fp64 when x64)
  int x64;
  int64_t ld;
  int count, pad;
  char* pk;
  char* lo;
  float* sc;
  double* norm;
  double* shift;
  const double* w;
  double* logw2;
};

// One half-iteration: update the potential of side R from side C.
hipError_t half_step(int prec, int D, int d, const Side& R, const Side& C, const double* c_pot,
                     const OnlineLayout& L, double reg, int S, OnlineMergeArgs& ma,
                     hipStream_t s) {
  const int nsteps = (C.count + kStep - 1) / kStep;
  const int sps = (nsteps + S - 1) / S;
  const int Seff = (nsteps + sps - 1) / sps;
  if (prec == POTAMD_ONLINE_F64) {
    F64PassArgs fa;
    fa.x64 = R.x64;
    fa.R = R.P;
    fa.ldr = R.ld;
    fa.C = C.P;
    fa.ldc = C.ld;
    fa.nr = R.count;
    fa.nc = C.count;
    fa.d = d;
    fa.c_pot = c_pot;
    fa.c_norm = C.norm;
    fa.cscale = L.cscale;
    fa.reg = reg;
    fa.steps_per_split = sps;
    fa.part = L.part;
    fa.nr_pad = R.pad;
    fa.ctrl = L.ctrl;
    hipLaunchKernelGGL(online_f64_pass_kernel, dim3(R.pad / 64, Seff), dim3(256), 0, s, fa);
  } else {
    PassArgs pa;
    pa.r_pk = R.pk;
    pa.r_lo = R.lo;
    pa.r_sc = R.sc;
    pa.c_pk = C.pk;
    pa.c_lo = C.lo;
    pa.c_sc = C.sc;
    pa.c_pot = c_pot;
    pa.c_norm = C.norm;
    pa.r_shift = R.shift;
    pa.nr = R.count;
    pa.nc = C.count;
    pa.steps_per_split = sps;
    pa.cscale = L.cscale;
    pa.reg = reg;
    pa.part = L.part;
    pa.part_idx = nullptr;
    pa.nr_pad = R.pad;
    pa.ctrl = L.ctrl;
    hipError_t e = launch_pass(prec, D, pa, Seff, s);
    if (e != hipSuccess) return e;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  ma.part = L.part;
  ma.S = Seff;
  ma.nr = R.count;
  ma.nr_pad = R.pad;
  ma.w = R.w;
  ma.logw2 = R.logw2;
  ma.norm = R.norm;
  ma.shift = R.shift;
  ma.cscale = L.cscale;
  ma.reg = reg;
  ma.errpart = L.errpart;
  ma.ctrl = L.ctrl;
  const int nblk = (R.count + kMergeRows - 1) / kMergeRows;
  hipLaunchKernelGGL(online_merge_kernel, dim3(nblk), dim3(kMergeRows), 0, s, ma);
  if (ma.decide) hipLaunchKernelGGL(online_decide_kernel, dim3(1), dim3(256), 0, s, ma, nblk);
  return hipGetLastError();
}

hipError_t fill(int64_t count, double val, double* y, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(online_fill_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                     count, val, y);
  return hipGetLastError();
}

}  // namespace
}  // namespace potamd

using namespace potamd;

extern "C" size_t potamd_online_workspace_size(int precision, int n, int m, int d) {
  if (n < 1 || m < 1 || !supported(precision, d)) return 0;
  return carve_online(nullptr, precision, n, m, d).bytes;
}

extern "C" int potamd_sinkhorn_online(int precision, int dtype, const void* X, int64_t ldx,
                                      const void* Y, int64_t ldy, int n, int m, int d, double reg,
                                      double cost_scale, const double* a, const double* b,
                                      double* u, double* v, int max_iter, double stop_thr,
                                      int semantics, double* err_hist, int err_cap, int splits,
                                      void* workspace,
                                      size_t ws_bytes, void* host_ctrl, void* stream, int* info,
                                      double* cost_scale_out) {
  if (!X || !Y || !a || !b || !u || !v || !info || n < 1 || m < 1 || d < 1 || ldx < d ||
      ldy < d || !(reg > 0.0) || cost_scale < 0.0 || max_iter < 1 || splits < 0 || splits > 64 ||
      (semantics != 0 && semantics != 1) || (err_cap > 0 && !err_hist) ||
      (dtype != POTAMD_F32 && dtype != POTAMD_F64))
    return POTAMD_EINVAL;
  if (!supported(precision, d)) return POTAMD_EUNSUPPORTED;
  // fp64 points feed the fp64 pass only (the MFMA operands are packed from fp32)
  if (dtype == POTAMD_F64 && precision != POTAMD_ONLINE_F64) return POTAMD_EUNSUPPORTED;
  const int x64 = dtype == POTAMD_F64;
  const int prec = precision;
  const int D = padded_dim(prec, d);
  OnlineLayout L = carve_online(workspace, prec, n, m, d);
  if (!workspace || ws_bytes < L.bytes) return POTAMD_EWORKSPACE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int np = round_up(n, kPad), mp = round_up(m, kPad);
  const double nan = __builtin_nan("");

  HIP_TRY(launch_ctrl_init(L.ctrl, nullptr, s));
  HIP_TRY(launch_log2(n, a, L.loga2, s));
  HIP_TRY(launch_log2(m, b, L.logb2, s));
  HIP_TRY(launch_scale_copy(n, u, kLog2e, L.U[0], s));
  HIP_TRY(fill(np, nan, L.shX, s));
  HIP_TRY(fill(mp, nan, L.shY, s));

  Side SX{X, x64, ldx, n, np, L.pkX, L.loX, L.scX, L.xn, L.shX, a, L.loga2};
  Side SY{Y, x64, ldy, m, mp, L.pkY, L.loY, L.scY, L.yn, L.shY, b, L.logb2};

  PackArgs pk;
  memset(&pk, 0, sizeof(pk));
  pk.cscale = L.cscale;
  pk.reg = reg;
  pk.x64 = x64;
  auto pack_side = [&](int p, const Side& S, int scaled) {
    pk.X = S.P;
    pk.ldx = S.ld;
    pk.rows = S.count;
    pk.d = d;
    pk.D = D;
    pk.rows_pad = S.pad;
    pk.scaled = scaled;
    pk.pk = S.pk;
    pk.pk_lo = S.lo;
    pk.pk_scale = S.sc;
    pk.norms = S.norm;
    return launch_pack(p, pk, s);
  };

  // Cost normalisation.
  if (cost_scale > 0.0) {
    HIP_TRY(fill(1, cost_scale, L.cscale, s));
  } else {
    const int Dm = round_up(d, 32);
    if (Dm > 128) return POTAMD_EUNSUPPORTED;
    pk.D = Dm;
    {
      pk.X = X; pk.ldx = ldx; pk.rows = n; pk.d = d; pk.D = Dm; pk.rows_pad = np; pk.scaled = 0;
      pk.pk = L.pkX; pk.pk_lo = nullptr; pk.pk_scale = L.scX; pk.norms = L.xn;
      HIP_TRY(launch_pack(POTAMD_ONLINE_F32, pk, s));
      pk.X = Y; pk.ldx = ldy; pk.rows = m; pk.rows_pad = mp;
      pk.pk = L.pkY; pk.pk_scale = L.scY; pk.norms = L.yn;
      HIP_TRY(launch_pack(POTAMD_ONLINE_F32, pk, s));
    }
    const int Sm = std::min(64, max_splits(n, m));
    const int nsteps = (m + kStep - 1) / kStep;
    const int sps = (nsteps + Sm - 1) / Sm;
    const int Seff = (nsteps + sps - 1) / sps;
    PassArgs pa;
    memset(&pa, 0, sizeof(pa));
    pa.r_pk = L.pkX;
    pa.c_pk = L.pkY;
    pa.c_norm = L.yn;
    pa.nr = n;
    pa.nc = m;
    pa.steps_per_split = sps;
    pa.cscale = L.cscale;
    pa.reg = reg;
    pa.part = L.part;
    pa.part_idx = L.pidx;
    pa.nr_pad = np;
    pa.ctrl = nullptr;
    HIP_TRY((launch_pass_d<POTAMD_ONLINE_F32, kModeMax>(Dm, pa, Seff, s)));
    MaxReduceArgs mr;
    mr.part = L.part;
    mr.part_idx = L.pidx;
    mr.S = Seff;
    mr.nr = n;
    mr.nr_pad = np;
    mr.x64 = x64;
    mr.X = X;
    mr.ldx = ldx;
    mr.Y = Y;
    mr.ldy = ldy;
    mr.d = d;
    mr.blockmax = L.blockmax;
    mr.ctrl = L.ctrl;
    mr.cscale = L.cscale;
    hipLaunchKernelGGL(online_max_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, s, mr);
    HIP_TRY(hipGetLastError());
  }
  // Operands of the chosen precision (norms again, identical).
  HIP_TRY(pack_side(prec, SX, 1));
  HIP_TRY(pack_side(prec, SY, 1));

  const int Sx = splits > 0 ? splits : std::min(64, max_splits(n, m));  // row update of x
  const int Sy = splits > 0 ? splits : std::min(64, max_splits(m, n));  // column update

  OnlineMergeArgs ma;
  memset(&ma, 0, sizeof(ma));
  ma.max_iter = max_iter;
  ma.stop_thr = stop_thr;
  ma.err_hist = err_hist;
  ma.err_cap = err_cap;
  ma.l1 = semantics;
  ma.inclusive = semantics;

  // Prologue: v_1 = log b - LSE_i(Mr + u_0).
  ma.check = 0;
  ma.decide = 0;
  ma.iter = -1;
  ma.pot_cur = L.V[0];
  ma.pot_new = L.V[1];
  HIP_TRY(half_step(prec, D, d, SY, SX, L.U[0], L, reg, Sy, ma, s));

  constexpr int kRing = 2;  // stop checks in flight: the host queues at most 2 past the device
  Ctrl* hc = static_cast<Ctrl*>(host_ctrl);
  Ctrl local;
  hipEvent_t ev[kRing];
  bool ev_used[kRing] = {};
  for (int k = 0; k < kRing; ++k) HIP_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
  int ring_pos = 0;
  int ring_head = 0;  // oldest check whose copy the host has not read
  bool stopped = false;
  int rc = 0;
  for (int ii = 0; ii < max_iter && !stopped; ++ii) {
    // Checks the device has already passed end the loop without a host wait.
    while (hc && ring_head < ring_pos) {
      const int k = ring_head % kRing;
      const hipError_t q = hipEventQuery(ev[k]);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) {
        rc = POTAMD_EHIP;
        break;
      }
      ev_used[k] = false;
      ++ring_head;
      if (hc[k].done) stopped = true;
    }
    if (rc || stopped) break;
    // u update from v_cur = V[(ii+1)&1]
    ma.check = 0;
    ma.decide = 0;
    ma.iter = ii;
    ma.pot_cur = L.U[ii & 1];
    ma.pot_new = L.U[(ii + 1) & 1];
    if (half_step(prec, D, d, SX, SY, L.V[(ii + 1) & 1], L, reg, Sx, ma, s) != hipSuccess) {
      rc = POTAMD_EHIP;
      break;
    }
    // v update from the new u; its column sums give the check of iteration ii
    // (sinkhorn_log checks at ii % 10 == 0, empirical_sinkhorn at
    // (ii + 1) % 10 == 0).
    const bool check = semantics ? (ii + 1) % 10 == 0 : ii % 10 == 0;
    ma.check = check;
    ma.decide = check || ii == max_iter - 1;
    ma.pot_cur = L.V[(ii + 1) & 1];
    ma.pot_new = L.V[ii & 1];
    if (half_step(prec, D, d, SY, SX, L.U[(ii + 1) & 1], L, reg, Sy, ma, s) != hipSuccess) {
      rc = POTAMD_EHIP;
      break;
    }
    if (check && ii + 1 < max_iter) {
      const int k = ring_pos % kRing;
      if (ev_used[k]) {
        if (hipEventSynchronize(ev[k]) != hipSuccess) { rc = POTAMD_EHIP; break; }
        ev_used[k] = false;
        ++ring_head;
        const Ctrl* snap = hc ? &hc[k] : &local;
        if (snap->done) stopped = true;
      }
      if (!stopped) {
        if (hc) {
          if (hipMemcpyAsync(&hc[k], L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipEventRecord(ev[k], s) != hipSuccess) {
            rc = POTAMD_EHIP;
            break;
          }
          ev_used[k] = true;
        } else {
          if (hipMemcpyAsync(&local, L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess) {
            rc = POTAMD_EHIP;
            break;
          }
          if (local.done) stopped = true;
        }
        ++ring_pos;
      }
    }
  }
  for (int k = 0; k < kRing; ++k) hipEventDestroy(ev[k]);
  if (rc) return rc;
  double cs_host = 0.0;
  HIP_TRY(hipMemcpyAsync(&local, L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&cs_host, L.cscale, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  int status = local.status;
  int niter = local.niter;
  if (!local.done) {
    status = POTAMD_MAX_ITER;
    niter = max_iter - 1;
  }
  const int f = (niter + 1) & 1;
  HIP_TRY(launch_scale_copy(n, L.U[f], kLn2, u, s));
  HIP_TRY(launch_scale_copy(m, L.V[f], kLn2, v, s));
  if (cost_scale_out) *cost_scale_out = cs_host;
  info[0] = niter;
  info[1] = local.n_err;
  info[2] = status;
  info[3] = 0;
  return status;
}

extern "C" size_t potamd_online_plan_workspace_size(int n, int m) {
  if (n < 1 || m < 1) return 0;
  Carver c(nullptr);
  c.take<double>(n);                 // xn
  c.take<double>(m);                 // yn
  c.take<double>(n);                 // rowsum
  c.take<double>(m);                 // colsum
  c.take<double>((n + 63) / 64);     // costpart
  c.take<double>(1);                 // cscale
  return c.off;
}

extern "C" int potamd_online_plan(int dtype, const void* X, int64_t ldx, const void* Y, int64_t ldy, int n,
                                  int m, int d, double reg, double cost_scale, const double* u,
                                  const double* v, const double* a, const double* b, double* P,
                                  int64_t ldp, double* sums, void* workspace, size_t ws_bytes,
                                  void* stream) {
  if (!X || !Y || !u || !v || !a || !b || !sums || n < 1 || m < 1 || d < 1 || ldx < d ||
      ldy < d || !(reg > 0.0) || !(cost_scale > 0.0) || (P && ldp < m) ||
      (dtype != POTAMD_F32 && dtype != POTAMD_F64))
    return POTAMD_EINVAL;
  const int x64 = dtype == POTAMD_F64;
  Carver c(workspace);
  double* xn = c.take<double>(n);
  double* yn = c.take<double>(m);
  double* rowsum = c.take<double>(n);
  double* colsum = c.take<double>(m);
  double* costpart = c.take<double>((n + 63) / 64);
  double* cscale = c.take<double>(1);
  if (!workspace || ws_bytes < c.off) return POTAMD_EWORKSPACE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(fill(1, cost_scale, cscale, s));
  HIP_TRY(fill(m, 0.0, colsum, s));
  PackArgs pk;
  memset(&pk, 0, sizeof(pk));
  pk.d = d;
  pk.D = d;
  pk.x64 = x64;
  pk.X = X;
  pk.ldx = ldx;
  pk.rows = n;
  pk.rows_pad = n;
  pk.norms = xn;
  HIP_TRY(launch_pack(POTAMD_ONLINE_F64, pk, s));
  pk.X = Y;
  pk.ldx = ldy;
  pk.rows = m;
  pk.rows_pad = m;
  pk.norms = yn;
  HIP_TRY(launch_pack(POTAMD_ONLINE_F64, pk, s));
  OnlinePlanArgs pa;
  pa.x64 = x64;
  pa.X = X;
  pa.ldx = ldx;
  pa.Y = Y;
  pa.ldy = ldy;
  pa.n = n;
  pa.m = m;
  pa.d = d;
  pa.reg = reg;
  pa.cscale = cscale;
  pa.u = u;
  pa.v = v;
  pa.xn = xn;
  pa.yn = yn;
  pa.P = P;
  pa.ldp = ldp;
  pa.rowsum = rowsum;
  pa.colsum = colsum;
  pa.costpart = costpart;
  const int nblk = (n + 63) / 64;
  hipLaunchKernelGGL(online_plan_kernel, dim3(nblk), dim3(256), 0, s, pa);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(online_plan_final_kernel, dim3(1), dim3(256), 0, s, n, m, nblk, rowsum,
                     colsum, costpart, a, b, sums);
  HIP_TRY(hipGetLastError());
  return POTAMD_CONVERGED;
}
