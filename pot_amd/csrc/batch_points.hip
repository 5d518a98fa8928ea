// Ragged batch of sinkhorn_log problems given as point clouds, with the
// squared Euclidean cost computed on the fly (never stored).
//
// Each workgroup solves one problem (n, m <= 1024, d <= 16) with the
// reference's sinkhorn_log loop (ot/bregman/_sinkhorn.py:880-906: v first,
// then u, the column-marginal check every `check_every` iterations, strict
// `err < stopThr`, niter = ii), exactly as batch_log_kernel does over a
// materialised cost, on the cost M / max M (ot.dist then M / M.max(), the
// reference's normalisation) or on M itself.
//
// Exponent of an element in base 2, with s = log2(e) / (reg * cmax) and the
// points centred on the midpoint of the two centroids (M is unchanged):
//     x_rc - shift_r = X_r . Y_c + (v_c - s|y_c|^2) - (shift_r + s|x_r|^2)
// where X = sqrt(2 s) x and Y = sqrt(2 s) y.  The dot product runs on
// 32x32x16 f16 MFMA with each coordinate split into f16 hi + lo (three
// products hi.hi + hi.lo + lo.hi, the lo.lo term is below fp32 rounding).
// The two potential terms are split into four bf16 parts each and added by
// one more (bf16) MFMA whose K slots hold [P1..P4, 1, 1, 1, 1] against
// [1, 1, 1, 1, -Q1..-Q4]; issued last, it cancels the large dot product
// inside the accumulator, so the fp32 result is the small exponent itself.
//
// One sweep per iteration, as in the materialised kernels: the row sums use
// the previous row log-sum-exp as shift (one exp2 per element), and the
// column sums of P(u_new, v) come from the same registers weighted by
// a_r / S_r.  Layout: the 8 waves own the column tiles (wave w owns tiles
// w, w+8, w+16, w+24 of 32 columns; its C operands stay in registers for the
// whole solve); all waves step through the rows 32 at a time, reading the
// row operands from LDS.  A row sum is complete once the 8 waves' partials
// meet in LDS (one barrier per block of 32 rows); the weighted column sums
// accumulate per lane over all row blocks and are reduced across lanes once
// per sweep.  A block whose row sums leave [2^-60, 2^60] is redone with its
// rows re-centred on their maximum, as in the materialised kernels.
//
// cmax (normalised cost) is found on the device: a max pass over all pairs
// on the same MFMA path picks each lane's largest pair, whose cost is then
// evaluated exactly in fp64 from the points.
#include <math.h>
#include <string.h>

#include <type_traits>

#include "../../include/pot_amd.h"
#include "host.h"
#include "internal.h"

namespace potamd {
namespace {

constexpr int kPWDefault = 8;                // wavefronts per problem (8 or 16)
constexpr int kPMax = 1024;                  // largest n and m
constexpr int kPD = 16;                      // coordinates per point (zero padded)
constexpr int kPBlocks = kPMax / 32;         // row blocks of 32
constexpr float kMaxOperand = 16384.f;       // largest |X|, |Y| accepted (f16 range)
constexpr unsigned kOnes = 0x3F803F80u;      // two bf16 1.0

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct PointsArgs {
  const float* X;        // packed points, problem p's rows at xoff[p] (row stride d)
  const float* Y;
  const int64_t* xoff;   // [B] also the offset of problem p in a and u
  const int64_t* yoff;   // [B] also the offset in b and v
  const int* ns;
  const int* ms;
  int d;
  int normalize;         // 0 raw cost, 1 cost / max cost (max found here), 2 cost / cmax[p]
  int max_only;          // 1: only find the cost maxima (into cmax)
  const double* regs;    // [B]
  const double* a;
  const double* b;
  double* u;             // in: u0 (natural log), out: log u
  double* v;             // out: log v
  const int* order;
  int max_iter;
  int check_every;
  double stop_thr;
  double fallback_thr;
  int* niter;
  int* status;
  double* err_last;
  double* err_first;
  double* cmax;          // [B] the cost maximum used (out; in when normalize == 2); may be null
  int batch;             // B
  int* queue;            // [2] work-queue counters (zero at launch, zero at exit) or null
  int items;             // entries of order: B, or more when problems are split
  const int* split;      // [B] 0 (whole problem) or parts | slot << 4; may be null
  char* split_ws;        // exchange slots of the split problems (kSplitSlot bytes each)
  int nsplit;            // number of slots
};

// Split problems: the parts of problem p (2 to 8 workgroups, each owning a
// contiguous range of row blocks) combine their column reductions through
// one exchange slot: a counter, then double-buffered per-part columns.
constexpr int kSplitMax = 8;
constexpr size_t kSplitCols = 256;                                    // byte offset
constexpr size_t kSplitDbl = kSplitCols + 2ull * kSplitMax * kPMax * 4;  // byte offset
constexpr size_t kSplitSlot = kSplitDbl + 256;
constexpr int kSplitSpin = 1 << 22;  // polls (with a sleep each) before giving up
static_assert(kSplitSlot == kPointsSplitSlot && kSplitCols == kPointsSplitCounters,
              "internal.h states the split slot layout");

__device__ __forceinline__ unsigned bf16_bits(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// x = p0 + p1 + p2 + p3 to 2^-32 relative, each part a bf16, packed two per
// word as (p0 | p1 << 16, p2 | p3 << 16); an infinite or NaN x gives
// (x, 0, 0, 0).  Straight-line code with selects (no local array, no early
// exit), so the parts stay in registers.
__device__ __forceinline__ void split4_bf16(double x, unsigned& w01, unsigned& w23) {
  const bool fin = __builtin_isfinite(x);
  double r = fin ? x : 0.0;
  const unsigned p0 = bf16_bits((float)r);
  r -= (double)__uint_as_float(p0 << 16);
  const unsigned p1 = bf16_bits((float)r);
  r -= (double)__uint_as_float(p1 << 16);
  const unsigned p2 = bf16_bits((float)r);
  r -= (double)__uint_as_float(p2 << 16);
  const unsigned p3 = bf16_bits((float)r);
  w01 = fin ? (p0 | (p1 << 16)) : bf16_bits((float)x);
  w23 = fin ? (p2 | (p3 << 16)) : 0u;
}

// K slots of the potential MFMA.  Column side (A operand, lanes 0-31):
// [P1, P2, P3, P4, 1, 1, 1, 1]; row side (B operand): [1, 1, 1, 1, -Q1..-Q4].
__device__ __forceinline__ u32x4 kslot_col(double P) {
  u32x4 k;
  unsigned w01, w23;
  split4_bf16(P, w01, w23);
  k[0] = w01;
  k[1] = w23;
  k[2] = kOnes;
  k[3] = kOnes;
  return k;
}
__device__ __forceinline__ u32x4 kslot_row(double Q) {
  u32x4 k;
  unsigned w01, w23;
  split4_bf16(-Q, w01, w23);
  k[0] = kOnes;
  k[1] = kOnes;
  k[2] = w01;
  k[3] = w23;
  return k;
}

typedef float f32x4v __attribute__((ext_vector_type(4)));

// Coordinates 8*hh .. 8*hh+7 of the point at p (those past d read
// coordinate 0 and are not used), loaded ahead of point_ops so that all the
// loads of a pass are issued before the first is used.  With V4 (d a
// multiple of 4 and 16-byte aligned clouds) as two 16-byte loads; groups of
// four past d read coordinates 0-3.
template <bool V4>
__device__ __forceinline__ void half_load(const float* p, int d, int hh, float (&v)[8]) {
  if constexpr (V4) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int k0 = 8 * hh + 4 * g;
      const f32x4v q = *reinterpret_cast<const f32x4v*>(p + (k0 < d ? k0 : 0));
#pragma unroll
      for (int e = 0; e < 4; ++e) v[4 * g + e] = q[e];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int kk = 8 * hh + k;
      v[k] = p[kk < d ? kk : 0];
    }
  }
}
// float -> f16 on v_cvt_f16_f32 (round to nearest even).  The value passes
// through an empty asm so that the compiler cannot merge a preceding
// double -> float conversion into one double -> f16 conversion, which it
// would emulate in about 40 integer instructions.
__device__ __forceinline__ _Float16 f16_of(float x) {
  asm("" : "+v"(x));
  return (_Float16)x;
}
// Half a point (coordinates 8*hh .. 8*hh+7, loaded by half_load) as f16 hi
// and lo operands of (p - cen) * sc, with cn the same coordinates of the
// centre; nrm accumulates the squared norm of the represented (hi + lo)
// coordinates, mx the largest |coordinate|.  Straight-line code: coordinates
// past d and invalid points give 0 by a select.
__device__ __forceinline__ void point_ops(const float (&pv)[8], bool valid, int d,
                                          const double (&cn)[8], double sc, int hh, u32x4& hi,
                                          u32x4& lo, double& nrm, float& mx) {
  unsigned hb[8], lb[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int kk = 8 * hh + k;
    const double x = (valid && kk < d) ? ((double)pv[k] - cn[k]) * sc : 0.0;
    const _Float16 xh = f16_of((float)x);
    const double xhd = (double)(float)xh;
    const _Float16 xl = f16_of((float)(x - xhd));
    const double xr = xhd + (double)(float)xl;
    nrm = fma(xr, xr, nrm);
    mx = fmaxf(mx, fabsf((float)x));
    hb[k] = __builtin_bit_cast(unsigned short, xh);
    lb[k] = __builtin_bit_cast(unsigned short, xl);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    hi[k] = hb[2 * k] | (hb[2 * k + 1] << 16);
    lo[k] = lb[2 * k] | (lb[2 * k + 1] << 16);
  }
}

// This lane's index in its wave, recomputed at the point of use: the asm
// result cannot be kept live from an earlier computation, so addresses built
// from it in the once-per-sweep sections are rebuilt there (two VALU
// instructions) instead of being spilled across the block loop and reloaded
// from scratch.
__device__ __forceinline__ int lane_now() {
  int l;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
  return l;
}

// Values that every lane of the workgroup holds (read back from LDS after a
// barrier), made visibly uniform to the compiler so that the control flow
// they steer -- around barriers -- is uniform.
__device__ __forceinline__ int uniform_i(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uniform_f(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
template <typename T>
__device__ __forceinline__ T* uniform_p(T* x) {
  const uintptr_t b = reinterpret_cast<uintptr_t>(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return reinterpret_cast<T*>(((uintptr_t)hi << 32) | lo);
}
__device__ __forceinline__ double uniform_d(double x) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ f32x16 mfma_h(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h16x8, a),
                                                __builtin_bit_cast(h16x8, b), c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma_b(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// x[l] + x[l ^ 32] in every lane (the two wave halves hold the same row of
// an MFMA tile), through v_permlane32_swap; the operands are added in the
// same order in both halves, so the result is identical there.
__device__ __forceinline__ float halves_sum(float x) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false,
                                                  false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float halves_max(float x) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false,
                                                  false);
  return fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

template <int PW>
struct PointsLds {
  u32x4 rop[kPBlocks][2][kWave];  // row operands: [block][hi, lo][lane]
  u32x4 rks[kPMax];               // row K slots
  double c2s[kPMax];              // row shifts log2 a - u2 (base 2)
  double la[kPMax];               // log2 a_r (+inf for rows of zero weight)
  double xn[kPMax];               // s |x_r|^2
  double vsh[kPMax];              // column potentials (base 2)
  double yn[kPMax];               // s |y_c|^2
  double lbs[kPMax];              // log2 b_c (-inf where b_c = 0)
  double bs[kPMax];               // b_c (read by the marginal check of every iteration)
  float af[kPMax];                // row weights
  float sh[kPMax];                // shifted row sums (prologue: column maxima)
  float colx[kPMax];              // column sums
  // Row-sum partials [buffer][row][wave]: a row's PW partials are adjacent,
  // so a lane reads them with 16-byte loads in one LDS round trip.  Rows are
  // PW + 4 floats apart (still 16-byte aligned), which spreads the 32 lanes'
  // stores of one wave over twice as many banks as a stride of PW.
  __attribute__((aligned(16))) float rowpart[2][32][PW + 4];
  float rowmax[PW][32];
  double cen[kPD];
  double red[PW];
  float redf[PW];
  int flag;
};

// Everything after the centring, for problems whose column tiles take NS
// slots per wave (NS = ceil(ceil(m / 32) / 8)); slots past m hold padding
// columns whose potential is -inf, so they add nothing.
template <int PW, int NS, bool SPLIT>
__device__ __forceinline__ void solve_points(const PointsArgs& A, PointsLds<PW>& L, int p, int n,
                                             int m, double ext, int part, int K, char* sws) {
  constexpr int PT = PW * kWave;
  // (The wave index is uniform: in a scalar register, the addresses and row
  // ranges built from it need no vector registers across the loops.)
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = uniform_i(tid / kWave);
  const int h = lane >> 5, l32 = lane & 31;
  const int d = A.d;
  const int64_t xo = A.xoff[p], yo = A.yoff[p];
  const float* Xp = A.X + xo * d;
  const float* Yp = A.Y + yo * d;
  // 16-byte point loads: d a multiple of 4 and both clouds 16-byte aligned
  // (every row then starts on a 16-byte boundary).
  const bool v4 = (d & 3) == 0 && ((reinterpret_cast<uintptr_t>(A.X) |
                                    reinterpret_cast<uintptr_t>(A.Y)) & 15) == 0;
  const double* bp = A.b + yo;
  double* up = A.u + xo;
  double* vp = A.v + yo;
  const int nblk = (n + 31) / 32;
  // Row blocks of this workgroup: all of them, or part `part` of K.
  const int blk0 = SPLIT ? part * nblk / K : 0;
  const int blk1 = SPLIT ? (part + 1) * nblk / K : nblk;
  // Split problems: wait until the K parts reach the same exchange (a
  // counter in the slot; false after a bounded wait), and combine a
  // per-column partial (max or sum, parts added in part order, so every part
  // gets the same values) or a double maximum.
  // The exchanged values are written and read with device-scope atomic
  // accesses (which bypass the caches that are not coherent across the
  // GPU), so the counter needs no cache write-back or invalidation around
  // it: each thread waits for its own stores to complete before the
  // workgroup barrier, and the counter is incremented after that barrier.
  int sgen = 0;
  auto group_sync = [&]() -> bool {
    __builtin_amdgcn_s_waitcnt(0);  // this thread's exchange stores have completed
    __syncthreads();
    if (tid == 0) {
      int* cnt = reinterpret_cast<int*>(sws);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int target = K * (sgen + 1);
      int ok = 1;
      for (int spin = 0;
           __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; ++spin) {
        if (spin >= kSplitSpin) {
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      L.flag = ok;
    }
    __syncthreads();
    ++sgen;
    return uniform_i(L.flag) != 0;
  };
  auto put = [](float* q, float x) {
    __hip_atomic_store(q, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  auto get = [](const float* q) {
    return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  auto exchange_cols = [&](float* col, bool is_max) -> bool {
    float* xf = reinterpret_cast<float*>(sws + kSplitCols);
    const int buf = sgen & 1;
    for (int c = tid; c < m; c += PT) put(&xf[(buf * kSplitMax + part) * kPMax + c], col[c]);
    if (!group_sync()) return false;
    for (int c = tid; c < m; c += PT) {
      float v = get(&xf[(buf * kSplitMax) * kPMax + c]);
      for (int k = 1; k < K; ++k) {
        const float x = get(&xf[(buf * kSplitMax + k) * kPMax + c]);
        v = is_max ? fmaxf(v, x) : v + x;
      }
      col[c] = v;
    }
    __syncthreads();
    return true;
  };
  auto exchange_max = [&](double& x) -> bool {
    double* xd = reinterpret_cast<double*>(sws + kSplitDbl);
    const int buf = sgen & 1;
    if (tid == 0)
      __hip_atomic_store(&xd[buf * kSplitMax + part], x, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (!group_sync()) return false;
    double v = __hip_atomic_load(&xd[buf * kSplitMax], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
    for (int k = 1; k < K; ++k)
      v = fmax(v, __hip_atomic_load(&xd[buf * kSplitMax + k], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT));
    x = uniform_d(v);
    __syncthreads();
    return true;
  };
  // A part that timed out reports the problem as not solved.
  auto abandon = [&]() {
    if (tid == 0 && part == 0) {
      A.niter[p] = 0;
      A.status[p] = kNumerical;
    }
  };
  // Tile j of this wave covers columns 32 * (wave + 8 j) + [0, 32).
  auto col_of = [&](int j) { return 32 * (wave + PW * j) + l32; };
  // The same column, and the half of the wave, from the recomputed lane index
  // (for the sections that run once per sweep).
  auto col_now = [&](int j) { return 32 * (wave + PW * j) + (lane_now() & 31); };
  auto half_now = [&]() { return lane_now() >> 5; };
  // Column of value i of tile j in the MFMA output layout.
  auto out_col = [&](int j, int i) {
    return 32 * (wave + PW * j) + 8 * (i / 4) + 4 * h + (i % 4);
  };

  // Operand packing: rows into rop / xn, this wave's columns into registers / yn.
  u32x4 ahi[NS], alo[NS], ak[NS];
  // Every coordinate a thread packs is loaded first (rows past n read row 0;
  // the results are not used), so the loads of a pass overlap instead of
  // waiting one point at a time.
  constexpr int RQ = (kPBlocks * kWave + PT - 1) / PT;  // row half-points per thread
  constexpr int RR = (kPMax + PT - 1) / PT;             // rows per thread
  auto pack = [&](double sc) -> float {
    float mx = 0.f;
    float rv[RQ][8], cv[NS][8];
    auto load_all = [&](auto v4) {
      constexpr bool V4 = decltype(v4)::value;
#pragma unroll
      for (int u = 0; u < RQ; ++u) {
        const int q = tid + u * PT, r = (q >> 6) * 32 + (q & 31);
        half_load<V4>(Xp + (int64_t)(r < n ? r : 0) * d, d, (q >> 5) & 1, rv[u]);
      }
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const int c = col_of(j);
        half_load<V4>(Yp + (int64_t)(c < m ? c : 0) * d, d, h, cv[j]);
      }
    };
    if (v4)
      load_all(std::true_type());
    else
      load_all(std::false_type());
    // A lane packs only its half of the coordinates (hh = h for the rows too,
    // since PT is a multiple of the wave size), so it needs 8 of the centre's.
    double cn[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) cn[k] = L.cen[8 * h + k];
#pragma unroll
    for (int u = 0; u < RQ; ++u) {
      const int q = tid + u * PT;
      if (q >= nblk * kWave) break;
      const int blk = q >> 6, ln = q & 63, r = blk * 32 + (ln & 31), hh = ln >> 5;
      u32x4 hi4, lo4;
      double nh = 0.0;
      point_ops(rv[u], r < n, d, cn, sc, hh, hi4, lo4, nh, mx);
      L.rop[blk][0][ln] = hi4;
      L.rop[blk][1][ln] = lo4;
      nh += __shfl_xor(nh, 32, kWave);
      if (hh == 0) L.xn[r] = r < n ? 0.5 * nh : 0.0;
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int c = col_of(j);
      u32x4 hi4, lo4;
      double nh = 0.0;
      point_ops(cv[j], c < m, d, cn, sc, h, hi4, lo4, nh, mx);
      ahi[j] = hi4;
      alo[j] = lo4;
      nh += __shfl_xor(nh, 32, kWave);
      if (h == 0) L.yn[c] = c < m ? 0.5 * nh : 0.0;
    }
    return mx;
  };
  const u32x4 zero4 = {0u, 0u, 0u, 0u};
  auto set_col_kslots = [&](auto P_of) {
#pragma unroll
    for (int j = 0; j < NS; ++j) ak[j] = h == 0 ? kslot_col(P_of(col_of(j))) : zero4;
  };
  // D[c][r] tiles of this wave's columns against row block blk; the small
  // products first, the potential last.
  f32x16 e[NS];
  auto tiles_ops = [&](u32x4 bh, u32x4 bl, u32x4 bk) {
    const f32x16 z = {};
#pragma unroll
    for (int j = 0; j < NS; ++j) e[j] = mfma_h(alo[j], bh, z);
#pragma unroll
    for (int j = 0; j < NS; ++j) e[j] = mfma_h(ahi[j], bl, e[j]);
#pragma unroll
    for (int j = 0; j < NS; ++j) e[j] = mfma_h(ahi[j], bh, e[j]);
#pragma unroll
    for (int j = 0; j < NS; ++j) e[j] = mfma_b(ak[j], bk, e[j]);
  };
  auto tiles = [&](int blk, u32x4 bk) { tiles_ops(L.rop[blk][0][lane], L.rop[blk][1][lane], bk); };
  // e <- exp2(e); returns the lane's row sum over this wave's columns (both
  // halves of the wave hold the same row), one partial per tile.
  auto exp_rowsum = [&]() {
    float part[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      part[j] = 0.f;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        e[j][i] = exp2_fast(e[j][i]);
        part[j] += e[j][i];
      }
    }
    float sl = part[0];
#pragma unroll
    for (int j = 1; j < NS; ++j) sl += part[j];
    return halves_sum(sl);
  };
  // phase_a + sweep_block compute one row block as tiles_ops followed by
  // exp_rowsum (same arithmetic, same order), scheduled so the matrix core
  // and the VALU work at the same time.  (Both waves of a SIMD meet the block
  // barrier in step, so the other wave cannot fill one unit while this one
  // works on the other.)
  constexpr int NA = NS < 2 ? NS : 2;  // tiles whose chains open a block
  constexpr int NB = NS - NA;          // tiles whose chains overlap exp2 work
  constexpr int NP = 1;                // opening chains issued a block early
  // The MFMA chain of tile 0 of a row block into dst.  The main sweep issues
  // it for the next block before the block barrier, so the matrix core works
  // while the waves meet and exchange row sums (one tile: more live tiles
  // across the barrier spill).
  auto phase_a = [&](u32x4 bh, u32x4 bl, u32x4 bk, f32x16* dst) {
    const f32x16 z = {};
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      dst[j] = mfma_h(alo[j], bh, z);
      dst[j] = mfma_h(ahi[j], bl, dst[j]);
      dst[j] = mfma_h(ahi[j], bh, dst[j]);
      dst[j] = mfma_b(ak[j], bk, dst[j]);
    }
  };
  // The rest of a row block, with e[0 .. NP-1] already holding their chains:
  // the other opening chains, then the remaining chains each followed by a
  // share of the exp2 / row-sum work of the first tiles, then the exp2 work
  // of the others.
  auto sweep_block = [&](u32x4 bh, u32x4 bl, u32x4 bk) {
    const f32x16 z = {};
    auto mf = [&](int j, int k) {
      if (k == 0) e[j] = mfma_h(alo[j], bh, z);
      if (k == 1) e[j] = mfma_h(ahi[j], bl, e[j]);
      if (k == 2) e[j] = mfma_h(ahi[j], bh, e[j]);
      if (k == 3) e[j] = mfma_b(ak[j], bk, e[j]);
    };
    float part[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) part[j] = 0.f;
    auto ex = [&](int j, int i) {
      e[j][i] = exp2_fast(e[j][i]);
      part[j] += e[j][i];
    };
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = NP; j < NA; ++j) mf(j, k);
    if constexpr (NB > 0) {
      constexpr int PER = 16 * NA / (4 * NB);  // exp2 elements per MFMA
#pragma unroll
      for (int q = 0; q < 4 * NB; ++q) {
        mf(NA + q % NB, q / NB);
#pragma unroll
        for (int t = 0; t < PER; ++t) {
          const int f = q * PER + t;
          ex(f / 16, f % 16);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 2 * PER, 0);
      }
    } else {
#pragma unroll
      for (int j = 0; j < NA; ++j)
#pragma unroll
        for (int i = 0; i < 16; ++i) ex(j, i);
    }
#pragma unroll
    for (int j = NA; j < NS; ++j)
#pragma unroll
      for (int i = 0; i < 16; ++i) ex(j, i);
    float sl = part[0];
#pragma unroll
    for (int j = 1; j < NS; ++j) sl += part[j];
    return halves_sum(sl);
  };
  // Sum of a row's PW wave partials, read as 16-byte vectors and added in
  // wave order.
  auto row_total = [&](const float* pr) {
    float S = 0.f;
#pragma unroll
    for (int w = 0; w < PW; w += 4) {
      const f32x4v v = *reinterpret_cast<const f32x4v*>(pr + w);
      S += v[0];
      S += v[1];
      S += v[2];
      S += v[3];
    }
    return S;
  };
  auto row_kslot = [&](int blk) {
    const u32x4 k = L.rks[blk * 32 + l32];
    return h == 0 ? k : zero4;
  };
  // One pass over this workgroup's row blocks: the tiles of each block into
  // e, then body(blk); the row operands are read one block ahead (the last
  // block reads itself again), so their LDS latency overlaps the work.
  auto rows_pass = [&](auto&& body) {
    u32x4 nbh = L.rop[blk0][0][lane], nbl = L.rop[blk0][1][lane], nbk = row_kslot(blk0);
    for (int blk = blk0; blk < blk1; ++blk) {
      const u32x4 bh = nbh, bl = nbl, bk = nbk;
      const int nb = blk + 1 < blk1 ? blk + 1 : blk;
      nbh = L.rop[nb][0][lane];
      nbl = L.rop[nb][1][lane];
      nbk = row_kslot(nb);
      tiles_ops(bh, bl, bk);
      body(blk);
    }
  };
  // The same pass on the hi parts only (x_hi . y_hi plus the potentials):
  // half the matrix work, exponents within about 2^-11 |X| |Y| of the full
  // ones -- enough for a shift, which only has to be near the maximum.
  auto rows_pass_hi = [&](auto&& body) {
    u32x4 nbh = L.rop[blk0][0][lane], nbk = row_kslot(blk0);
    for (int blk = blk0; blk < blk1; ++blk) {
      const u32x4 bh = nbh, bk = nbk;
      const int nb = blk + 1 < blk1 ? blk + 1 : blk;
      nbh = L.rop[nb][0][lane];
      nbk = row_kslot(nb);
      const f32x16 z = {};
#pragma unroll
      for (int j = 0; j < NS; ++j) e[j] = mfma_h(ahi[j], bh, z);
#pragma unroll
      for (int j = 0; j < NS; ++j) e[j] = mfma_b(ak[j], bk, e[j]);
      body(blk);
    }
  };
  // Per-column reduction of acc over the rows (sum or max) into dst.
  f32x16 acc[NS];
  // Recursive halving: an exchange with lane ^ 1 pairs values i and i+1
  // (each lane keeps one of the two, summed over the pair), one with lane ^ 2
  // pairs the results again, so 4 values per lane remain, value 4r + (lane &
  // 3) in slot r; rotations by 4 and 8 within the 16-lane row sum each slot
  // over the row, and a swap of rows 0 and 1 (2 and 3) adds the other row.
  // Lanes l32 < 16 then store value 4 (l32 >> 2) + (l32 & 3) of the tile.
  auto columns_out = [&](auto max_tag, float* dst) {
    constexpr bool MAX = decltype(max_tag)::value;
    auto op = [](float x, float y) { return MAX ? __builtin_elementwise_maximum(x, y) : x + y; };
    const bool b0 = lane & 1, b1 = (lane >> 1) & 1;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      float s1[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float x0 = acc[j][2 * q], x1 = acc[j][2 * q + 1];
        s1[q] = op(b0 ? x1 : x0, dpp_f<0xB1>(b0 ? x0 : x1));
      }
      float s2[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float x0 = s1[2 * r], x1 = s1[2 * r + 1];
        s2[r] = op(b1 ? x1 : x0, dpp_f<0x4E>(b1 ? x0 : x1));
      }
      float w = 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float y = op(s2[r], dpp_f<0x124>(s2[r]));  // row_ror:4
        y = op(y, dpp_f<0x128>(y));                // row_ror:8
        const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(y), __float_as_uint(y),
                                                         false, false);
        const float tot = op(__uint_as_float(sw[0]), __uint_as_float(sw[1]));
        w = (l32 >> 2) == r ? tot : w;
      }
      const int ln = lane_now(), lw = ln & 31;
      if (lw < 16) dst[32 * (wave + PW * j) + 8 * (lw >> 2) + 4 * (ln >> 5) + (lw & 3)] = w;
    }
  };

  // ---- cost maximum ----
  double cmax = A.normalize == 2 ? uniform_d(A.cmax[p]) : 1.0;
  if (A.normalize == 1) {
    const double sc0 = 1.4142135623730951 * (ext > 0.0 ? ldexp(1.0, -ilogb(ext) - 1) : 1.0);
    pack(sc0);
    __syncthreads();
    // Padding is +inf so that it never wins the minimum of -cost.
    set_col_kslots([&](int c) { return c < m ? -L.yn[c] : __builtin_inf(); });
    for (int r = tid; r < nblk * 32; r += PT)
      L.rks[r] = kslot_row(r < n ? L.xn[r] : -__builtin_inf());
    __syncthreads();
    // Per column tile, each lane keeps the smallest value over the rows and
    // the first block reaching it (the minimum over a block's 16 values is a
    // v_minimum3 chain: no NaN arises here, padding rows and columns give
    // +inf, so the IEEE minimum equals fminf without the canonicalising
    // instructions that fminf costs).  Lanes l and l ^ 32 hold the same row slot (other
    // columns): they agree on one block per tile, the one with the smaller
    // value (the earlier on a tie), whose tile is recomputed (the same MFMA
    // chain, so the same values) to find the column: the first occurrence of
    // the pair's minimum in (block, tile, value) order wins.
    float bestj[NS];
    int blkj[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      bestj[j] = __builtin_inff();
      blkj[j] = -1;
    }
    rows_pass([&](int blk) {
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        float mn = e[j][0];
#pragma unroll
        for (int i = 1; i < 16; ++i) mn = __builtin_elementwise_minimum(mn, e[j][i]);
        const bool lt = mn < bestj[j];
        bestj[j] = lt ? mn : bestj[j];
        blkj[j] = lt ? blk : blkj[j];
      }
    });
    float best = __builtin_inff();
    int code = -1;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const auto sv = __builtin_amdgcn_permlane32_swap(__float_as_uint(bestj[j]),
                                                       __float_as_uint(bestj[j]), false, false);
      const auto sb = __builtin_amdgcn_permlane32_swap((unsigned)blkj[j], (unsigned)blkj[j],
                                                       false, false);
      // sv / sb: [0] the value of lanes 0-31, [1] of lanes 32-63, in every lane.
      const float v0 = __uint_as_float(sv[0]), v1 = __uint_as_float(sv[1]);
      const int b0 = (int)sb[0], b1 = (int)sb[1];
      const bool one = v1 < v0 || (v1 == v0 && (unsigned)b1 < (unsigned)b0);
      const float pv = one ? v1 : v0;
      const int pb = one ? b1 : b0;
      const int rb = pb >= 0 ? pb : blk0;
      const f32x16 z = {};
      const u32x4 rh = L.rop[rb][0][lane], rl = L.rop[rb][1][lane], rk = row_kslot(rb);
      f32x16 t = mfma_h(alo[j], rh, z);
      t = mfma_h(ahi[j], rl, t);
      t = mfma_h(ahi[j], rh, t);
      t = mfma_b(ak[j], rk, t);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int cand = (pb << 8) | (j << 4) | i;
        const bool take = pb >= 0 && t[i] == pv && (pv < best || (pv == best && cand < code));
        best = take ? pv : best;
        code = take ? cand : code;
      }
    }
    double mexact = 0.0;
    if (code >= 0) {
      const int blk = code >> 8, j = (code >> 4) & 15, i = code & 15;
      const int r = blk * 32 + l32, c = out_col(j, i);
      if (r < n && c < m) {
        // All coordinates loaded first (d <= kPD, checked at launch; the
        // index past d is clamped to 0), then summed in order.
        float xs[kPD], ys[kPD];
        if (v4) {
#pragma unroll
          for (int g = 0; g < kPD / 4; ++g) {
            const int k0 = 4 * g < d ? 4 * g : 0;
            const f32x4v qx = *reinterpret_cast<const f32x4v*>(Xp + (int64_t)r * d + k0);
            const f32x4v qy = *reinterpret_cast<const f32x4v*>(Yp + (int64_t)c * d + k0);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              xs[4 * g + e] = qx[e];
              ys[4 * g + e] = qy[e];
            }
          }
        } else {
#pragma unroll
          for (int k = 0; k < kPD; ++k) {
            const int kk = k < d ? k : 0;
            xs[k] = Xp[(int64_t)r * d + kk];
            ys[k] = Yp[(int64_t)c * d + kk];
          }
        }
#pragma unroll
        for (int k = 0; k < kPD; ++k) {
          if (k < d) {
            const double t = (double)xs[k] - (double)ys[k];
            mexact = fma(t, t, mexact);
          }
        }
      }
    }
    mexact = wave_max(mexact);
    if (lane == 0) L.red[wave] = mexact;
    __syncthreads();
    cmax = 0.0;
    for (int w = 0; w < PW; ++w) cmax = fmax(cmax, L.red[w]);
    cmax = uniform_d(cmax);
    __syncthreads();
    if constexpr (SPLIT) {
      if (!exchange_max(cmax)) return abandon();
    }
    if (A.max_only) {
      if (tid == 0) A.cmax[p] = cmax;
      return;
    }
  }
  const double reg = uniform_d(A.regs[p]) * cmax;
  const double s = kLog2e / reg;
  {
    const float wm = wave_max(pack(sqrt(2.0 * s)));
    if (lane == 0) L.redf[wave] = wm;
  }
  __syncthreads();
  float allmax = 0.f;
  for (int w = 0; w < PW; ++w) allmax = fmaxf(allmax, L.redf[w]);
  allmax = uniform_f(allmax);
  if (!(allmax <= kMaxOperand) || !(cmax > 0.0) || !(reg > 0.0)) {
    // Outside the operand range of the f16 MFMA path (reg far below the
    // cost scale) or a degenerate cost: report instead of solving.
    for (int i = tid; i < n; i += PT) up[i] = __builtin_nan("");
    for (int j = tid; j < m; j += PT) vp[j] = __builtin_nan("");
    if (tid == 0) {
      A.niter[p] = 0;
      A.status[p] = kNumerical;
      if (A.cmax && A.normalize != 2) A.cmax[p] = cmax;
    }
    return;
  }

  int status = kMaxIter, niter = A.max_iter - 1, restarted = 0;
  double err_last = -1.0, err_first = -1.0;
  const bool warm = A.fallback_thr > 0.0;
  bool cold = false;

  // b_c and log2 b_c of this lane's columns, once per problem.  (No double
  // logarithm is left inside the attempt loop below: the compiler would keep
  // its polynomial's constants in registers across the whole loop.)
  {
    double bv[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int c = col_of(j);
      bv[j] = bp[c < m ? c : 0];
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int c = col_of(j);
      if (h == 0 && c < m) {
        L.bs[c] = bv[j];
        L.lbs[c] = bv[j] > 0.0 ? log2(bv[j]) : -__builtin_inf();
      }
    }
  }

  // One attempt from the given start; a rejected warm start repeats the
  // prologue and the iterations once from the cold start.
  for (;;) {
  bool again = false;
  // ---- prologue: v_1 = log2 b - LSE2_r(x_rc + u2_r), a max pass then a sum pass ----
  // (The per-row and per-column values are loaded before any is used; rows
  // and columns past the end read entry 0, which is not used.)
  {
    double uv[RR];
#pragma unroll
    for (int t = 0; t < RR; ++t) {
      const int r = tid + t * PT;
      uv[t] = up[r < n ? r : 0];
    }
#pragma unroll
    for (int t = 0; t < RR; ++t) {
      const int r = tid + t * PT;
      if (r < nblk * 32) {
        const double u2 = (r < n && !cold) ? uv[t] * kLog2e : 0.0;
        L.rks[r] = kslot_row(r < n ? L.xn[r] - u2 : __builtin_inf());
      }
    }
  }
  set_col_kslots([&](int c) { return c < m ? -L.yn[c] : -__builtin_inf(); });
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NS; ++j)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[j][i] = -__builtin_inff();
  // (IEEE maximum, one v_maximum3 per value: padding gives -inf, never NaN.
  // The maxima only shift the sum pass below, whose log-sum-exp is exact for
  // any shift near the maximum, so the pass runs on the hi parts.)
  rows_pass_hi([&](int) {
#pragma unroll
    for (int j = 0; j < NS; ++j) acc[j] = __builtin_elementwise_maximum(acc[j], e[j]);
  });
  columns_out(std::integral_constant<bool, true>(), L.sh);
  __syncthreads();
  if constexpr (SPLIT) {
    if (!exchange_cols(L.sh, true)) return abandon();
  }
  set_col_kslots([&](int c) {
    const float g = L.sh[c < m ? c : 0];
    return (c < m && __builtin_isfinite(g)) ? -L.yn[c] - (double)g : -__builtin_inf();
  });
#pragma unroll
  for (int j = 0; j < NS; ++j) acc[j] = f32x16{};
  rows_pass([&](int) {
#pragma unroll
    for (int j = 0; j < NS; ++j)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[j][i] += exp2_fast(e[j][i]);
  });
  columns_out(std::integral_constant<bool, false>(), L.colx);
  __syncthreads();
  if constexpr (SPLIT) {
    if (!exchange_cols(L.colx, false)) return abandon();
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int c = col_of(j);
    if (h == 0 && c < m) {
      // (log2 of the column sum on the fp32 log, as in the iterations: a
      // finite shift g makes the sum at least about 1.)
      const double g = (double)L.sh[c];
      L.vsh[c] = (L.bs[c] > 0.0 && __builtin_isfinite(g))
                     ? L.lbs[c] - (g + (double)__builtin_amdgcn_logf(L.colx[c]))
                     : -__builtin_inf();
    }
  }
  set_col_kslots([&](int c) { return c < m ? L.vsh[c] - L.yn[c] : -__builtin_inf(); });
  {
    double uv[RR];
#pragma unroll
    for (int t = 0; t < RR; ++t) {
      const int r = tid + t * PT;
      uv[t] = up[r < n ? r : 0];
    }
#pragma unroll
    for (int t = 0; t < RR; ++t) {
      const int r = tid + t * PT;
      if (r < nblk * 32) {
        double c2 = __builtin_inf();
        if (r < n && L.af[r] > 0.f) {
          const double u2 = cold ? 0.0 : uv[t] * kLog2e;
          c2 = L.la[r] - u2;
        }
        L.c2s[r] = c2;
        L.rks[r] = kslot_row(r < n ? c2 + L.xn[r] : __builtin_inf());
      }
    }
  }
  __syncthreads();

  for (int ii = 0; ii < A.max_iter; ++ii) {
    // ---- fused sweep: row sums (u update) and column sums of P(u_new, v) ----
#pragma unroll
    for (int j = 0; j < NS; ++j) acc[j] = f32x16{};
    // The row operands of the next block are read one block ahead, so their
    // LDS latency overlaps this block's work.
    u32x4 nbh = L.rop[blk0][0][lane], nbl = L.rop[blk0][1][lane], nbk = row_kslot(blk0);
    phase_a(nbh, nbl, nbk, e);
    for (int blk = blk0; blk < blk1; ++blk) {
      u32x4 bk = nbk;
      const u32x4 bh = nbh, bl = nbl;
      // Unconditional (the last block re-reads itself), so the block's MFMA
      // and exp2 work stay in one scheduling region.
      const int nb = blk + 1 < blk1 ? blk + 1 : blk;
      nbh = L.rop[nb][0][lane];
      nbl = L.rop[nb][1][lane];
      nbk = row_kslot(nb);
      const int r = blk * 32 + l32;
      const float ar = L.af[r];  // read before the barrier, off the wait chain
      float sl = sweep_block(bh, bl, bk);
      const int buf = blk & 1;
      if (h == 0) L.rowpart[buf][l32][wave] = sl;
      // The next block's tile-0 chain (after the last block, this block's
      // again, unused).
      f32x16 na[NP];
      phase_a(nbh, nbl, nbk, na);
      __builtin_amdgcn_sched_barrier(0);  // keep the chain ahead of the barrier
      __syncthreads();
      float S = row_total(L.rowpart[buf][l32]);
      const bool bad = ar > 0.f && !in_safe_range((double)S);
      if (__ballot(bad)) {
        // Re-centre the rows that left the range on their maximum exponent.
        tiles(blk, bk);
        float mx = -__builtin_inff();
#pragma unroll
        for (int j = 0; j < NS; ++j)
#pragma unroll
          for (int i = 0; i < 16; ++i) mx = fmaxf(mx, e[j][i]);
        mx = halves_max(mx);
        if (h == 0) L.rowmax[wave][l32] = mx;
        __syncthreads();
        float G = -__builtin_inff();
#pragma unroll
        for (int w = 0; w < PW; ++w) G = fmaxf(G, L.rowmax[w][l32]);
        double c2 = L.c2s[r];
        if (bad && __builtin_isfinite(G)) {
          c2 += (double)G;
          const u32x4 k = kslot_row(c2 + L.xn[r]);
          bk = h == 0 ? k : zero4;
        }
        tiles(blk, bk);
        sl = exp_rowsum();
        if (h == 0) L.rowpart[buf][l32][wave] = sl;
        __syncthreads();
        S = row_total(L.rowpart[buf][l32]);
        if (wave == 0 && h == 0 && bad) {
          L.c2s[r] = c2;  // the shift actually used
          L.rks[r] = bk;
        }
      }
      const float wr = ar > 0.f ? ar * __builtin_amdgcn_rcpf(S) : 0.f;
#pragma unroll
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[j][i] = __builtin_fmaf(e[j][i], wr, acc[j][i]);
      if (wave == 0 && h == 0) L.sh[r] = S;
#pragma unroll
      for (int j = 0; j < NP; ++j) e[j] = na[j];
    }
    columns_out(std::integral_constant<bool, false>(), L.colx);
    __syncthreads();
    if constexpr (SPLIT) {
      if (!exchange_cols(L.colx, false)) return abandon();
    }

    // ---- row potentials, marginal error, stop decision ----
    // (Every LDS value of a section is read before any is used -- indexes
    // clamped, results selected -- so each section waits for LDS once.)
    const int rend = SPLIT ? (blk1 * 32 < n ? blk1 * 32 : n) : n;
    {
      const int r0 = blk0 * 32 + wave * kWave + lane_now();
      float afv[RR], shv[RR];
      double c2v[RR], xnv[RR];
#pragma unroll
      for (int t = 0; t < RR; ++t) {
        const int r = r0 + t * PT;
        const int rr = r < rend ? r : 0;
        afv[t] = L.af[rr];
        shv[t] = L.sh[rr];
        c2v[t] = L.c2s[rr];
        xnv[t] = L.xn[rr];
      }
#pragma unroll
      for (int t = 0; t < RR; ++t) {
        const int r = r0 + t * PT;
        if (r < rend && afv[t] > 0.f) {
          // sh[r] lies in [2^-60, 2^60] (the safe range of the sweep).
          const double c = c2v[t] + (double)__builtin_amdgcn_logf(shv[t]);
          L.c2s[r] = c;
          L.rks[r] = kslot_row(c + xnv[t]);
        }
      }
    }
    float csv[NS];
    {
      double bsv[NS];
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const int c = col_now(j);
        const int cc = c < m ? c : 0;
        csv[j] = L.colx[cc];
        bsv[j] = L.bs[cc];
      }
      double e2 = 0.0;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const double dd = (double)csv[j] - bsv[j];
        e2 += (half_now() == 0 && col_now(j) < m) ? dd * dd : 0.0;
      }
      e2 = wave_sum_dpp_d(e2);
      if (lane == 0) L.red[wave] = e2;
    }
    __syncthreads();
    // Every thread forms the same error from the PW wave sums (added in wave
    // order), so the decision needs no second barrier.
    int fl = 0;
    {
      double rv[PW];
#pragma unroll
      for (int w = 0; w < PW; ++w) rv[w] = L.red[w];
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < PW; ++w) t += rv[w];
      const double err = sqrt(uniform_d(t));
      if (ii % A.check_every == 0) {
        err_last = err;
        if (ii == 0) {
          err_first = err;
          if (warm && !restarted && err > A.fallback_thr) fl = 2;
        }
        if (fl == 0 && err < A.stop_thr) fl = 1;
      }
      fl = uniform_i(fl);
    }
    if (fl == 2) {
      // Warm start rejected: restart from the cold start u0 = 0.
      restarted = 1;
      cold = true;
      again = true;
      __syncthreads();
      break;
    }
    if (fl == 1) {
      status = kConverged;
      niter = ii;
      break;
    }
    if (ii == A.max_iter - 1) break;
    // ---- column update v_c += log2 b_c - log2 colsum_c (own columns) ----
    {
      double lbv[NS], vsv[NS], ynv[NS];
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const int c = col_now(j);
        const int cc = c < m ? c : 0;
        lbv[j] = L.lbs[cc];
        vsv[j] = L.vsh[cc];
        ynv[j] = L.yn[cc];
      }
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const int c = col_now(j);
        if (half_now() == 0 && c < m) {
          // log2 of the column sum on the fp32 log (the fp32 sums carry no
          // more than that); a subnormal sum is scaled into the normal range
          // first
          const float cs = csv[j];
          const bool tiny = cs < 1.1754944e-38f;
          const double lcs = (double)__builtin_amdgcn_logf(tiny ? cs * 0x1p64f : cs) -
                             (tiny ? 64.0 : 0.0);
          const double vn = lbv[j] > -__builtin_inf() ? vsv[j] + lbv[j] - lcs
                                                      : -__builtin_inf();
          L.vsh[c] = vn;
          ak[j] = kslot_col(vn - ynv[j]);
        }
      }
    }
    __syncthreads();
  }
  if (!again) break;
  }

  // ---- outputs (natural-log potentials; a split problem's parts write their
  // own rows, part 0 the columns and the status) ----
  __syncthreads();
  {
    const int rend = SPLIT ? (blk1 * 32 < n ? blk1 * 32 : n) : n;
#pragma unroll
    for (int t = 0; t < RR; ++t) {
      const int i = blk0 * 32 + tid + t * PT;
      if (i < rend) up[i] = L.af[i] > 0.f ? (L.la[i] - L.c2s[i]) * kLn2 : -__builtin_inf();
    }
  }
  if (part != 0) return;
  for (int j = tid; j < m; j += PT) vp[j] = L.vsh[j] * kLn2;
  if (tid == 0) {
    A.niter[p] = niter;
    A.status[p] = status | (restarted ? 4 : 0);
    if (A.err_last) A.err_last[p] = err_last;
    if (A.err_first) A.err_first[p] = err_first;
    if (A.cmax && A.normalize != 2) A.cmax[p] = cmax;
  }
}

// One problem, from the centring to the outputs (every thread of the
// workgroup calls it; returns are uniform).  Kept out of line, with its LDS
// declared here, so that the persistent kernel's loop around it is a plain
// uniform loop.
template <int PW>
__device__ __attribute__((noinline)) void one_problem(const PointsArgs Ain, int item) {
  __shared__ PointsLds<PW> L;
  // Arguments of an out-of-line function count as per-lane values for the
  // compiler: the ones that steer barriers are made visibly uniform.
  PointsArgs A = Ain;
  // All of them are made so, which keeps the pointers and scalars in scalar
  // registers instead of vector registers (or scratch) across the solve.
  A.X = uniform_p(Ain.X);
  A.Y = uniform_p(Ain.Y);
  A.xoff = uniform_p(Ain.xoff);
  A.yoff = uniform_p(Ain.yoff);
  A.ns = uniform_p(Ain.ns);
  A.ms = uniform_p(Ain.ms);
  A.d = uniform_i(Ain.d);
  A.normalize = uniform_i(Ain.normalize);
  A.max_only = uniform_i(Ain.max_only);
  A.regs = uniform_p(Ain.regs);
  A.a = uniform_p(Ain.a);
  A.b = uniform_p(Ain.b);
  A.u = uniform_p(Ain.u);
  A.v = uniform_p(Ain.v);
  A.order = uniform_p(Ain.order);
  A.max_iter = uniform_i(Ain.max_iter);
  A.check_every = uniform_i(Ain.check_every);
  A.stop_thr = uniform_d(Ain.stop_thr);
  A.fallback_thr = uniform_d(Ain.fallback_thr);
  A.niter = uniform_p(Ain.niter);
  A.status = uniform_p(Ain.status);
  A.err_last = uniform_p(Ain.err_last);
  A.err_first = uniform_p(Ain.err_first);
  A.cmax = uniform_p(Ain.cmax);
  A.batch = uniform_i(Ain.batch);
  A.queue = uniform_p(Ain.queue);
  A.items = uniform_i(Ain.items);
  A.split = uniform_p(Ain.split);
  A.split_ws = uniform_p(Ain.split_ws);
  A.nsplit = uniform_i(Ain.nsplit);
  item = uniform_i(item);
  // An item is problem p, part `part` of a split problem (part 0 otherwise).
  const int p = item & 0xFFFFFF;
  const int part = item >> 24;
  const int sp = A.split ? uniform_i(A.split[p]) : 0;
  const int K = sp ? (sp & 15) : 1;
  if (p >= A.batch || part >= K || (sp && (K < 2 || K > kSplitMax || (sp >> 4) >= A.nsplit))) {
    // A malformed item or split entry: nothing is touched but the status.
    if (p < A.batch && threadIdx.x == 0 && !A.max_only) {
      A.niter[p] = 0;
      A.status[p] = kNumerical;
    }
    return;
  }
  char* sws = sp ? A.split_ws + (size_t)(sp >> 4) * kSplitSlot : nullptr;
  constexpr int PT = PW * kWave;
  constexpr int NG = PT / 16;  // point groups of the centring reduction
  const int tid = threadIdx.x;
  const int n = uniform_i(A.ns[p]), m = uniform_i(A.ms[p]), d = A.d;
  if (n < 1 || n > kPMax || m < 1 || m > kPMax) {
    if (tid == 0) {
      if (A.max_only) {
        A.cmax[p] = __builtin_nan("");
      } else {
        A.niter[p] = 0;
        A.status[p] = kNumerical;
      }
    }
    return;
  }
  const float* Xp = A.X + A.xoff[p] * d;
  const float* Yp = A.Y + A.yoff[p] * d;
  if (!A.max_only) {
    const double* ap = A.a + A.xoff[p];
    constexpr int RR = kPMax / PT;
    double av[RR];
#pragma unroll
    for (int t = 0; t < RR; ++t) {
      const int i = tid + t * PT;
      av[t] = ap[i < n ? i : 0];
    }
#pragma unroll
    for (int t = 0; t < RR; ++t) {
      const int i = tid + t * PT;
      const float af = i < n ? (float)av[t] : 0.f;
      L.af[i] = af;
      L.la[i] = af > 0.f ? log2(av[t]) : __builtin_inf();
    }
  }

  // ---- centre (midpoint of the centroids) and coordinate extent ----
  double ext = 0.0;  // largest |coordinate - centre| over both clouds
  {
    double* scr = reinterpret_cast<double*>(&L.rop[0][0][0]);
    const int k = tid & 15, g = tid >> 4;  // g < NG
    double sx = 0.0, sy = 0.0;
    float lo = __builtin_inff(), hi = -__builtin_inff();
    if (k < d) {
      // Every point of both clouds this thread reads (kU of each) is loaded
      // before any is used (indices past the end read point 0; those values
      // are not used), so all the latencies overlap; they are summed in order
      // of the point index.
      constexpr int kU = kPMax / NG;
      float xv[kU], yv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int i = g + NG * u;
        xv[u] = Xp[(int64_t)(i < n ? i : 0) * d + k];
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int i = g + NG * u;
        yv[u] = Yp[(int64_t)(i < m ? i : 0) * d + k];
      }
      // (Selects rather than branches; adding +0 changes a sum only from -0
      // to +0, which gives the same centre.)
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const bool in = g + NG * u < n;
        sx += in ? (double)xv[u] : 0.0;
        lo = in ? fminf(lo, xv[u]) : lo;
        hi = in ? fmaxf(hi, xv[u]) : hi;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const bool in = g + NG * u < m;
        sy += in ? (double)yv[u] : 0.0;
        lo = in ? fminf(lo, yv[u]) : lo;
        hi = in ? fmaxf(hi, yv[u]) : hi;
      }
    }
    scr[tid] = sx;
    scr[PT + tid] = sy;
    scr[2 * PT + tid] = lo;
    scr[3 * PT + tid] = hi;
    __syncthreads();
    if (tid < kPD) {
      double cx = 0.0, cy = 0.0, l = __builtin_inf(), u = -__builtin_inf();
      for (int gg = 0; gg < NG; ++gg) {
        cx += scr[gg * 16 + tid];
        cy += scr[PT + gg * 16 + tid];
        l = fmin(l, scr[2 * PT + gg * 16 + tid]);
        u = fmax(u, scr[3 * PT + gg * 16 + tid]);
      }
      const double c = tid < d ? 0.5 * (cx / n + cy / m) : 0.0;
      L.cen[tid] = c;
      scr[4 * PT + tid] = tid < d ? fmax(u - c, c - l) : 0.0;
    }
    __syncthreads();
    for (int kk = 0; kk < kPD; ++kk) ext = fmax(ext, scr[4 * PT + kk]);
    __syncthreads();  // the scratch is the row-operand area, packed next
  }
  const int ns = ((m + 31) / 32 + PW - 1) / PW;
  if (K > 1) {
    // Split problems (8 waves only): the same solve over the part's rows.
    if constexpr (PW == 8) {
      if (ns <= 2)
        solve_points<PW, 2, true>(A, L, p, n, m, ext, part, K, sws);
      else
        solve_points<PW, 4, true>(A, L, p, n, m, ext, part, K, sws);
    }
    return;
  }
  if constexpr (PW == 16) {
    if (ns <= 1)
      solve_points<PW, 1, false>(A, L, p, n, m, ext, 0, 1, nullptr);
    else
      solve_points<PW, 2, false>(A, L, p, n, m, ext, 0, 1, nullptr);
  } else {
    if (ns <= 1)
      solve_points<PW, 1, false>(A, L, p, n, m, ext, 0, 1, nullptr);
    else if (ns == 2)
      solve_points<PW, 2, false>(A, L, p, n, m, ext, 0, 1, nullptr);
    else if (ns == 3)
      solve_points<PW, 3, false>(A, L, p, n, m, ext, 0, 1, nullptr);
    else
      solve_points<PW, 4, false>(A, L, p, n, m, ext, 0, 1, nullptr);
  }
}

// With a queue the launch is persistent, one workgroup per CU: each takes
// the next problem of `order` from a device counter until the batch is
// exhausted, so the CUs never wait for the dispatch of a new workgroup and
// the longest-first order is kept across the whole GPU.  The last workgroup
// to leave resets the counters to zero, which leaves the queue ready for
// the next launch (also inside a graph).  Without a queue, workgroup i
// solves problem order[i].
template <int PW>
__global__ __launch_bounds__(PW * kWave) void batch_points_kernel(PointsArgs A) {
  __shared__ int qslot;
  if (!A.queue) {
    one_problem<PW>(A, A.order ? uniform_i(A.order[blockIdx.x]) : (int)blockIdx.x);
    return;
  }
  for (;;) {
    if (threadIdx.x == 0) qslot = atomicAdd(&A.queue[0], 1);
    __syncthreads();
    const int q = uniform_i(qslot);
    __syncthreads();  // everyone has the slot before it is taken again
    if (q < 0 || q >= A.items) break;
    one_problem<PW>(A, A.order ? uniform_i(A.order[q]) : q);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&A.queue[1], 1) == (int)gridDim.x - 1) {
      // Every part has left its problem: the split counters go back to zero
      // with the queue.
      for (int k = 0; k < A.nsplit; ++k)
        atomicExch(reinterpret_cast<int*>(A.split_ws + (size_t)k * kSplitSlot), 0);
      atomicExch(&A.queue[0], 0);
      atomicExch(&A.queue[1], 0);
    }
  }
}

}  // namespace
}  // namespace potamd

using namespace potamd;

namespace potamd {
// Argument checks shared by the point-cloud entry points (0 or an error).
int points_validate(int batch, int max_n, int max_m, int d, const float* X, const float* Y,
                    const int64_t* xoff, const int64_t* yoff, const int* ns, const int* ms,
                    const double* regs, int normalize, const double* a, const double* b,
                    const double* u, const double* v, int max_iter, int check_every,
                    const int* niter, const int* status, const double* cmax, int waves,
                    bool queued) {
  if (batch < 1 || max_n < 1 || max_m < 1 || d < 1 || !X || !Y || !xoff || !yoff || !ns ||
      !ms || !regs || !a || !b || !u || !v || max_iter < 1 || !niter || !status ||
      check_every < 0 || normalize < 0 || normalize > 2 || (normalize == 2 && !cmax))
    return POTAMD_EINVAL;
  if (max_n > kPMax || max_m > kPMax || d > kPD || (queued && batch > (1 << 30)))
    return POTAMD_EUNSUPPORTED;
  if (waves != 0 && waves != 8 && waves != 16) return POTAMD_EINVAL;
  return 0;
}

int points_launch(int batch, int max_n, int max_m, int d, const float* X, const float* Y,
                  const int64_t* xoff, const int64_t* yoff, const int* ns, const int* ms,
                  const double* regs, int normalize, const double* a, const double* b, double* u,
                  double* v, int max_iter, int check_every, double stop_thr,
                  double fallback_thr, const int* order, int* niter, int* status,
                  double* err_last, double* err_first, double* cmax, int waves, int* queue,
                  int items, const int* split, char* split_ws, int nsplit, void* stream) {
  const int rc = points_validate(batch, max_n, max_m, d, X, Y, xoff, yoff, ns, ms, regs,
                                 normalize, a, b, u, v, max_iter, check_every, niter, status,
                                 cmax, waves, queue != nullptr);
  if (rc != 0) return rc;
  if (waves == 0) waves = kPWDefault;
  PointsArgs pa;
  memset(&pa, 0, sizeof(pa));
  pa.X = X;
  pa.Y = Y;
  pa.xoff = xoff;
  pa.yoff = yoff;
  pa.ns = ns;
  pa.ms = ms;
  pa.d = d;
  pa.normalize = normalize;
  pa.regs = regs;
  pa.a = a;
  pa.b = b;
  pa.u = u;
  pa.v = v;
  pa.order = order;
  pa.max_iter = max_iter;
  pa.check_every = check_every > 0 ? check_every : 10;
  pa.stop_thr = stop_thr;
  pa.fallback_thr = fallback_thr;
  pa.niter = niter;
  pa.status = status;
  pa.err_last = err_last;
  pa.err_first = err_first;
  pa.cmax = cmax;
  pa.batch = batch;
  pa.queue = queue;
  pa.items = items;
  pa.split = split;
  pa.split_ws = split_ws;
  pa.nsplit = nsplit;
  // Persistent launch with the queue: one workgroup per CU (the LDS of one
  // problem fills most of a CU).
  const int grid = queue ? (items < device_cus() ? items : device_cus()) : items;
  if (waves == 16)
    hipLaunchKernelGGL(batch_points_kernel<16>, dim3(grid), dim3(16 * kWave), 0,
                       static_cast<hipStream_t>(stream), pa);
  else
    hipLaunchKernelGGL(batch_points_kernel<8>, dim3(grid), dim3(8 * kWave), 0,
                       static_cast<hipStream_t>(stream), pa);
  HIP_TRY(hipGetLastError());
  return 0;
}
}  // namespace potamd

extern "C" int potamd_sinkhorn_log_batch_points(
    int batch, int max_n, int max_m, int d, const float* X, const float* Y, const int64_t* xoff,
    const int64_t* yoff, const int* ns, const int* ms, const double* regs, int normalize,
    const double* a, const double* b, double* u, double* v, int max_iter, int check_every,
    double stop_thr, double fallback_thr, const int* order, int* niter, int* status,
    double* err_last, double* err_first, double* cmax, int waves, int* queue, void* stream) {
  return points_launch(batch, max_n, max_m, d, X, Y, xoff, yoff, ns, ms, regs, normalize, a, b,
                       u, v, max_iter, check_every, stop_thr, fallback_thr, order, niter, status,
                       err_last, err_first, cmax, waves, queue, batch, nullptr, nullptr, 0,
                       stream);
}

extern "C" size_t potamd_points_split_workspace_size(int nsplit) {
  return nsplit > 0 ? (size_t)nsplit * kSplitSlot : 0;
}

extern "C" int potamd_sinkhorn_log_batch_points_split(
    int batch, int max_n, int max_m, int d, const float* X, const float* Y, const int64_t* xoff,
    const int64_t* yoff, const int* ns, const int* ms, const double* regs, int normalize,
    const double* a, const double* b, double* u, double* v, int max_iter, int check_every,
    double stop_thr, double fallback_thr, int items, const int* order, const int* split,
    int nsplit, void* split_ws, size_t split_ws_bytes, int* niter, int* status,
    double* err_last, double* err_first, double* cmax, int* queue, void* stream) {
  if (items < batch || !order || !split || !queue || nsplit < 1 || !split_ws ||
      split_ws_bytes < potamd_points_split_workspace_size(nsplit) ||
      (reinterpret_cast<uintptr_t>(split_ws) & 255) != 0)
    return POTAMD_EINVAL;
  if (batch >= (1 << 24) || items > (1 << 30)) return POTAMD_EUNSUPPORTED;
  return points_launch(batch, max_n, max_m, d, X, Y, xoff, yoff, ns, ms, regs, normalize, a, b,
                       u, v, max_iter, check_every, stop_thr, fallback_thr, order, niter, status,
                       err_last, err_first, cmax, 8, queue, items, split,
                       static_cast<char*>(split_ws), nsplit, stream);
}

extern "C" int potamd_ragged_points_cmax(int batch, int max_n, int max_m, int d, const float* X,
                                         const float* Y, const int64_t* xoff, const int64_t* yoff,
                                         const int* ns, const int* ms, double* cmax,
                                         void* stream) {
  if (batch < 1 || max_n < 1 || max_m < 1 || d < 1 || !X || !Y || !xoff || !yoff || !ns ||
      !ms || !cmax)
    return POTAMD_EINVAL;
  if (max_n > kPMax || max_m > kPMax || d > kPD) return POTAMD_EUNSUPPORTED;
  PointsArgs pa;
  memset(&pa, 0, sizeof(pa));
  pa.X = X;
  pa.Y = Y;
  pa.xoff = xoff;
  pa.yoff = yoff;
  pa.ns = ns;
  pa.ms = ms;
  pa.d = d;
  pa.normalize = 1;
  pa.max_only = 1;
  pa.cmax = cmax;
  pa.batch = batch;
  pa.items = batch;
  hipLaunchKernelGGL(batch_points_kernel<kPWDefault>, dim3(batch), dim3(kPWDefault * kWave), 0,
                     static_cast<hipStream_t>(stream), pa);
  HIP_TRY(hipGetLastError());
  return 0;
}
