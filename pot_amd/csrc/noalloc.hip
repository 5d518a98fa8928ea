// potamd_sinkhorn_log_batch_points_split_noalloc: the split point-cloud batch
// solve with its scratch (the work-queue counters and the exchange slots of
// the split problems) in a fixed device array of this library instead of
// caller-provided buffers, for callers that cannot allocate device memory
// (code under graph capture, serving loops, harnesses that forbid it).
//
// Scratch layout (one copy per GPU, alive until the process ends):
//   [0, 256)                       int queue[2] (rest of the line unused)
//   [256, 256 + kMaxSlots * slot)  exchange slots, kPointsSplitSlot bytes each
// Every call clears the queue and the slots it uses with a fill kernel on the
// caller's stream before the solve (no hipMemsetAsync, so the pair can be
// captured into a graph), so a call never depends on what an earlier call
// left.  Calls that overlap on one GPU share the array and race.
#include <algorithm>
#include <mutex>

#include "../../include/pot_amd.h"
#include "host.h"
#include "internal.h"

namespace {

constexpr int kMaxSlots = POTAMD_NOALLOC_MAX_SPLIT;
constexpr size_t kQueueBytes = 256;
constexpr size_t kScratchBytes = kQueueBytes + (size_t)kMaxSlots * potamd::kPointsSplitSlot;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up once per device).
hipError_t scratch_ptr(char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = static_cast<char*>(p);
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

// Zeroes `words` 16-byte words from base (grid-stride).
__global__ __launch_bounds__(256) void clear_scratch_kernel(uint4* base, int64_t words) {
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * blockDim.x)
    base[i] = z;
}

}  // namespace

static_assert(kQueueBytes % 16 == 0 && potamd::kPointsSplitSlot % 16 == 0,
              "the clear kernel writes 16-byte words");

extern "C" size_t potamd_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" int potamd_sinkhorn_log_batch_points_split_noalloc(
    int batch, int max_n, int max_m, int d, const float* X, const float* Y, const int64_t* xoff,
    const int64_t* yoff, const int* ns, const int* ms, const double* regs, int normalize,
    const double* a, const double* b, double* u, double* v, int max_iter, int check_every,
    double stop_thr, double fallback_thr, int items, const int* order, const int* split,
    int nsplit, int* niter, int* status, double* err_last, double* err_first, double* cmax,
    void* stream) {
  if (items < batch || !order || !split || nsplit < 0) return POTAMD_EINVAL;
  if (nsplit > kMaxSlots) return POTAMD_EWORKSPACE;
  if (batch >= (1 << 24) || items > (1 << 30)) return POTAMD_EUNSUPPORTED;
  const int rc = potamd::points_validate(batch, max_n, max_m, d, X, Y, xoff, yoff, ns, ms, regs,
                                         normalize, a, b, u, v, max_iter, check_every, niter,
                                         status, cmax, 8, true);
  if (rc != 0) return rc;
  char* base = nullptr;
  HIP_TRY(scratch_ptr(&base));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t words = (int64_t)(kQueueBytes + (size_t)nsplit * potamd::kPointsSplitSlot) / 16;
  const int grid = (int)std::min<int64_t>(1024, (words + 255) / 256);
  hipLaunchKernelGGL(clear_scratch_kernel, dim3(grid), dim3(256), 0, s,
                     reinterpret_cast<uint4*>(base), words);
  HIP_TRY(hipGetLastError());
  int* queue = reinterpret_cast<int*>(base);
  char* slots = base + kQueueBytes;
  return potamd::points_launch(batch, max_n, max_m, d, X, Y, xoff, yoff, ns, ms, regs,
                               normalize, a, b, u, v, max_iter, check_every, stop_thr,
                               fallback_thr, order, niter, status, err_last, err_first, cmax, 8,
                               queue, items, split, nsplit > 0 ? slots : nullptr, nsplit, s);
}
