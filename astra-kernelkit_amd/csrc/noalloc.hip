// Allocation-free variant of the exact transpose, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).
//
// kk_cone_fp_transpose needs a workspace of about n_angles * rows * cols
// floats (the weighted projections, column-major per angle), which its
// callers allocate.  Here that workspace is a file-scope device array of fixed size,
// `scratch`.  Each device the library runs on holds its own copy of it, at
// its own address, until the process ends.  In host code the name `scratch`
// is a host-side stand-in, not a device address, so it is never passed to a
// kernel: the first call on a device resolves the device address with
// hipGetSymbolAddress (with that device current) and keeps it per device.
//
// Every call on one device shares that device's array and leaves the last
// call's contents in it.  The transpose overwrites every workspace element
// it reads (the ray-weight pass writes every workspace value, pads
// included, before the gather reads them, in stream order), so nothing needs clearing
// between calls.  Calls that overlap on one device (other streams, other
// host threads) race on the array.
#include <mutex>

#include "kk_internal.h"

namespace {

constexpr int64_t kScratchFloats = int64_t(1) << 26;  // 64 Mi floats, 256 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kScratchFloats];

std::mutex g_mutex;
float* g_scratch[kMaxDevices] = {};

// Device address of this device's `scratch` (resolved once per device).
int scratch_ptr(float** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return kk::hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return kk::fail(KK_ERR_UNSUPPORTED, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return kk::hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (p == nullptr)
      return kk::fail(KK_ERR_HIP, "hipGetSymbolAddress(scratch): null");
    g_scratch[dev] = static_cast<float*>(p);
  }
  *out = g_scratch[dev];
  return KK_OK;
}

}  // namespace

extern "C" int64_t kk_cone_fp_transpose_static_capacity(void) {
  return kScratchFloats;
}

extern "C" int kk_cone_fp_transpose_static(const kk_plan* plan,
                                           const float* proj,
                                           const int64_t proj_strides[3],
                                           int32_t rows, int32_t cols,
                                           float* vol,
                                           const int32_t vol_shape[3],
                                           const int64_t vol_strides[3],
                                           int32_t additive, void* stream) {
  if (plan == nullptr || !plan->fp_set)
    return kk::fail(KK_ERR_NOT_SET,
                    "Please set the projection geometry with `set_params()`.");
  const int64_t need = kk::fp_transpose_work_floats(plan, rows, cols);
  if (need < 0)
    return kk::fail(KK_ERR_INVALID,
                    "kk_cone_fp_transpose_static: bad detector");
  if (need > kScratchFloats)
    return kk::fail(KK_ERR_UNSUPPORTED,
                    "kk_cone_fp_transpose_static: %lld angles of %d x %d "
                    "detector pixels need a workspace of %lld floats, more "
                    "than the static scratch of %lld floats; use "
                    "kk_cone_fp_transpose with a caller workspace",
                    static_cast<long long>(plan->fp_n), rows, cols,
                    static_cast<long long>(need),
                    static_cast<long long>(kScratchFloats));
  float* work = nullptr;
  const int rc = scratch_ptr(&work);
  if (rc != KK_OK) return rc;
  return kk::fp_transpose_launch(plan, proj, proj_strides, rows, cols, vol,
                                 vol_shape, vol_strides, additive, work,
                                 static_cast<hipStream_t>(stream));
}
