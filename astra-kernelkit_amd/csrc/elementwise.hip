// Small elementwise kernels used by the SIRT set-up.
#include "kk_internal.h"

namespace {

// x <- 1/x, with +inf (from +0) mapped to 0 (kernelkit/algo.py:324-325,
// :331-332).
__global__ __launch_bounds__(256) void recip_or_zero_kernel(
    float* __restrict__ p, int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * 256;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += stride) {
    const float r = 1.0f / p[i];
    p[i] = (r == __builtin_inff()) ? 0.0f : r;
  }
}

}  // namespace

extern "C" int kk_recip_or_zero(float* ptr, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && ptr == nullptr))
    return kk::fail(KK_ERR_INVALID, "kk_recip_or_zero: bad arguments");
  if (n == 0) return KK_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(recip_or_zero_kernel, dim3(unsigned(blocks)), dim3(256),
                     0, static_cast<hipStream_t>(stream), ptr, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "recip_or_zero_kernel");
  return KK_OK;
}
