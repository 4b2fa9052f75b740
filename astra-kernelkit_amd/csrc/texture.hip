// Padded "texture" copies and fills.
//
// The reference samples CUDA texture objects with border addressing (zero
// outside, kernelkit/kernel.py:22-32) and copies volumes/projections into
// CUDA arrays on every set (kernel.py:91-94, projector.py:280-283, :529-542).
// Here the equivalent is one pass that writes a C-contiguous copy of the
// array, in the array's own axis order, with a zero halo of KK_HALO_LO
// elements before and KK_HALO_HI after every axis.  The projector kernels
// then clamp coordinates into the halo instead of testing bounds per tap.
#include "kk_internal.h"

namespace {

// One thread per texture element: x covers the fastest axis in blocks of
// 256, grid y/z enumerate the middle and slowest axes.
__global__ __launch_bounds__(256) void array_to_tex_kernel(
    const float* __restrict__ src, int64_t s0, int64_t s1, int64_t s2,
    int32_t n0, int32_t n1, int32_t n2, float* __restrict__ tex) {
  const int32_t t2n = n2 + kk::kHalo;
  const int32_t t1n = n1 + kk::kHalo;
  const int32_t t2 = blockIdx.x * 256 + threadIdx.x;
  const int32_t t1 = blockIdx.y;
  const int32_t t0 = blockIdx.z;
  if (t2 >= t2n) return;
  const int32_t i0 = t0 - kk::kHaloLo, i1 = t1 - kk::kHaloLo,
                i2 = t2 - kk::kHaloLo;
  float v = 0.f;
  if (i0 >= 0 && i0 < n0 && i1 >= 0 && i1 < n1 && i2 >= 0 && i2 < n2)
    v = src[i0 * s0 + i1 * s1 + i2 * s2];
  tex[(int64_t(t0) * t1n + t1) * t2n + t2] = v;
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ p,
                                                   int64_t n, float value) {
  const int64_t stride = int64_t(gridDim.x) * 256;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += stride)
    p[i] = value;
}

}  // namespace

extern "C" int kk_array_to_texture(const float* src, const int64_t sh[3],
                                   const int64_t st[3], float* tex,
                                   void* stream) {
  if (src == nullptr || tex == nullptr || sh == nullptr || st == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_array_to_texture: null argument");
  if (sh[0] <= 0 || sh[1] <= 0 || sh[2] <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_array_to_texture: bad shape");
  if (sh[0] + kk::kHalo > 65535 || sh[1] + kk::kHalo > 65535 ||
      sh[2] + kk::kHalo > (int64_t(1) << 30))
    return kk::fail(KK_ERR_UNSUPPORTED, "kk_array_to_texture: too large");
  const dim3 grid(unsigned((sh[2] + kk::kHalo + 255) / 256),
                  unsigned(sh[1] + kk::kHalo), unsigned(sh[0] + kk::kHalo));
  hipLaunchKernelGGL(array_to_tex_kernel, grid, dim3(256), 0,
                     static_cast<hipStream_t>(stream), src, st[0], st[1],
                     st[2], int32_t(sh[0]), int32_t(sh[1]), int32_t(sh[2]),
                     tex);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "array_to_tex_kernel");
  return KK_OK;
}

extern "C" int kk_fill(float* ptr, int64_t n, float value, void* stream) {
  if (n < 0 || (n > 0 && ptr == nullptr))
    return kk::fail(KK_ERR_INVALID, "kk_fill: bad arguments");
  if (n == 0) return KK_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel, dim3(unsigned(blocks)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), ptr, n, value);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "fill_kernel");
  return KK_OK;
}
