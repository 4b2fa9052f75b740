// Device memory for callers without PyTorch: allocation, copies, strided
// copies and stream synchronisation.
//
// The reference's arrays are CuPy arrays (kernelkit/kernel.py:47-119,
// kernelkit/projector.py:266-283), so its callers never touch the runtime
// directly.  These entry points let the Python drop-in own device buffers
// (kernelkit.DeviceArray) when no torch tensor is involved, and let any
// C caller do the same through this library alone.
#include "kk_internal.h"

namespace {

// dst[i, j, k] = src[i, j, k] for a 3-D index space of n0 x n1 x n2
// elements (element strides, any signs and orders).  One thread per
// element of the innermost axis, grid-stride over the rest.
template <class T>
__global__ __launch_bounds__(256) void copy3_kernel(
    const T* __restrict__ src, int64_t s0, int64_t s1, int64_t s2,
    T* __restrict__ dst, int64_t d0, int64_t d1, int64_t d2, int64_t n0,
    int64_t n1, int64_t n2) {
  const int64_t n = n0 * n1 * n2;
  const int64_t stride = int64_t(gridDim.x) * 256;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < n;
       t += stride) {
    const int64_t k = t % n2;
    const int64_t r = t / n2;
    const int64_t j = r % n1;
    const int64_t i = r / n1;
    dst[i * d0 + j * d1 + k * d2] = src[i * s0 + j * s1 + k * s2];
  }
}

}  // namespace

extern "C" int kk_device_count(int32_t* count) {
  if (count == nullptr) return kk::fail(KK_ERR_INVALID, "kk_device_count");
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return kk::hip_fail(e, "hipGetDeviceCount");
  }
  *count = n;
  return KK_OK;
}

extern "C" int kk_get_device(int32_t* device) {
  if (device == nullptr) return kk::fail(KK_ERR_INVALID, "kk_get_device");
  int d = 0;
  const hipError_t e = hipGetDevice(&d);
  if (e != hipSuccess) return kk::hip_fail(e, "hipGetDevice");
  *device = d;
  return KK_OK;
}

extern "C" int kk_set_device(int32_t device) {
  const hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return kk::hip_fail(e, "hipSetDevice");
  return KK_OK;
}

extern "C" int kk_malloc(void** ptr, int64_t bytes) {
  if (ptr == nullptr || bytes < 0)
    return kk::fail(KK_ERR_INVALID, "kk_malloc: bad arguments");
  *ptr = nullptr;
  if (bytes == 0) return KK_OK;
  const hipError_t e = hipMalloc(ptr, static_cast<size_t>(bytes));
  if (e != hipSuccess) {
    *ptr = nullptr;
    return kk::hip_fail(e, "hipMalloc");
  }
  return KK_OK;
}

extern "C" int kk_free(void* ptr) {
  if (ptr == nullptr) return KK_OK;
  const hipError_t e = hipFree(ptr);
  if (e != hipSuccess) return kk::hip_fail(e, "hipFree");
  return KK_OK;
}

extern "C" int kk_memcpy(void* dst, const void* src, int64_t bytes,
                         int32_t kind, void* stream) {
  if (bytes < 0 || (bytes > 0 && (dst == nullptr || src == nullptr)))
    return kk::fail(KK_ERR_INVALID, "kk_memcpy: bad arguments");
  if (bytes == 0) return KK_OK;
  hipMemcpyKind k;
  switch (kind) {
    case KK_COPY_H2D: k = hipMemcpyHostToDevice; break;
    case KK_COPY_D2H: k = hipMemcpyDeviceToHost; break;
    case KK_COPY_D2D: k = hipMemcpyDeviceToDevice; break;
    default: return kk::fail(KK_ERR_INVALID, "kk_memcpy: kind %d", kind);
  }
  const hipError_t e = hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), k,
                                      static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return kk::hip_fail(e, "hipMemcpyAsync");
  return KK_OK;
}

extern "C" int kk_stream_synchronize(void* stream) {
  const hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return kk::hip_fail(e, "hipStreamSynchronize");
  return KK_OK;
}

extern "C" int kk_copy_strided(const void* src, const int64_t src_strides[3],
                               void* dst, const int64_t dst_strides[3],
                               const int64_t shape[3], int32_t elem_bytes,
                               void* stream) {
  if (src_strides == nullptr || dst_strides == nullptr || shape == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_copy_strided: null argument");
  for (int t = 0; t < 3; ++t)
    if (shape[t] < 0)
      return kk::fail(KK_ERR_INVALID, "kk_copy_strided: negative shape");
  const int64_t n = shape[0] * shape[1] * shape[2];
  if (n == 0) return KK_OK;
  if (src == nullptr || dst == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_copy_strided: null pointer");
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (elem_bytes == 4) {
    hipLaunchKernelGGL((copy3_kernel<float>), dim3(unsigned(blocks)),
                       dim3(256), 0, st, static_cast<const float*>(src),
                       src_strides[0], src_strides[1], src_strides[2],
                       static_cast<float*>(dst), dst_strides[0],
                       dst_strides[1], dst_strides[2], shape[0], shape[1],
                       shape[2]);
  } else if (elem_bytes == 8) {
    hipLaunchKernelGGL((copy3_kernel<int64_t>), dim3(unsigned(blocks)),
                       dim3(256), 0, st, static_cast<const int64_t*>(src),
                       src_strides[0], src_strides[1], src_strides[2],
                       static_cast<int64_t*>(dst), dst_strides[0],
                       dst_strides[1], dst_strides[2], shape[0], shape[1],
                       shape[2]);
  } else {
    return kk::fail(KK_ERR_INVALID, "kk_copy_strided: element size %d",
                    elem_bytes);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "copy3_kernel");
  return KK_OK;
}
