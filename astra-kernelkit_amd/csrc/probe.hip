// Measurement probe: aggregate L2 read bandwidth of the device.
//
// Not part of the reference's interface.  Every workgroup streams float4
// loads over a buffer small enough to stay resident in each XCD's L2
// (e.g. 2 MiB against 4 MiB) but far larger than a CU's L1, so the
// loads are served by L2.  bench.py divides the bytes read by the time to
// obtain the L2 read rate that bounds the gather kernels.
#include "kk_internal.h"

namespace {

constexpr int kUnroll = 8;

__global__ __launch_bounds__(256) void l2_read_kernel(
    const float4* __restrict__ buf, uint32_t mask, uint32_t iters,
    float* __restrict__ out) {
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float4 acc[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) acc[u] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t it = 0; it < iters; ++it) {
    float4 v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      v[u] = buf[(i + static_cast<uint32_t>(u) * stride) & mask];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      acc[u].x += v[u].x;
      acc[u].y += v[u].y;
      acc[u].z += v[u].z;
      acc[u].w += v[u].w;
    }
    i += kUnroll * stride;
  }
  float s = 0.f;
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) s += acc[u].x + acc[u].y + acc[u].z + acc[u].w;
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" int kk_probe_l2_read(const float* buf, int64_t n_floats,
                                int32_t blocks, int32_t iters, float* out,
                                void* stream) {
  if (buf == nullptr || out == nullptr || blocks <= 0 || iters <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_probe_l2_read: bad argument");
  const int64_t n_vec = n_floats / 4;
  if (n_vec < 1024 || (n_vec & (n_vec - 1)) != 0 || n_vec > (int64_t(1) << 31))
    return kk::fail(KK_ERR_INVALID,
                    "kk_probe_l2_read: n_floats / 4 must be a power of two "
                    "in [1024, 2^31]");
  if (reinterpret_cast<uintptr_t>(buf) % 16 != 0)
    return kk::fail(KK_ERR_INVALID, "kk_probe_l2_read: buffer not 16-byte "
                                    "aligned");
  hipLaunchKernelGGL(l2_read_kernel, dim3(static_cast<unsigned>(blocks)),
                     dim3(256), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const float4*>(buf),
                     static_cast<uint32_t>(n_vec - 1),
                     static_cast<uint32_t>(iters), out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "l2_read_kernel");
  return KK_OK;
}
