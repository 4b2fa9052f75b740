// Measurement probes: read rates that bound the gather kernels.
//
// Not part of the reference's interface.
//  * l2_read_kernel: every workgroup streams float4 loads over a buffer
//    small enough to stay resident in each XCD's L2 (e.g. 2 MiB against
//    4 MiB) but far larger than a CU's L1, so the loads are served by L2.
//  * gather_kernel: the projectors' own access shape.  Every lane issues
//    16-byte loads at 8-byte alignment (two neighbouring float2 texels, the
//    four bilinear taps of one sample); the 64 lanes of a wave form an 8 x 8
//    tile with lane steps sx (across) and sy (down) in float2 units; every
//    load moves the wave's tile to a pseudo-random base inside a footprint
//    of 2^k float2.  A footprint that fits a CU's L1 measures the L1 gather
//    ceiling, one that fits each XCD's L2 the L2 gather ceiling.
// bench.py divides the bytes read by the time.
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

constexpr int kGatherLoads = 8;   // loads in flight per lane and iteration

typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));

__global__ __launch_bounds__(256) void gather_kernel(
    const float2* __restrict__ buf, uint32_t mask, uint32_t sx, uint32_t sy,
    uint32_t iters, float* __restrict__ out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t toff = (lane & 7u) * sx + (lane >> 3) * sy;
  // wave-uniform pseudo-random walk (an LCG), as the projectors' per-wave
  // texel base
  uint32_t h = __builtin_amdgcn_readfirstlane(
      ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 2654435761u + 12345u);
  float acc = 0.f;
  for (uint32_t it = 0; it < iters; ++it) {
    f4a8 v[kGatherLoads];
#pragma unroll
    for (int u = 0; u < kGatherLoads; ++u) {
      h = h * 1664525u + 1013904223u;
      const uint32_t idx = ((h >> 7) + toff) & mask;   // idx + 1 <= mask + 1
      v[u] = *reinterpret_cast<const f4a8*>(buf + idx);
    }
#pragma unroll
    for (int u = 0; u < kGatherLoads; ++u)
      acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace

extern "C" int kk_probe_gather(const float* buf, int64_t n_floats,
                               int32_t footprint_log2, int32_t sx, int32_t sy,
                               int32_t blocks, int32_t iters, float* out,
                               void* stream) {
  if (buf == nullptr || out == nullptr || blocks <= 0 || iters <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_probe_gather: bad argument");
  if (footprint_log2 < 6 || footprint_log2 > 30)
    return kk::fail(KK_ERR_INVALID,
                    "kk_probe_gather: footprint_log2 must be in [6, 30]");
  const int64_t fp2 = int64_t(1) << footprint_log2;   // float2 texels
  if (n_floats / 2 < fp2 + 1)
    return kk::fail(KK_ERR_INVALID,
                    "kk_probe_gather: buffer must hold 2^footprint_log2 + 1 "
                    "float2");
  if (sx < 0 || sy < 0 || sx > (1 << 20) || sy > (1 << 20))
    return kk::fail(KK_ERR_INVALID, "kk_probe_gather: steps must be in "
                                    "[0, 2^20]");
  if (reinterpret_cast<uintptr_t>(buf) % 8 != 0)
    return kk::fail(KK_ERR_INVALID, "kk_probe_gather: buffer not 8-byte "
                                    "aligned");
  hipLaunchKernelGGL(gather_kernel, dim3(static_cast<unsigned>(blocks)),
                     dim3(256), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const float2*>(buf),
                     static_cast<uint32_t>(fp2 - 1), static_cast<uint32_t>(sx),
                     static_cast<uint32_t>(sy), static_cast<uint32_t>(iters),
                     out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "gather_kernel");
  return KK_OK;
}

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
