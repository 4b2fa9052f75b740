// Texture builders and fills.
//
// The reference samples CUDA texture objects with border addressing (zero
// outside, kernelkit/kernel.py:22-32) and copies volumes/projections into
// CUDA arrays whenever they are set (kernel.py:91-94, projector.py:280-283,
// :529-542).  Here the equivalent copies build the layouts the projector
// kernels read (see kk_internal.h): zero-halo textures of float2 pairs
// (value, neighbour), so that the four taps of a bilinear sample are one
// 16-byte load and coordinates are clamped into the halo instead of being
// bounds-tested per tap.
#include <new>
#include <vector>

#include "kk_internal.h"

namespace {

// Texture rows per workgroup of the two builders below: a thread walks one
// texture column down kTexRows rows, carrying the previous source row's
// value (each source element is read once; a row's texels need rows r - 1
// and r).
constexpr int32_t kTexRows = 16;

// Frame texture of ray axis a: element (i, jb, jc) of a [na][nb+3][nc+3]
// float2 array = (V(i, jb-1, jc-1), V(i, jb, jc-1)), zero outside.
__global__ __launch_bounds__(256) void fp_frame_kernel(
    const float* __restrict__ vol, int64_t sa, int64_t sb, int64_t sc,
    int32_t nb, int32_t nc, float2* __restrict__ tex) {
  const int32_t tc = nc + kk::kHalo;
  const int32_t tb = nb + kk::kHalo;
  const int32_t jc = blockIdx.x * 256 + threadIdx.x;
  const int32_t jb0 = blockIdx.y * kTexRows;
  const int32_t i = blockIdx.z;
  if (jc >= tc) return;
  const int32_t c = jc - kk::kHaloLo;
  const bool col = c >= 0 && c < nc;
  const float* p = vol + i * sa + (col ? c : 0) * sc;
  auto value = [&](int32_t b) {
    return col && b >= 0 && b < nb ? p[b * sb] : 0.f;
  };
  const int32_t jb1 = min(jb0 + kTexRows, tb);
  float prev = value(jb0 - kk::kHaloLo);
  float2* out = tex + (int64_t(i) * tb + jb0) * tc + jc;
  for (int32_t jb = jb0; jb < jb1; ++jb, out += tc) {
    const float cur = value(jb - kk::kHaloLo + 1);
    *out = make_float2(prev, cur);
    prev = cur;
  }
}

// Projection texture: element (a, jv, ju) of a [n][rows+3][cols+3] float2
// array = (P(a, jv-1, ju-1), P(a, jv, ju-1)), zero outside.
// RESID: the texture of (proj - y) * R instead (the SIRT residual), with y
// and R addressed by their own strides.
struct ResidArgs {
  const float* y;
  int64_t y0, y1, y2;
  const float* R;
  int64_t r0, r1, r2;
};

template <bool RESID>
__global__ __launch_bounds__(256) void bp_tex_kernel(
    const float* __restrict__ proj, int64_t s0, int64_t s1, int64_t s2,
    int32_t rows, int32_t cols, float2* __restrict__ tex, ResidArgs res) {
  const int32_t tc = cols + kk::kHalo;
  const int32_t tr = rows + kk::kHalo;
  const int32_t ju = blockIdx.x * 256 + threadIdx.x;
  const int32_t jv0 = blockIdx.y * kTexRows;
  const int32_t a = blockIdx.z;
  if (ju >= tc) return;
  const int32_t c = ju - kk::kHaloLo;
  const bool col = c >= 0 && c < cols;
  const int32_t cc = col ? c : 0;
  auto value = [&](int32_t r) {
    if (!col || r < 0 || r >= rows) return 0.f;
    float v = proj[a * s0 + r * s1 + cc * s2];
    if constexpr (RESID)
      v = (v - res.y[a * res.y0 + r * res.y1 + cc * res.y2]) *
          res.R[a * res.r0 + r * res.r1 + cc * res.r2];
    return v;
  };
  const int32_t jv1 = min(jv0 + kTexRows, tr);
  float prev = value(jv0 - kk::kHaloLo);
  float2* out = tex + (int64_t(a) * tr + jv0) * tc + ju;
  for (int32_t jv = jv0; jv < jv1; ++jv, out += tc) {
    const float cur = value(jv - kk::kHaloLo + 1);
    *out = make_float2(prev, cur);
    prev = cur;
  }
}

// Many texture builds in one launch: a grid-stride loop over the jobs'
// tiles (one image's band of kTexRows texture rows by 256 columns, column
// chunks fastest), each found by binary search over the jobs' first tile
// ids (wave-uniform).  A thread walks one column of its tile like
// fp_frame_kernel / bp_tex_kernel, with the same arithmetic per element.
__global__ __launch_bounds__(256) void tex_jobs_kernel(
    const kk::TexJob* __restrict__ jobs, int32_t n_jobs, int64_t total) {
  for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
    int32_t lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (jobs[mid].first <= g)
        lo = mid;
      else
        hi = mid - 1;
    }
    const kk::TexJob& J = jobs[lo];
    const int32_t tr = J.n1 + kk::kHalo, tc = J.n2 + kk::kHalo;
    const int32_t nch = (tc + 255) / 256;
    const int32_t nbands = (tr + kTexRows - 1) / kTexRows;
    const int64_t local = g - J.first;
    const int32_t chunk = static_cast<int32_t>(local % nch);
    const int64_t t = local / nch;
    const int32_t jv0 = static_cast<int32_t>(t % nbands) * kTexRows;
    const int32_t i = static_cast<int32_t>(t / nbands);
    const int32_t ju = chunk * 256 + static_cast<int32_t>(threadIdx.x);
    if (ju >= tc) continue;
    const int32_t c = ju - kk::kHaloLo;
    const bool col = c >= 0 && c < J.n2;
    const int64_t cc = col ? c : 0;
    const bool res = J.y != nullptr;
    const float* sp = J.src + i * J.s0 + cc * J.s2;
    const float* yp = res ? J.y + i * J.y0 + cc * J.y2 : nullptr;
    const float* rp = res ? J.R + i * J.r0 + cc * J.r2 : nullptr;
    auto value = [&](int32_t r) {
      if (!col || r < 0 || r >= J.n1) return 0.f;
      float v = sp[r * J.s1];
      if (res) v = (v - yp[r * J.y1]) * rp[r * J.r1];
      return v;
    };
    const int32_t jv1 = min(jv0 + kTexRows, tr);
    float prev = value(jv0 - kk::kHaloLo);
    float2* out = J.tex + (int64_t(i) * tr + jv0) * tc + ju;
    for (int32_t jv = jv0; jv < jv1; ++jv, out += tc) {
      const float cur = value(jv - kk::kHaloLo + 1);
      *out = make_float2(prev, cur);
      prev = cur;
    }
  }
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ p,
                                                   int64_t n, float value) {
  const int64_t stride = int64_t(gridDim.x) * 256;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += stride)
    p[i] = value;
}

}  // namespace

extern "C" int kk_volume_to_fp_texture(const float* vol,
                                       const int32_t shape[3],
                                       const int64_t st[3], int32_t mask,
                                       float* tex, void* stream) {
  if (vol == nullptr || tex == nullptr || shape == nullptr || st == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_volume_to_fp_texture: null");
  if (shape[0] <= 0 || shape[1] <= 0 || shape[2] <= 0 || mask <= 0 ||
      mask > 7)
    return kk::fail(KK_ERR_INVALID, "kk_volume_to_fp_texture: bad args");
  int64_t off = 0;  // in float2 elements
  for (int a = 0; a < 3; ++a) {
    if (!(mask & (1 << a))) continue;
    const int b = kk::kFrameB[a], c = kk::kFrameC[a];
    if (shape[b] + kk::kHalo > 65535 || shape[a] > 65535)
      return kk::fail(KK_ERR_UNSUPPORTED, "kk_volume_to_fp_texture: large");
    const dim3 grid(unsigned((shape[c] + kk::kHalo + 255) / 256),
                    unsigned((shape[b] + kk::kHalo + kTexRows - 1) /
                             kTexRows),
                    unsigned(shape[a]));
    hipLaunchKernelGGL(fp_frame_kernel, grid, dim3(256), 0,
                       static_cast<hipStream_t>(stream), vol, st[a], st[b],
                       st[c], shape[b], shape[c],
                       reinterpret_cast<float2*>(tex) + off);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kk::hip_fail(e, "fp_frame_kernel");
    off += kk::frame_elems(shape, a);
  }
  return KK_OK;
}

extern "C" int kk_projections_to_bp_texture(const float* proj, int32_t n,
                                            int32_t rows, int32_t cols,
                                            const int64_t st[3], float* tex,
                                            void* stream) {
  if (proj == nullptr || tex == nullptr || st == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_projections_to_bp_texture: null");
  if (n <= 0 || rows <= 0 || cols <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_projections_to_bp_texture: shape");
  if (rows + kk::kHalo > 65535 || n > 65535)
    return kk::fail(KK_ERR_UNSUPPORTED,
                    "kk_projections_to_bp_texture: too large");
  const dim3 grid(unsigned((cols + kk::kHalo + 255) / 256),
                  unsigned((rows + kk::kHalo + kTexRows - 1) / kTexRows),
                  unsigned(n));
  hipLaunchKernelGGL(bp_tex_kernel<false>, grid, dim3(256), 0,
                     static_cast<hipStream_t>(stream), proj, st[0], st[1],
                     st[2], rows, cols, reinterpret_cast<float2*>(tex),
                     ResidArgs{});
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "bp_tex_kernel");
  return KK_OK;
}

extern "C" int kk_residual_to_bp_texture(
    const float* proj, const int64_t ps[3], const float* y,
    const int64_t ys[3], const float* R, const int64_t rs[3],
    int32_t n_angles, int32_t rows, int32_t cols, float* tex, void* stream) {
  if (proj == nullptr || ps == nullptr || y == nullptr || ys == nullptr ||
      R == nullptr || rs == nullptr || tex == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_residual_to_bp_texture: null");
  if (n_angles <= 0 || rows <= 0 || cols <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_residual_to_bp_texture: bad sizes");
  if (n_angles > 65535 || rows + kk::kHalo > 65535)
    return kk::fail(KK_ERR_UNSUPPORTED,
                    "kk_residual_to_bp_texture: too large");
  const dim3 grid(unsigned((cols + kk::kHalo + 255) / 256),
                  unsigned((rows + kk::kHalo + kTexRows - 1) / kTexRows),
                  unsigned(n_angles));
  const ResidArgs res{y, ys[0], ys[1], ys[2], R, rs[0], rs[1], rs[2]};
  hipLaunchKernelGGL(bp_tex_kernel<true>, grid, dim3(256), 0,
                     static_cast<hipStream_t>(stream), proj, ps[0], ps[1],
                     ps[2], rows, cols, reinterpret_cast<float2*>(tex), res);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "bp_tex_kernel");
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

namespace kk {

struct TexJobs {
  TexJob* jobs = nullptr;
  int32_t n = 0;
  int64_t total = 0;  // tiles of all jobs (see tex_jobs_kernel)
};

int tex_jobs_build(std::vector<TexJob> jobs, TexJobs** out) {
  if (jobs.empty()) return fail(KK_ERR_INVALID, "empty texture batch");
  int64_t total = 0;
  for (TexJob& J : jobs) {
    if (J.src == nullptr || J.tex == nullptr || J.n0 <= 0 || J.n1 <= 0 ||
        J.n2 <= 0)
      return fail(KK_ERR_INVALID, "texture batch: bad job");
    if (int64_t(J.n0) * (J.n1 + kHalo) >= (int64_t(1) << 31))
      return fail(KK_ERR_UNSUPPORTED, "texture batch: job too large");
    J.first = total;
    total += int64_t(J.n0) * ((J.n1 + kHalo + kTexRows - 1) / kTexRows) *
             ((J.n2 + kHalo + 255) / 256);
  }
  auto* b = new (std::nothrow) TexJobs();
  if (b == nullptr) return fail(KK_ERR_INVALID, "out of host memory");
  b->n = static_cast<int32_t>(jobs.size());
  b->total = total;
  hipError_t e = hipMalloc(&b->jobs, sizeof(TexJob) * jobs.size());
  if (e == hipSuccess)
    e = hipMemcpy(b->jobs, jobs.data(), sizeof(TexJob) * jobs.size(),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tex_jobs_free(b);
    return hip_fail(e, "texture batch descriptors");
  }
  *out = b;
  return KK_OK;
}

void tex_jobs_free(TexJobs* b) {
  if (b == nullptr) return;
  if (b->jobs) hipFree(b->jobs);
  delete b;
}

int tex_jobs_launch(const TexJobs* b, hipStream_t stream) {
  if (b == nullptr) return fail(KK_ERR_NOT_SET, "texture batch not built");
  // a tile per workgroup, at most 1024 workgroups per CU of the 256
  const int64_t blocks = b->total < 262144 ? b->total : 262144;
  hipLaunchKernelGGL(tex_jobs_kernel, dim3(static_cast<unsigned>(blocks)),
                     dim3(256), 0, stream, b->jobs, b->n, b->total);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "tex_jobs_kernel");
  return KK_OK;
}

}  // namespace kk
