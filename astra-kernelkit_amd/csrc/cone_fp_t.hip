// Exact transpose of the ray-driven (Joseph) forward projection: the
// matched adjoint A^T of kk_cone_fp, for gradients.
//
// The reference has no such operator: its autograd backward uses the
// voxel-driven backprojector (kernelkit/torch_support.py:29-33), which is
// not the transpose of the forward projector.  Here every (ray, slice)
// sample of the forward kernel, with the same arithmetic (kk_fp_ray,
// kk_fp_weight, kk_fp_coord: ray slopes, intercepts, clamping into the zero
// halo, tap index and weights), adds value * scale * bilinear weight to
// each of its four taps that lie inside the volume, so
// <A x, y> = <x, A^T y> up to rounding.
//
// MI355X design: a scatter with one global float atomic per tap would put
// ~4 atomics per sample on the L2.  Instead each workgroup owns a tile of
// the volume in the ray frame of one axis (kTS slices x kTB x kTC voxels
// plus a halo, 37 KiB of LDS) and accumulates it in LDS over every angle
// of that axis: per angle it projects the tile's corners onto the detector,
// walks the rays of the covering pixel rectangle through the tile's slices
// and adds their taps with LDS float atomics (ds_add_f32).  The tile is
// written to HBM once.  Angles of different ray axes run as separate
// launches over tiles of their own frame, the later ones adding.
#include "kk_internal.h"

namespace {

constexpr int kTS = 8;    // slices (ray axis a) per tile
constexpr int kTB = 32;   // voxels along b per tile
constexpr int kTC = 32;   // voxels along c per tile
// The LDS tile has a one-voxel halo around the b and c extents: taps that
// fall outside the tile are clamped into it and never written back, so the
// sample loop has no bounds tests.  Row stride kTCP (odd): the lanes of a
// wavefront (rays of consecutive detector columns, i.e. consecutive b rows)
// hit different LDS banks.
constexpr int kTBH = kTB + 2;
constexpr int kTCP = kTC + 3;
constexpr int kThreads = 256;

struct FptArgs {
  const float* proj;
  int64_t ps0, ps1, ps2;  // projection strides: angle, row, col
  float* vol;
  int64_t sa, sb, sc;     // volume strides along the frame axes a, b, c
  int32_t rows, cols;
  int32_t na, nb, nc;
  int32_t tiles_b, tiles_c;
  const int32_t* angles;  // angle ids of this ray axis
  int32_t n_angles;
  int32_t mode;           // 0: vol = A^T y, 1: vol += A^T y
};

// Detector coordinates (col, row) in pixel-centre units of the ray from
// the source through the source-relative point X (frame components): the
// solution of q + col U + row V = lam X; ok iff finite with lam > 0.
__device__ __forceinline__ void project(const kk::FpAngle& P, double xa,
                                        double xb, double xc, float& col,
                                        float& row, bool& ok) {
  // V x X, q x X, V x q
  const double vx0 = P.vb * xc - P.vc * xb;
  const double vx1 = P.vc * xa - P.va * xc;
  const double vx2 = P.va * xb - P.vb * xa;
  const double qx0 = P.qb * xc - P.qc * xb;
  const double qx1 = P.qc * xa - P.qa * xc;
  const double qx2 = P.qa * xb - P.qb * xa;
  const double vq0 = P.vb * P.qc - P.vc * P.qb;
  const double vq1 = P.vc * P.qa - P.va * P.qc;
  const double vq2 = P.va * P.qb - P.vb * P.qa;
  const double den = P.ua * vx0 + P.ub * vx1 + P.uc * vx2;
  const double c = -(P.qa * vx0 + P.qb * vx1 + P.qc * vx2) / den;
  const double r = -(P.ua * qx0 + P.ub * qx1 + P.uc * qx2) / den;
  const double lam = (P.ua * vq0 + P.ub * vq1 + P.uc * vq2) / den;
  ok = isfinite(c) && isfinite(r) && lam > 0.0 && fabs(c) < 1e8 &&
       fabs(r) < 1e8;
  col = static_cast<float>(ok ? c : 0.0);
  row = static_cast<float>(ok ? r : 0.0);
}

__global__ __launch_bounds__(kThreads) void cone_fp_t_kernel(
    const kk::FpAngle* __restrict__ tab, FptArgs A) {
  __shared__ float acc[kTS * kTBH * kTCP];
  int64_t bid = blockIdx.x;
  const int32_t tc = static_cast<int32_t>(bid % A.tiles_c);
  bid /= A.tiles_c;
  const int32_t tb = static_cast<int32_t>(bid % A.tiles_b);
  const int32_t ta = static_cast<int32_t>(bid / A.tiles_b);
  const int32_t i0 = ta * kTS, j0 = tb * kTB, k0 = tc * kTC;
  const int32_t ns = min(kTS, A.na - i0);
  const int32_t nbt = min(kTB, A.nb - j0);
  const int32_t nct = min(kTC, A.nc - k0);
  const int tid = threadIdx.x;
  const int lane = tid & 63;

  for (int e = tid; e < kTS * kTBH * kTCP; e += kThreads) acc[e] = 0.f;
  __syncthreads();

  for (int32_t q = 0; q < A.n_angles; ++q) {
    const int32_t ang = A.angles[q];
    const kk::FpAngle& P = tab[ang];
    // Pixel rectangle covering the rays whose samples can reach the tile:
    // samples at slices [i0, i0+ns) with padded coordinates in
    // [j0, j0+nbt+1] x [k0, k0+nct+1].  Lanes 0..7 of each wave take one
    // corner of that box each.
    float cmin, cmax, rmin, rmax;
    bool ok;
    {
      const int cl = lane & 7;
      const double xa =
          static_cast<double>(cl & 1 ? i0 + ns - 1 : i0) + P.sig;
      const double xb =
          static_cast<double>(cl & 2 ? j0 + nbt + 1 : j0) - P.taub;
      const double xc =
          static_cast<double>(cl & 4 ? k0 + nct + 1 : k0) - P.tauc;
      float c, r;
      project(P, xa, xb, xc, c, r, ok);
      cmin = cmax = c;
      rmin = rmax = r;
#pragma unroll
      for (int o = 1; o < 8; o <<= 1) {
        cmin = fminf(cmin, __shfl_xor(cmin, o));
        cmax = fmaxf(cmax, __shfl_xor(cmax, o));
        rmin = fminf(rmin, __shfl_xor(rmin, o));
        rmax = fmaxf(rmax, __shfl_xor(rmax, o));
        ok = ok && __shfl_xor(static_cast<int>(ok), o);
      }
    }
    int32_t c_lo = 0, c_hi = A.cols - 1, r_lo = 0, r_hi = A.rows - 1;
    if (__builtin_amdgcn_readfirstlane(static_cast<int>(ok))) {
      const float fcmin = __int_as_float(
          __builtin_amdgcn_readfirstlane(__float_as_int(cmin)));
      const float fcmax = __int_as_float(
          __builtin_amdgcn_readfirstlane(__float_as_int(cmax)));
      const float frmin = __int_as_float(
          __builtin_amdgcn_readfirstlane(__float_as_int(rmin)));
      const float frmax = __int_as_float(
          __builtin_amdgcn_readfirstlane(__float_as_int(rmax)));
      // one pixel of margin against rounding in the corner projection;
      // the clamps keep the float -> int conversions in range
      c_lo = max(c_lo, static_cast<int32_t>(floorf(kk_clamp(fcmin, -4.f,
                                                            1e8f))) - 1);
      c_hi = min(c_hi, static_cast<int32_t>(ceilf(kk_clamp(fcmax, -4.f,
                                                           1e8f))) + 1);
      r_lo = max(r_lo, static_cast<int32_t>(floorf(kk_clamp(frmin, -4.f,
                                                            1e8f))) - 1);
      r_hi = min(r_hi, static_cast<int32_t>(ceilf(kk_clamp(frmax, -4.f,
                                                           1e8f))) + 1);
    }
    if (c_lo > c_hi || r_lo > r_hi) continue;
    const int32_t W = c_hi - c_lo + 1;
    const int32_t npx = W * (r_hi - r_lo + 1);
    const double nb1 = static_cast<double>(A.nb + 1);
    const double nc1 = static_cast<double>(A.nc + 1);
    for (int32_t t = tid; t < npx; t += kThreads) {
      const int32_t row = r_lo + t / W;
      const int32_t col = c_lo + t % W;
      // ray parameters and samples exactly as in the forward kernel
      double a1, b1, a2, b2;
      if (!kk_fp_ray(P, row, col, a1, b1, a2, b2)) continue;
      const float val = A.proj[ang * A.ps0 + row * A.ps1 + col * A.ps2] *
                        kk_fp_weight(P, a1, a2);
      if (val == 0.f) continue;
      for (int32_t s = 0; s < ns; ++s) {
        const double fi = static_cast<double>(i0 + s);
        uint32_t ixb, ixc;
        float fb, fc;
        kk_fp_coord(a1, b1, fi, nb1, ixb, fb);
        kk_fp_coord(a2, b2, fi, nc1, ixc, fc);
        // tile-local index of the taps along b and c, +1 for the halo; a
        // tap outside the tile is clamped into the halo (index 0 or
        // kTB + 1 / kTC + 1), which is never written back (valid voxels are
        // 1..nbt, 1..nct)
        const int32_t ub = static_cast<int32_t>(ixb) - kk::kHaloLo - j0 + 1;
        const int32_t uc = static_cast<int32_t>(ixc) - kk::kHaloLo - k0 + 1;
        const int32_t tb0 = min(max(ub, 0), kTB + 1);
        const int32_t tb1 = min(max(ub + 1, 0), kTB + 1);
        const int32_t tc0 = min(max(uc, 0), kTC + 1);
        const int32_t tc1 = min(max(uc + 1, 0), kTC + 1);
        const float vb0 = val * (1.f - fb), vb1 = val * fb;
        float* base = acc + s * (kTBH * kTCP);
        atomicAdd(base + tb0 * kTCP + tc0, vb0 * (1.f - fc));
        atomicAdd(base + tb1 * kTCP + tc0, vb1 * (1.f - fc));
        atomicAdd(base + tb0 * kTCP + tc1, vb0 * fc);
        atomicAdd(base + tb1 * kTCP + tc1, vb1 * fc);
      }
    }
  }
  __syncthreads();

  const int32_t plane = nbt * nct;
  for (int32_t e = tid; e < ns * plane; e += kThreads) {
    const int32_t s = e / plane;
    const int32_t r = e - s * plane;
    const int32_t j = r / nct;
    const int32_t k = r - j * nct;
    const float v = acc[(s * kTBH + j + 1) * kTCP + k + 1];
    const int64_t o = (i0 + s) * A.sa + (j0 + j) * A.sb + (k0 + k) * A.sc;
    if (A.mode == 1)
      A.vol[o] += v;
    else
      A.vol[o] = v;
  }
}

}  // namespace

extern "C" int kk_cone_fp_transpose(const kk_plan* plan, const float* proj,
                                    const int64_t ps[3], int32_t rows,
                                    int32_t cols, float* vol,
                                    const int32_t vol_shape[3],
                                    const int64_t vs[3], int32_t additive,
                                    void* stream) {
  if (plan == nullptr || !plan->fp_set)
    return kk::fail(KK_ERR_NOT_SET,
                    "Please set the projection geometry with `set_params()`.");
  if (proj == nullptr || ps == nullptr || vol == nullptr ||
      vol_shape == nullptr || vs == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_cone_fp_transpose: null argument");
  if (rows <= 0 || cols <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_cone_fp_transpose: bad detector");
  for (int t = 0; t < 3; ++t)
    if (vol_shape[t] != plan->fp_shape[t])
      return kk::fail(KK_ERR_SHAPE,
                      "kk_cone_fp_transpose: volume shape differs from the "
                      "one the parameters were set for");
  bool first = true;
  for (int a = 0; a < 3; ++a) {
    const int32_t n = plan->fr_start[a + 1] - plan->fr_start[a];
    if (n == 0) continue;
    const int b = kk::kFrameB[a], c = kk::kFrameC[a];
    FptArgs A{};
    A.proj = proj;
    A.ps0 = ps[0];
    A.ps1 = ps[1];
    A.ps2 = ps[2];
    A.vol = vol;
    A.sa = vs[a];
    A.sb = vs[b];
    A.sc = vs[c];
    A.rows = rows;
    A.cols = cols;
    A.na = vol_shape[a];
    A.nb = vol_shape[b];
    A.nc = vol_shape[c];
    A.tiles_b = (A.nb + kTB - 1) / kTB;
    A.tiles_c = (A.nc + kTC - 1) / kTC;
    const int64_t tiles_a = (A.na + kTS - 1) / kTS;
    const int64_t blocks = tiles_a * A.tiles_b * A.tiles_c;
    if (blocks >= (int64_t(1) << 31))
      return kk::fail(KK_ERR_UNSUPPORTED, "kk_cone_fp_transpose: grid");
    A.angles = plan->fr_idx + plan->fr_start[a];
    A.n_angles = n;
    A.mode = (first && !additive) ? 0 : 1;
    first = false;
    hipLaunchKernelGGL(cone_fp_t_kernel, dim3(unsigned(blocks)),
                       dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                       plan->fp, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kk::hip_fail(e, "cone_fp_t_kernel");
  }
  return KK_OK;
}
