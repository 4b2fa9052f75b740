// Voxel-driven cone-beam backprojection for gfx950.
//
// Behaviour follows cone_bp of the reference (kernelkit/cuda/cone_bp.cu:
// 35-159, launched by kernelkit/kernels/cone_bp.py:292-302): for every voxel
// centre X and angle, r = 1/den(X), (u, v) = (numU(X), numV(X)) * r, the
// projection is sampled bilinearly at (u, v) with zero outside, and
// r^2 * sample is accumulated; the sum is scaled by the voxel volume.
//
// MI355X design:
// * one launch covers all angles (the reference launches per batch of 32
//   angles and read-modify-writes the volume each time); each voxel is
//   written once, or updated in place by the fused SIRT epilogue;
// * a thread owns K consecutive voxels along z in registers.  For a circular
//   orbit about z the column coordinate u and the weight 1/den do not depend
//   on z, so the reciprocal, u and its interpolation weights are computed
//   once per (x, y, angle) and shared by the K voxels; rows advance by a
//   constant step.  Other geometries take the general per-voxel path (a
//   wave-uniform branch per angle);
// * projections are read from a zero-halo texture [angle][row][col]: a
//   sample is a clamp, two 8-byte loads (column pairs of two rows) and
//   three lerps, with no bounds tests;
// * lanes of a wavefront are a 16 x 4 (x, y) patch, which projects onto a
//   compact detector patch; workgroups are ordered z-chunk -> y -> x and
//   grouped per XCD so each L2 sees a contiguous part of the (x, y) plane.
#include "kk_internal.h"

namespace {

struct BpArgs {
  const float* tex;  // padded projection texture, element [0][0][0]
  int64_t tas;       // texture stride of one angle
  int32_t trs, tcs;  // texture strides of one row / one column
  int32_t nx, ny, nz, n_angles, rows, cols;
  float* out;
  int64_t vs0, vs1, vs2;  // output world strides
  float scale;
  int32_t mode;  // 0: out = b, 1: out += b, 2: SIRT update
  int32_t tiles_x, tiles_y, tiles_z;
  int32_t xcd_group;
  int64_t total_blocks;
  // mode 2: x <- clip(mask * (x - C * b)), optionally mirrored into xtex
  const float* C;
  int64_t cs0, cs1, cs2;
  const float* mask;
  int64_t ms0, ms1, ms2;
  float lo, hi;
  int32_t clip_lo, clip_hi;
  float* xtex;  // FP texture of x at voxel (0,0,0) (interior), or null
  int64_t xts0, xts1, xts2;  // its world strides
};

// Two horizontally adjacent detector taps (columns u0, u0+1) at a
// wave-uniform base plus a per-lane byte offset.
template <bool PAIR>
__device__ __forceinline__ kk_float2u col_pair(const char* base, uint32_t off,
                                               uint32_t tcb) {
  if constexpr (PAIR) {
    return kk_ld2(base, off);
  } else {
    kk_float2u t;
    t.x = kk_ld1(base, off);
    t.y = kk_ld1(base + tcb, off);
    return t;
  }
}

template <int K, bool PAIR>
__global__ __launch_bounds__(256) void cone_bp_kernel(
    const kk::BpAngle* __restrict__ tab, BpArgs A) {
  const int64_t bid =
      kk_xcd_remap(blockIdx.x, A.total_blocks, A.xcd_group);
  const int32_t tx = static_cast<int32_t>(bid % A.tiles_x);
  const int64_t t = bid / A.tiles_x;
  const int32_t ty = static_cast<int32_t>(t % A.tiles_y);
  const int32_t tz = static_cast<int32_t>(t / A.tiles_y);
  const int32_t ix = tx * 16 + int32_t(threadIdx.x & 15);
  const int32_t iy = ty * 16 + int32_t(threadIdx.x >> 4);
  const int32_t iz0 = tz * K;
  if (ix >= A.nx || iy >= A.ny) return;

  const float X = static_cast<float>(ix) - 0.5f * A.nx + 0.5f;
  const float Y = static_cast<float>(iy) - 0.5f * A.ny + 0.5f;
  const float Z0 = static_cast<float>(iz0) - 0.5f * A.nz + 0.5f;
  const float rmax = static_cast<float>(A.rows + 1);
  const float cmax = static_cast<float>(A.cols + 1);
  const uint32_t trb = 4u * static_cast<uint32_t>(A.trs);
  const uint32_t tcb = PAIR ? 4u : 4u * static_cast<uint32_t>(A.tcs);

  float acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.f;

  // padded image of angle 0; the halo offset is folded into the forms
  const char* img = reinterpret_cast<const char*>(A.tex + A.tas);
  const size_t tasb = 4 * static_cast<size_t>(A.tas);
  for (int32_t a = 0; a < A.n_angles; ++a, img += tasb) {
    const kk::BpAngle& P = tab[a];
    const char* img_r = img + trb;  // next detector row, uniform
    const float den = fmaf(Y, P.dy, fmaf(X, P.dx, P.dw));
    const float nu = fmaf(Y, P.nuy, fmaf(X, P.nux, P.nuw));
    const float nv = fmaf(Y, P.nvy, fmaf(X, P.nvx, P.nvw));
    if (P.fast) {
      const float r = __builtin_amdgcn_rcpf(den);
      const float un = nu * r;
      const float v0 = fmaf(Z0, P.nvz, nv) * r;
      const float gv = P.nvz * r;
      // a voxel in the plane of the source parallel to the detector
      // (den = 0) has no projection: skip it
      if (!(isfinite(un) && isfinite(v0) && isfinite(gv))) continue;
      const float w = r * r;
      const float u = kk_clamp(un, 0.f, cmax);
      const float fu = __builtin_amdgcn_fractf(u);
      const uint32_t offu = __umul24(static_cast<uint32_t>(u), tcb);
      uint32_t off[K];
      float fv[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float v = kk_clamp(fmaf(static_cast<float>(k), gv, v0), 0.f,
                                 rmax);
        fv[k] = __builtin_amdgcn_fractf(v);
        off[k] = __umul24(static_cast<uint32_t>(v), trb) + offu;
      }
      kk_float2u t0[K], t1[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        t0[k] = col_pair<PAIR>(img, off[k], tcb);
        t1[k] = col_pair<PAIR>(img_r, off[k], tcb);
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float h0 = fmaf(fu, t0[k].y - t0[k].x, t0[k].x);
        const float h1 = fmaf(fu, t1[k].y - t1[k].x, t1[k].x);
        acc[k] = fmaf(w, fmaf(fv[k], h1 - h0, h0), acc[k]);
      }
    } else {
      uint32_t off[K];
      float fu[K], fv[K], w[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float Z = Z0 + static_cast<float>(k);
        const float r = __builtin_amdgcn_rcpf(fmaf(Z, P.dz, den));
        const float un = fmaf(Z, P.nuz, nu) * r;
        const float vn = fmaf(Z, P.nvz, nv) * r;
        const bool ok = isfinite(un) && isfinite(vn);
        const float u = kk_clamp(ok ? un : 0.f, 0.f, cmax);
        const float v = kk_clamp(ok ? vn : 0.f, 0.f, rmax);
        fu[k] = __builtin_amdgcn_fractf(u);
        fv[k] = __builtin_amdgcn_fractf(v);
        off[k] = __umul24(static_cast<uint32_t>(v), trb) +
                 __umul24(static_cast<uint32_t>(u), tcb);
        w[k] = ok ? r * r : 0.f;
      }
      kk_float2u t0[K], t1[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        t0[k] = col_pair<PAIR>(img, off[k], tcb);
        t1[k] = col_pair<PAIR>(img_r, off[k], tcb);
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float h0 = fmaf(fu[k], t0[k].y - t0[k].x, t0[k].x);
        const float h1 = fmaf(fu[k], t1[k].y - t1[k].x, t1[k].x);
        acc[k] = fmaf(w[k], fmaf(fv[k], h1 - h0, h0), acc[k]);
      }
    }
  }

  const int64_t base = ix * A.vs0 + iy * A.vs1;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int32_t iz = iz0 + k;
    if (iz >= A.nz) break;
    const int64_t o = base + iz * A.vs2;
    const float b = acc[k] * A.scale;
    if (A.mode == 0) {
      A.out[o] = b;
    } else if (A.mode == 1) {
      A.out[o] += b;
    } else {
      float x = A.out[o] - A.C[ix * A.cs0 + iy * A.cs1 + iz * A.cs2] * b;
      if (A.mask) x *= A.mask[ix * A.ms0 + iy * A.ms1 + iz * A.ms2];
      if (A.clip_lo) x = fmaxf(x, A.lo);
      if (A.clip_hi) x = fminf(x, A.hi);
      A.out[o] = x;
      if (A.xtex) A.xtex[ix * A.xts0 + iy * A.xts1 + iz * A.xts2] = x;
    }
  }
}

}  // namespace

namespace kk {

struct BpEpilogue {
  const float* C = nullptr;
  const int64_t* cs = nullptr;
  const float* mask = nullptr;
  const int64_t* ms = nullptr;
  float lo = 0.f, hi = 0.f;
  int32_t clip_lo = 0, clip_hi = 0;
  float* xtex = nullptr;
  const int64_t* xts = nullptr;
};

int launch_cone_bp(const kk_plan* plan, const float* tex,
                   const int32_t pax[3], int32_t n_angles, int32_t rows,
                   int32_t cols, float* vol, const int32_t shape[3],
                   const int64_t vs[3], float scale, int32_t mode,
                   const BpEpilogue& ep, int32_t zrun, hipStream_t stream) {
  if (plan == nullptr || !plan->bp_set)
    return fail(KK_ERR_NOT_SET,
                "Please set the parameters with `set_params()`.");
  if (tex == nullptr || vol == nullptr || shape == nullptr || vs == nullptr ||
      pax == nullptr)
    return fail(KK_ERR_INVALID, "kk_cone_bp: null argument");
  if (n_angles != plan->bp_n)
    return fail(KK_ERR_SHAPE,
                "kk_cone_bp: %d projections but parameters for %d angles",
                n_angles, plan->bp_n);
  if (rows <= 0 || cols <= 0 || shape[0] <= 0 || shape[1] <= 0 ||
      shape[2] <= 0)
    return fail(KK_ERR_INVALID, "kk_cone_bp: bad sizes");
  int seen = 0;
  for (int k = 0; k < 3; ++k)
    if (pax[k] >= 0 && pax[k] < 3) seen |= 1 << pax[k];
  if (seen != 7)
    return fail(KK_ERR_INVALID,
                "projection_axes must be a permutation of (0, 1, 2)");
  // texture = padded projection array in its own axis order
  const int64_t logical[3] = {n_angles, rows, cols};
  int64_t text[3];
  for (int k = 0; k < 3; ++k) text[k] = logical[pax[k]] + kHalo;
  const int64_t astride[3] = {text[1] * text[2], text[2], 1};
  int64_t lstride[3];
  for (int k = 0; k < 3; ++k) lstride[pax[k]] = astride[k];
  // in-image byte offsets (row <= rows + 2, col <= cols + 2) must fit 32
  // bits and the row/column byte strides 24 bits (v_mul_u32_u24)
  if (4 * ((rows + 2) * lstride[1] + (cols + 2) * lstride[2] + 2) >=
          (int64_t(1) << 32) ||
      4 * lstride[1] >= (int64_t(1) << 24) ||
      4 * lstride[2] >= (int64_t(1) << 24))
    return fail(KK_ERR_UNSUPPORTED,
                "kk_cone_bp: projection layout too large for 32-bit "
                "in-image offsets");
  if (mode == 2 && (ep.C == nullptr || ep.cs == nullptr))
    return fail(KK_ERR_INVALID, "kk_cone_bp: SIRT update needs C");
  BpArgs A{};
  A.tex = tex;
  A.tas = lstride[0];
  A.trs = int32_t(lstride[1]);
  A.tcs = int32_t(lstride[2]);
  A.nx = shape[0];
  A.ny = shape[1];
  A.nz = shape[2];
  A.n_angles = plan->bp_n;
  A.rows = rows;
  A.cols = cols;
  A.out = vol;
  A.vs0 = vs[0];
  A.vs1 = vs[1];
  A.vs2 = vs[2];
  A.scale = scale;
  A.mode = mode;
  if (mode == 2) {
    A.C = ep.C;
    A.cs0 = ep.cs[0];
    A.cs1 = ep.cs[1];
    A.cs2 = ep.cs[2];
    A.mask = ep.mask;
    if (ep.mask) {
      A.ms0 = ep.ms[0];
      A.ms1 = ep.ms[1];
      A.ms2 = ep.ms[2];
    }
    A.lo = ep.lo;
    A.hi = ep.hi;
    A.clip_lo = ep.clip_lo;
    A.clip_hi = ep.clip_hi;
    A.xtex = ep.xtex;
    if (ep.xtex) {
      if (ep.xts == nullptr)
        return fail(KK_ERR_INVALID, "kk_cone_bp: texture strides missing");
      A.xts0 = ep.xts[0];
      A.xts1 = ep.xts[1];
      A.xts2 = ep.xts[2];
    }
  }
  A.tiles_x = (A.nx + 15) / 16;
  A.tiles_y = (A.ny + 15) / 16;
  A.tiles_z = (A.nz + zrun - 1) / zrun;
  A.total_blocks = int64_t(A.tiles_x) * A.tiles_y * A.tiles_z;
  if (A.total_blocks >= (int64_t(1) << 31))
    return fail(KK_ERR_UNSUPPORTED, "kk_cone_bp: grid too large");
  const int32_t plane = A.tiles_x * A.tiles_y;
  A.xcd_group = (plane % kNumXcd == 0) ? plane / kNumXcd : 1;
  const dim3 grid(unsigned(A.total_blocks)), block(256);
  const bool pair = (A.tcs == 1);
#define KK_BP_LAUNCH(KV)                                                   \
  if (pair)                                                                \
    hipLaunchKernelGGL((cone_bp_kernel<KV, true>), grid, block, 0, stream, \
                       plan->bp, A);                                       \
  else                                                                     \
    hipLaunchKernelGGL((cone_bp_kernel<KV, false>), grid, block, 0, stream,\
                       plan->bp, A);
  switch (zrun) {
    case 4:
      KK_BP_LAUNCH(4);
      break;
    case 8:
      KK_BP_LAUNCH(8);
      break;
    case 16:
      KK_BP_LAUNCH(16);
      break;
    default:
      return fail(KK_ERR_INVALID, "kk_cone_bp: z-run must be 4, 8 or 16");
  }
#undef KK_BP_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "cone_bp_kernel");
  return KK_OK;
}

}  // namespace kk

extern "C" int kk_cone_bp(const kk_plan* plan, const float* proj_tex,
                          const int32_t pax[3], int32_t n_angles,
                          int32_t rows, int32_t cols, float* vol,
                          const int32_t shape[3], const int64_t vs[3],
                          float output_scale, int32_t additive,
                          void* stream) {
  kk::BpEpilogue ep;
  return kk::launch_cone_bp(plan, proj_tex, pax, n_angles, rows, cols, vol,
                            shape, vs, output_scale, additive ? 1 : 0, ep, 8,
                            static_cast<hipStream_t>(stream));
}

extern "C" int kk_sirt_bp_update(const kk_plan* plan, const float* res_tex,
                                 const int32_t pax[3], int32_t n_angles,
                                 int32_t rows, int32_t cols, float* x,
                                 const int32_t shape[3],
                                 const int64_t xs[3], const float* C,
                                 const int64_t cs[3], const float* mask,
                                 const int64_t ms[3], float lo, float hi,
                                 int32_t clip_lo, int32_t clip_hi,
                                 float* x_tex, const int64_t xts[3],
                                 float output_scale, void* stream) {
  kk::BpEpilogue ep;
  ep.C = C;
  ep.cs = cs;
  ep.mask = mask;
  ep.ms = ms;
  if (mask != nullptr && ms == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_sirt_bp_update: mask strides");
  ep.lo = lo;
  ep.hi = hi;
  ep.clip_lo = clip_lo;
  ep.clip_hi = clip_hi;
  ep.xtex = x_tex;
  ep.xts = xts;
  return kk::launch_cone_bp(plan, res_tex, pax, n_angles, rows, cols, x,
                            shape, xs, output_scale, 2, ep, 8,
                            static_cast<hipStream_t>(stream));
}
