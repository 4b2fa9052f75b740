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
//   constant step.  A detector tilted in its own plane keeps 1/den shared
//   (u and v both affine along the run); other geometries take the
//   general per-voxel path (a wave-uniform branch per angle);
// * projections are read from a zero-halo texture of (value, next-row)
//   float2 pairs (kk_internal.h): a sample is a clamp, ONE 16-byte load for
//   the four taps and three lerps, with no bounds tests (the
//   texture-address unit, not the ALU, limits a gather kernel);
// * lanes of a wavefront are a 16 x 4 (x, y) patch, which projects onto a
//   compact detector patch; on circular orbits workgroups go to the XCDs
//   in groups of 4 x 4 x 4 neighbouring tiles, whose detector patches
//   overlap, otherwise in runs of 16 tiles along x.
#include <algorithm>

#include "kk_internal.h"

namespace {

constexpr int kMaxFrames = 3;

struct BpArgs {
  const char* tex;   // projection texture (byte base)
  uint32_t tab_b;    // bytes per angle image ((rows+3) * (cols+3) * 8)
  uint32_t trb;      // bytes per texture row ((cols+3) * 8)
  int32_t nx, ny, nz, n_angles, rows, cols;
  float* out;
  int64_t vs0, vs1, vs2;  // output world strides
  float scale;
  int32_t mode;  // 0: out = b, 1: out += b, 2: SIRT update
  int32_t tiles_x, tiles_y, tiles_z;
  int32_t xcd_group;
  int32_t gx, gy, gz;  // tiles per workgroup group along x, y, z
  int64_t total_blocks;
  // mode 2: x <- clip(mask * (x - C * b)), mirrored into the FP textures
  const float* C;
  int64_t cs0, cs1, cs2;
  const float* mask;
  int64_t ms0, ms1, ms2;
  float lo, hi;
  int32_t clip_lo, clip_hi;
  float* xtex;
  int32_t n_frames;
  int64_t fr_origin[kMaxFrames], fr_s0[kMaxFrames], fr_s1[kMaxFrames],
      fr_s2[kMaxFrames], fr_dual[kMaxFrames];
};

// Bilinear value from the 16-byte tap record (p[r][c], p[r+1][c],
// p[r][c+1], p[r+1][c+1]): lerp along rows (fv), then columns (fu).
__device__ __forceinline__ float bilerp(kk_float4u t, float fu, float fv) {
  const float l0 = fmaf(fv, t.y - t.x, t.x);
  const float l1 = fmaf(fv, t.w - t.z, t.z);
  return fmaf(fu, l1 - l0, l0);
}

// 1/den from a float32 estimate and one float64 Newton step (relative
// error ~1e-14).
__device__ __forceinline__ double refine_rcp(double den, float rf) {
  const double r = static_cast<double>(rf);
  return fma(r, fma(-den, r, 1.0), r);
}

// Coordinate base + off (base an integer, off a small float) clamped into
// [0, hi] (the zero halo), split into element index and weight: the same
// result as clamping the float64 coordinate, then floor and fraction.
// The clamp is applied to off, into [lo, up] = [-base, hi - base] (integers,
// exact in float32 for |base|, hi < 2^24; the caller passes them
// converted), so at either bound the weight is 0 and the index 0 or hi.
__device__ __forceinline__ void split_clamped(float off, int32_t base,
                                              float lo, float up,
                                              int32_t& idx, float& frac) {
  const float o = __builtin_amdgcn_fmed3f(off, lo, up);
  const float fl = floorf(o);
  idx = base + static_cast<int32_t>(fl);
  frac = o - fl;
}

// split_clamped for a coordinate known to lie in [0, hi): no clamps, the
// same index and weight.
__device__ __forceinline__ void split_inside(float off, int32_t base,
                                             int32_t& idx, float& frac) {
  const float fl = floorf(off);
  idx = base + static_cast<int32_t>(fl);
  frac = off - fl;
}

// Whether base + [lo, hi_off] lies in [1, hi - 1]: a run of coordinates
// there never reaches split_clamped's clamps (one texel of margin covers
// the float32 evaluation of this test).
__device__ __forceinline__ bool run_inside(int32_t base, float lo,
                                           float hi_off, int32_t hi) {
  const float b = static_cast<float>(base);
  return b + lo >= 1.f && b + hi_off <= static_cast<float>(hi - 1);
}

// One workgroup's tile (16 x 16 columns of K voxels) of logical block id
// `bid`.  TABLE_IN_MEMORY: the angle table pointer was read from a
// descriptor (batch kernel), so each record is copied through
// constant-memory (scalar) loads; a kernel-argument table is read in place.
template <int K, bool TABLE_IN_MEMORY>
__device__ __forceinline__ void bp_tile(const kk::BpAngle* __restrict__ tab,
                                        const BpArgs& A, int64_t bid) {
  // groups of gx x gy x gz tiles (x fastest inside), groups ordered
  // x -> y -> z (the group sides divide the tile counts; 1 x 1 x 1 is
  // plain x -> y -> z order)
  const int32_t gx = A.gx, gy = A.gy, gz = A.gz;
  const int64_t inner = bid % (gx * gy * gz);
  const int64_t outer = bid / (gx * gy * gz);
  const int32_t ngx = (A.tiles_x + gx - 1) / gx;
  const int32_t ngy = (A.tiles_y + gy - 1) / gy;
  const int32_t tx = static_cast<int32_t>(outer % ngx) * gx +
                     static_cast<int32_t>(inner % gx);
  const int64_t o2 = outer / ngx;
  const int32_t ty = static_cast<int32_t>(o2 % ngy) * gy +
                     static_cast<int32_t>((inner / gx) % gy);
  const int32_t tz = static_cast<int32_t>(o2 / ngy) * gz +
                     static_cast<int32_t>(inner / (gx * gy));
  if (tx >= A.tiles_x || ty >= A.tiles_y || tz >= A.tiles_z) return;
  const int32_t ix = tx * 16 + int32_t(threadIdx.x & 15);
  const int32_t iy = ty * 16 + int32_t(threadIdx.x >> 4);
  const int32_t iz0 = tz * K;
  if (ix >= A.nx || iy >= A.ny) return;

  // voxel centres; the geometry is evaluated in float64 per (x, y, angle):
  // detector coordinates reach ~cols + 1 (up to ~1.5e3), where float32
  // would round them by ~1e-4 pixel.  Per voxel only small offsets from an
  // integer base are float32.
  const double X = static_cast<double>(ix) - 0.5 * A.nx + 0.5;
  const double Y = static_cast<double>(iy) - 0.5 * A.ny + 0.5;
  const double Z0 = static_cast<double>(iz0) - 0.5 * A.nz + 0.5;
  const int32_t rmax = A.rows + 1;
  const int32_t cmax = A.cols + 1;
  const uint32_t trb = A.trb;

  float acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.f;

  const char* img = A.tex;
  for (int32_t a = 0; a < A.n_angles; ++a, img += A.tab_b) {
    kk::BpAngle rec;
    const kk::BpAngle& P =
        TABLE_IN_MEMORY ? (rec = kk_uniform_load(tab + a), rec) : tab[a];
    const double den = fma(Y, P.dy, fma(X, P.dx, P.dw));
    const double nu = fma(Y, P.nuy, fma(X, P.nux, P.nuw));
    const double nv = fma(Y, P.nvy, fma(X, P.nvx, P.nvw));
    if (P.fast == 2) {
      // 1/den and the weight do not depend on z; u and v are affine in z
      const float rf = __builtin_amdgcn_rcpf(static_cast<float>(den));
      const double r = refine_rcp(den, rf);
      const double u0 = fma(Z0, P.nuz, nu) * r;
      const double v0 = fma(Z0, P.nvz, nv) * r;
      const float gu = static_cast<float>(P.nuz * r);
      const float gv = static_cast<float>(P.nvz * r);
      if (!(isfinite(u0) && isfinite(v0) && isfinite(gu) && isfinite(gv)))
        continue;
      const float w = rf * rf;
      const double uc0 = fmin(fmax(u0, -1e6), 1e6);
      const double ufl = floor(uc0);
      const int32_t iub = static_cast<int32_t>(ufl);
      const float fu0 = static_cast<float>(uc0 - ufl);
      const double vc0 = fmin(fmax(v0, -1e6), 1e6);
      const double vfl = floor(vc0);
      const int32_t ivb = static_cast<int32_t>(vfl);
      const float fv0 = static_cast<float>(vc0 - vfl);
      uint32_t off[K];
      float fu[K], fv[K];
      const float ku = static_cast<float>(K - 1) * gu;
      const float kv = static_cast<float>(K - 1) * gv;
      const bool inside =
          run_inside(iub, fu0 + fminf(ku, 0.f), fu0 + fmaxf(ku, 0.f), cmax) &&
          run_inside(ivb, fv0 + fminf(kv, 0.f), fv0 + fmaxf(kv, 0.f), rmax);
      if (__all(inside)) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float fk = static_cast<float>(k);
          int32_t iu, iv;
          split_inside(fmaf(fk, gu, fu0), iub, iu, fu[k]);
          split_inside(fmaf(fk, gv, fv0), ivb, iv, fv[k]);
          off[k] = __umul24(static_cast<uint32_t>(iv), trb) +
                   (static_cast<uint32_t>(iu) << 3);
        }
      } else {
        const float ulo = static_cast<float>(-iub);
        const float uup = static_cast<float>(cmax - iub);
        const float vlo = static_cast<float>(-ivb);
        const float vup = static_cast<float>(rmax - ivb);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float fk = static_cast<float>(k);
          int32_t iu, iv;
          split_clamped(fmaf(fk, gu, fu0), iub, ulo, uup, iu, fu[k]);
          split_clamped(fmaf(fk, gv, fv0), ivb, vlo, vup, iv, fv[k]);
          off[k] = __umul24(static_cast<uint32_t>(iv), trb) +
                   (static_cast<uint32_t>(iu) << 3);
        }
      }
      kk_float4u tp[K];
#pragma unroll
      for (int k = 0; k < K; ++k) tp[k] = kk_ld4(img, off[k]);
#pragma unroll
      for (int k = 0; k < K; ++k)
        acc[k] = fmaf(w, bilerp(tp[k], fu[k], fv[k]), acc[k]);
    } else if (P.fast) {
      // u and 1/den do not depend on z; v is affine in z
      const float rf = __builtin_amdgcn_rcpf(static_cast<float>(den));
      const double r = refine_rcp(den, rf);
      const double u = nu * r;
      const double v0 = fma(Z0, P.nvz, nv) * r;
      const float gv = static_cast<float>(P.nvz * r);
      // a voxel in the plane of the source parallel to the detector
      // (den = 0) has no projection: skip it
      if (!(isfinite(u) && isfinite(v0) && isfinite(gv))) continue;
      const float w = rf * rf;
      uint32_t iu;
      float fu;
      kk_split_nonneg(fmin(fmax(u, 0.0), static_cast<double>(cmax)), iu, fu);
      const uint32_t offu = iu << 3;
      const double vc0 = fmin(fmax(v0, -1e6), 1e6);
      const double vfl = floor(vc0);
      const int32_t ivb = static_cast<int32_t>(vfl);
      const float fv0 = static_cast<float>(vc0 - vfl);
      const float vlo = static_cast<float>(-ivb);
      const float vup = static_cast<float>(rmax - ivb);
      uint32_t off[K];
      float fv[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        int32_t iv;
        split_clamped(fmaf(static_cast<float>(k), gv, fv0), ivb, vlo, vup, iv,
                      fv[k]);
        off[k] = __umul24(static_cast<uint32_t>(iv), trb) + offu;
      }
      kk_float4u tp[K];
#pragma unroll
      for (int k = 0; k < K; ++k) tp[k] = kk_ld4(img, off[k]);
#pragma unroll
      for (int k = 0; k < K; ++k)
        acc[k] = fmaf(w, bilerp(tp[k], fu, fv[k]), acc[k]);
    } else {
      // integer bases at the run's first voxel; per voxel,
      // u(k) - ub = (e0u + k e1u) / den(k) with float32 numerators
      const double den0 = fma(Z0, P.dz, den);
      const double nu0 = fma(Z0, P.nuz, nu);
      const double nv0 = fma(Z0, P.nvz, nv);
      const double r0 = refine_rcp(
          den0, __builtin_amdgcn_rcpf(static_cast<float>(den0)));
      double ub = 0.0, vb = 0.0;
      const double u0 = nu0 * r0, v0 = nv0 * r0;
      if (isfinite(u0) && isfinite(v0)) {
        ub = floor(fmin(fmax(u0, -1e6), 1e6));
        vb = floor(fmin(fmax(v0, -1e6), 1e6));
      }
      const float e0u = static_cast<float>(fma(-ub, den0, nu0));
      const float e1u = static_cast<float>(fma(-ub, P.dz, P.nuz));
      const float e0v = static_cast<float>(fma(-vb, den0, nv0));
      const float e1v = static_cast<float>(fma(-vb, P.dz, P.nvz));
      const float d0f = static_cast<float>(den0);
      const float dzf = static_cast<float>(P.dz);
      const int32_t iub = static_cast<int32_t>(ub);
      const int32_t ivb = static_cast<int32_t>(vb);
      uint32_t off[K];
      float fu[K], fv[K], w[K];
      // the run's end points: with den of one sign along the run, u and v
      // are monotonic in k (ratios of affine forms), so end points inside
      // the detector put the whole run inside (non-finite values fail the
      // comparisons)
      const float kl = static_cast<float>(K - 1);
      const float dK = fmaf(kl, dzf, d0f);
      const float rf0 = __builtin_amdgcn_rcpf(d0f);
      const float rfK = __builtin_amdgcn_rcpf(dK);
      const float du0 = e0u * rf0, duK = fmaf(kl, e1u, e0u) * rfK;
      const float dv0 = e0v * rf0, dvK = fmaf(kl, e1v, e0v) * rfK;
      const bool inside = d0f * dK > 0.f &&
                          run_inside(iub, fminf(du0, duK), fmaxf(du0, duK),
                                     cmax) &&
                          run_inside(ivb, fminf(dv0, dvK), fmaxf(dv0, dvK),
                                     rmax);
      if (__all(inside)) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float fk = static_cast<float>(k);
          const float r = __builtin_amdgcn_rcpf(fmaf(fk, dzf, d0f));
          int32_t iu, iv;
          split_inside(fmaf(fk, e1u, e0u) * r, iub, iu, fu[k]);
          split_inside(fmaf(fk, e1v, e0v) * r, ivb, iv, fv[k]);
          off[k] = __umul24(static_cast<uint32_t>(iv), trb) +
                   (static_cast<uint32_t>(iu) << 3);
          w[k] = r * r;
        }
      } else {
        // a voxel without a finite projection reads index 0 with weight 0
        // (its offset is put at the lower bound)
        const float ulo = static_cast<float>(-iub);
        const float uup = static_cast<float>(cmax - iub);
        const float vlo = static_cast<float>(-ivb);
        const float vup = static_cast<float>(rmax - ivb);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float fk = static_cast<float>(k);
          const float r = __builtin_amdgcn_rcpf(fmaf(fk, dzf, d0f));
          const float du = fmaf(fk, e1u, e0u) * r;
          const float dv = fmaf(fk, e1v, e0v) * r;
          const bool ok = isfinite(du) && isfinite(dv);
          int32_t iu, iv;
          split_clamped(ok ? du : ulo, iub, ulo, uup, iu, fu[k]);
          split_clamped(ok ? dv : vlo, ivb, vlo, vup, iv, fv[k]);
          off[k] = __umul24(static_cast<uint32_t>(iv), trb) +
                   (static_cast<uint32_t>(iu) << 3);
          w[k] = ok ? r * r : 0.f;
        }
      }
      kk_float4u tp[K];
#pragma unroll
      for (int k = 0; k < K; ++k) tp[k] = kk_ld4(img, off[k]);
#pragma unroll
      for (int k = 0; k < K; ++k)
        acc[k] = fmaf(w[k], bilerp(tp[k], fu[k], fv[k]), acc[k]);
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
      for (int f = 0; f < A.n_frames; ++f) {
        const int64_t q = A.fr_origin[f] + ix * A.fr_s0[f] +
                          iy * A.fr_s1[f] + iz * A.fr_s2[f];
        A.xtex[q] = x;
        A.xtex[q + A.fr_dual[f]] = x;
      }
    }
  }
}

// Register caps without spills: 5 wavefronts per SIMD up to 8 voxels per
// thread, 3 at 16 (the unclamped and clamped branches together would
// otherwise take 102 and 176 VGPRs, i.e. 4 and 2 waves).
constexpr int bp_min_blocks(int k) { return k <= 8 ? 5 : 3; }

template <int K>
__global__ __launch_bounds__(256, bp_min_blocks(K)) void cone_bp_kernel(
    const kk::BpAngle* __restrict__ tab, BpArgs A) {
  bp_tile<K, false>(tab, A,
                    kk_xcd_remap(blockIdx.x, A.total_blocks, A.xcd_group));
}

struct BpItem {
  const kk::BpAngle* tab;
  BpArgs args;
};

// Many problems, one persistent launch (see cone_fp_batch_kernel).
template <int K>
__global__ __launch_bounds__(256, bp_min_blocks(K)) void cone_bp_batch_kernel(
    const BpItem* __restrict__ items, const int64_t* __restrict__ first,
    int32_t n, int64_t total) {
  for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (first[mid] <= g)
        lo = mid;
      else
        hi = mid - 1;
    }
    const BpItem& it = items[lo];
    const int64_t local = g - first[lo];
    bp_tile<K, true>(it.tab, it.args,
               kk_xcd_remap(local, it.args.total_blocks, it.args.xcd_group));
  }
}

}  // namespace

namespace {

// Validate one backprojection problem and fill its kernel arguments for
// z-runs of `zrun` voxels.
int make_bp_args(const kk::BpSpec& S, int32_t zrun, BpArgs& A) {
  using namespace kk;
  if (S.plan == nullptr || !S.plan->bp_set)
    return fail(KK_ERR_NOT_SET,
                "Please set the parameters with `set_params()`.");
  if (S.tex == nullptr || S.out == nullptr)
    return fail(KK_ERR_INVALID, "kk_cone_bp: null argument");
  if (S.n_angles != S.plan->bp_n)
    return fail(KK_ERR_SHAPE,
                "kk_cone_bp: %d projections but parameters for %d angles",
                S.n_angles, S.plan->bp_n);
  if (S.rows <= 0 || S.cols <= 0 || S.shape[0] <= 0 || S.shape[1] <= 0 ||
      S.shape[2] <= 0)
    return fail(KK_ERR_INVALID, "kk_cone_bp: bad sizes");
  if (zrun != 4 && zrun != 8 && zrun != 16)
    return fail(KK_ERR_INVALID, "kk_cone_bp: z-run must be 4, 8 or 16");
  const int64_t trb = int64_t(S.cols + kHalo) * 8;
  const int64_t tab_b = int64_t(S.rows + kHalo) * trb;
  if (tab_b >= (int64_t(1) << 32) || trb >= (int64_t(1) << 24))
    return fail(KK_ERR_UNSUPPORTED, "kk_cone_bp: detector too large");
  if (S.mode == 2 && S.C == nullptr)
    return fail(KK_ERR_INVALID, "kk_cone_bp: SIRT update needs C");
  A = BpArgs{};
  A.tex = reinterpret_cast<const char*>(S.tex);
  A.tab_b = uint32_t(tab_b);
  A.trb = uint32_t(trb);
  A.nx = S.shape[0];
  A.ny = S.shape[1];
  A.nz = S.shape[2];
  A.n_angles = S.plan->bp_n;
  A.rows = S.rows;
  A.cols = S.cols;
  A.out = S.out;
  A.vs0 = S.vs[0];
  A.vs1 = S.vs[1];
  A.vs2 = S.vs[2];
  A.scale = S.scale;
  A.mode = S.mode;
  if (S.mode == 2) {
    A.C = S.C;
    A.cs0 = S.cs[0];
    A.cs1 = S.cs[1];
    A.cs2 = S.cs[2];
    A.mask = S.mask;
    if (S.mask) {
      A.ms0 = S.ms[0];
      A.ms1 = S.ms[1];
      A.ms2 = S.ms[2];
    }
    A.lo = S.lo;
    A.hi = S.hi;
    A.clip_lo = S.clip_lo;
    A.clip_hi = S.clip_hi;
    A.xtex = S.xtex;
    if (S.xtex) {
      // voxel (x, y, z) in frame a: element (c_a, c_b + 1, c_c + 1)
      // component 0 and element (c_a, c_b, c_c + 1) component 1
      int64_t off = 0;  // floats
      for (int a = 0; a < 3; ++a) {
        if (!(S.xtex_mask & (1 << a))) continue;
        const int b = kFrameB[a], c = kFrameC[a];
        const int64_t tcn = S.shape[c] + kHalo, tbn = S.shape[b] + kHalo;
        int64_t ws[3];
        ws[a] = 2 * tbn * tcn;
        ws[b] = 2 * tcn;
        ws[c] = 2;
        const int f = A.n_frames++;
        A.fr_origin[f] = off + 2 * (kHaloLo * tcn + kHaloLo);
        A.fr_s0[f] = ws[0];
        A.fr_s1[f] = ws[1];
        A.fr_s2[f] = ws[2];
        A.fr_dual[f] = 1 - 2 * tcn;
        off += 2 * frame_elems(S.shape, a);
      }
    }
  }
  A.tiles_x = (A.nx + 15) / 16;
  A.tiles_y = (A.ny + 15) / 16;
  A.tiles_z = (A.nz + zrun - 1) / zrun;
  // Circular orbits, tile counts divisible by 4: groups of 4 x 4 x 4
  // tiles, each group's ids on one XCD; neighbouring tiles project onto
  // overlapping detector patches, which that XCD's L2 then serves
  // (measured at 512^3 / 720 angles: 51.5 ms, against 53.3 ms for runs of
  // 16 tiles along x; 4 x 4 x 8, 8 x 4 x 4, 4 x 4 x 2 and 4 x 4 x 16 took
  // 52.2, 51.7, 52.1 and 52.6 ms).  Other geometries keep runs of 16
  // tiles along x per XCD (1.5% faster than round-robin at 512^3): the
  // groups made the tilted-detector batches of the batching workload
  // slower (its batched backprojection 57.1 -> 65.8 ms).
  const bool groups = S.plan->bp_all_circ && A.tiles_x % 4 == 0 &&
                      A.tiles_y % 4 == 0 && A.tiles_z % 4 == 0;
  A.gx = A.gy = A.gz = groups ? 4 : 1;
  A.xcd_group = groups ? 64 : 16;
  A.total_blocks = int64_t(A.tiles_x) * A.tiles_y * A.tiles_z;
  if (A.total_blocks >= (int64_t(1) << 31))
    return fail(KK_ERR_UNSUPPORTED, "kk_cone_bp: grid too large");
  return KK_OK;
}

int launch_bp_spec(const kk::BpSpec& S, int32_t zrun, hipStream_t stream) {
  BpArgs A;
  const int rc = make_bp_args(S, zrun, A);
  if (rc != KK_OK) return rc;
  const dim3 grid(unsigned(A.total_blocks)), block(256);
  switch (zrun) {
    case 4:
      hipLaunchKernelGGL(cone_bp_kernel<4>, grid, block, 0, stream,
                         S.plan->bp, A);
      break;
    case 8:
      hipLaunchKernelGGL(cone_bp_kernel<8>, grid, block, 0, stream,
                         S.plan->bp, A);
      break;
    default:
      hipLaunchKernelGGL(cone_bp_kernel<16>, grid, block, 0, stream,
                         S.plan->bp, A);
      break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return kk::hip_fail(e, "cone_bp_kernel");
  return KK_OK;
}

}  // namespace

namespace kk {

struct BpBatch {
  BpItem* items = nullptr;
  int64_t* first = nullptr;
  int32_t n = 0;
  int64_t total = 0;
  int32_t zrun = 8;
};

int bp_batch_build(const std::vector<BpSpec>& specs, int32_t zrun,
                   BpBatch** out) {
  if (specs.empty()) return fail(KK_ERR_INVALID, "empty backprojection batch");
  if (zrun != 4 && zrun != 8 && zrun != 16) zrun = 8;
  std::vector<BpItem> items(specs.size());
  std::vector<int64_t> first(specs.size());
  int64_t total = 0;
  for (size_t p = 0; p < specs.size(); ++p) {
    const int rc = make_bp_args(specs[p], zrun, items[p].args);
    if (rc != KK_OK) return rc;
    items[p].tab = specs[p].plan->bp;
    first[p] = total;
    total += items[p].args.total_blocks;
  }
  auto* b = new (std::nothrow) BpBatch();
  if (b == nullptr) return fail(KK_ERR_INVALID, "out of host memory");
  b->n = static_cast<int32_t>(specs.size());
  b->total = total;
  b->zrun = zrun;
  hipError_t e = hipMalloc(&b->items, sizeof(BpItem) * items.size());
  if (e == hipSuccess) e = hipMalloc(&b->first, sizeof(int64_t) * first.size());
  if (e == hipSuccess)
    e = hipMemcpy(b->items, items.data(), sizeof(BpItem) * items.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->first, first.data(), sizeof(int64_t) * first.size(),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    bp_batch_free(b);
    return hip_fail(e, "backprojection batch descriptors");
  }
  *out = b;
  return KK_OK;
}

void bp_batch_free(BpBatch* b) {
  if (b == nullptr) return;
  if (b->items) hipFree(b->items);
  if (b->first) hipFree(b->first);
  delete b;
}

int bp_batch_launch(const BpBatch* b, int32_t grid_blocks,
                    hipStream_t stream) {
  if (b == nullptr) return fail(KK_ERR_NOT_SET, "backprojection batch not built");
  int64_t g = grid_blocks > 0 ? grid_blocks : b->total;
  if (g > b->total) g = b->total;
  if (g <= 0) return KK_OK;
  const dim3 grid(static_cast<unsigned>(g)), block(256);
  switch (b->zrun) {
    case 4:
      hipLaunchKernelGGL(cone_bp_batch_kernel<4>, grid, block, 0, stream,
                         b->items, b->first, b->n, b->total);
      break;
    case 8:
      hipLaunchKernelGGL(cone_bp_batch_kernel<8>, grid, block, 0, stream,
                         b->items, b->first, b->n, b->total);
      break;
    default:
      hipLaunchKernelGGL(cone_bp_batch_kernel<16>, grid, block, 0, stream,
                         b->items, b->first, b->n, b->total);
      break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "cone_bp_batch_kernel");
  return KK_OK;
}

}  // namespace kk

namespace {

int fill_bp_spec(kk::BpSpec& S, const kk_plan* plan, const float* tex,
                 int32_t n_angles, int32_t rows, int32_t cols, float* vol,
                 const int32_t shape[3], const int64_t vs[3], float scale,
                 int32_t mode) {
  if (shape == nullptr || vs == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_cone_bp: null argument");
  S.plan = plan;
  S.tex = tex;
  S.n_angles = n_angles;
  S.rows = rows;
  S.cols = cols;
  S.out = vol;
  for (int t = 0; t < 3; ++t) {
    S.shape[t] = shape[t];
    S.vs[t] = vs[t];
  }
  S.scale = scale;
  S.mode = mode;
  return KK_OK;
}

int fill_sirt_update(kk::BpSpec& S, const float* C, const int64_t cs[3],
                     const float* mask, const int64_t ms[3], float lo,
                     float hi, int32_t clip_lo, int32_t clip_hi, float* x_tex,
                     int32_t x_tex_mask) {
  if (C == nullptr || cs == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_sirt_bp_update: C is required");
  if (mask != nullptr && ms == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_sirt_bp_update: mask strides");
  if (x_tex != nullptr && (x_tex_mask <= 0 || x_tex_mask > 7))
    return kk::fail(KK_ERR_INVALID, "kk_sirt_bp_update: texture mask");
  S.C = C;
  S.mask = mask;
  for (int t = 0; t < 3; ++t) {
    S.cs[t] = cs[t];
    S.ms[t] = mask ? ms[t] : 0;
  }
  S.lo = lo;
  S.hi = hi;
  S.clip_lo = clip_lo;
  S.clip_hi = clip_hi;
  S.xtex = x_tex;
  S.xtex_mask = x_tex ? x_tex_mask : 0;
  return KK_OK;
}

}  // namespace

namespace kk {

int bp_spec(BpSpec& S, const kk_plan* plan, const float* tex,
            int32_t n_angles, int32_t rows, int32_t cols, float* vol,
            const int32_t shape[3], const int64_t vs[3], float scale,
            int32_t additive) {
  return fill_bp_spec(S, plan, tex, n_angles, rows, cols, vol, shape, vs,
                      scale, additive ? 1 : 0);
}

int sirt_update_spec(BpSpec& S, const kk_plan* plan, const float* res_tex,
                     int32_t n_angles, int32_t rows, int32_t cols, float* x,
                     const int32_t shape[3], const int64_t xs[3],
                     const float* C, const int64_t cs[3], const float* mask,
                     const int64_t ms[3], float lo, float hi,
                     int32_t clip_lo, int32_t clip_hi, float* x_tex,
                     int32_t x_tex_mask, float scale) {
  int rc = fill_bp_spec(S, plan, res_tex, n_angles, rows, cols, x, shape, xs,
                        scale, 2);
  if (rc != KK_OK) return rc;
  return fill_sirt_update(S, C, cs, mask, ms, lo, hi, clip_lo, clip_hi,
                          x_tex, x_tex_mask);
}

}  // namespace kk

extern "C" int kk_cone_bp(const kk_plan* plan, const float* proj_tex,
                          int32_t n_angles, int32_t rows, int32_t cols,
                          float* vol, const int32_t shape[3],
                          const int64_t vs[3], float output_scale,
                          int32_t additive, void* stream) {
  kk::BpSpec S;
  const int rc = kk::bp_spec(S, plan, proj_tex, n_angles, rows, cols, vol,
                             shape, vs, output_scale, additive);
  if (rc != KK_OK) return rc;
  return launch_bp_spec(S, 8, static_cast<hipStream_t>(stream));
}

extern "C" int kk_sirt_bp_update(const kk_plan* plan, const float* res_tex,
                                 int32_t n_angles, int32_t rows,
                                 int32_t cols, float* x,
                                 const int32_t shape[3], const int64_t xs[3],
                                 const float* C, const int64_t cs[3],
                                 const float* mask, const int64_t ms[3],
                                 float lo, float hi, int32_t clip_lo,
                                 int32_t clip_hi, float* x_tex,
                                 int32_t x_tex_mask, float output_scale,
                                 void* stream) {
  kk::BpSpec S;
  const int rc = kk::sirt_update_spec(S, plan, res_tex, n_angles, rows, cols,
                                      x, shape, xs, C, cs, mask, ms, lo, hi,
                                      clip_lo, clip_hi, x_tex, x_tex_mask,
                                      output_scale);
  if (rc != KK_OK) return rc;
  return launch_bp_spec(S, 8, static_cast<hipStream_t>(stream));
}
