// Ray-driven (Joseph) cone-beam forward projection for gfx950.
//
// Behaviour follows cone_fp<DIR_X|Y|Z> of the reference
// (kernelkit/cuda/cone_fp.cu:73-207) as launched by
// kernelkit/kernels/cone_fp.py:195-262: for every angle a ray axis a is
// fixed; each ray through a pixel centre is sampled at the centre of every
// slice along a, bilinearly in the two other axes with zero outside the
// volume, and the sum is scaled by sqrt(1 + a1^2 sc1 + a2^2 sc2) * oscale.
//
// MI355X design:
// * one launch covers all angles and all slices (the reference issues one
//   launch per angle group and 16-slice chunk and accumulates in global
//   memory); each output pixel is written exactly once;
// * the volume is read from a zero-halo texture [x][y][z] (z fastest), so a
//   sample is: two FMAs for its coordinates, a clamp into the halo, and four
//   taps with no bounds tests; when the c axis is z the two c-neighbours are
//   one 8-byte load;
// * sample coordinates are evaluated directly, x_b(i) = a1*i + beta, instead
//   of the reference's running sums, so the per-sample error does not grow
//   along the ray;
// * each ray loops only over the slices where it can touch the volume;
// * lanes of a wavefront are consecutive detector rows of one column: for
//   rotation about z they sample consecutive z voxels of the same (x or y)
//   row, i.e. contiguous memory;
// * per-angle parameters are wave-uniform (scalar loads); workgroups are
//   ordered row band -> angle -> column tile, and grouped so that each XCD
//   keeps the same column range across consecutive angles (L2 reuse).
#include "kk_internal.h"

namespace {

struct FpArgs {
  const float* tex;  // padded volume texture, element [0][0][0]
  float* out;
  int64_t os0, os1, os2;  // output strides: angle, row, col
  const float* y;         // mode 2: measured projections
  int64_t ys0, ys1, ys2;
  const float* R;  // mode 2: ray weights
  int64_t rs0, rs1, rs2;
  int32_t rows, cols, n_angles;
  int32_t tile_rows, tile_cols, row_tiles, col_tiles;
  int32_t mode;  // 0: out = p, 1: out += p, 2: out = (p - y) * R
  int32_t xcd_group;
  int64_t total_blocks;
};

template <bool PAIR>
__device__ __forceinline__ float joseph_sample(const float* __restrict__ base,
                                               int32_t sa, int32_t sb,
                                               int32_t sc, float a1, float a2,
                                               float b1, float b2, float nbf,
                                               float ncf, int32_t i) {
  const float fi = static_cast<float>(i);
  const float xb = kk_clamp(fmaf(a1, fi, b1), -1.f, nbf);
  const float xc = kk_clamp(fmaf(a2, fi, b2), -1.f, ncf);
  const float jbf = floorf(xb);
  const float jcf = floorf(xc);
  const float fb = xb - jbf;
  const float fc = xc - jcf;
  const float* p = base + (i * sa + static_cast<int32_t>(jbf) * sb +
                           static_cast<int32_t>(jcf) * sc);
  float p00, p01, p10, p11;
  if constexpr (PAIR) {
    const kk_float2u lo = *reinterpret_cast<const kk_float2u*>(p);
    const kk_float2u hi = *reinterpret_cast<const kk_float2u*>(p + sb);
    p00 = lo.x;
    p01 = lo.y;
    p10 = hi.x;
    p11 = hi.y;
  } else {
    p00 = p[0];
    p01 = p[sc];
    p10 = p[sb];
    p11 = p[sb + sc];
  }
  const float v0 = fmaf(fc, p01 - p00, p00);
  const float v1 = fmaf(fc, p11 - p10, p10);
  return fmaf(fb, v1 - v0, v0);
}

template <bool PAIR>
__device__ __forceinline__ float joseph_ray(const float* __restrict__ base,
                                            int32_t sa, int32_t sb,
                                            int32_t sc, float a1, float a2,
                                            float b1, float b2, float nbf,
                                            float ncf, int32_t i0,
                                            int32_t i1) {
  float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
  int32_t i = i0;
  for (; i + 3 <= i1; i += 4) {
    acc0 += joseph_sample<PAIR>(base, sa, sb, sc, a1, a2, b1, b2, nbf, ncf, i);
    acc1 += joseph_sample<PAIR>(base, sa, sb, sc, a1, a2, b1, b2, nbf, ncf,
                                i + 1);
    acc2 += joseph_sample<PAIR>(base, sa, sb, sc, a1, a2, b1, b2, nbf, ncf,
                                i + 2);
    acc3 += joseph_sample<PAIR>(base, sa, sb, sc, a1, a2, b1, b2, nbf, ncf,
                                i + 3);
  }
  for (; i <= i1; ++i)
    acc0 += joseph_sample<PAIR>(base, sa, sb, sc, a1, a2, b1, b2, nbf, ncf, i);
  return (acc0 + acc1) + (acc2 + acc3);
}

// Narrow [lo, hi] (slice indices, initially [0, n_a - 1]) to the slices i
// with -1 < a*i + beta < n.  For slopes below 1e-3 the division is not
// trusted; the ray is then only dropped when it stays outside over the
// whole slice range.  Over-inclusion is harmless: clamped samples outside
// the volume read the zero halo.
__device__ __forceinline__ void clip_range(float a, float beta, float n,
                                           float last, float& lo,
                                           float& hi) {
  if (fabsf(a) < 1e-3f) {
    const float e0 = beta, e1 = fmaf(a, last, beta);
    if (fmaxf(e0, e1) <= -1.f || fminf(e0, e1) >= n) {
      lo = 3e38f;
      hi = -3e38f;
    }
    return;
  }
  const float inv = 1.f / a;
  const float t0 = (-1.f - beta) * inv;
  const float t1 = (n - beta) * inv;
  lo = fmaxf(lo, fminf(t0, t1));
  hi = fminf(hi, fmaxf(t0, t1));
}

__global__ __launch_bounds__(256) void cone_fp_kernel(
    const kk::FpAngle* __restrict__ tab, FpArgs A) {
  const int64_t bid =
      kk_xcd_remap(blockIdx.x, A.total_blocks, A.xcd_group);
  const int32_t ct = static_cast<int32_t>(bid % A.col_tiles);
  const int64_t t = bid / A.col_tiles;
  const int32_t ang = static_cast<int32_t>(t % A.n_angles);
  const int32_t rt = static_cast<int32_t>(t / A.n_angles);
  const int32_t row = rt * A.tile_rows + int32_t(threadIdx.x) % A.tile_rows;
  const int32_t col = ct * A.tile_cols + int32_t(threadIdx.x) / A.tile_rows;
  if (row >= A.rows || col >= A.cols) return;

  const kk::FpAngle& P = tab[ang];
  const float fr = static_cast<float>(row);
  const float fcl = static_cast<float>(col);
  const float Da = fmaf(fr, P.va, fmaf(fcl, P.ua, P.qa));
  const float Db = fmaf(fr, P.vb, fmaf(fcl, P.ub, P.qb));
  const float Dc = fmaf(fr, P.vc, fmaf(fcl, P.uc, P.qc));
  const float a1 = Db / Da;
  const float a2 = Dc / Da;
  const float nbf = static_cast<float>(P.nb);
  const float ncf = static_cast<float>(P.nc);

  const float b1 = fmaf(a1, P.sig, P.taub);
  const float b2 = fmaf(a2, P.sig, P.tauc);
  float val = 0.f;
  // Non-finite slopes/intercepts (a ray parallel to the slices) contribute
  // nothing; finite ones keep every clamped coordinate finite.
  if (isfinite(a1) && isfinite(a2) && isfinite(b1) && isfinite(b2)) {
    const float last = static_cast<float>(P.na - 1);
    float lo = 0.f, hi = last;
    clip_range(a1, b1, nbf, last, lo, hi);
    clip_range(a2, b2, ncf, last, lo, hi);
    if (lo <= hi + 2.f) {
      const int32_t i0 = max(0, static_cast<int32_t>(floorf(lo)) - 1);
      const int32_t i1 = min(P.na - 1, static_cast<int32_t>(ceilf(hi)) + 1);
      const float* base = A.tex + kk::kHaloLo * (P.sa + P.sb + P.sc);
      const float sum =
          P.pair ? joseph_ray<true>(base, P.sa, P.sb, P.sc, a1, a2, b1, b2,
                                    nbf, ncf, i0, i1)
                 : joseph_ray<false>(base, P.sa, P.sb, P.sc, a1, a2, b1, b2,
                                     nbf, ncf, i0, i1);
      val = sum * sqrtf(fmaf(a1 * a1, P.sc1, fmaf(a2 * a2, P.sc2, 1.f))) *
            P.oscale;
    }
  }
  const int64_t o = ang * A.os0 + row * A.os1 + col * A.os2;
  if (A.mode == 0) {
    A.out[o] = val;
  } else if (A.mode == 1) {
    A.out[o] += val;
  } else {
    const float yv = A.y[ang * A.ys0 + row * A.ys1 + col * A.ys2];
    const float rv = A.R[ang * A.rs0 + row * A.rs1 + col * A.rs2];
    A.out[o] = (val - yv) * rv;
  }
}

}  // namespace

namespace kk {

int launch_cone_fp(const kk_plan* plan, const float* tex, float* out,
                   const int64_t os[3], int32_t rows, int32_t cols,
                   int32_t mode, const float* y, const int64_t ys[3],
                   const float* R, const int64_t rs[3], int32_t tile_rows,
                   hipStream_t stream) {
  if (plan == nullptr || !plan->fp_set)
    return fail(KK_ERR_NOT_SET,
                "Please set the projection geometry with `set_params()`.");
  if (tex == nullptr || out == nullptr || os == nullptr)
    return fail(KK_ERR_INVALID, "kk_cone_fp: null argument");
  if (rows <= 0 || cols <= 0)
    return fail(KK_ERR_INVALID, "kk_cone_fp: bad detector size");
  if (mode == 2 && (y == nullptr || R == nullptr || ys == nullptr ||
                    rs == nullptr))
    return fail(KK_ERR_INVALID, "kk_cone_fp: residual mode needs y and R");
  if (tile_rows <= 0 || tile_rows > 256 || 256 % tile_rows != 0)
    return fail(KK_ERR_INVALID, "kk_cone_fp: tile_rows must divide 256");
  FpArgs A{};
  A.tex = tex;
  A.out = out;
  A.os0 = os[0];
  A.os1 = os[1];
  A.os2 = os[2];
  if (mode == 2) {
    A.y = y;
    A.ys0 = ys[0];
    A.ys1 = ys[1];
    A.ys2 = ys[2];
    A.R = R;
    A.rs0 = rs[0];
    A.rs1 = rs[1];
    A.rs2 = rs[2];
  }
  A.rows = rows;
  A.cols = cols;
  A.n_angles = plan->fp_n;
  A.tile_rows = tile_rows;
  A.tile_cols = 256 / tile_rows;
  A.row_tiles = (rows + A.tile_rows - 1) / A.tile_rows;
  A.col_tiles = (cols + A.tile_cols - 1) / A.tile_cols;
  A.mode = mode;
  A.total_blocks = int64_t(A.row_tiles) * A.n_angles * A.col_tiles;
  if (A.total_blocks >= (int64_t(1) << 31))
    return fail(KK_ERR_UNSUPPORTED, "kk_cone_fp: grid too large");
  A.xcd_group = (A.col_tiles % kNumXcd == 0) ? A.col_tiles / kNumXcd : 1;
  hipLaunchKernelGGL(cone_fp_kernel, dim3(unsigned(A.total_blocks)),
                     dim3(256), 0, stream, plan->fp, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "cone_fp_kernel");
  return KK_OK;
}

}  // namespace kk

extern "C" int kk_cone_fp(const kk_plan* plan, const float* vol_tex,
                          float* proj, const int64_t ps[3], int32_t rows,
                          int32_t cols, int32_t additive, void* stream) {
  return kk::launch_cone_fp(plan, vol_tex, proj, ps, rows, cols,
                            additive ? 1 : 0, nullptr, nullptr, nullptr,
                            nullptr, 64, static_cast<hipStream_t>(stream));
}

extern "C" int kk_sirt_fp_residual(const kk_plan* plan, const float* x_tex,
                                   const float* y, const int64_t ys[3],
                                   const float* R, const int64_t rs[3],
                                   float* out, const int64_t os[3],
                                   int32_t rows, int32_t cols, void* stream) {
  return kk::launch_cone_fp(plan, x_tex, out, os, rows, cols, 2, y, ys, R,
                            rs, 64, static_cast<hipStream_t>(stream));
}
