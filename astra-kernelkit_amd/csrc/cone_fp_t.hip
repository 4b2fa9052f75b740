// Exact transpose of the ray-driven (Joseph) forward projection: the
// matched adjoint A^T of kk_cone_fp, for gradients.
//
// The reference has no such operator: its autograd backward uses the
// voxel-driven backprojector (kernelkit/torch_support.py:29-33), which is
// not the transpose of the forward projector.  Here every voxel sums, over
// the (ray, slice) samples of the forward kernel that touch it, the ray's
// value times the ray weight times the sample's bilinear weight for that
// voxel.  Samples are recomputed with the forward kernel's arithmetic
// (kk_fp_ray, kk_fp_weight, kk_fp_coord), so <A x, y> = <x, A^T y> up to
// rounding.
//
// MI355X design: a gather, so that no atomics are needed (an LDS float
// atomic add costs ~190 cycles per wavefront instruction on this part,
// against ~8 for a plain LDS read-modify-write).  A lane owns kJB voxels
// along b at one (slice, c) position, the 64 lanes of a wavefront
// consecutive c positions, accumulated in registers over all angles of the
// launch's ray axis and written once.  Per angle:
// * circular orbits ("column-coherent" angles: the row step has no a/b
//   part): the b coordinate of a sample depends on the detector column
//   only, so the columns reaching the wavefront's b-run are a
//   wave-uniform interval (inverse of the column -> b map).  The lanes
//   compute 64 columns' ray data at a time into LDS; per column, the c
//   coordinate is affine in the row, so each lane sums the rows reaching
//   its voxel and the sum enters the b-run with the column's two b taps;
// * other geometries: each lane projects the support rectangle of its
//   voxels (at its slice) onto the detector and walks the pixels of the
//   bounding box.
// Candidate intervals are widened by 1e-6 voxel in coordinate space, far
// above the float64 rounding of the sample coordinates; samples outside
// the support get zero weight.
#include "kk_internal.h"

namespace {

constexpr int kJB = 8;       // voxels along b per lane
constexpr int kWaves = 4;    // wavefronts per workgroup (b-runs)
constexpr int kThreads = 64 * kWaves;
constexpr double kEps = 1e-6;

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

// Detector position (col, row) of the ray from the source through the
// source-relative point X (frame components): q + col U + row V = lam X;
// ok iff finite with lam > 0.
__device__ __forceinline__ bool project(const kk::FpAngle& P, double xa,
                                        double xb, double xc, double& col,
                                        double& row) {
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
  col = -(P.qa * vx0 + P.qb * vx1 + P.qc * vx2) / den;
  row = -(P.ua * qx0 + P.ub * qx1 + P.uc * qx2) / den;
  const double lam = (P.ua * vq0 + P.ub * vq1 + P.uc * vq2) / den;
  return isfinite(col) && isfinite(row) && lam > 0.0;
}

// Integer interval [ceil(lo), floor(hi)] clipped to [0, n - 1].
__device__ __forceinline__ void int_range(double lo, double hi, int32_t n,
                                          int32_t& a, int32_t& b) {
  lo = fmax(lo, -1.0);
  hi = fmin(hi, static_cast<double>(n));
  a = max(0, static_cast<int32_t>(ceil(lo)));
  b = min(n - 1, static_cast<int32_t>(floor(hi)));
}
__device__ __forceinline__ void int_range(float lo, float hi, int32_t n,
                                          int32_t& a, int32_t& b) {
  lo = fmaxf(lo, -1.f);
  hi = fminf(hi, static_cast<float>(n));
  a = max(0, static_cast<int32_t>(ceilf(lo)));
  b = min(n - 1, static_cast<int32_t>(floorf(hi)));
}

// Weight of voxel `v` (unpadded index) in a sample whose taps are the
// padded element `idx` (voxels idx - 1 and idx, weights 1 - f and f).
__device__ __forceinline__ float tap_weight(uint32_t idx, float f,
                                            int32_t v) {
  const int32_t t = static_cast<int32_t>(idx) - kk::kHaloLo;
  return t == v ? 1.f - f : (t + 1 == v ? f : 0.f);
}

// Orders one wavefront's LDS accesses across its lanes (no workgroup
// barrier: wavefronts leave the kernel independently).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-column data of a column-coherent angle, computed lane-parallel and
// read wave-uniformly from LDS.
struct ColData {
  double inv, a1, D0c;
  float fb, alpha, beta;
  int32_t t0;
};

__global__ __launch_bounds__(kThreads) void cone_fp_t_kernel(
    const kk::FpAngle* __restrict__ tab, FptArgs A) {
  int64_t bid = blockIdx.x;
  const int32_t tc = static_cast<int32_t>(bid % A.tiles_c);
  bid /= A.tiles_c;
  const int32_t tb = static_cast<int32_t>(bid % A.tiles_b);
  const int32_t i = static_cast<int32_t>(bid / A.tiles_b);
  const int32_t lane = threadIdx.x & 63;
  const int32_t k = tc * 64 + lane;
  const int32_t jw = (tb * kWaves + static_cast<int32_t>(threadIdx.x >> 6)) *
                     kJB;
  if (jw >= A.nb) return;  // no barriers below: idle wavefronts may leave

  const double fi = static_cast<double>(i);
  const double nb1 = static_cast<double>(A.nb + 1);
  const double nc1 = static_cast<double>(A.nc + 1);
  const double kd = static_cast<double>(k);  // general path
  __shared__ ColData cols_lds[kWaves][64];
  float acc[kJB];
#pragma unroll
  for (int t = 0; t < kJB; ++t) acc[t] = 0.f;

  for (int32_t q = 0; q < A.n_angles; ++q) {
    const int32_t ang = A.angles[q];
    const kk::FpAngle& P = tab[ang];
    const float* img = A.proj + ang * A.ps0;
    const double arel = fi + P.sig;  // slice centre, relative to the source
    if (P.colc) {
      // columns whose b coordinate at this slice lies in
      // [jw, jw + kJB + 1): a1 = (x_b - taub) / arel, and for these
      // angles a1 = (col ub + qb) / (col ua + qa)
      const double t1 = (jw - kEps - P.taub) / arel;
      const double t2 = (jw + kJB + 1 + kEps - P.taub) / arel;
      const double d1 = fma(t1, P.ua, -P.ub), d2 = fma(t2, P.ua, -P.ub);
      int32_t c_lo = 0, c_hi = A.cols - 1;
      if (d1 * d2 > 0.0) {
        const double c1 = fma(-t1, P.qa, P.qb) / d1;
        const double c2 = fma(-t2, P.qa, P.qb) / d2;
        if (isfinite(c1) && isfinite(c2))
          int_range(fmin(c1, c2), fmax(c1, c2), A.cols, c_lo, c_hi);
      }
      ColData* cd = cols_lds[threadIdx.x >> 6];
      for (int32_t cb = c_lo; cb <= c_hi; cb += 64) {
        const int32_t ncol = min(64, c_hi - cb + 1);
        // the wavefront's columns, one per lane (exactly kk_fp_ray's Da,
        // Db: their row terms are exact zeros for these angles)
        if (lane < ncol) {
          const double fc = static_cast<double>(cb + lane);
          const double Da = fma(fc, P.ua, P.qa);
          const double Db = fma(fc, P.ub, P.qb);
          const double inv = 1.0 / Da;
          const double a1 = Db * inv;
          const double b1 = fma(a1, P.sig, P.taub);
          ColData c;
          c.t0 = -99;
          if (isfinite(a1) && isfinite(b1)) {
            uint32_t ib;
            kk_fp_coord(a1, b1, fi, nb1, ib, c.fb);
            c.t0 = static_cast<int32_t>(ib) - kk::kHaloLo - jw;
          }
          c.inv = inv;
          c.a1 = a1;
          c.D0c = fma(fc, P.uc, P.qc);
          // c coordinate of the column's rays: alpha * row + beta
          c.alpha = static_cast<float>(P.vc * inv * arel);
          c.beta = static_cast<float>(fma(c.D0c * inv, arel, P.tauc));
          cd[lane] = c;
        }
        wave_sync();
        for (int32_t m = 0; m < ncol; ++m) {
          const ColData c = cd[m];  // wave-uniform
          if (c.t0 < -1 || c.t0 >= kJB) continue;
          // rows whose c coordinate lies in [k, k + 2), found in float32
          // and widened by 1e-2 row (float32 error ~1e-4 row); the
          // samples themselves are evaluated exactly below
          int32_t r_lo = 0, r_hi = A.rows - 1;
          if (fabsf(c.alpha) > 1e-9f) {
            const float ia = 1.f / c.alpha;
            const float r1 = (static_cast<float>(k) - c.beta) * ia;
            const float r2 = (static_cast<float>(k) + 2.f - c.beta) * ia;
            int_range(fminf(r1, r2) - 1e-2f, fmaxf(r1, r2) + 1e-2f, A.rows,
                      r_lo, r_hi);
          }
          const int32_t col = cb + m;
          const float* ycol = img + col * A.ps2;
          float colsum = 0.f;
          for (int32_t r = r_lo; r <= r_hi; ++r) {
            const double Dc = fma(static_cast<double>(r), P.vc, c.D0c);
            const double a2 = Dc * c.inv;
            const double b2 = fma(a2, P.sig, P.tauc);
            if (!(isfinite(a2) && isfinite(b2))) continue;
            uint32_t ic;
            float fcw;
            kk_fp_coord(a2, b2, fi, nc1, ic, fcw);
            const float wk = tap_weight(ic, fcw, k);
            if (wk == 0.f) continue;
            colsum = fmaf(ycol[r * A.ps1] * kk_fp_weight(P, c.a1, a2), wk,
                          colsum);
          }
#pragma unroll
          for (int t = 0; t < kJB; ++t) {
            const float w =
                t == c.t0 ? 1.f - c.fb : (t == c.t0 + 1 ? c.fb : 0.f);
            acc[t] = fmaf(w, colsum, acc[t]);
          }
        }
        wave_sync();  // reads done before the next batch's writes
      }
    } else {
      // pixels whose rays meet the support rectangle of this lane's voxels
      // at this slice: bounding box of its projected corners
      double cmin = 1e300, cmax = -1e300, rmin = 1e300, rmax = -1e300;
      bool ok = true;
#pragma unroll
      for (int corner = 0; corner < 4; ++corner) {
        const double xb = (corner & 1 ? jw + kJB + 1 + kEps : jw - kEps) -
                          P.taub;
        const double xc = (corner & 2 ? kd + 2.0 + kEps : kd - kEps) -
                          P.tauc;
        double pc, pr;
        ok = ok && project(P, arel, xb, xc, pc, pr);
        cmin = fmin(cmin, pc);
        cmax = fmax(cmax, pc);
        rmin = fmin(rmin, pr);
        rmax = fmax(rmax, pr);
      }
      int32_t c_lo = 0, c_hi = A.cols - 1, r_lo = 0, r_hi = A.rows - 1;
      if (ok) {
        int_range(cmin, cmax, A.cols, c_lo, c_hi);
        int_range(rmin, rmax, A.rows, r_lo, r_hi);
      }
      for (int32_t r = r_lo; r <= r_hi; ++r) {
        for (int32_t col = c_lo; col <= c_hi; ++col) {
          double a1, b1, a2, b2;
          if (!kk_fp_ray(P, r, col, a1, b1, a2, b2)) continue;
          uint32_t ic;
          float fcw;
          kk_fp_coord(a2, b2, fi, nc1, ic, fcw);
          const float wk = tap_weight(ic, fcw, k);
          if (wk == 0.f) continue;
          uint32_t ib;
          float fb;
          kk_fp_coord(a1, b1, fi, nb1, ib, fb);
          const int32_t t0 = static_cast<int32_t>(ib) - kk::kHaloLo - jw;
          if (t0 < -1 || t0 >= kJB) continue;
          const float val = img[r * A.ps1 + col * A.ps2] *
                            kk_fp_weight(P, a1, a2) * wk;
#pragma unroll
          for (int t = 0; t < kJB; ++t) {
            const float w = t == t0 ? 1.f - fb : (t == t0 + 1 ? fb : 0.f);
            acc[t] = fmaf(w, val, acc[t]);
          }
        }
      }
    }
  }

  if (k >= A.nc) return;
#pragma unroll
  for (int t = 0; t < kJB; ++t) {
    const int32_t j = jw + t;
    if (j >= A.nb) break;
    const int64_t o = i * A.sa + j * A.sb + k * A.sc;
    if (A.mode == 1)
      A.vol[o] += acc[t];
    else
      A.vol[o] = acc[t];
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
    A.tiles_b = (A.nb + kJB * kWaves - 1) / (kJB * kWaves);
    A.tiles_c = (A.nc + 63) / 64;
    const int64_t blocks = int64_t(A.na) * A.tiles_b * A.tiles_c;
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
