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
// MI355X design, two launches per ray axis group:
// 1. ray_weight_kernel: y * (ray weight), 0 for rays the forward kernel
//    skips, written column-major ([angle][col][row]) into the caller's
//    workspace through an LDS tile.  The gather below then reads, for one
//    column, the rows of 64 neighbouring voxels: a few cache lines per
//    wavefront instead of 64.
// 2. cone_fp_t_kernel: a gather, so that no atomics are needed (an LDS
//    float atomic add costs ~190 cycles per wavefront instruction on this
//    part, against ~8 for a plain LDS read-modify-write).  A lane owns kJB x kJC
//    voxels (kJB along b, kJC consecutive along c) of one slice, the 64
//    lanes of a wavefront consecutive c pairs, accumulated in the lane's
//    own LDS words over all angles of the launch's ray axis and written
//    once.  Owning
//    two c voxels, a lane evaluates each of its samples once for both
//    (a sample has two c taps): 3/2 evaluations per sample instead of 2.  Per angle:
//    * circular orbits ("column-coherent" angles: the row step has no a/b
//      part): the b coordinate of a sample depends on the detector column
//      only, so the columns reaching the wavefront's b-run are a
//      wave-uniform interval (inverse of the column -> b map).  The lanes
//      compute 64 columns' ray data at a time and broadcast them with
//      readlane (scalar registers); per column, the c coordinate is affine
//      in the row, so each lane sums the rows reaching its voxel and the
//      sum enters the b-run with the column's two b taps (a wave-uniform
//      LDS row);
//    * other geometries: each lane projects the support rectangle of its
//      voxels (at its slice) onto the detector and walks the pixels of the
//      bounding box; the b taps, whose index differs per lane, are summed
//      in the lane's own LDS words (one read-modify-write pair per sample
//      instead of a select over all kJB registers).
// Candidate intervals are widened by 1e-6 voxel in coordinate space, far
// above the float64 rounding of the sample coordinates; samples outside
// the support get zero weight.
#include "kk_internal.h"

namespace {

constexpr int kJB = 8;       // voxels along b per lane
constexpr int kJC = 2;       // consecutive voxels along c per lane
constexpr int kWaves = 4;    // wavefronts per workgroup (b-runs)
constexpr int kThreads = 64 * kWaves;
constexpr double kEps = 1e-6;
constexpr int kRU = 4;       // rows per load group (column-coherent path)

struct FptArgs {
  const float* yw;        // weighted projections [angle][col][row]
  float* vol;
  int64_t sa, sb, sc;     // volume strides along the frame axes a, b, c
  int32_t rows, cols;
  int32_t na, nb, nc;
  int32_t tiles_b, tiles_c;
  const int32_t* angles;  // angle ids of this ray axis
  int32_t n_angles;
  int32_t mode;           // 0: vol = A^T y, 1: vol += A^T y
};

constexpr int kTile = 64;  // ray_weight_kernel: 64 x 64 pixels per block
// Zero rows before and after every column of the workspace: a lane's row
// window of up to kRowPad rows may start above or end below the detector
// and still reads in-bounds zeros, so the column loop needs no per-row
// bounds test.
constexpr int kRowPad = 16;

// yw[angle][col][kRowPad + row] = proj(angle, row, col) * kk_fp_weight of
// the ray, 0 where kk_fp_ray rejects the ray, and kRowPad zeros before and
// after each column (column stride rows + 2 kRowPad).  Reads along columns
// and writes along rows, both coalesced, through an LDS tile.
__global__ __launch_bounds__(256) void ray_weight_kernel(
    const kk::FpAngle* __restrict__ tab, const float* __restrict__ proj,
    int64_t ps0, int64_t ps1, int64_t ps2, int32_t rows, int32_t cols,
    int32_t tiles_r, int32_t tiles_c, float* __restrict__ yw) {
  __shared__ float tile[kTile][kTile + 1];
  int64_t bid = blockIdx.x;
  const int32_t tcol = static_cast<int32_t>(bid % tiles_c);
  bid /= tiles_c;
  const int32_t trow = static_cast<int32_t>(bid % tiles_r);
  const int32_t ang = static_cast<int32_t>(bid / tiles_r);
  const kk::FpAngle& P = tab[ang];
  const int32_t lane = threadIdx.x & 63;
  const int32_t w = threadIdx.x >> 6;
  const int32_t r0 = trow * kTile, c0 = tcol * kTile;
  const float* img = proj + ang * ps0;
  for (int32_t t = w; t < kTile; t += 4) {
    const int32_t r = r0 + t, col = c0 + lane;
    float v = 0.f;
    if (r < rows && col < cols) {
      double a1, b1, a2, b2;
      if (kk_fp_ray(P, r, col, a1, b1, a2, b2))
        v = img[r * ps1 + col * ps2] * kk_fp_weight(P, a1, a2);
    }
    tile[t][lane] = v;
  }
  __syncthreads();
  const int64_t cs = rows + 2 * kRowPad;
  float* out = yw + static_cast<int64_t>(ang) * cols * cs + kRowPad;
  for (int32_t t = w; t < kTile; t += 4) {
    const int32_t col = c0 + t, r = r0 + lane;
    if (col >= cols) continue;
    if (r < rows) out[col * cs + r] = tile[lane][t];
    // the pads: above the first row band, below the last
    if (trow == 0 && lane < kRowPad) out[col * cs - kRowPad + lane] = 0.f;
    if (trow == tiles_r - 1 && lane < kRowPad)
      out[col * cs + rows + lane] = 0.f;
  }
}

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

// Value of lane `l` of a wavefront (l wave-uniform), in scalar registers.
__device__ __forceinline__ int32_t bcast(int32_t v, int32_t l) {
  return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ float bcast(float v, int32_t l) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, v), l));
}
__device__ __forceinline__ double bcast(double v, int32_t l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int32_t>(u), l));
  const uint32_t hi = static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int32_t>(u >> 32), l));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

__global__ __launch_bounds__(kThreads, 6) void cone_fp_t_kernel(
    const kk::FpAngle* __restrict__ tab, FptArgs A) {
  int64_t bid = blockIdx.x;
  const int32_t tc = static_cast<int32_t>(bid % A.tiles_c);
  bid /= A.tiles_c;
  const int32_t tb = static_cast<int32_t>(bid % A.tiles_b);
  const int32_t i = static_cast<int32_t>(bid / A.tiles_b);
  const int32_t lane = threadIdx.x & 63;
  const int32_t k = (tc * 64 + lane) * kJC;  // first of the lane's c voxels
  const int32_t jw = (tb * kWaves + static_cast<int32_t>(threadIdx.x >> 6)) *
                     kJB;
  if (jw >= A.nb) return;  // no barriers below: idle wavefronts may leave

  const double fi = static_cast<double>(i);
  const double nb1 = static_cast<double>(A.nb + 1);
  const double kd = static_cast<double>(k);
  const float kf = static_cast<float>(k);
  const int32_t cstride = A.rows + 2 * kRowPad;  // workspace column stride
  const int64_t angle_elems = static_cast<int64_t>(cstride) * A.cols;
  // b-tap sums of this lane's voxels at t = -1 .. kJB (row t + 1; rows 0
  // and kJB + 1 catch taps just outside the b-run and are discarded),
  // indexed per lane: every lane owns its words, so no atomics and no
  // cross-lane ordering are needed.  (A register array indexed by the
  // wave-uniform tap row was compiled into private scratch memory.)
  __shared__ float gacc[kWaves][kJB + 2][kJC][64];
  float(*g)[kJC][64] = gacc[threadIdx.x >> 6];
#pragma unroll
  for (int t = 0; t < kJB + 2; ++t)
#pragma unroll
    for (int cc = 0; cc < kJC; ++cc) g[t][cc][lane] = 0.f;

  for (int32_t q = 0; q < A.n_angles; ++q) {
    const int32_t ang = A.angles[q];
    const kk::FpAngle& P = tab[ang];
    const float* img = A.yw + ang * angle_elems + kRowPad;
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
      for (int32_t cb = c_lo; cb <= c_hi; cb += 64) {
        const int32_t ncol = min(64, c_hi - cb + 1);
        // ray data of column cb + lane (kk_fp_ray's terms: the row terms
        // of a and b are exact zeros for these angles).  The c coordinate
        // of the column's rays at this slice is affine in the row:
        // x_c(r) = beta + alpha r (padded coordinates, float64).
        double alpha = 0.0, beta = 0.0;
        float fb = 0.f, ia = 0.f, betaf = 0.f, alphaf = 0.f, off = 0.f;
        int32_t t0 = -99, nrow = 0;
        if (lane < ncol) {
          const double fc = static_cast<double>(cb + lane);
          const double Da = fma(fc, P.ua, P.qa);
          const double Db = fma(fc, P.ub, P.qb);
          const double inv = 1.0 / Da;
          const double a1 = Db * inv;
          const double b1 = fma(a1, P.sig, P.taub);
          if (isfinite(a1) && isfinite(b1)) {
            uint32_t ib;
            kk_fp_coord(a1, b1, fi, nb1, ib, fb);
            t0 = static_cast<int32_t>(ib) - kk::kHaloLo - jw;
          }
          const double D0c = fma(fc, P.uc, P.qc);
          alpha = P.vc * inv * arel;
          beta = fma(D0c * inv, arel, P.tauc);
          if (!isfinite(alpha) || !isfinite(beta)) t0 = -99;
          alphaf = static_cast<float>(alpha);
          ia = fabsf(alphaf) > 1e-9f ? 1.f / alphaf : 0.f;
          betaf = static_cast<float>(beta);
          // a lane's window: the rows whose c coordinate lies in the open
          // padded span (k, k + kJC + 1) of its voxels' tents, (kJC + 1)
          // |1 / alpha| rows long, i.e. at most floor(len) + 1 integer
          // rows.  Windows of at most kRowPad - 3 rows are read as nrow
          // rows from the ceiling of the first row, moved 1e-3 row earlier
          // (float32 rounding of the start is ~1e-5 row; rows outside get
          // zero tent weight); longer ones take the exact range below.
          const float len = static_cast<float>(kJC + 1) * fabsf(ia);
          if (ia != 0.f && len + 3.f <= static_cast<float>(kRowPad)) {
            nrow = static_cast<int32_t>(floorf(len + 2e-3f)) + 1;
            off = fminf(0.f, static_cast<float>(kJC + 1) * ia) - 1e-3f;
          }
        }
        // the column's tap row and window length, one broadcast
        const int32_t tpk = (min(max(t0, -99), 100) + 128) | (nrow << 8);
        for (int32_t m = 0; m < ncol; ++m) {
          const int32_t pk = bcast(tpk, m);
          const int32_t tm = (pk & 255) - 128;
          if (tm < -1 || tm >= kJB) continue;
          const float iam = bcast(ia, m);
          const int32_t nr = pk >> 8;
          const float* ycol = img + (cb + m) * cstride;
          float colsum[kJC];
#pragma unroll
          for (int cc = 0; cc < kJC; ++cc) colsum[cc] = 0.f;
          if (nr > 0) {
            // nr rows (wave-uniform: no per-lane loop control) from the
            // ceiling of the window's first row, clamped into the padded
            // column (a window off the detector reads zeros)
            const float r1 =
                fmaf(kf - bcast(betaf, m), iam, bcast(off, m));
            const int32_t r_lo =
                min(max(static_cast<int32_t>(ceilf(r1)), -kRowPad),
                    A.rows + kRowPad - nr);
            // the window's first row in float64, relative to the centre
            // of the lane's first voxel (padded coordinate k + 1); the
            // rows after it in float32 offsets (|alpha u| is a few voxels,
            // so the offsets carry ~1e-7 voxel of rounding).  A voxel's
            // weight is the tent 1 - |x_c - centre| of the sample, equal
            // to the forward kernel's bilinear tap weight.
            const float alm = bcast(alphaf, m);
            const float loc = static_cast<float>(
                fma(bcast(alpha, m), static_cast<double>(r_lo),
                    bcast(beta, m)) - (kd + 1.0));
            const float* yw = ycol + r_lo;
            int32_t u = 0;
            float uf = 0.f;   // u as a float, counted (exact)
            for (; u + kRU <= nr; u += kRU, uf += static_cast<float>(kRU)) {
              float y[kRU];
#pragma unroll
              for (int t = 0; t < kRU; ++t) y[t] = yw[u + t];
#pragma unroll
              for (int t = 0; t < kRU; ++t) {
                const float x =
                    fmaf(alm, uf + static_cast<float>(t), loc);
#pragma unroll
                for (int cc = 0; cc < kJC; ++cc)
                  colsum[cc] = fmaf(
                      y[t],
                      fmaxf(0.f, 1.f - fabsf(x - static_cast<float>(cc))),
                      colsum[cc]);
              }
            }
            for (; u < nr; ++u, uf += 1.f) {
              const float yv = yw[u];
              const float x = fmaf(alm, uf, loc);
#pragma unroll
              for (int cc = 0; cc < kJC; ++cc)
                colsum[cc] = fmaf(
                    yv, fmaxf(0.f, 1.f - fabsf(x - static_cast<float>(cc))),
                    colsum[cc]);
            }
          } else {
          // rows whose c coordinate lies in [k, k + kJC + 1) (the padded
          // span of the lane's voxels' tents), found in float32 and
          // widened by 1e-2 row; samples outside get zero weight below
          int32_t r_lo = 0, r_hi = A.rows - 1;
          if (iam != 0.f) {
            const float r1 = (kf - bcast(betaf, m)) * iam;
            const float r2 = fmaf(static_cast<float>(kJC + 1), iam, r1);
            int_range(fminf(r1, r2) - 1e-2f, fmaxf(r1, r2) + 1e-2f, A.rows,
                      r_lo, r_hi);
          }
          // the lane's first row in float64 (the first voxel's centre is
          // at padded coordinate k + 1), float32 offsets after it, tents
          // as above; rays the forward kernel skips have yw = 0
          const float alm = bcast(alphaf, m);
          const float loc = static_cast<float>(
              fma(bcast(alpha, m), static_cast<double>(r_lo),
                  bcast(beta, m)) - (kd + 1.0));
          int32_t r = r_lo;
          float du = 0.f;
          for (; r + (kRU - 1) <= r_hi; r += kRU) {
            float y[kRU];
#pragma unroll
            for (int u = 0; u < kRU; ++u) y[u] = ycol[r + u];
#pragma unroll
            for (int u = 0; u < kRU; ++u) {
              const float x = fmaf(alm, du + static_cast<float>(u), loc);
#pragma unroll
              for (int cc = 0; cc < kJC; ++cc)
                colsum[cc] = fmaf(
                    y[u], fmaxf(0.f, 1.f - fabsf(x - static_cast<float>(cc))),
                    colsum[cc]);
            }
            du += static_cast<float>(kRU);
          }
          for (; r <= r_hi; ++r) {
            const float yv = ycol[r];
            const float x = fmaf(alm, du, loc);
#pragma unroll
            for (int cc = 0; cc < kJC; ++cc)
              colsum[cc] = fmaf(
                  yv, fmaxf(0.f, 1.f - fabsf(x - static_cast<float>(cc))),
                  colsum[cc]);
            du += 1.f;
          }
          }
          // the column's two b taps (tm wave-uniform): read-modify-write
          // of the lane's own LDS words
          const float fbm = bcast(fb, m);
#pragma unroll
          for (int cc = 0; cc < kJC; ++cc) {
            g[tm + 1][cc][lane] = fmaf(1.f - fbm, colsum[cc],
                                       g[tm + 1][cc][lane]);
            g[tm + 2][cc][lane] = fmaf(fbm, colsum[cc], g[tm + 2][cc][lane]);
          }
        }
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
        const double xc = (corner & 2 ? kd + (kJC + 1) + kEps : kd - kEps) -
                          P.tauc;
        double pc = 0.0, pr = 0.0;
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
      // columns in a wave-uniform loop (the union of the lanes' ranges),
      // rows per lane.  Sample coordinates: float64 at an anchor pixel
      // (the wave's first column, the lane's first row), float32 offsets
      // from it.  With D = q + col U + row V (kk_fp_ray's terms) and
      // t = D_b / D_a at the anchor, x_b = arel D_b / D_a + tau_b moves by
      // arel (dD_b - t dD_a) / D_a between the anchor and a pixel: a
      // quantity of the box's size, so its float32 rounding is ~1e-6
      // voxel.  Weights are the tents 1 - |x - centre| of the forward
      // kernel's bilinear taps.
      const int32_t wc_lo = kk_wave_min(c_lo <= c_hi ? c_lo : A.cols);
      const int32_t wc_hi = kk_wave_max(c_lo <= c_hi ? c_hi : -1);
      const double fwc = static_cast<double>(wc_lo);
      const double fr0 = static_cast<double>(r_lo);
      const double Da0 = fma(fr0, P.va, fma(fwc, P.ua, P.qa));
      const double Db0 = fma(fr0, P.vb, fma(fwc, P.ub, P.qb));
      const double Dc0 = fma(fr0, P.vc, fma(fwc, P.uc, P.qc));
      const double inv0 = 1.0 / Da0;
      const double tb = Db0 * inv0, tcv = Dc0 * inv0;
      const float ob =
          static_cast<float>(fma(tb, arel, P.taub) - (jw + kk::kHaloLo));
      const float oc =
          static_cast<float>(fma(tcv, arel, P.tauc) - (kd + kk::kHaloLo));
      const float da0 = static_cast<float>(Da0);
      const float tbf = static_cast<float>(tb), tcf = static_cast<float>(tcv);
      const float arf = static_cast<float>(arel);
      const float uaf = static_cast<float>(P.ua), ubf = static_cast<float>(P.ub),
                  ucf = static_cast<float>(P.uc);
      const float vaf = static_cast<float>(P.va), vbf = static_cast<float>(P.vb),
                  vcf = static_cast<float>(P.vc);
      const bool anchor_ok = ok && isfinite(tb) && isfinite(tcv);
      for (int32_t col = wc_lo; col <= wc_hi; ++col) {
        const float* ycol = img + col * cstride;
        const float dcf = static_cast<float>(col - wc_lo);
        const float dAc = dcf * uaf, dBc = dcf * ubf, dCc = dcf * ucf;
        const int32_t rl = r_lo;
        const int32_t rh = (anchor_ok && col >= c_lo && col <= c_hi) ? r_hi
                                                                    : r_lo - 1;
        // b-tap sums over a run of samples with the same tap index, added
        // to the lane's LDS words when the index changes (along a column
        // it rarely does)
        int32_t tcur = -2;  // no run yet
        float s0[kJC], s1[kJC];
#pragma unroll
        for (int cc = 0; cc < kJC; ++cc) s0[cc] = s1[cc] = 0.f;
        float drf = 0.f;
        for (int32_t r = rl; r <= rh; ++r, drf += 1.f) {
          const float dA = fmaf(drf, vaf, dAc);
          const float dB = fmaf(drf, vbf, dBc);
          const float dC = fmaf(drf, vcf, dCc);
          const float sc = arf * __builtin_amdgcn_rcpf(da0 + dA);
          const float xb = fmaf(fmaf(-tbf, dA, dB), sc, ob);
          const float xc = fmaf(fmaf(-tcf, dA, dC), sc, oc);
          float wk[kJC];
          bool any = false;
#pragma unroll
          for (int cc = 0; cc < kJC; ++cc) {
            wk[cc] = fmaxf(0.f, 1.f - fabsf(xc - static_cast<float>(cc)));
            any = any || wk[cc] > 0.f;
          }
          if (!any) continue;
          const float fl = floorf(xb);
          if (!(fl >= -1.f && fl < static_cast<float>(kJB))) continue;
          const int32_t t0 = static_cast<int32_t>(fl);
          const float fb = xb - fl;
          if (t0 != tcur) {
            if (tcur >= -1) {
#pragma unroll
              for (int cc = 0; cc < kJC; ++cc) {
                g[tcur + 1][cc][lane] += s0[cc];
                g[tcur + 2][cc][lane] += s1[cc];
                s0[cc] = s1[cc] = 0.f;
              }
            }
            tcur = t0;
          }
          const float yv = ycol[r];
#pragma unroll
          for (int cc = 0; cc < kJC; ++cc) {
            const float val = yv * wk[cc];
            s0[cc] = fmaf(1.f - fb, val, s0[cc]);
            s1[cc] = fmaf(fb, val, s1[cc]);
          }
        }
        if (tcur >= -1) {
#pragma unroll
          for (int cc = 0; cc < kJC; ++cc) {
            g[tcur + 1][cc][lane] += s0[cc];
            g[tcur + 2][cc][lane] += s1[cc];
          }
        }
      }
    }
  }

#pragma unroll
  for (int cc = 0; cc < kJC; ++cc) {
    const int32_t kc = k + cc;
    if (kc >= A.nc) break;
#pragma unroll
    for (int t = 0; t < kJB; ++t) {
      const int32_t j = jw + t;
      if (j >= A.nb) break;
      const float v = g[t + 1][cc][lane];
      const int64_t o = i * A.sa + j * A.sb + kc * A.sc;
      if (A.mode == 1)
        A.vol[o] += v;
      else
        A.vol[o] = v;
    }
  }
}

}  // namespace

int64_t kk::fp_transpose_work_floats(const kk_plan_s* plan, int32_t rows,
                                     int32_t cols) {
  if (plan == nullptr || rows <= 0 || cols <= 0) return -1;
  return int64_t(plan->fp_n) * cols * (rows + 2 * kRowPad);
}

int kk::fp_transpose_launch(const kk_plan_s* plan, const float* proj,
                            const int64_t ps[3], int32_t rows, int32_t cols,
                            float* vol, const int32_t vol_shape[3],
                            const int64_t vs[3], int32_t additive, float* work,
                            hipStream_t st) {
  if (plan == nullptr || !plan->fp_set)
    return kk::fail(KK_ERR_NOT_SET,
                    "Please set the projection geometry with `set_params()`.");
  if (proj == nullptr || ps == nullptr || vol == nullptr ||
      vol_shape == nullptr || vs == nullptr || work == nullptr)
    return kk::fail(KK_ERR_INVALID, "kk_cone_fp_transpose: null argument");
  if (rows <= 0 || cols <= 0)
    return kk::fail(KK_ERR_INVALID, "kk_cone_fp_transpose: bad detector");
  if (int64_t(rows) * cols >= (int64_t(1) << 31))
    return kk::fail(KK_ERR_UNSUPPORTED,
                    "kk_cone_fp_transpose: detector of 2^31 pixels or more");
  for (int t = 0; t < 3; ++t)
    if (vol_shape[t] != plan->fp_shape[t])
      return kk::fail(KK_ERR_SHAPE,
                      "kk_cone_fp_transpose: volume shape differs from the "
                      "one the parameters were set for");
  {
    const int32_t tr = (rows + kTile - 1) / kTile;
    const int32_t tcl = (cols + kTile - 1) / kTile;
    const int64_t blocks = int64_t(plan->fp_n) * tr * tcl;
    if (blocks >= (int64_t(1) << 31))
      return kk::fail(KK_ERR_UNSUPPORTED, "kk_cone_fp_transpose: grid");
    hipLaunchKernelGGL(ray_weight_kernel, dim3(unsigned(blocks)), dim3(256),
                       0, st, plan->fp, proj, ps[0], ps[1], ps[2], rows, cols,
                       tr, tcl, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kk::hip_fail(e, "ray_weight_kernel");
  }
  bool first = true;
  for (int a = 0; a < 3; ++a) {
    const int32_t n = plan->fr_start[a + 1] - plan->fr_start[a];
    if (n == 0) continue;
    const int b = kk::kFrameB[a], c = kk::kFrameC[a];
    FptArgs A{};
    A.yw = work;
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
    A.tiles_c = (A.nc + 64 * kJC - 1) / (64 * kJC);
    const int64_t blocks = int64_t(A.na) * A.tiles_b * A.tiles_c;
    if (blocks >= (int64_t(1) << 31))
      return kk::fail(KK_ERR_UNSUPPORTED, "kk_cone_fp_transpose: grid");
    A.angles = plan->fr_idx + plan->fr_start[a];
    A.n_angles = n;
    A.mode = (first && !additive) ? 0 : 1;
    first = false;
    hipLaunchKernelGGL(cone_fp_t_kernel, dim3(unsigned(blocks)),
                       dim3(kThreads), 0, st, plan->fp, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kk::hip_fail(e, "cone_fp_t_kernel");
  }
  return KK_OK;
}

extern "C" int kk_cone_fp_transpose(const kk_plan* plan, const float* proj,
                                    const int64_t ps[3], int32_t rows,
                                    int32_t cols, float* vol,
                                    const int32_t vol_shape[3],
                                    const int64_t vs[3], int32_t additive,
                                    float* work, void* stream) {
  return kk::fp_transpose_launch(plan, proj, ps, rows, cols, vol, vol_shape,
                                 vs, additive, work,
                                 static_cast<hipStream_t>(stream));
}

extern "C" int64_t kk_cone_fp_transpose_work_elems(const kk_plan* plan,
                                                   int32_t rows,
                                                   int32_t cols) {
  return kk::fp_transpose_work_floats(plan, rows, cols);
}
