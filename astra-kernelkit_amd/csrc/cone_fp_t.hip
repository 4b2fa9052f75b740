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
#include <algorithm>

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

// Angles per cone_fp_t_kernel launch.  A workgroup walks every angle of
// its launch, so the resident workgroups (which start and retire at
// different times) are spread over all of them; neighbouring slices read
// nearly the same pixels of an angle, but share them in the caches only
// if few angles are in flight.  A launch takes the angles whose workspace
// fits this budget (three quarters of MI355X's 256 MiB last-level cache;
// each launch after the first adds to the volume, one read-modify-write
// of it per launch).
constexpr int64_t kChunkBytes = int64_t(192) << 20;
// Consecutive slices of one tile per XCD run (measured on MI355X at
// 1024^3 / 1440 wobbling angles: 4.41 s with 32, 4.38-4.56 s with 16 and
// 64, 5.24 s with 128, 5.44 s with workgroups dealt round-robin in
// c-tile-fastest order).
constexpr int32_t kSliceRun = 32;

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

// 1 / d in float64 for the detector boxes and the walk's anchor: the
// hardware estimate and one Newton step, which squares its relative error
// (to ~1e-14 from the estimate's ~1e-7: far below the 1e-6 voxel widening
// of the boxes and the float32 rounding of the anchor offsets); non-finite
// for d = 0.
__device__ __forceinline__ double rcp64(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

// Least / greatest value over a wavefront whose 64 lanes are all active
// (DPP steps and a readlane, no LDS round trips), as a scalar.
__device__ __forceinline__ int32_t wave_min_all(int32_t v) {
  return __builtin_amdgcn_readfirstlane(__reduce_min_sync(~0ull, v));
}
__device__ __forceinline__ int32_t wave_max_all(int32_t v) {
  return __builtin_amdgcn_readfirstlane(__reduce_max_sync(~0ull, v));
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

// A sample at c offset x from the centre of a lane's first voxel, carrying
// value y: adds y times the tents of the lane's two voxels, 1 - |x| and
// 1 - |x - 1|, written max(0, 1 - |x|) and med3(0, x, 2 - x) (exact for
// x in [0, 2]).
static_assert(kJC == 2, "tent_pair covers two c voxels per lane");
__device__ __forceinline__ void tent_pair(float x, float y, float* colsum) {
  colsum[0] = fmaf(y, fmaxf(0.f, 1.f - fabsf(x)), colsum[0]);
  colsum[1] = fmaf(y, __builtin_amdgcn_fmed3f(0.f, x, 2.f - x), colsum[1]);
}

// Ends a run of samples whose b tap index is the float t: S - T goes to
// tap row t and T to row t + 1 of the lane's b-tap sums, then the run is
// cleared.  g has a spare row on either side (rows -1 and kJB, never
// written out), so the four read-modify-writes need no bounds branches
// and are issued together (two paired LDS reads, two paired writes, one
// round trip).  A run whose index is outside [-1, kJB) (outside the b-run,
// the "no run yet" marker -2, non-finite coordinates) adds zeros to rows
// -1 and 0.
__device__ __forceinline__ void close_run(float (*g)[kJC][64], int32_t lane,
                                          float t, float* S, float* T) {
  static_assert(kJC == 2, "close_run covers two c voxels per lane");
  const bool in = t >= -1.f && t < static_cast<float>(kJB);
  const int32_t row = in ? static_cast<int32_t>(t) : -1;
  float* p = &g[row][0][lane];
  const float a0 = in ? S[0] - T[0] : 0.f, a1 = in ? S[1] - T[1] : 0.f;
  const float b0 = in ? T[0] : 0.f, b1 = in ? T[1] : 0.f;
  // rows are kJC * 64 words apart, c voxels 64 words
  const float v0 = p[0], v1 = p[64], v2 = p[128], v3 = p[192];
  p[0] = v0 + a0;
  p[64] = v1 + a1;
  p[128] = v2 + b0;
  p[192] = v3 + b1;
#pragma unroll
  for (int cc = 0; cc < kJC; ++cc) S[cc] = T[cc] = 0.f;
}

// Value of lane `l` of a wavefront (l wave-uniform), in scalar registers.
__device__ __forceinline__ int32_t bcast(int32_t v, int32_t l) {
  return __builtin_amdgcn_readlane(v, l);
}

// Paths a launch compiles (its plan's angles need): the column-coherent
// one, the general one, or both.  A launch without column-coherent angles
// does not reserve the column records' LDS, which keeps a sixth workgroup
// per CU resident.
constexpr int kPathColumn = 1;
constexpr int kPathGeneral = 2;

template <int PATHS>
__global__ __launch_bounds__(kThreads, PATHS == 3 ? 5 : 6) void cone_fp_t_kernel(
    const kk::FpAngle* __restrict__ tab, FptArgs A) {
  // slice fastest, then b tile, then c tile; runs of kSliceRun
  // consecutive ids (neighbouring slices of one tile, which read nearly
  // the same pixels of every angle) go to one XCD, so its L2 serves them
  const int64_t total = int64_t(A.na) * A.tiles_b * A.tiles_c;
  int64_t bid = kk_xcd_remap(blockIdx.x, total, kSliceRun);
  const int32_t i = static_cast<int32_t>(bid % A.na);
  bid /= A.na;
  const int32_t tb = static_cast<int32_t>(bid % A.tiles_b);
  const int32_t tc = static_cast<int32_t>(bid / A.tiles_b);
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
  // b-tap sums of this lane's voxels (row t: b voxel jw + t), indexed per
  // lane: every lane owns its words, so no atomics and no cross-lane
  // ordering are needed; taps just outside the b-run are not written out.
  // (A register array indexed by the wave-uniform tap row was compiled
  // into private scratch memory.)  The general path's run flushes use a
  // spare row on either side of the b-run (g[-1], g[kJB]); launches
  // without that path keep the LDS smaller.
  constexpr int kSpare = (PATHS & kPathGeneral) != 0 ? 1 : 0;
  __shared__ float gacc[kWaves][kJB + 2 * kSpare][kJC][64];
  float(*g)[kJC][64] = gacc[threadIdx.x >> 6] + kSpare;
  // column records of the column-coherent path, written lane-parallel and
  // read back per column at a wave-uniform address (an LDS broadcast, which
  // costs no vector-ALU slot, unlike a readlane): {alpha, beta} in float64
  // and {alpha, 1 / alpha, window origin, b fraction} in float32.  Each
  // wavefront uses its own records, in program order: no barrier.
  constexpr int kRec = (PATHS & kPathColumn) != 0 ? 64 : 1;
  __shared__ double2 crd_s[kWaves][kRec];
  __shared__ float4 crf_s[kWaves][kRec];
  double2* crd = crd_s[threadIdx.x >> 6];
  float4* crf = crf_s[threadIdx.x >> 6];
#pragma unroll
  for (int t = -kSpare; t < kJB + kSpare; ++t)
#pragma unroll
    for (int cc = 0; cc < kJC; ++cc) g[t][cc][lane] = 0.f;

  for (int32_t q = 0; q < A.n_angles; ++q) {
    const int32_t ang = A.angles[q];
    const kk::FpAngle& P = tab[ang];
    const float* img = A.yw + ang * angle_elems + kRowPad;
    const double arel = fi + P.sig;  // slice centre, relative to the source
    if (P.colc) { if constexpr ((PATHS & kPathColumn) != 0) {
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
      // computed from wave-uniform values: scalar loop control below
      c_lo = __builtin_amdgcn_readfirstlane(c_lo);
      c_hi = __builtin_amdgcn_readfirstlane(c_hi);
      for (int32_t cb = c_lo; cb <= c_hi; cb += 64) {
        const int32_t ncol = min(64, c_hi - cb + 1);
        // ray data of column cb + lane (kk_fp_ray's terms: the row terms
        // of a and b are exact zeros for these angles).  The c coordinate
        // of the column's rays at this slice is affine in the row:
        // x_c(r) = beta + alpha r (padded coordinates, float64).
        double alpha = 0.0, beta = 0.0;
        float fb = 0.f, ia = 0.f, alphaf = 0.f, off = 0.f;
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
        // window origin: the lane's first row is k / alpha + origin
        const float org = fmaf(-static_cast<float>(beta), ia, off);
        crd[lane] = make_double2(alpha, beta);
        crf[lane] = make_float4(alphaf, ia, org, fb);
        // the column's tap row and window length, one broadcast
        const int32_t tpk = (min(max(t0, -99), 100) + 128) | (nrow << 8);
        for (int32_t m = 0; m < ncol; ++m) {
          const int32_t pk = bcast(tpk, m);
          const int32_t tm = (pk & 255) - 128;
          if (tm < -1 || tm >= kJB) continue;
          const int32_t nr = pk >> 8;
          const float4 cf = crf[m];
          const double2 cd = crd[m];
          const float alm = cf.x, iam = cf.y;
          const float* ycol = img + (cb + m) * cstride;
          float colsum[kJC];
#pragma unroll
          for (int cc = 0; cc < kJC; ++cc) colsum[cc] = 0.f;
          if (nr > 0) {
            // nr rows (wave-uniform: no per-lane loop control) from the
            // ceiling of the window's first row, clamped into the padded
            // column (a window off the detector reads zeros)
            const float r1 = fmaf(kf, iam, cf.z);
            const int32_t r_lo =
                min(max(static_cast<int32_t>(ceilf(r1)), -kRowPad),
                    A.rows + kRowPad - nr);
            // the window's first row in float64, relative to the centre
            // of the lane's first voxel (padded coordinate k + 1); the
            // rows after it in float32 offsets (|alpha u| is a few voxels,
            // so the offsets carry ~1e-7 voxel of rounding).  A voxel's
            // weight is the tent 1 - |x_c - centre| of the sample, equal
            // to the forward kernel's bilinear tap weight.
            const float loc = static_cast<float>(
                fma(cd.x, static_cast<double>(r_lo), cd.y) - (kd + 1.0));
            const float* yw = ycol + r_lo;
            int32_t u = 0;
            float uf = 0.f;   // u as a float, counted (exact)
            for (; u + kRU <= nr; u += kRU, uf += static_cast<float>(kRU)) {
              float y[kRU];
#pragma unroll
              for (int t = 0; t < kRU; ++t) y[t] = yw[u + t];
              const float xg = fmaf(alm, uf, loc);
#pragma unroll
              for (int t = 0; t < kRU; ++t)
                tent_pair(t == 0 ? xg : fmaf(alm, static_cast<float>(t), xg),
                          y[t], colsum);
            }
            for (; u < nr; ++u, uf += 1.f)
              tent_pair(fmaf(alm, uf, loc), yw[u], colsum);
          } else {
          // rows whose c coordinate lies in [k, k + kJC + 1) (the padded
          // span of the lane's voxels' tents), found in float32 and
          // widened by 1e-2 row; samples outside get zero weight below
          int32_t r_lo = 0, r_hi = A.rows - 1;
          if (iam != 0.f) {
            const float r1 = fmaf(kf, iam, cf.z);
            const float r2 = fmaf(static_cast<float>(kJC + 1), iam, r1);
            int_range(fminf(r1, r2) - 1e-2f, fmaxf(r1, r2) + 1e-2f, A.rows,
                      r_lo, r_hi);
          }
          // the lane's first row in float64 (the first voxel's centre is
          // at padded coordinate k + 1), float32 offsets after it, tents
          // as above; rays the forward kernel skips have yw = 0
          const float loc = static_cast<float>(
              fma(cd.x, static_cast<double>(r_lo), cd.y) - (kd + 1.0));
          float du = 0.f;
          for (int32_t r = r_lo; r <= r_hi; ++r, du += 1.f)
            tent_pair(fmaf(alm, du, loc), ycol[r], colsum);
          }
          // the column's two b taps tm, tm + 1 (wave-uniform, so the
          // tests are scalar branches): read-modify-write of the lane's
          // own LDS words
          const float fbm = cf.w;
          if (tm >= 0) {
#pragma unroll
            for (int cc = 0; cc < kJC; ++cc)
              g[tm][cc][lane] = fmaf(1.f - fbm, colsum[cc], g[tm][cc][lane]);
          }
          if (tm + 1 < kJB) {
#pragma unroll
            for (int cc = 0; cc < kJC; ++cc)
              g[tm + 1][cc][lane] = fmaf(fbm, colsum[cc], g[tm + 1][cc][lane]);
          }
        }
      }
    } } else if constexpr ((PATHS & kPathGeneral) != 0) {
      // pixels whose rays meet the support rectangle of this lane's voxels
      // at this slice: bounding box of its projected corners.  The ray
      // through the source-relative point X meets the detector at
      // col = (V x Q).X / N.X, row = (Q x U).X / N.X with N = U x V
      // (q + col U + row V = lam X, lam = U.(V x Q) / N.X > 0 in front of
      // the source): linear forms in X, whose a and b parts (the slice and
      // the wave's b-run) are shared by the wave; each lane adds its c part.
      const double nva = P.ub * P.vc - P.uc * P.vb;
      const double nvb = P.uc * P.va - P.ua * P.vc;
      const double nvc = P.ua * P.vb - P.ub * P.va;
      const double mva = P.vb * P.qc - P.vc * P.qb;
      const double mvb = P.vc * P.qa - P.va * P.qc;
      const double mvc = P.va * P.qb - P.vb * P.qa;
      const double rva = P.qb * P.uc - P.qc * P.ub;
      const double rvb = P.qc * P.ua - P.qa * P.uc;
      const double rvc = P.qa * P.ub - P.qb * P.ua;
      const double lamn = P.ua * mva + P.ub * mvb + P.uc * mvc;
      double cmin = 1e300, cmax = -1e300, rmin = 1e300, rmax = -1e300;
      bool ok = true;
#pragma unroll
      for (int eb = 0; eb < 2; ++eb) {
        const double xb = (eb ? jw + kJB + 1 + kEps : jw - kEps) - P.taub;
        const double dn = fma(nvb, xb, nva * arel);
        const double cn = fma(mvb, xb, mva * arel);
        const double rn = fma(rvb, xb, rva * arel);
#pragma unroll
        for (int ec = 0; ec < 2; ++ec) {
          const double xc =
              (ec ? kd + (kJC + 1) + kEps : kd - kEps) - P.tauc;
          const double inv = rcp64(fma(nvc, xc, dn));
          const double pc = fma(mvc, xc, cn) * inv;
          const double pr = fma(rvc, xc, rn) * inv;
          ok = ok && isfinite(pc) && isfinite(pr) && lamn * inv > 0.0;
          cmin = fmin(cmin, pc);
          cmax = fmax(cmax, pc);
          rmin = fmin(rmin, pr);
          rmax = fmax(rmax, pr);
        }
      }
      int32_t c_lo = 0, c_hi = A.cols - 1, r_lo = 0, r_hi = A.rows - 1;
      if (ok) {
        int_range(cmin, cmax, A.cols, c_lo, c_hi);
        int_range(rmin, rmax, A.rows, r_lo, r_hi);
      }
      // columns in a loop of wave-uniform length, each lane walking its
      // own column range from its first column (a tilted rotation axis
      // shifts the lanes' ranges against each other: a loop over their
      // union would leave a lane idle for the difference), rows per lane.
      // Sample coordinates: float64 at an anchor pixel
      // (the wave's first column, the lane's first row), float32 offsets
      // from it.  With D = q + col U + row V (kk_fp_ray's terms) and
      // t = D_b / D_a at the anchor, x_b = arel D_b / D_a + tau_b moves by
      // arel (dD_b - t dD_a) / D_a between the anchor and a pixel: a
      // quantity of the box's size, so its float32 rounding is ~1e-6
      // voxel.  Weights are the tents 1 - |x - centre| of the forward
      // kernel's bilinear taps.
      const int32_t wc_lo = wave_min_all(c_lo <= c_hi ? c_lo : A.cols);
      const double fwc = static_cast<double>(wc_lo);
      const double fr0 = static_cast<double>(r_lo);
      const double Da0 = fma(fr0, P.va, fma(fwc, P.ua, P.qa));
      const double Db0 = fma(fr0, P.vb, fma(fwc, P.ub, P.qb));
      const double Dc0 = fma(fr0, P.vc, fma(fwc, P.uc, P.qc));
      const double inv0 = rcp64(Da0);
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
      // With u the row offset from the lane's first row and dc the column
      // offset from the wave's first column, dA = dc ua + u va (b, c
      // alike) and x_b = ob + arel (dB - tb dA) / (Da0 + dA): per column
      // the numerator arel (dB - tb dA) is affine in u with slope
      // arel (vb - tb va), and the denominator with slope va (c alike).
      const float kvb = arf * fmaf(-tbf, vaf, vbf);
      const float kvc = arf * fmaf(-tcf, vaf, vcf);
      const int32_t ncol = (anchor_ok && c_lo <= c_hi) ? c_hi - c_lo + 1 : 0;
      const int32_t wncol = wave_max_all(ncol);
      // b-tap sums over a run of samples with the same tap index, added to
      // the lane's LDS words when the index changes: S = sum of y w_c,
      // T = sum of y w_c fb, so the taps t and t + 1 receive S - T and T.
      // A run stays open from column to column while the index holds (the
      // sums of one index add up whatever the column), so along the walk
      // a run closes about once per b voxel crossed.  The index is kept as
      // a float (the floor of x_b); a run reaches the LDS words only if
      // its index lies in [-1, kJB) and is dropped otherwise.  Rows
      // outside the lane's tents add zeros; a non-finite coordinate (a ray
      // parallel to the slices) starts a run of its own that is dropped.
      float tcur = -2.f;  // no run yet
      float S[kJC], T[kJC];
#pragma unroll
      for (int cc = 0; cc < kJC; ++cc) S[cc] = T[cc] = 0.f;
      // the workspace of this angle as bytes (a wave-uniform base) and a
      // 32-bit byte offset per lane: the launcher keeps an angle's
      // workspace below 4 GiB
      const char* imgb = reinterpret_cast<const char*>(img);
      const uint32_t nrow =
          static_cast<uint32_t>(max(0, r_hi - r_lo + 1));
      for (int32_t j = 0; j < wncol; ++j) {
        const int32_t col = c_lo + j;  // this lane's column (in range if
                                       // j < ncol)
        const float dcf = static_cast<float>(col - wc_lo);
        const float dAc = dcf * uaf;
        const float nb0 = arf * fmaf(-tbf, dAc, dcf * ubf);
        const float nc0 = arf * fmaf(-tcf, dAc, dcf * ucf);
        const float den0 = da0 + dAc;
        uint32_t o = 4u * (static_cast<uint32_t>(col) *
                               static_cast<uint32_t>(cstride) +
                           static_cast<uint32_t>(r_lo));
        const uint32_t oend = o + 4u * (j < ncol ? nrow : 0u);
        float drf = 0.f;
        for (; o < oend; o += 4u, drf += 1.f) {
          // the load first: its latency overlaps the coordinates below.
          // The empty asm keeps the offset a 32-bit register (an SGPR base
          // plus a VGPR offset), not a 64-bit pointer stepped per row.
          asm volatile("" : "+v"(o));
          const float yv = *reinterpret_cast<const float*>(imgb + o);
          const float rc = __builtin_amdgcn_rcpf(fmaf(drf, vaf, den0));
          const float xc = fmaf(fmaf(drf, kvc, nc0), rc, oc);
          const float xb = fmaf(fmaf(drf, kvb, nb0), rc, ob);
          const float fl = floorf(xb);
          if (fl != tcur) {
            close_run(g, lane, tcur, S, T);
            tcur = fl;
          }
          const float yfb = yv * (xb - fl);
          const float w0 = fmaxf(0.f, 1.f - fabsf(xc));
          const float w1 = __builtin_amdgcn_fmed3f(0.f, xc, 2.f - xc);
          S[0] = fmaf(yv, w0, S[0]);
          S[1] = fmaf(yv, w1, S[1]);
          T[0] = fmaf(yfb, w0, T[0]);
          T[1] = fmaf(yfb, w1, T[1]);
        }
      }
      close_run(g, lane, tcur, S, T);
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
      const float v = g[t][cc][lane];
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
  // the gather addresses one angle's padded workspace with 32-bit byte
  // offsets
  if (int64_t(cols) * (rows + 2 * kRowPad) * 4 >= (int64_t(1) << 32))
    return kk::fail(KK_ERR_UNSUPPORTED,
                    "kk_cone_fp_transpose: detector of 2^30 padded pixels "
                    "or more");
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
  const int64_t angle_bytes = int64_t(cols) * (rows + 2 * kRowPad) * 4;
  const int32_t chunk = static_cast<int32_t>(std::min<int64_t>(
      INT32_MAX, std::max<int64_t>(1, kChunkBytes / angle_bytes)));
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
    const dim3 grid(static_cast<unsigned>(blocks));
    for (int32_t q0 = 0; q0 < n; q0 += chunk) {
      A.angles = plan->fr_idx + plan->fr_start[a] + q0;
      A.n_angles = std::min(chunk, n - q0);
      A.mode = (first && !additive) ? 0 : 1;
      first = false;
      if (plan->fp_all_colc)
        hipLaunchKernelGGL(cone_fp_t_kernel<kPathColumn>, grid,
                           dim3(kThreads), 0, st, plan->fp, A);
      else if (!plan->fp_any_colc)
        hipLaunchKernelGGL(cone_fp_t_kernel<kPathGeneral>, grid,
                           dim3(kThreads), 0, st, plan->fp, A);
      else
        hipLaunchKernelGGL((cone_fp_t_kernel<kPathColumn | kPathGeneral>),
                           grid, dim3(kThreads), 0, st, plan->fp, A);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return kk::hip_fail(e, "cone_fp_t_kernel");
    }
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
